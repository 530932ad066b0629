#include "sampler.hpp"

#include <algorithm>
#include <chrono>
#include <thread>

#include <cmath>

namespace rbmdx {
namespace {
const float kPiF = 3.14159265358979323846f;  // RBMD_PI in float builds
}

RbeSampler::RbeSampler(float alpha, const float L[3], int P)
    : alpha_(alpha), P_(P) {
  for (int d = 0; d < 3; ++d) L_[d] = L[d];
  for (int d = 0; d < 3; ++d) {
    // rbmd_parallel_until.cpp:710-720: sqrt(alpha L L / (2 pi pi))
    sigma_[d] = std::sqrt(alpha_ * L_[d] * L_[d] / (2 * kPiF * kPiF));
    H_[d] = H(d);
  }
  for (int d = 0; d < 3; ++d) {
    ptab_[d].resize(kTab);
    qtab_[d].resize(kTab);
    for (int a = 0; a < kTab; ++a) {
      ptab_[d][a] = Pm((float)a, d);
      qtab_[d][a] = Qm((float)a, d);
    }
  }
}

// Compute_H, rbmd_parallel_until.cpp:609-624.
float RbeSampler::H(int d) const {
  const float factor = -(alpha_ * L_[d] * L_[d]);
  float h = 0.0f;
  for (int m = -10; m <= 10; ++m) h += std::exp((float)(m * m) * factor);
  return h * std::sqrt(-(factor) / kPiF);
}

// Distribution_P, :645-652.  H is the cached Compute_H of the same box.
float RbeSampler::Pm(float m, int d) const {
  const float t = 2 * kPiF * m / L_[d];
  return std::exp(-(t * t) / (4 * alpha_)) / H_[d];
}

// Distribution_q, :654-670: mass of the rounded Gaussian proposal at m.
float RbeSampler::Qm(float m, int d) const {
  const float s = std::sqrt(alpha_ * (L_[d] * L_[d]) / (kPiF * kPiF));
  if (m == 0) return std::erf((float)(0.5 / (double)s));
  const double a = (double)std::fabs(m);
  const float hi = std::erf((float)((0.5 + a) / (double)s));
  const float lo = std::erf((float)((a - 0.5) / (double)s));
  return (hi - lo) / 2;
}

void RbeSampler::draw(uint32_t seed, int32_t* m_out) const {
  std::mt19937 gen(seed);
  auto uniform01 = [&gen]() {
    std::uniform_real_distribution<float> dis(0.0f, 1.0f);
    return dis(gen);
  };
  // FetchSample_1D, :672-692 (mu = 0)
  auto gauss = [&](float sigma) {
    float u1;
    do {
      u1 = uniform01();
    } while (u1 < 1e-6f);
    const float u2 = uniform01();
    const float r = std::sqrt((float)(-2.0 * (double)std::log(u1)));
    const float c = std::cos(2 * kPiF * u2);
    return sigma * r * c + 0.0f;
  };
  // MH_Algorithm, :627-643
  auto mh = [&](float m, int d) {
    const float mw = std::round(gauss(sigma_[d]));
    float prob = (P_of(mw, d) / P_of(m, d)) * (Q_of(m, d) / Q_of(mw, d));
    prob = prob < 1.0f ? prob : 1.0f;
    const float u = uniform01();
    return u <= prob ? mw : m;
  };
  const float tiny = 1e-6f;
  float x0[3];
  do {  // Fetch_P_Sample, :697-704: first sample rejects (0,0,0)
    x0[0] = std::round(gauss(sigma_[0]));
    x0[1] = std::round(gauss(sigma_[1]));
    x0[2] = std::round(gauss(sigma_[2]));
  } while (std::fabs(x0[0]) < tiny && std::fabs(x0[1]) < tiny &&
           std::fabs(x0[2]) < tiny);
  for (int d = 0; d < 3; ++d) m_out[d] = (int32_t)x0[d];
  for (int i = 1; i < P_; ++i) {  // :709-723
    float x1[3];
    x1[0] = mh(x0[0], 0);
    x1[1] = mh(x0[1], 1);
    x1[2] = mh(x0[2], 2);
    for (int d = 0; d < 3; ++d) {
      m_out[3 * i + d] = (int32_t)x1[d];
      x0[d] = x1[d];
    }
    if (std::fabs(x1[0]) < tiny && std::fabs(x1[1]) < tiny &&
        std::fabs(x1[2]) < tiny)
      --i;  // slot re-drawn; the chain itself stays at (0,0,0)
  }
}

uint32_t rbe_seed_of_call(uint32_t base_seed, int64_t call_index) {
  std::mt19937 base(base_seed);
  if (call_index > 0) base.discard((unsigned long long)call_index);
  return (uint32_t)base();
}

SamplePipeline::SamplePipeline(const KSampler* s, uint32_t base_seed,
                               int threads, int window)
    : sampler_(s), base_(base_seed), window_(window < 1 ? 1 : window),
      slots_(new Slot[(size_t)(window < 1 ? 1 : window)]) {
  for (int k = 0; k < window_; ++k) slots_[k].data.resize((size_t)sampler_->out_size());
  if (threads < 1) threads = 1;
  for (int t = 0; t < threads; ++t) threads_.emplace_back([this] { worker(); });
}

SamplePipeline::~SamplePipeline() {
  stop_.store(true);
  for (auto& t : threads_) t.join();
}

void SamplePipeline::worker() {
  for (;;) {
    int64_t call;
    uint32_t seed;
    {
      std::lock_guard<std::mutex> g(mu_);
      if (stop_.load()) return;
      call = next_seeded_++;
      seed = (uint32_t)base_();
    }
    // the slot is free once the consumer has taken call - window
    while (call >= consumer_.load(std::memory_order_acquire) + window_) {
      if (stop_.load()) return;
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    Slot& sl = slots_[(size_t)(call % window_)];
    sampler_->draw(seed, sl.data.data());
    sl.tag.store(call, std::memory_order_release);
  }
}

void SamplePipeline::take(int64_t call, int32_t* m_out) {
  Slot& sl = slots_[(size_t)(call % window_)];
  for (int spin = 0; sl.tag.load(std::memory_order_acquire) != call; ++spin)
    if (spin > 64) std::this_thread::yield();
  std::copy(sl.data.begin(), sl.data.end(), m_out);
  sl.tag.store(-1, std::memory_order_relaxed);
  consumer_.store(call + 1, std::memory_order_release);
}

}  // namespace rbmdx
