// RBE importance sampler of integer k-vectors m (k = 2 pi m / L) and the
// per-evaluation seed stream.
//
// Follows RbmdParallelUntil::GetRbeRadomM (src/parallel/src/rbmd_parallel_until.cpp:557-724)
// with the seed chain of :305-325 and :484-501: a std::mt19937 seeded with
// coulomb.random_seed yields one 32-bit seed per force evaluation; each seed
// starts a std::mt19937 driving an independent Metropolis-Hastings chain per
// dimension with target P_d(m) ~ exp(-(2 pi m / L_d)^2 / (4 alpha)) and a
// rounded Box-Muller proposal of sigma_d = sqrt(alpha L_d^2 / (2 pi^2)).  The
// arithmetic is float with the double promotions of the reference source, so
// the stream is identical for a given libstdc++/libm.  The target and
// proposal densities depend on |m| only and are tabulated once per box.
//
// Evaluations are independent given their seeds, so a small thread pool
// draws the samples of upcoming evaluations ahead of the GPU.
#pragma once
#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

namespace rbmdx {

// A per-evaluation sampler: draw(seed) writes out_size() integers.
class KSampler {
 public:
  virtual ~KSampler() = default;
  virtual void draw(uint32_t seed, int32_t* out) const = 0;
  virtual int out_size() const = 0;
};

class RbeSampler : public KSampler {
 public:
  RbeSampler(float alpha, const float L[3], int P);
  // The k-sample of one evaluation started from `seed`; m_out holds P*3.
  void draw(uint32_t seed, int32_t* m_out) const override;
  int out_size() const override { return 3 * P_; }
  int P() const { return P_; }

 private:
  static constexpr int kTab = 256;
  float alpha_, L_[3], sigma_[3], H_[3];
  int P_;
  std::vector<float> ptab_[3], qtab_[3];
  float H(int d) const;
  float Pm(float m, int d) const;  // direct evaluation
  float Qm(float m, int d) const;
  float P_of(float m, int d) const {
    const float a = m < 0 ? -m : m;
    return a < kTab ? ptab_[d][(int)a] : Pm(m, d);
  }
  float Q_of(float m, int d) const {
    const float a = m < 0 ? -m : m;
    return a < kTab ? qtab_[d][(int)a] : Qm(m, d);
  }
};

// Seeds of successive evaluations: outputs of std::mt19937(base_seed).
uint32_t rbe_seed_of_call(uint32_t base_seed, int64_t call_index);

// Draws samples for upcoming evaluations on worker threads.  A ring of
// `window` slots holds them: the worker that draws evaluation `call` (seeds
// are drawn in call order under the workers' lock) fills slot call % window
// once the consumer has taken call - window, then publishes the call index
// in the slot's tag.  The consumer's take is lock-free and makes no system
// call when the sample is ready, which it normally is: the workers run up to
// `window` evaluations ahead and poll while the ring is full.
class SamplePipeline {
 public:
  SamplePipeline(const KSampler* s, uint32_t base_seed, int threads,
                 int window);
  ~SamplePipeline();
  // Waits until the sample of evaluation `call` is ready and copies it out
  // (calls are taken in order 0, 1, 2, ...).
  void take(int64_t call, int32_t* m_out);

 private:
  struct Slot {
    std::atomic<int64_t> tag{-1};  // the call whose sample the slot holds
    std::vector<int32_t> data;
  };
  void worker();
  const KSampler* sampler_;
  std::mt19937 base_;
  std::mutex mu_;                  // call claims and seed draws (workers)
  int64_t next_seeded_ = 0;        // next call index whose seed is drawn
  std::atomic<int64_t> consumer_{0};  // next call the consumer will take
  int window_;
  std::atomic<bool> stop_{false};
  std::unique_ptr<Slot[]> slots_;
  std::vector<std::thread> threads_;
};

}  // namespace rbmdx
