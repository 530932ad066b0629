// RBSOG set-up and sampler (see sog.hpp for the reference citations).
#include "sog.hpp"

#include <algorithm>
#include <cmath>
#include <random>

namespace rbmdx {
namespace {

const double kPi = 3.14159265358979323846;

// rbsog_level_preset.h:15-35, stored as float there
const SogPresetValues kPresets[5] = {
    {2.0f, 5.027010924194599f, 2.289e-3f, 6},
    {1.62976708826776469f, 3.633717409009413f, 1.158e-4f, 16},
    {1.48783512395703226f, 2.662784519725113f, 1.142e-5f, 30},
    {1.32070036405934420f, 2.277149356440992f, 5.583e-8f, 64},
    {1.21812525709410644f, 1.774456369233284f, 3.389e-11f, 102},
};

// Unit-width Gaussian density in float (G_sigma with sigma = 1,
// kspace_calculator.cpp:1545-1547).
float unit_gauss_f(float r) {
  return std::exp(-r * r / (2 * 1.0f * 1.0f)) /
         std::sqrt((float)(2 * kPi * 1.0f * 1.0f));
}

// w0 of the u-series (Compute_W0, :1549-1557) in float and in double.
float w0_float(float r0, float b) {
  float acc = 0.0f;
  for (int i = 1; i < 200; i++) {
    const float bi = std::pow(b, (float)(-i));
    acc = acc + bi * unit_gauss_f(bi * r0);
  }
  return (float)((1.0 / unit_gauss_f(r0)) * ((1.0 / (2 * std::log(b) * r0)) - (double)acc));
}
double w0_double(double r0, double b) {
  auto g = [](double r) { return std::exp(-0.5 * r * r) / std::sqrt(2 * kPi); };
  double acc = 0;
  for (int i = 1; i < 200; i++) {
    const double bi = std::pow(b, -i);
    acc += bi * g(bi * r0);
  }
  return (1.0 / g(r0)) * (1.0 / (2 * std::log(b) * r0) - acc);
}

// f(m) for integer m as the reference's Gaussian() evaluates it (:1559-1588):
// k^2 through double then powf, each level through expf.
float gauss_m(int mx, int my, int mz, const float L[3], float sigma, float b,
              int mmax, const float* coef) {
  const int m[3] = {mx, my, mz};
  float k2 = 0.0f;
  for (int d = 0; d < 3; ++d)
    k2 = k2 + std::pow((float)((m[d] + 0.00) * 2 * kPi / L[d]), 2.0f);
  const float b2 = b * b;
  float f = 0.0f;
  for (int l = 0; l < mmax; l++) {
    const float arg = -(std::pow(b2, (float)(l + 0.0)) * sigma * sigma) * k2;
    f = f + coef[l] * std::exp((float)(arg / 2.0));
  }
  return f;
}

void set_reference(SogSet& s, double cut_off, const double L[3]) {
  const SogPresetValues& p = kPresets[s.level];
  const float rc = (float)cut_off;
  const float Lf[3] = {(float)L[0], (float)L[1], (float)L[2]};
  const float b = p.b;
  const float sigma = (float)(p.sigma_rc10 * rc / 10.0);
  const int M = p.mmax;
  const float w0 = w0_float(rc / sigma, b);
  std::vector<float> coef(M), coefn(M), sl(M);
  const float sigma2 = sigma * sigma, sigma4 = sigma2 * sigma2;
  const float b2 = b * b, b4 = b2 * b2;
  sl[0] = std::sqrt(2.0f) * sigma;
  coef[0] = (float)(4 * kPi * std::log(b) * w0 * sigma2);
  coefn[0] = (float)(8 * kPi * std::log(b) * w0 * sigma4);
  for (int l = 1; l < M; ++l) {  // every level scaled from level 0 (w0 kept)
    sl[l] = sl[l - 1] * b;
    coef[l] = coef[l - 1] * b2;
    coefn[l] = coefn[l - 1] * b4;
  }
  // low-k part: |m|^2 <= kcut on the non-negative octant, doubled
  float low = 0, lown = 0;
  for (int i = 0; i <= s.kcut; i++) {
    const float Kx = (float)((2 * kPi / Lf[0]) * (i + 0.00));
    for (int j = 0; j <= s.kcut; j++) {
      const float Ky = (float)((2 * kPi / Lf[1]) * (j + 0.00));
      for (int k = 0; k <= s.kcut; k++) {
        const float Kz = (float)((2 * kPi / Lf[2]) * (k + 0.00));
        if (i * i + j * j + k * k > s.kcut) continue;
        const float K2 = Kx * Kx + Ky * Ky + Kz * Kz, K4 = K2 * K2;
        low += 2 * gauss_m(i, j, k, Lf, sigma, b, M, coef.data()) * K2;
        lown += 2 * gauss_m(i, j, k, Lf, sigma, b, M, coefn.data()) * K4;
      }
    }
  }
  // high-k analytic series; its stopping flag is shared by all levels, so
  // only level 0 contributes
  const float prd = std::max(Lf[0], std::max(Lf[1], Lf[2]));
  const float log_b = std::log(b);
  float high = 0, highn = 0, bound = 1.0f;
  for (int l = 0; l < M; ++l) {
    float H[3] = {0, 0, 0}, Y[3] = {0, 0, 0}, Z[3] = {0, 0, 0};
    int j = 1;
    while (bound > 1e-12) {
      const float kmin = (float)(kPi * j / prd);
      const float kmin2 = kmin * kmin, kmin4 = kmin2 * kmin2;
      const float s_l = sl[l], s_l2 = s_l * s_l;
      for (int d = 0; d < 3; ++d) {
        const float kd = (float)(kPi * j / Lf[d]);
        const float e = std::exp(-(kd * s_l * kd * s_l));
        const float kd2 = 4 * kd * kd, kd4 = 16 * kd * kd * kd * kd;
        H[d] += 2 * e;
        Y[d] += 2 * e * kd2;
        Z[d] += 2 * e * kd4;
      }
      ++j;
      bound = std::exp(4 * l * log_b - kmin2 * s_l2) * kmin4 * sigma4;
    }
    for (int d = 0; d < 3; ++d) H[d] += 1;
    high += coef[l] * (Y[0] * H[1] * H[2] + H[0] * Y[1] * H[2] + H[0] * H[1] * Y[2]);
    highn += coefn[l] * (Z[0] * H[1] * H[2] + H[0] * Z[1] * H[2] + H[0] * H[1] * Z[2] +
                         2 * Y[0] * Y[1] * H[2] + 2 * Y[0] * H[1] * Y[2] +
                         2 * H[0] * Y[1] * Y[2]);
  }
  s.b = b;
  s.sigma = sigma;
  s.w0 = w0;
  s.S = high - low;
  s.S_npt = highn - lown;
  s.mmax = M;
  s.b_f = b;
  s.sigma_f = sigma;
  s.coef_f = coef;
  s.coef_npt_f = coefn;
  s.coef.assign(coef.begin(), coef.end());
  s.coef_npt.assign(coefn.begin(), coefn.end());
}

// sum over all integers j of e^{-k_j^2 s2/2} (H), k_j^2 e (Y), k_j^4 e (Z)
void series_1d(double L, double s2, double& H, double& Y, double& Z) {
  H = 1, Y = 0, Z = 0;
  for (int j = 1;; ++j) {
    const double k = 2 * kPi * j / L, k2 = k * k;
    const double e = std::exp(-0.5 * k2 * s2);
    H += 2 * e;
    Y += 2 * e * k2;
    Z += 2 * e * k2 * k2;
    if (e * k2 * k2 < 1e-300 || (j > 4 && e < 1e-30)) break;
  }
}

void set_exact(SogSet& s, double cut_off, const double L[3]) {
  const SogPresetValues& p = kPresets[s.level];
  s.b = p.b;
  s.sigma = (double)p.sigma_rc10 * cut_off / 10.0;
  s.mmax = p.mmax;
  s.w0 = w0_double(cut_off / s.sigma, s.b);
  const double lb = std::log(s.b), s2 = s.sigma * s.sigma;
  s.coef.assign(s.mmax, 0.0);
  s.coef_npt.assign(s.mmax, 0.0);
  for (int l = 0; l < s.mmax; ++l) {
    const double b2l = std::pow(s.b, 2.0 * l);
    s.coef[l] = 4 * kPi * lb * s2 * (l ? b2l : s.w0);
    s.coef_npt[l] = 8 * kPi * lb * s2 * s2 * (l ? b2l * b2l : s.w0);
  }
  double S = 0, Sn = 0;
  for (int l = 0; l < s.mmax; ++l) {
    double H[3], Y[3], Z[3];
    const double sl2 = s2 * std::pow(s.b, 2.0 * l);
    for (int d = 0; d < 3; ++d) series_1d(L[d], sl2, H[d], Y[d], Z[d]);
    S += s.coef[l] * (Y[0] * H[1] * H[2] + H[0] * Y[1] * H[2] + H[0] * H[1] * Y[2]);
    Sn += s.coef_npt[l] * (Z[0] * H[1] * H[2] + H[0] * Z[1] * H[2] + H[0] * H[1] * Z[2] +
                           2 * (Y[0] * Y[1] * H[2] + Y[0] * H[1] * Y[2] + H[0] * Y[1] * Y[2]));
  }
  const int e = (int)std::ceil(std::sqrt((double)s.kcut));
  for (int i = -e; i <= e; ++i)
    for (int j = -e; j <= e; ++j)
      for (int k = -e; k <= e; ++k) {
        if ((!i && !j && !k) || i * i + j * j + k * k > s.kcut) continue;
        const double kx = 2 * kPi * i / L[0], ky = 2 * kPi * j / L[1],
                     kz = 2 * kPi * k / L[2];
        const double k2 = kx * kx + ky * ky + kz * kz;
        S -= sog_f(s, s.coef, k2) * k2;
        Sn -= sog_f(s, s.coef_npt, k2) * k2 * k2;
      }
  s.S = S;
  s.S_npt = Sn;
  s.b_f = (float)s.b;
  s.sigma_f = (float)s.sigma;
  s.coef_f.assign(s.coef.begin(), s.coef.end());
  s.coef_npt_f.assign(s.coef_npt.begin(), s.coef_npt.end());
}

}  // namespace

bool sog_preset(int level, SogPresetValues* out) {
  if (level < 0 || level > 4) return false;
  *out = kPresets[level];
  return true;
}

double sog_f(const SogSet& s, const std::vector<double>& coef, double k2) {
  double f = 0;
  for (int l = 0; l < s.mmax; ++l)
    f += coef[l] * std::exp(-std::pow(s.b, 2.0 * l) * s.sigma * s.sigma * k2 / 2.0);
  return f;
}

SogSet sog_set(int level, double cut_off, const double L[3], int kcut,
               bool exact) {
  SogSet s;
  s.level = level;
  s.kcut = kcut;
  s.exact = exact;
  if (exact)
    set_exact(s, cut_off, L);
  else
    set_reference(s, cut_off, L);
  return s;
}

double sog_self_energy(const SogSet& s, double sum_q2, double qqr2e) {
  const double c = (std::log(s.b) / (std::sqrt(2 * kPi) * s.sigma)) *
                   (s.w0 + (1 - std::pow(s.b, -s.mmax)) / (s.b - 1));
  return -c * sum_q2 * qqr2e;
}

SogNear sog_near(const SogSet& s) {
  SogNear n;
  n.w0 = s.w0;
  const double sig2 = s.sigma * s.sigma;
  const double c = std::log(s.b) / (sig2 * std::sqrt(2 * kPi * sig2));
  double fact = 1;
  for (int p = 0; p < 6; ++p) {
    if (p) fact *= p;
    double acc = 0;
    for (int l = 0; l < s.mmax; ++l) {
      const double v = l ? std::pow(s.b, -3.0 * l) : s.w0;
      acc += v * std::pow(1.0 / (2 * sig2 * std::pow(s.b, 2.0 * l)), p) / fact;
    }
    n.T[p] = ((p & 1) ? 2.0 : -2.0) * c * acc;
  }
  const double norm = 2.0 * std::log(s.b) / std::sqrt(2.0 * kPi * sig2);
  n.omega.resize(s.mmax);
  n.inv2s2.resize(s.mmax);
  for (int l = 0; l < s.mmax; ++l) {
    n.omega[l] = norm * (l ? std::pow(s.b, -l) : s.w0);
    n.inv2s2[l] = 1.0 / (2 * sig2 * std::pow(s.b, 2.0 * l));
  }
  return n;
}

// ---- sampler ---------------------------------------------------------------
SogSampler::SogSampler(const SogSet& s, const double L[3], int P)
    : P_(P), mmax_(s.mmax), kcut_(s.kcut), sigma_(s.sigma_f), b_(s.b_f),
      coef_(s.coef_f), coef_npt_(s.coef_npt_f) {
  for (int d = 0; d < 3; ++d) {
    L_[d] = (float)L[d];
    pxyz_[d] = (float)(2 * kPi / L_[d]);
    factor_[d] = (float)(L_[d] / (2 * kPi * sigma_));
    mhd_[d] = std::sqrt((float)(2 * sigma_ * sigma_ * kPi * kPi)) / L_[d];
  }
}

float SogSampler::gauss(int mx, int my, int mz,
                        const std::vector<float>& coef) const {
  return gauss_m(mx, my, mz, L_, sigma_, b_, mmax_, coef.data());
}

namespace {
// Proposal probability of the rounded normal (MH_D_Modify, :1643-1651).
float proposal_q(int m, float f) {
  if (m == 0) return std::erf((float)(0.5 * f));
  const float a = std::fabs((float)m);
  return (float)(0.5 * (std::erf((float)((a + 0.5) * f)) - std::erf((float)((a - 0.5) * f))));
}
}  // namespace

void SogSampler::draw(uint32_t seed, int32_t* out) const {
  std::mt19937 gen(seed);
  auto uniform = [&] {  // UniformRandomValue(0.0, 1.0)
    std::uniform_real_distribution<float> dis(0.0f, 1.0f);
    return dis(gen);
  };
  auto normal = [&](float scale) {  // randn_box_muller(0, scale)
    const float two_pi = (float)(2.0 * kPi);
    float u1, u2;
    do {
      u1 = uniform();
      u2 = uniform();
    } while (u1 <= 1.17549e-038f);
    return std::sqrt((float)(-2.0 * (double)std::log(u1))) * std::cos(two_pi * u2) *
               scale + 0.0f;
  };
  const int n2 = 2 * P_;
  std::vector<int> cx(n2), cy(n2), cz(n2);
  do {
    cx[0] = (int)std::round(normal(factor_[0]));
    cy[0] = (int)std::round(normal(factor_[1]));
    cz[0] = (int)std::round(normal(factor_[2]));
  } while (cx[0] == 0 && cy[0] == 0 && cz[0] == 0);
  for (int i = 0; i < n2 - 1; i++) {
    const float nx = std::round(normal(factor_[0]));
    const float ny = std::round(normal(factor_[1]));
    const float nz = std::round(normal(factor_[2]));
    const float ox = (float)cx[i], oy = (float)cy[i], oz = (float)cz[i];
    auto kk = [&](float a, float bb, float c) {
      const float kx = (float)(pxyz_[0] * (a + 0.00));
      const float ky = (float)(pxyz_[1] * (bb + 0.00));
      const float kz = (float)(pxyz_[2] * (c + 0.00));
      return kx * kx + ky * ky + kz * kz;
    };
    const float pup = gauss((int)nx, (int)ny, (int)nz, coef_) * kk(nx, ny, nz);
    const float qup = proposal_q((int)ox, mhd_[0]) * proposal_q((int)oy, mhd_[1]) *
                      proposal_q((int)oz, mhd_[2]);
    const float pdown = gauss((int)ox, (int)oy, (int)oz, coef_) * kk(ox, oy, oz);
    const float qdown = proposal_q((int)nx, mhd_[0]) * proposal_q((int)ny, mhd_[1]) *
                        proposal_q((int)nz, mhd_[2]);
    const float r = pup * qup / (pdown * qdown);
    const float acc = r > 1.0 ? 1.0f : r;
    const bool take = uniform() < acc;
    cx[i + 1] = (int)(take ? nx : ox);
    cy[i + 1] = (int)(take ? ny : oy);
    cz[i + 1] = (int)(take ? nz : oz);
    const int msq = cx[i + 1] * cx[i + 1] + cy[i + 1] * cy[i + 1] + cz[i + 1] * cz[i + 1];
    if (msq == 0 || msq <= kcut_) --i;  // the slot is drawn again
  }
  // keep the odd entries
  std::vector<float> kx(P_), ky(P_), kz(P_);
  for (int i = 0; i < P_; ++i) {
    out[3 * i] = cx[2 * i + 1];
    out[3 * i + 1] = cy[2 * i + 1];
    out[3 * i + 2] = cz[2 * i + 1];
    kx[i] = (float)(cx[2 * i + 1] + 0.00);
    ky[i] = (float)(cy[2 * i + 1] + 0.00);
    kz[i] = (float)(cz[2 * i + 1] + 0.00);
  }
  // virial chain over the kept samples: target f_npt(k) k^4, proposal the
  // sample itself (independence sampler with density f(k) k^2)
  int32_t* idx = out + 3 * P_;
  idx[0] = 0;
  for (int i = 1; i < P_; i++) {
    const int o = idx[i - 1];
    auto k2 = [&](int t) {
      const float a = pxyz_[0] * kx[t], bb = pxyz_[1] * ky[t], c = pxyz_[2] * kz[t];
      return a * a + bb * bb + c * c;
    };
    const float k2n = k2(i), k2o = k2(o);
    const float pup = gauss((int)kx[i], (int)ky[i], (int)kz[i], coef_npt_) * (k2n * k2n);
    const float qup = gauss((int)kx[o], (int)ky[o], (int)kz[o], coef_) * k2o;
    const float pdown = gauss((int)kx[o], (int)ky[o], (int)kz[o], coef_npt_) * (k2o * k2o);
    const float qdown = gauss((int)kx[i], (int)ky[i], (int)kz[i], coef_) * k2n;
    const float r = pup * qup / (pdown * qdown);
    const float acc = r > 1.0 ? 1.0f : r;
    idx[i] = uniform() < acc ? i : o;
  }
}

}  // namespace rbmdx
