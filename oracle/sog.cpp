// CPU oracle, part 2: RBSOG k-space and the VERLET-SOG near field.
//
// TEST INFRASTRUCTURE ONLY (see oracle.cpp): loaded by tests/ and bench.py's
// CPU leg, never by the product library.
//
// Restates (paths relative to the SOG-AITech/rbmd source root):
//   * the level presets                 src/force/include/rbsog_level_preset.h:15-35
//   * w0, coefficients, normalisation S  src/force/src/kspace_calculator.cpp:988-1140,
//                                        :1545-1557 (G_sigma, Compute_W0)
//   * the two Metropolis chains          kspace_calculator.cpp:1142-1329, :1559-1647
//   * sampled + direct force/energy/virial kspace_calculator_op.hip.cu:955-1383,
//                                        driven by kspace_calculator.cpp:1331-1543
//   * the VERLET-SOG near field          src/force/src/lj_cut_coul_kspace.cpp:651-711,
//                                        lj_cut_coul_kspace_op.hip.cu:196-252
//
// Two coefficient sets are produced.  "reference": the float arithmetic of the
// reference, including two of its defects: every level l >= 1 of the k-space
// coefficient carries an extra factor w0 (the i == 1 special case is
// overwritten), and the high-k analytic series of S runs for level 0 only (the
// loop bound flag is never reset).  "exact": the same definitions evaluated in
// double with both defects removed: coef_l = 4 pi ln(b) sigma^2 (w0 for l = 0,
// b^{2l} otherwise), S = sum over |m|^2 > Kcut of f(k) k^2 summed to
// convergence for every level.
//
// Pinning: S against a brute-force lattice sum (exact set) and against an
// independent recomputation of the reference's level-0-only series; the far
// field f against 1/r beyond the cutoff to the preset's stated accuracy; the
// Taylor near-field force against the derivative of the Gaussian sum; k-space
// force = -grad of k-space energy; chi-square of the sampled m-marginals.
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;

struct SogPreset {
  float b, sigma_rc10, accuracy;
  int mmax;
};
// rbsog_level_preset.h:15-35 (stored as rbmd::Real = float there)
const SogPreset kSogPresets[5] = {
    {2.0f, 5.027010924194599f, 2.289e-3f, 6},
    {1.62976708826776469f, 3.633717409009413f, 1.158e-4f, 16},
    {1.48783512395703226f, 2.662784519725113f, 1.142e-5f, 30},
    {1.32070036405934420f, 2.277149356440992f, 5.583e-8f, 64},
    {1.21812525709410644f, 1.774456369233284f, 3.389e-11f, 102},
};

// ---- float mirror of the reference set-up ---------------------------------
// G_sigma, kspace_calculator.cpp:1545-1547 (float Real, expf / sqrtf)
float g_sigma_f(float sigma, float r) {
  return std::exp(-r * r / (2 * sigma * sigma)) /
         std::sqrt((float)(2 * kPi * sigma * sigma));
}
// Compute_W0, kspace_calculator.cpp:1549-1557
float w0_f(float r0, float b) {
  float sum = 0.0f;
  for (int i = 1; i < 200; i++)
    sum = sum + std::pow(b, (float)(-i)) *
                    g_sigma_f(1.0f, std::pow(b, (float)(-i)) * r0);
  return (float)((1.0 / g_sigma_f(1.0f, r0)) *
                 ((1.0 / (2 * std::log(b) * r0)) - (double)sum));
}
// Gaussian / Gaussian_modify, kspace_calculator.cpp:1559-1588: f(m) with the
// k^2 of an integer m formed through double then powf.
float gauss_f(int kx, int ky, int kz, const float L[3], float sigma, float b,
              int mmax, const float* coef) {
  const float k2 = std::pow((float)((kx + 0.00) * 2 * kPi / L[0]), 2.0f) +
                   std::pow((float)((ky + 0.00) * 2 * kPi / L[1]), 2.0f) +
                   std::pow((float)((kz + 0.00) * 2 * kPi / L[2]), 2.0f);
  float sum = 0.0f;
  const float b2 = b * b;
  for (int i = 0; i < mmax; i++)
    sum = sum + coef[i] * std::exp((float)(-(std::pow(b2, (float)(i + 0.0)) *
                                             sigma * sigma) *
                                           k2 / 2.0));
  return sum;
}

struct SogRef {
  float b, sigma, w0, S, S_npt;
  int mmax;
  std::vector<float> coef, coef_npt, sl;
};

// RBSOGInit (:988-1042) and RBSOGSetup (:1044-1140) in float.
SogRef sog_reference(int level, float cut_off, const float L[3], int kcut) {
  SogRef s;
  const SogPreset& p = kSogPresets[level];
  s.b = p.b;
  s.sigma = (float)(p.sigma_rc10 * cut_off / 10.0);
  s.mmax = p.mmax;
  const float r0 = cut_off / s.sigma;
  s.w0 = w0_f(r0, s.b);
  s.coef.assign(s.mmax, 0.f);
  s.coef_npt.assign(s.mmax, 0.f);
  s.sl.assign(s.mmax, 0.f);
  s.sl[0] = std::sqrt(2.0f) * s.sigma;
  const float sigma2 = s.sigma * s.sigma, sigma4 = sigma2 * sigma2;
  const float b2 = s.b * s.b, b4 = b2 * b2;
  s.coef[0] = (float)(4 * kPi * std::log(s.b) * s.w0 * sigma2);
  s.coef_npt[0] = (float)(8 * kPi * std::log(s.b) * s.w0 * sigma4);
  for (int i = 1; i < s.mmax; ++i) {
    s.sl[i] = s.sl[i - 1] * s.b;
    s.coef[i] = s.coef[i - 1] * b2;  // the i == 1 special case is overwritten
    s.coef_npt[i] = s.coef_npt[i - 1] * b4;
  }
  const float prd = std::max(L[0], std::max(L[1], L[2]));
  float sum = 0, sum_npt = 0, sum_1 = 0, sum_npt_1 = 0;
  const float log_b = std::log(s.b);
  float bool_bound = 1.0f;  // never reset between levels (reference)
  for (int i = 0; i < kcut + 1; i++) {
    const float Kx = (float)((2 * kPi / L[0]) * (i + 0.00));
    for (int j = 0; j < kcut + 1; j++) {
      const float Ky = (float)((2 * kPi / L[1]) * (j + 0.00));
      for (int k = 0; k < kcut + 1; k++) {
        const float Kz = (float)((2 * kPi / L[2]) * (k + 0.00));
        if ((i * i + j * j + k * k) <= kcut) {
          const float K2 = Kx * Kx + Ky * Ky + Kz * Kz;
          const float K4 = K2 * K2;
          sum_1 += 2 * gauss_f(i, j, k, L, s.sigma, s.b, s.mmax, s.coef.data()) * K2;
          sum_npt_1 +=
              2 * gauss_f(i, j, k, L, s.sigma, s.b, s.mmax, s.coef_npt.data()) * K4;
        }
      }
    }
  }
  for (int i = 0; i < s.mmax; ++i) {
    float Hx = 0, Hy = 0, Hz = 0, Yx = 0, Yy = 0, Yz = 0, Yxx = 0, Yyy = 0, Yzz = 0;
    int j = 1;
    while (bool_bound > 1e-12) {
      const float kx = (float)(kPi * j / L[0]);
      const float ky = (float)(kPi * j / L[1]);
      const float kz = (float)(kPi * j / L[2]);
      const float k_min = (float)(kPi * j / prd);
      const float k_min2 = k_min * k_min, k_min4 = k_min2 * k_min2;
      const float sl_i = s.sl[i], sl_i_sq = sl_i * sl_i;
      const float kx2_s = kx * sl_i * kx * sl_i;
      const float ky2_s = ky * sl_i * ky * sl_i;
      const float kz2_s = kz * sl_i * kz * sl_i;
      const float kx2 = 4 * kx * kx, ky2 = 4 * ky * ky, kz2 = 4 * kz * kz;
      const float kx4 = 16 * kx * kx * kx * kx, ky4 = 16 * ky * ky * ky * ky,
                  kz4 = 16 * kz * kz * kz * kz;
      Hx += 2 * std::exp(-kx2_s);
      Yx += 2 * std::exp(-kx2_s) * kx2;
      Yxx += 2 * std::exp(-kx2_s) * kx4;
      Hy += 2 * std::exp(-ky2_s);
      Yy += 2 * std::exp(-ky2_s) * ky2;
      Yyy += 2 * std::exp(-ky2_s) * ky4;
      Hz += 2 * std::exp(-kz2_s);
      Yz += 2 * std::exp(-kz2_s) * kz2;
      Yzz += 2 * std::exp(-kz2_s) * kz4;
      ++j;
      bool_bound = std::exp(4 * i * log_b - k_min2 * sl_i_sq) * k_min4 * sigma4;
    }
    Hx += 1;
    Hy += 1;
    Hz += 1;
    sum += s.coef[i] * (Yx * Hy * Hz + Hx * Yy * Hz + Hx * Hy * Yz);
    sum_npt += s.coef_npt[i] * (Yxx * Hy * Hz + Hx * Yyy * Hz + Hx * Hy * Yzz +
                                2 * Yx * Yy * Hz + 2 * Yx * Hy * Yz + 2 * Hx * Yy * Yz);
  }
  s.S = sum - sum_1;
  s.S_npt = sum_npt - sum_npt_1;
  return s;
}

// ---- exact set (double) ----------------------------------------------------
double w0_d(double r0, double b) {
  auto g = [](double r) { return std::exp(-r * r / 2) / std::sqrt(2 * kPi); };
  double sum = 0;
  for (int i = 1; i < 200; i++) sum += std::pow(b, -i) * g(std::pow(b, -i) * r0);
  return (1.0 / g(r0)) * (1.0 / (2 * std::log(b) * r0) - sum);
}

struct SogExact {
  double b, sigma, w0, S, S_npt;
  int mmax;
  std::vector<double> coef, coef_npt;
};

// Per dimension: H = sum_j e^{-k_j^2 s^2/2}, Y = sum_j k_j^2 e^{..},
// Z = sum_j k_j^4 e^{..} over all integers j, k_j = 2 pi j / L.
void lattice_sums(double L, double s2, double& H, double& Y, double& Z) {
  H = 1, Y = 0, Z = 0;
  for (int j = 1;; ++j) {
    const double k = 2 * kPi * j / L, k2 = k * k;
    const double e = std::exp(-k2 * s2 / 2);
    H += 2 * e;
    Y += 2 * e * k2;
    Z += 2 * e * k2 * k2;
    if (e * k2 * k2 < 1e-300 || (j > 4 && e < 1e-30)) break;
  }
}

SogExact sog_exact(int level, double cut_off, const double L[3], int kcut) {
  SogExact s;
  const SogPreset& p = kSogPresets[level];
  s.b = p.b;
  s.sigma = (double)p.sigma_rc10 * cut_off / 10.0;
  s.mmax = p.mmax;
  s.w0 = w0_d(cut_off / s.sigma, s.b);
  const double lb = std::log(s.b), s2 = s.sigma * s.sigma;
  s.coef.resize(s.mmax);
  s.coef_npt.resize(s.mmax);
  for (int l = 0; l < s.mmax; ++l) {
    const double bl2 = std::pow(s.b, 2.0 * l);
    s.coef[l] = 4 * kPi * lb * s2 * (l == 0 ? s.w0 : bl2);
    s.coef_npt[l] = 8 * kPi * lb * s2 * s2 * (l == 0 ? s.w0 : bl2 * bl2);
  }
  double S = 0, Sn = 0;
  for (int l = 0; l < s.mmax; ++l) {
    const double sl2 = s2 * std::pow(s.b, 2.0 * l);
    double H[3], Y[3], Z[3];
    for (int d = 0; d < 3; ++d) lattice_sums(L[d], sl2, H[d], Y[d], Z[d]);
    S += s.coef[l] * (Y[0] * H[1] * H[2] + H[0] * Y[1] * H[2] + H[0] * H[1] * Y[2]);
    // sum of k^4 = sum (kx^2 + ky^2 + kz^2)^2
    Sn += s.coef_npt[l] * (Z[0] * H[1] * H[2] + H[0] * Z[1] * H[2] + H[0] * H[1] * Z[2] +
                           2 * Y[0] * Y[1] * H[2] + 2 * Y[0] * H[1] * Y[2] +
                           2 * H[0] * Y[1] * Y[2]);
  }
  // remove |m|^2 <= kcut
  const int bnd = (int)std::ceil(std::sqrt((double)kcut));
  for (int i = -bnd; i <= bnd; ++i)
    for (int j = -bnd; j <= bnd; ++j)
      for (int k = -bnd; k <= bnd; ++k) {
        if (i * i + j * j + k * k > kcut || (i == 0 && j == 0 && k == 0)) continue;
        const double kx = 2 * kPi * i / L[0], ky = 2 * kPi * j / L[1],
                     kz = 2 * kPi * k / L[2];
        const double k2 = kx * kx + ky * ky + kz * kz;
        double f = 0, fn = 0;
        for (int l = 0; l < s.mmax; ++l) {
          const double e = std::exp(-std::pow(s.b, 2.0 * l) * s2 * k2 / 2);
          f += s.coef[l] * e;
          fn += s.coef_npt[l] * e;
        }
        S -= f * k2;
        Sn -= fn * k2 * k2;
      }
  s.S = S;
  s.S_npt = Sn;
  return s;
}

// ---- the two Metropolis chains (float mirror) ------------------------------
struct SogChain {
  std::mt19937& gen;
  float uniform01() {  // UniformRandomValue(0.0, 1.0), :1620-1623
    std::uniform_real_distribution<float> dis(0.0f, 1.0f);
    return dis(gen);
  }
  float randn(float mean, float margin) {  // randn_box_muller, :1625-1641
    const float epsilon = 1.17549e-038f;
    const float two_pi = (float)(2.0 * kPi);
    float u1, u2;
    do {
      u1 = uniform01();
      u2 = uniform01();
    } while (u1 <= epsilon);
    const float z0 = std::sqrt((float)(-2.0 * (double)std::log(u1))) *
                     std::cos(two_pi * u2);
    return z0 * margin + mean;
  }
};
// MH_D_Modify, :1643-1651
float mhd_f(int xx, float factor) {
  if (xx == 0) return std::erf((float)(0.5 * factor));
  const float ax = std::fabs((float)xx);
  return (float)(0.5 * (std::erf((float)((ax + 0.5) * factor)) -
                        std::erf((float)((ax - 0.5) * factor))));
}

}  // namespace

extern "C" {

// Level preset: out = {b, sigma_rc10, accuracy, mmax}.
void orc_sog_preset(int level, double* out) {
  const SogPreset& p = kSogPresets[level];
  out[0] = p.b;
  out[1] = p.sigma_rc10;
  out[2] = p.accuracy;
  out[3] = p.mmax;
}

// Coefficients and normalisation.  exact = 0: the reference's float set-up;
// 1: the corrected double set-up.  scal = {b, sigma, w0, S, S_npt, mmax};
// coef / coef_npt receive mmax values each.
void orc_sog_setup(int level, double cut_off, const double* L, int kcut,
                   int exact, double* scal, double* coef, double* coef_npt) {
  if (!exact) {
    const float Lf[3] = {(float)L[0], (float)L[1], (float)L[2]};
    const SogRef s = sog_reference(level, (float)cut_off, Lf, kcut);
    scal[0] = s.b;
    scal[1] = s.sigma;
    scal[2] = s.w0;
    scal[3] = s.S;
    scal[4] = s.S_npt;
    scal[5] = s.mmax;
    for (int l = 0; l < s.mmax; ++l) {
      coef[l] = s.coef[l];
      coef_npt[l] = s.coef_npt[l];
    }
  } else {
    const SogExact s = sog_exact(level, cut_off, L, kcut);
    scal[0] = s.b;
    scal[1] = s.sigma;
    scal[2] = s.w0;
    scal[3] = s.S;
    scal[4] = s.S_npt;
    scal[5] = s.mmax;
    for (int l = 0; l < s.mmax; ++l) {
      coef[l] = s.coef[l];
      coef_npt[l] = s.coef_npt[l];
    }
  }
}

// The k-sample of force evaluation `call_index`: RBSOGSampleKSpace
// (kspace_calculator.cpp:1142-1329) with the evaluation's seed from the base
// stream (rbmd_parallel_until.cpp:484-501: the c-th output of
// std::mt19937(base_seed)).  coef / coef_npt are the float coefficients.
// Writes P integer m-vectors and the P indices of the second (virial) chain.
void orc_sog_sample(uint32_t base_seed, long call_index, int P, const double* L,
                    double sigma_d, double b_d, int mmax, const double* coef_d,
                    const double* coef_npt_d, int kcut, int* m_out, int* idx_out) {
  std::mt19937 base(base_seed);
  if (call_index > 0) base.discard((unsigned long long)call_index);
  std::mt19937 gen((uint32_t)base());
  SogChain ch{gen};
  const float Lf[3] = {(float)L[0], (float)L[1], (float)L[2]};
  const float sigma = (float)sigma_d, b = (float)b_d;
  std::vector<float> coef(mmax), coef_npt(mmax);
  for (int l = 0; l < mmax; ++l) {
    coef[l] = (float)coef_d[l];
    coef_npt[l] = (float)coef_npt_d[l];
  }
  float pxyz[3], factor[3], mhd[3];
  for (int d = 0; d < 3; ++d) {
    pxyz[d] = (float)(2 * kPi / Lf[d]);
    factor[d] = (float)(Lf[d] / (2 * kPi * sigma));
    mhd[d] = std::sqrt((float)(2 * sigma * sigma * kPi * kPi)) / Lf[d];
  }
  std::vector<int> mx(2 * P), my(2 * P), mz(2 * P);
  do {
    mx[0] = (int)std::round(ch.randn(0, factor[0]));
    my[0] = (int)std::round(ch.randn(0, factor[1]));
    mz[0] = (int)std::round(ch.randn(0, factor[2]));
  } while (mx[0] == 0 && my[0] == 0 && mz[0] == 0);
  auto G = [&](int a, int bb, int c) {
    return gauss_f(a, bb, c, Lf, sigma, b, mmax, coef.data());
  };
  auto Gn = [&](int a, int bb, int c) {
    return gauss_f(a, bb, c, Lf, sigma, b, mmax, coef_npt.data());
  };
  for (int i = 0; i < 2 * P - 1; i++) {
    const float x = ch.randn(0, factor[0]);
    const float mold_x = (float)mx[i], mnew_x = std::round(x);
    const float xx = ch.randn(0, factor[1]);
    const float mold_y = (float)my[i], mnew_y = std::round(xx);
    const float xxx = ch.randn(0, factor[2]);
    const float mold_z = (float)mz[i], mnew_z = std::round(xxx);
    const float Kx_new = (float)(pxyz[0] * (mnew_x + 0.00));
    const float Ky_new = (float)(pxyz[1] * (mnew_y + 0.00));
    const float Kz_new = (float)(pxyz[2] * (mnew_z + 0.00));
    const float K2_new = Kx_new * Kx_new + Ky_new * Ky_new + Kz_new * Kz_new;
    const float Kx_old = (float)(pxyz[0] * (mold_x + 0.00));
    const float Ky_old = (float)(pxyz[1] * (mold_y + 0.00));
    const float Kz_old = (float)(pxyz[2] * (mold_z + 0.00));
    const float K2_old = Kx_old * Kx_old + Ky_old * Ky_old + Kz_old * Kz_old;
    const float pup = G((int)mnew_x, (int)mnew_y, (int)mnew_z) * K2_new;
    const float qup = mhd_f((int)mold_x, mhd[0]) * mhd_f((int)mold_y, mhd[1]) *
                      mhd_f((int)mold_z, mhd[2]);
    const float pdown = G((int)mold_x, (int)mold_y, (int)mold_z) * K2_old;
    const float qdown = mhd_f((int)mnew_x, mhd[0]) * mhd_f((int)mnew_y, mhd[1]) *
                        mhd_f((int)mnew_z, mhd[2]);
    const float ratio = pup * qup / (pdown * qdown);
    const float acce = ratio > 1.0 ? 1.0f : ratio;
    const float yyy = ch.uniform01();
    if (yyy < acce) {
      mx[i + 1] = (int)mnew_x;
      my[i + 1] = (int)mnew_y;
      mz[i + 1] = (int)mnew_z;
    } else {
      mx[i + 1] = (int)mold_x;
      my[i + 1] = (int)mold_y;
      mz[i + 1] = (int)mold_z;
    }
    const int m_sq = mx[i + 1] * mx[i + 1] + my[i + 1] * my[i + 1] + mz[i + 1] * mz[i + 1];
    if ((mx[i + 1] == 0 && my[i + 1] == 0 && mz[i + 1] == 0) || (m_sq <= kcut))
      i = i - 1;
  }
  std::vector<float> ks[3];
  for (int d = 0; d < 3; ++d) ks[d].resize(P);
  for (int i = 0; i < P; i++) {
    ks[0][i] = (float)(mx[2 * i + 1] + 0.00);
    ks[1][i] = (float)(my[2 * i + 1] + 0.00);
    ks[2][i] = (float)(mz[2 * i + 1] + 0.00);
    m_out[3 * i] = mx[2 * i + 1];
    m_out[3 * i + 1] = my[2 * i + 1];
    m_out[3 * i + 2] = mz[2 * i + 1];
  }
  std::vector<int> idx(P, 0);
  for (int i = 1; i < P; i++) {
    const int id = idx[i - 1];
    const float ox = ks[0][id], oy = ks[1][id], oz = ks[2][id];
    const float nx = ks[0][i], ny = ks[1][i], nz = ks[2][i];
    const float Kxn = pxyz[0] * nx, Kyn = pxyz[1] * ny, Kzn = pxyz[2] * nz;
    const float K2n = Kxn * Kxn + Kyn * Kyn + Kzn * Kzn, K4n = K2n * K2n;
    const float Kxo = pxyz[0] * ox, Kyo = pxyz[1] * oy, Kzo = pxyz[2] * oz;
    const float K2o = Kxo * Kxo + Kyo * Kyo + Kzo * Kzo, K4o = K2o * K2o;
    const float pup = Gn((int)nx, (int)ny, (int)nz) * K4n;
    const float qup = G((int)ox, (int)oy, (int)oz) * K2o;
    const float pdown = Gn((int)ox, (int)oy, (int)oz) * K4o;
    const float qdown = G((int)nx, (int)ny, (int)nz) * K2n;
    const float ratio = pup * qup / (pdown * qdown);
    const float acce = ratio > 1.0 ? 1.0f : ratio;
    const float yyy = ch.uniform01();
    idx[i] = (yyy < acce) ? i : id;
  }
  for (int i = 0; i < P; ++i) idx_out[i] = idx[i];
}

// f(k) = sum_l coef_l exp(-b^{2l} sigma^2 k^2 / 2) in double.
double orc_sog_f(double k2, double sigma, double b, int mmax, const double* coef) {
  double f = 0;
  for (int l = 0; l < mmax; ++l)
    f += coef[l] * std::exp(-std::pow(b, 2.0 * l) * sigma * sigma * k2 / 2.0);
  return f;
}

// RBSOG k-space in double: sampled part (P vectors m, virial chain idx) and the
// direct part (|m|^2 <= kcut, m != 0).  Forces fx..fz (n), energy[0] =
// sampled + direct (self term excluded), virial[6] (xx yy zz xy xz yz).
// ComputeRBSOG (kspace_calculator.cpp:1331-1543) with its kernels
// (kspace_calculator_op.hip.cu:955-1060, 1133-1383); the volume-rescale
// factors fac, fac_npt are identically 1 there (xprd0 = xprd).
void orc_sog_kspace_f64(int n, const double* x, const double* y, const double* z,
                        const double* q, const double* L, int P, const int* m,
                        const int* idx, double S, double S_npt, double sigma,
                        double b, int mmax, const double* coef,
                        const double* coef_npt, int kcut, double qqr2e,
                        double* fx, double* fy, double* fz, double* energy,
                        double* virial) {
  const double V = L[0] * L[1] * L[2];
  for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0;
  for (int a = 0; a < 6; ++a) virial[a] = 0;
  double e = 0;
  auto rho_of = [&](double kx, double ky, double kz, double& re, double& im) {
    re = im = 0;
    for (int i = 0; i < n; ++i) {
      const double t = kx * x[i] + ky * y[i] + kz * z[i];
      re += q[i] * std::cos(t);
      im += q[i] * std::sin(t);
    }
  };
  auto force_add = [&](double kx, double ky, double kz, double re, double im,
                       double scale) {
    for (int i = 0; i < n; ++i) {
      const double t = kx * x[i] + ky * y[i] + kz * z[i];
      const double g = (std::cos(t) * im - std::sin(t) * re) * scale * q[i];
      fx[i] += g * kx;
      fy[i] += g * ky;
      fz[i] += g * kz;
    }
  };
  // sampled
  std::vector<double> K(3 * (size_t)P), re(P), im(P);
  for (int j = 0; j < P; ++j) {
    for (int d = 0; d < 3; ++d) K[3 * j + d] = 2 * kPi * m[3 * j + d] / L[d];
    rho_of(K[3 * j], K[3 * j + 1], K[3 * j + 2], re[j], im[j]);
  }
  const double Mt = (S / P) * qqr2e / (2.0 * V);
  const double Mn = -(S_npt / P) * qqr2e / (4.0 * V);
  for (int j = 0; j < P; ++j) {
    const double kx = K[3 * j], ky = K[3 * j + 1], kz = K[3 * j + 2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    force_add(kx, ky, kz, re[j], im[j], -(S / P) * qqr2e / V / k2);
    const double r2 = re[j] * re[j] + im[j] * im[j];
    e += Mt * r2 / k2;
    const int s = idx[j];
    const double nx = K[3 * s], ny = K[3 * s + 1], nz = K[3 * s + 2];
    const double n2 = nx * nx + ny * ny + nz * nz;
    const double c2 = (re[s] * re[s] + im[s] * im[s]) / (n2 * n2);
    virial[0] += r2 / k2 * Mt + c2 * Mn * nx * nx;
    virial[1] += r2 / k2 * Mt + c2 * Mn * ny * ny;
    virial[2] += r2 / k2 * Mt + c2 * Mn * nz * nz;
    virial[3] += c2 * Mn * nx * ny;
    virial[4] += c2 * Mn * nx * nz;
    virial[5] += c2 * Mn * ny * nz;
  }
  // direct
  const int bnd = (int)std::ceil(std::sqrt((double)kcut));
  const double c1 = qqr2e / (2.0 * V), c2 = -qqr2e / (4.0 * V);
  for (int i = -bnd; i <= bnd; ++i)
    for (int j = -bnd; j <= bnd; ++j)
      for (int k = -bnd; k <= bnd; ++k) {
        if ((i == 0 && j == 0 && k == 0) || i * i + j * j + k * k > kcut) continue;
        const double kx = 2 * kPi * i / L[0], ky = 2 * kPi * j / L[1],
                     kz = 2 * kPi * k / L[2];
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double fb = orc_sog_f(k2, sigma, b, mmax, coef);
        const double fbn = orc_sog_f(k2, sigma, b, mmax, coef_npt);
        double rr, ii;
        rho_of(kx, ky, kz, rr, ii);
        force_add(kx, ky, kz, rr, ii, fb * (-qqr2e / V));
        const double r2 = rr * rr + ii * ii;
        e += c1 * fb * r2;
        virial[0] += c1 * fb * r2 + c2 * fbn * r2 * kx * kx;
        virial[1] += c1 * fb * r2 + c2 * fbn * r2 * ky * ky;
        virial[2] += c1 * fb * r2 + c2 * fbn * r2 * kz * kz;
        virial[3] += c2 * fbn * r2 * kx * ky;
        virial[4] += c2 * fbn * r2 * kx * kz;
        virial[5] += c2 * fbn * r2 * ky * kz;
      }
  *energy = e;
}

// Self energy, kspace_calculator.cpp:1524-1528.
double orc_sog_self(double sum_q2, double sigma, double b, int mmax, double w0,
                    double qqr2e) {
  const double coeff = (std::log(b) / (std::sqrt(2 * kPi) * sigma)) *
                       (w0 + (1 - std::pow(b, -mmax)) / (b - 1));
  return -coeff * sum_q2 * qqr2e;
}

// Far-field kernel in real space: sum_l omega_l exp(-r^2 / (2 s_l^2)) with
// omega_l = 2 ln(b) / sqrt(2 pi sigma^2) (w0 for l = 0, b^-l otherwise), the
// Gaussian sum that the near field subtracts (lj_cut_coul_kspace_op.hip.cu:
// 229-241).
double orc_sog_far_r(double r, double sigma, double b, int mmax, double w0) {
  const double norm = 2.0 * std::log(b) / std::sqrt(2.0 * kPi * sigma * sigma);
  double s = norm * w0 * std::exp(-r * r / (2 * sigma * sigma));
  for (int l = 1; l < mmax; ++l)
    s += norm * std::pow(b, -l) *
         std::exp(-r * r / (2 * sigma * sigma * std::pow(b, 2.0 * l)));
  return s;
}

// Taylor coefficients of the near-field force, LJCoulSOGInit
// (lj_cut_coul_kspace.cpp:651-711): T_n = (-1)^{n+1} 2 c sum_l v_l a_l^n / n!
// with c = ln b / (sigma^2 sqrt(2 pi sigma^2)), v_0 = w0, v_l = b^{-3l},
// a_l = 1 / (2 sigma^2 b^{2l}).  In double.
void orc_sog_taylor(double sigma, double b, int mmax, double w0, double* T) {
  const double c = std::log(b) / (sigma * sigma * std::sqrt(2 * kPi * sigma * sigma));
  double fact = 1;
  for (int nn = 0; nn < 6; ++nn) {
    if (nn > 0) fact *= nn;
    double s = 0;
    for (int l = 0; l < mmax; ++l) {
      const double v = l == 0 ? w0 : std::pow(b, -3.0 * l);
      const double a = 1.0 / (2 * sigma * sigma * std::pow(b, 2.0 * l));
      s += v * std::pow(a, nn) / fact;
    }
    T[nn] = std::pow(-1.0, nn + 1.0) * 2.0 * c * s;
  }
}

// VERLET-SOG pair forces in double over a full list within rc (LJ cut at
// rc_lj <= rc): LJ 12-6 plus the SOG near-field Coulomb
// f = -qqr2e q_i q_j (1/r^3 + sum_n T_n r^{2n}),
// E = 1/2 qqr2e q_i q_j (1/r - far(r)) per directed pair.
// ComputeLJCutCoulForceUserKernel, lj_cut_coul_kspace_op.hip.cu:357-460.
void orc_sog_verlet_f64(int n, const double* x, const double* y, const double* z,
                        const double* q, const int* type, int ntypes,
                        const double* eps, const double* sig, const double* L,
                        double rc, double rc_lj, double qqr2e, double sigma,
                        double b, int mmax, double w0, const double* T,
                        double* fx, double* fy, double* fz, double* energies,
                        double* virial) {
  (void)ntypes;
  for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0;
  for (int a = 0; a < 6; ++a) virial[a] = 0;
  double elj = 0, ec = 0;
  auto mi = [](double d, double Ld) {
    if (std::fabs(d) > 0.5 * Ld) d = d > 0 ? d - Ld : d + Ld;
    return d;
  };
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (i == j) continue;
      const double dx = mi(x[j] - x[i], L[0]), dy = mi(y[j] - y[i], L[1]),
                   dz = mi(z[j] - z[i], L[2]);
      const double d2 = dz * dz + dy * dy + dx * dx;
      if (!(d2 < rc * rc)) continue;
      double f = 0;
      if (d2 < rc_lj * rc_lj) {
        const double e_ij = std::sqrt(eps[type[i]] * eps[type[j]]);
        const double s_ij = 0.5 * (sig[type[i]] + sig[type[j]]);
        const double s6 = std::pow(s_ij * s_ij / d2, 3);
        f += -24 * e_ij * (2 * s6 - 1) * s6 / d2;
        elj += 2 * e_ij * (s6 - 1) * s6;
      }
      const double qq = q[i] * q[j];
      if (qq != 0) {
        const double r = std::sqrt(d2);
        double poly = 0, p = 1;
        for (int nn = 0; nn < 6; ++nn) {
          poly += T[nn] * p;
          p *= d2;
        }
        f += -qqr2e * qq * (1 / (r * d2) + poly);
        ec += 0.5 * qqr2e * qq * (1 / r - orc_sog_far_r(r, sigma, b, mmax, w0));
      }
      fx[i] += f * dx;
      fy[i] += f * dy;
      fz[i] += f * dz;
      virial[0] += -0.5 * dx * dx * f;
      virial[1] += -0.5 * dy * dy * f;
      virial[2] += -0.5 * dz * dz * f;
      virial[3] += -0.5 * dx * dy * f;
      virial[4] += -0.5 * dx * dz * f;
      virial[5] += -0.5 * dy * dz * f;
    }
  energies[0] = elj;
  energies[1] = ec;
}

}  // extern "C"
