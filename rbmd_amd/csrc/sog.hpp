// RBSOG (random batch sum-of-Gaussians) k-space and the VERLET-SOG near field:
// host-side set-up and the per-evaluation k-sampler.
//
// Follows (paths relative to the SOG-AITech/rbmd source root):
//   * level presets        src/force/include/rbsog_level_preset.h:15-35
//   * RBSOGInit / Setup    src/force/src/kspace_calculator.cpp:988-1140, :1545-1588
//   * the two chains       kspace_calculator.cpp:1142-1329, :1620-1651
//   * ComputeRBSOG         kspace_calculator.cpp:1331-1543
//   * LJCoulSOGInit        src/force/src/lj_cut_coul_kspace.cpp:651-711
//
// The far field is f(k) = sum_{l<M} coef_l exp(-b^{2l} sigma^2 k^2 / 2).  Two
// coefficient sets exist.  The reference set reproduces the reference's float
// arithmetic exactly, including two defects: levels l >= 1 of coef carry an
// extra factor w0, and the analytic high-k series of the normalisation S runs
// for level 0 only.  The exact set (coulomb.rbsog_exact = true, an extension
// key) evaluates the same definitions in double without those defects.
#pragma once
#include <cstdint>
#include <vector>

#include "sampler.hpp"

namespace rbmdx {

struct SogPresetValues {
  float b, sigma_rc10, accuracy;
  int mmax;
};
// Level 0..4; returns false for a level out of range.
bool sog_preset(int level, SogPresetValues* out);

struct SogSet {
  int level = 0, mmax = 0, kcut = 1;
  bool exact = false;
  double b = 0, sigma = 0, w0 = 0, S = 0, S_npt = 0;
  std::vector<double> coef, coef_npt;  // as used by the force path
  // float values the sampler's acceptance ratios use
  float b_f = 0, sigma_f = 0;
  std::vector<float> coef_f, coef_npt_f;
};

// Coefficients and normalisation for a box of edge lengths L and cutoff rc.
SogSet sog_set(int level, double cut_off, const double L[3], int kcut,
               bool exact);

// f(k^2) in double with the given coefficients.
double sog_f(const SogSet& s, const std::vector<double>& coef, double k2);

// Self energy -(ln b / (sqrt(2 pi) sigma)) (w0 + (1 - b^-M)/(b - 1)) qqr2e sum q^2.
double sog_self_energy(const SogSet& s, double sum_q2, double qqr2e);

// Near field consistent with the set's far field: Taylor coefficients
// T[0..5] of the force polynomial in r^2,
//   f = -qqr2e q_i q_j (1/r^3 + sum_n T_n r^{2n}),
// and the Gaussian-sum weights omega_l, 1/(2 s_l^2) of the energy
//   E = 1/2 qqr2e q_i q_j (1/r - sum_l omega_l exp(-r^2/(2 s_l^2))).
struct SogNear {
  double w0 = 0, T[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> omega, inv2s2;
};
SogNear sog_near(const SogSet& s);

// Per-evaluation sampler: writes P m-vectors (3P ints) followed by the P
// indices of the virial chain.
class SogSampler : public KSampler {
 public:
  SogSampler(const SogSet& s, const double L[3], int P);
  void draw(uint32_t seed, int32_t* out) const override;
  int out_size() const override { return 4 * P_; }
  int P() const { return P_; }

 private:
  float gauss(int mx, int my, int mz, const std::vector<float>& coef) const;
  int P_, mmax_, kcut_;
  float L_[3], sigma_, b_, pxyz_[3], factor_[3], mhd_[3];
  std::vector<float> coef_, coef_npt_;
};

}  // namespace rbmdx
