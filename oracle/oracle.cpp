// CPU oracle for the RBMD per-step force path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
// product library (librbmdx.so) never links, loads or calls anything here.
//
// It restates, as plain CPU code, the algorithm of the reference files cited
// at each function (paths relative to the SOG-AITech/rbmd source root).  Two
// arithmetic modes are provided:
//   * FP64    - the error reference: every force/energy/virial term in double.
//   * FP32    - a "mirror" that evaluates pair terms in float, in the order the
//               product defines them.
// Neighbour classification (which pairs exist, which shell pairs are kept)
// is always evaluated with the FP32 mirror distance of float positions (or the
// FP64 distance of double positions for the all-double LJ path), so that FP32
// and FP64 force evaluations share one pair set, one RBL selection and one
// k-sample, which is how force RMS error against FP64 is defined.
//
// Pinning (see tests/test_oracle_*.py): brute-force O(N^2) pair sets, closed
// form LJ / erfc-Coulomb pair values, finite differences of energy vs force
// and of energy vs box strain (virial), Newton's third law, the NaCl Madelung
// constant through the Ewald sum, alpha-independence of the total Ewald
// energy, the Poisson dual of the RBE normalisation S, the SplitMix64 test
// vector, the C++-standard mt19937 check value and chi-square tests of the
// k-sampler against its target distribution.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Min-image and distance.  box.h:51-73 (single-shift MinImageDistance);
// full_neighbor_list_op.h:98-106 (squared distance of the shifted vector).
// The FP32 mirror is d2 = fma(dz,dz, fma(dy,dy, dx*dx)): the product's kernels
// evaluate exactly this expression, so pair sets agree bit for bit.
// ---------------------------------------------------------------------------
template <typename T>
inline T min_image_1(T d, T L) {
  const T half = T(0.5) * L;
  if (std::fabs(d) > half) d = d > T(0) ? d - L : d + L;
  return d;
}

inline float dist2_f32(float xi, float yi, float zi, float xj, float yj,
                       float zj, const float L[3]) {
  float dx = min_image_1<float>(xj - xi, L[0]);
  float dy = min_image_1<float>(yj - yi, L[1]);
  float dz = min_image_1<float>(zj - zi, L[2]);
  return std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
}

inline double dist2_f64(double xi, double yi, double zi, double xj, double yj,
                        double zj, const double L[3]) {
  double dx = min_image_1<double>(xj - xi, L[0]);
  double dy = min_image_1<double>(yj - yi, L[1]);
  double dz = min_image_1<double>(zj - zi, L[2]);
  return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

// Cell-binned enumeration of every ordered pair (i != j) whose distance is
// below a radius.  A plain linked-cell restatement (linked_cell.cpp:251-411,
// full_neighbor_list_op.hip.cu:177-288); any dimension with fewer than three
// cells is scanned without a stencil, so no pair is visited twice (the
// reference duplicates pairs there).  Callback gets (i, j).
template <typename P, typename F>
void for_each_candidate(int n, const P* x, const P* y, const P* z,
                        const double L[3], double radius, F&& fn) {
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    nc[d] = (int)std::floor(L[d] / radius);
    if (nc[d] < 3) nc[d] = 1;  // scan the whole dimension
  }
  const long ncell = (long)nc[0] * nc[1] * nc[2];
  std::vector<int> head(ncell, -1), next(n, -1), cell_of(n);
  for (int i = n - 1; i >= 0; --i) {
    int c[3];
    const P p[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; ++d) {
      double s = (double)p[d] / L[d];
      s -= std::floor(s);
      c[d] = std::min(nc[d] - 1, std::max(0, (int)(s * nc[d])));
    }
    const long id = ((long)c[2] * nc[1] + c[1]) * nc[0] + c[0];
    cell_of[i] = (int)id;
    next[i] = head[id];
    head[id] = i;
  }
  for (int i = 0; i < n; ++i) {
    const long id = cell_of[i];
    const int cx = (int)(id % nc[0]);
    const int cy = (int)((id / nc[0]) % nc[1]);
    const int cz = (int)(id / ((long)nc[0] * nc[1]));
    const int ex = nc[0] >= 3 ? 1 : 0, ey = nc[1] >= 3 ? 1 : 0,
              ez = nc[2] >= 3 ? 1 : 0;
    for (int oz = -ez; oz <= ez; ++oz)
      for (int oy = -ey; oy <= ey; ++oy)
        for (int ox = -ex; ox <= ex; ++ox) {
          const int jx = (cx + ox + nc[0]) % nc[0];
          const int jy = (cy + oy + nc[1]) % nc[1];
          const int jz = (cz + oz + nc[2]) % nc[2];
          const long jc = ((long)jz * nc[1] + jy) * nc[0] + jx;
          for (int j = head[jc]; j >= 0; j = next[j])
            if (j != i) fn(i, j);
        }
  }
}

// SplitMix64 finaliser, rbl_full_neighbor_list_op.hip.cu:6-11.
inline uint64_t mix64(uint64_t v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31);
}

// RblShellSampleSelected, rbl_full_neighbor_list_op.hip.cu:13-30: keep the
// directed shell pair (i -> j) of epoch e iff
//   mix(e ^ mix(gid_i) ^ mix(gid_j + 0xd1b54a32d192ed03)) mod f == 0.
inline bool rbl_selected(int f, long long epoch, int gid_i, int gid_j) {
  if (f <= 1) return true;
  uint64_t key = (uint64_t)epoch;
  key ^= mix64((uint64_t)(long long)gid_i);
  key ^= mix64((uint64_t)(long long)gid_j + 0xd1b54a32d192ed03ULL);
  return mix64(key) % (uint64_t)f == 0ULL;
}

// Pair force parameters, passed from Python as a flat double array.
struct PairParams {
  double r_core, rc, rc_lj, alpha, qqr2e;
  int freq;        // RBL selection frequency (pice_num)
  long long epoch; // RBL sampling epoch (the MD step)
  int rbl;         // 1 = RBL split, 0 = full (Verlet) list
  int coulomb;     // 1 = erfc real-space Coulomb on
  int mean_fix;    // 1 = subtract the mean pair force (RBL)
};

PairParams unpack(const double* p) {
  PairParams q;
  q.r_core = p[0];
  q.rc = p[1];
  q.rc_lj = p[2];
  q.alpha = p[3];
  q.qqr2e = p[4];
  q.freq = (int)p[5];
  q.epoch = (long long)p[6];
  q.rbl = (int)p[7];
  q.coulomb = (int)p[8];
  q.mean_fix = (int)p[9];
  return q;
}

// Float thresholds exactly as the reference forms them:
// rbl_full_neighbor_list_builder.cpp:63,92 (r_core^2 - EPSILON) and :302
// (cut_off^2 - EPSILON), EPSILON = 1e-4f (rbmd_define.h:51).
struct Thresholds {
  float core, shell, rs2, rc2, rclj2;
};
Thresholds thresholds(const PairParams& p) {
  Thresholds t;
  const float rs = (float)p.r_core, rc = (float)p.rc, rl = (float)p.rc_lj;
  t.rs2 = rs * rs;
  t.rc2 = rc * rc;
  t.rclj2 = rl * rl;
  t.core = t.rs2 - 1e-4f;
  t.shell = t.rc2 - 1e-4f;
  return t;
}

// Class of a directed pair under RBL (rbl_full_neighbor_list_op.hip.cu:237-304):
// 1 = core (d2 < r_core^2 - eps), 2 = selected shell, 0 = dropped.
inline int rbl_class(float d2, const Thresholds& t, const PairParams& p,
                     int gi, int gj) {
  if (d2 < t.core) return 1;
  if (d2 < t.shell && rbl_selected(p.freq, p.epoch, gi, gj)) return 2;
  return 0;
}

// Pair force weight of a directed pair.  Core entries: lj126_rs /
// CoulCutForce_rs_fix (lj_op.hip.cu:40-53, lj_cut_coul_kspace_op.hip.cu:134-157)
// require eps < d2 < rs^2.  Shell entries: lj126_rcs / CoulCutForce_rcs_fix
// (lj_op.hip.cu:56-73, lj_cut_coul_kspace_op.hip.cu:159-194) require
// rs^2 < d2 < rc^2 and carry the factor pice_num.  Full (Verlet) list:
// lj126 / CoulCutForce (lj_op.hip.cu:18-37) require d2 < rc^2.
inline double pair_weight(float d2, int cls, const Thresholds& t,
                          const PairParams& p) {
  if (!p.rbl) return d2 < t.rc2 ? 1.0 : 0.0;
  if (cls == 1) return (d2 < t.rs2 && d2 > 1e-4f) ? 1.0 : 0.0;
  if (cls == 2) return (d2 < t.rc2 && d2 > t.rs2) ? (double)p.freq : 0.0;
  return 0.0;
}

// LJ 12-6 (lj_op.hip.cu:18-37) with Lorentz-Berthelot mixing
// (lj_op.hip.cu:130-131).  Returns the scalar f such that F_i += f * (r_j - r_i);
// e = 0.5 * 4 eps (s6 - 1) s6, the half-counted entry energy.
template <typename T>
inline void lj_pair(T d2, T eps_ij, T sig_ij, T& f, T& e) {
  const T s2 = sig_ij * sig_ij / d2;
  const T s6 = s2 * s2 * s2;
  f = -T(24) * eps_ij * (T(2) * s6 - T(1)) * s6 / d2;
  e = T(0.5) * T(4) * eps_ij * (s6 - T(1)) * s6;
}

// erfc-screened Coulomb (lj_cut_coul_kspace_op.hip.cu:108-132):
// f = -qqr2e qi qj [erfc(g r)/r^2 + 2 g exp(-g^2 r^2)/(sqrt(pi) r)] / r,
// e = 0.5 qqr2e qi qj erfc(g r)/r, g = sqrt(alpha).
template <typename T>
inline void coul_pair(T d2, T qiqj, T alpha, T qqr2e, T& f, T& e) {
  const T r = std::sqrt(d2);
  const T g = std::sqrt(alpha);
  const T ec = std::erfc(g * r);
  const T gn = ec / d2 + T(2) * g * std::exp(-alpha * d2) /
                             (T(1.7724538509055160273) * r);
  f = -qqr2e * qiqj * gn / r;
  e = T(0.5) * qqr2e * qiqj * ec / r;
}

// SOG near field (LJCoulSOGInit / CoulCutForceUser, lj_cut_coul_kspace.cpp:
// 651-711, lj_cut_coul_kspace_op.hip.cu:196-252) in place of erfc:
// f = -qqr2e qi qj (1/r^3 + sum_n T_n r^{2n}), e = 0.5 qqr2e qi qj (1/r -
// far(r)), far from the SOG set (sigma, b, mmax, w0).
struct SogNearO {
  bool on = false;
  double T[6] = {0, 0, 0, 0, 0, 0};
  double sigma = 0, b = 0, w0 = 0;
  int mmax = 0;
};

}  // namespace
extern "C" double orc_sog_far_r(double r, double sigma, double b, int mmax, double w0);
namespace {

// CVFF special-bond topology of the caller-order atoms: cumulative counts
// (n12, n13, n14) and partner global ids, weights w12, w13, w14.
struct Special {
  int ms = 0;
  const int* ns = nullptr;
  const int* ids = nullptr;
  double w[3] = {1.0, 1.0, 1.0};
};

// Weight of the pair (i, gid_j): LoadSpecialCounts' clamping, then the first
// matching partner decides (cvff_op.hip.cu:51-110); 1 when not special.
inline double special_weight(const Special& sp, int i, int gid_j) {
  if (sp.ms <= 0 || !sp.ns || !sp.ids) return 1.0;
  int n12 = std::max(0, sp.ns[3 * i]), n13 = std::max(0, sp.ns[3 * i + 1]),
      n14 = std::max(0, sp.ns[3 * i + 2]);
  if (n12 > n13) n13 = n12;
  if (n13 > n14) n14 = n13;
  if (n14 > sp.ms) n14 = sp.ms;
  if (n13 > n14) n13 = n14;
  if (n12 > n13) n12 = n13;
  const int* row = sp.ids + (size_t)i * sp.ms;
  for (int k = 0; k < n14; ++k)
    if (row[k] == gid_j) return k < n12 ? sp.w[0] : (k < n13 ? sp.w[1] : sp.w[2]);
  return 1.0;
}

// Pair forces over float positions: classification in the FP32 mirror,
// arithmetic in T (double for the error reference, float for the mirror).
// With special bonds (CVFF, cvff_op.hip.cu:572-599 and :401-520) a pair of
// weight ws contributes ws f_LJ + f_coul - (1 - ws) f_full, f_full =
// -qqr2e q_i q_j / r^3 the full Coulomb force, where the RBL multiplier
// scales ws f_LJ + f_coul only; energies ws E_LJ + E_coul - (1 - ws) E_full.
template <typename T>
void pair_force_f32pos(int n, const float* x, const float* y, const float* z,
                       const float* q, const int* type, const int* gid,
                       int ntypes, const double* eps, const double* sig,
                       const double* L64, const double* pp, double* fx,
                       double* fy, double* fz, double* energy, double* virial,
                       double* mean, const Special& sp = Special(),
                       const SogNearO& sn = SogNearO()) {
  const PairParams p = unpack(pp);
  const Thresholds t = thresholds(p);
  const float L[3] = {(float)L64[0], (float)L64[1], (float)L64[2]};
  std::vector<T> ax(n, T(0)), ay(n, T(0)), az(n, T(0));
  double evdwl = 0, ecoul = 0, vir[6] = {0, 0, 0, 0, 0, 0};
  const double radius = std::sqrt((double)t.rc2) * 1.0001 + 1e-6;
  for_each_candidate<float>(n, x, y, z, L64, radius, [&](int i, int j) {
    const float d2f = dist2_f32(x[i], y[i], z[i], x[j], y[j], z[j], L);
    if (!(d2f < t.rc2)) return;
    const int cls = p.rbl ? rbl_class(d2f, t, p, gid[i], gid[j]) : 1;
    const double w = pair_weight(d2f, cls, t, p);
    // force: RBL-selected/weighted pairs.  The FP32 mirror reuses the float
    // displacement; the FP64 reference recomputes it in double from the same
    // (float) positions, so no FP32 rounding leaks into the error reference.
    T dx, dy, dz;
    if (std::is_same<T, float>::value) {
      dx = (T)min_image_1<float>(x[j] - x[i], L[0]);
      dy = (T)min_image_1<float>(y[j] - y[i], L[1]);
      dz = (T)min_image_1<float>(z[j] - z[i], L[2]);
    } else {
      dx = (T)min_image_1<double>((double)x[j] - (double)x[i], L64[0]);
      dy = (T)min_image_1<double>((double)y[j] - (double)y[i], L64[1]);
      dz = (T)min_image_1<double>((double)z[j] - (double)z[i], L64[2]);
    }
    const T d2 = std::is_same<T, float>::value ? (T)d2f : dx * dx + dy * dy + dz * dz;
    const int ti = type[i], tj = type[j];
    const T eij = (T)std::sqrt(eps[ti] * eps[tj]);
    const T sij = (T)((sig[ti] + sig[tj]) / 2.0);
    T flj = 0, elj = 0, fc = 0, ec = 0;
    if (d2f < t.rclj2) lj_pair<T>(d2, eij, sij, flj, elj);
    const T qq = (T)q[i] * (T)q[j];
    if (p.coulomb && qq != T(0)) {
      if (sn.on) {
        const T r = std::sqrt(d2);
        T poly = 0, pw = 1;
        for (int k = 0; k < 6; ++k) {
          poly += (T)sn.T[k] * pw;
          pw *= d2;
        }
        fc = -(T)p.qqr2e * qq * (T(1) / (r * d2) + poly);
        ec = T(0.5) * (T)p.qqr2e * qq *
             (T(1) / r - (T)orc_sog_far_r((double)r, sn.sigma, sn.b, sn.mmax, sn.w0));
      } else {
        coul_pair<T>(d2, qq, (T)p.alpha, (T)p.qqr2e, fc, ec);
      }
    }
    const T ws = (T)special_weight(sp, i, gid[j]);
    T ffull = 0, efull = 0;
    if (p.coulomb && qq != T(0) && ws != T(1)) {
      const T r = std::sqrt(d2);
      ffull = -(T)p.qqr2e * qq / (d2 * r);
      efull = T(0.5) * (T)p.qqr2e * qq / r;
    }
    if (w != 0.0) {
      const T f = (T)w * (ws * flj + fc) - (T(1) - ws) * ffull;
      ax[i] += f * dx;
      ay[i] += f * dy;
      az[i] += f * dz;
    }
    // energy / virial: the full (non-sampled) list within cut_off, as the
    // reference's energy path (lj_cut_coul_kspace_op.hip.cu:475-565).
    if (d2f < t.rc2) {
      evdwl += (double)(ws * elj);
      ecoul += (double)(ec - (T(1) - ws) * efull);
      const double ff = (double)(ws * flj + fc - (T(1) - ws) * ffull);
      const double d[3] = {(double)dx, (double)dy, (double)dz};
      vir[0] += -0.5 * d[0] * d[0] * ff;
      vir[1] += -0.5 * d[1] * d[1] * ff;
      vir[2] += -0.5 * d[2] * d[2] * ff;
      vir[3] += -0.5 * d[0] * d[1] * ff;
      vir[4] += -0.5 * d[0] * d[2] * ff;
      vir[5] += -0.5 * d[1] * d[2] * ff;
    }
  });
  double m[3] = {0, 0, 0};
  if (p.mean_fix) {
    for (int i = 0; i < n; ++i) {
      m[0] += (double)ax[i];
      m[1] += (double)ay[i];
      m[2] += (double)az[i];
    }
    for (int d = 0; d < 3; ++d) m[d] /= (double)n;
  }
  for (int i = 0; i < n; ++i) {
    fx[i] = (double)ax[i] - m[0];
    fy[i] = (double)ay[i] - m[1];
    fz[i] = (double)az[i] - m[2];
  }
  if (energy) {
    energy[0] = evdwl;
    energy[1] = ecoul;
  }
  if (virial)
    for (int k = 0; k < 6; ++k) virial[k] = vir[k];
  if (mean)
    for (int d = 0; d < 3; ++d) mean[d] = m[d];
}

}  // namespace

extern "C" {

// ---- pair sets ------------------------------------------------------------
// All ordered pairs i != j with FP32-mirror d2 < R2, as CSR sorted by (i, j).
// offsets has n+1 entries; js may be NULL to only count.  Returns the total.
long orc_pairs_f32(int n, const float* x, const float* y, const float* z,
                   const double* L, float R2, long* offsets, int* js) {
  const float Lf[3] = {(float)L[0], (float)L[1], (float)L[2]};
  std::vector<std::vector<int>> nb(n);
  const double radius = std::sqrt((double)R2) * 1.0001 + 1e-6;
  for_each_candidate<float>(n, x, y, z, L, radius, [&](int i, int j) {
    if (dist2_f32(x[i], y[i], z[i], x[j], y[j], z[j], Lf) < R2)
      nb[i].push_back(j);
  });
  long tot = 0;
  for (int i = 0; i < n; ++i) {
    std::sort(nb[i].begin(), nb[i].end());
    if (offsets) offsets[i] = tot;
    if (js) std::copy(nb[i].begin(), nb[i].end(), js + tot);
    tot += (long)nb[i].size();
  }
  if (offsets) offsets[n] = tot;
  return tot;
}

// RBL classification of every directed pair inside cut_off^2 - eps:
// cls = 1 core, 2 selected shell.  Dropped shell pairs are not listed.
long orc_rbl_pairs_f32(int n, const float* x, const float* y, const float* z,
                       const int* gid, const double* L, const double* pp,
                       long* offsets, int* js, int* cls) {
  const PairParams p = unpack(pp);
  const Thresholds t = thresholds(p);
  const float Lf[3] = {(float)L[0], (float)L[1], (float)L[2]};
  std::vector<std::vector<std::pair<int, int>>> nb(n);
  const double radius = std::sqrt((double)t.rc2) * 1.0001 + 1e-6;
  for_each_candidate<float>(n, x, y, z, L, radius, [&](int i, int j) {
    const float d2 = dist2_f32(x[i], y[i], z[i], x[j], y[j], z[j], Lf);
    const int c = rbl_class(d2, t, p, gid[i], gid[j]);
    if (c) nb[i].push_back({j, c});
  });
  long tot = 0;
  for (int i = 0; i < n; ++i) {
    std::sort(nb[i].begin(), nb[i].end());
    if (offsets) offsets[i] = tot;
    for (size_t k = 0; k < nb[i].size(); ++k) {
      if (js) js[tot + k] = nb[i][k].first;
      if (cls) cls[tot + k] = nb[i][k].second;
    }
    tot += (long)nb[i].size();
  }
  if (offsets) offsets[n] = tot;
  return tot;
}

uint64_t orc_mix64(uint64_t v) { return mix64(v); }
int orc_rbl_selected(int f, long long epoch, int gi, int gj) {
  return rbl_selected(f, epoch, gi, gj) ? 1 : 0;
}

// RblFullNeighborListBuilder::GetRblParams, rbl_full_neighbor_list_builder.cpp:193-228,
// in the reference's float/double mix.
int orc_rbl_frequency(long n_global, const double* L, double r_core, double rc,
                      int sample_num) {
  const float volume = (float)L[0] * (float)L[1] * (float)L[2];
  const float rho = (float)n_global / volume;
  const float coeff = (float)(1.0 + (0.05 / rho - 0.05));
  const int icoeff = (int)std::round((double)coeff);
  const float rsf = (float)r_core, rcf = (float)rc;
  const int rs_num = (int)(icoeff * rho *
                               std::ceil(4.0 / 3.0 * M_PI * std::pow((double)rsf, 3)) +
                           1);
  const int rc_num = (int)(icoeff * rho *
                               std::ceil(4.0 / 3.0 * M_PI * std::pow((double)rcf, 3)) +
                           1);
  const int shell = rc_num - rs_num;
  if (shell <= sample_num) return 1;
  return (int)std::ceil((float)shell / (float)sample_num);
}

// ---- pair forces ------------------------------------------------------------
// prec: 0 = FP64 arithmetic (error reference), 1 = FP32 mirror.
void orc_pair_force(int prec, int n, const float* x, const float* y,
                    const float* z, const float* q, const int* type,
                    const int* gid, int ntypes, const double* eps,
                    const double* sig, const double* L, const double* pp,
                    double* fx, double* fy, double* fz, double* energy,
                    double* virial, double* mean) {
  if (prec == 1)
    pair_force_f32pos<float>(n, x, y, z, q, type, gid, ntypes, eps, sig, L, pp,
                             fx, fy, fz, energy, virial, mean);
  else
    pair_force_f32pos<double>(n, x, y, z, q, type, gid, ntypes, eps, sig, L,
                              pp, fx, fy, fz, energy, virial, mean);
}

// The same with CVFF special bonds: nspecial[3n], special[n * maxspecial]
// (global ids), w[3] = w12, w13, w14.
void orc_pair_force_special(int prec, int n, const float* x, const float* y,
                            const float* z, const float* q, const int* type,
                            const int* gid, int ntypes, const double* eps,
                            const double* sig, const double* L,
                            const double* pp, int maxspecial,
                            const int* nspecial, const int* special,
                            const double* w, double* fx, double* fy,
                            double* fz, double* energy, double* virial,
                            double* mean) {
  Special sp;
  sp.ms = maxspecial;
  sp.ns = nspecial;
  sp.ids = special;
  for (int k = 0; k < 3; ++k) sp.w[k] = w[k];
  if (prec == 1)
    pair_force_f32pos<float>(n, x, y, z, q, type, gid, ntypes, eps, sig, L, pp,
                             fx, fy, fz, energy, virial, mean, sp);
  else
    pair_force_f32pos<double>(n, x, y, z, q, type, gid, ntypes, eps, sig, L,
                              pp, fx, fy, fz, energy, virial, mean, sp);
}

// The pair force with the SOG near field in place of erfc (RBL or full list,
// optional special bonds: maxspecial 0 for none).
void orc_pair_force_sog(int prec, int n, const float* x, const float* y, const float* z,
                        const float* q, const int* type, const int* gid, int ntypes,
                        const double* eps, const double* sig, const double* L,
                        const double* pp, int maxspecial, const int* nspecial,
                        const int* special, const double* w, const double* T,
                        double sigma, double b, int mmax, double w0, double* fx,
                        double* fy, double* fz, double* energy, double* virial,
                        double* mean) {
  Special sp;
  sp.ms = maxspecial;
  sp.ns = nspecial;
  sp.ids = special;
  for (int k = 0; k < 3; ++k) sp.w[k] = w ? w[k] : 1.0;
  SogNearO sn;
  sn.on = true;
  for (int k = 0; k < 6; ++k) sn.T[k] = T[k];
  sn.sigma = sigma;
  sn.b = b;
  sn.mmax = mmax;
  sn.w0 = w0;
  if (prec == 1)
    pair_force_f32pos<float>(n, x, y, z, q, type, gid, ntypes, eps, sig, L, pp, fx, fy,
                             fz, energy, virial, mean, sp, sn);
  else
    pair_force_f32pos<double>(n, x, y, z, q, type, gid, ntypes, eps, sig, L, pp, fx, fy,
                              fz, energy, virial, mean, sp, sn);
}

// All-double full-list LJ (+ optional erfc Coulomb): the LJ/CUT Verlet path
// (lj.cpp:160-205, lj_op.hip.cu:92-176) for FP64 positions.  Energies are the
// half-counted entry sums; virial[6] = xx yy zz xy xz yz of -1/2 d d f.
void orc_verlet_force_f64(int n, const double* x, const double* y,
                          const double* z, const double* q, const int* type,
                          int ntypes, const double* eps, const double* sig,
                          const double* L, double rc, double rc_lj,
                          double alpha, double qqr2e, int coulomb, double* fx,
                          double* fy, double* fz, double* energy,
                          double* virial) {
  const double rc2 = rc * rc, rl2 = rc_lj * rc_lj;
  std::fill(fx, fx + n, 0.0);
  std::fill(fy, fy + n, 0.0);
  std::fill(fz, fz + n, 0.0);
  double ev = 0, ecl = 0, vir[6] = {0, 0, 0, 0, 0, 0};
  for_each_candidate<double>(n, x, y, z, L, rc * 1.0001 + 1e-9, [&](int i, int j) {
    const double dx = min_image_1<double>(x[j] - x[i], L[0]);
    const double dy = min_image_1<double>(y[j] - y[i], L[1]);
    const double dz = min_image_1<double>(z[j] - z[i], L[2]);
    const double d2 = std::fma(dz, dz, std::fma(dy, dy, dx * dx));
    if (!(d2 < rc2)) return;
    double flj = 0, elj = 0, fc = 0, ec = 0;
    const int ti = type[i], tj = type[j];
    if (d2 < rl2)
      lj_pair<double>(d2, std::sqrt(eps[ti] * eps[tj]), (sig[ti] + sig[tj]) / 2.0,
                      flj, elj);
    if (coulomb && q[i] * q[j] != 0.0)
      coul_pair<double>(d2, q[i] * q[j], alpha, qqr2e, fc, ec);
    const double f = flj + fc;
    fx[i] += f * dx;
    fy[i] += f * dy;
    fz[i] += f * dz;
    ev += elj;
    ecl += ec;
    vir[0] += -0.5 * dx * dx * f;
    vir[1] += -0.5 * dy * dy * f;
    vir[2] += -0.5 * dz * dz * f;
    vir[3] += -0.5 * dx * dy * f;
    vir[4] += -0.5 * dx * dz * f;
    vir[5] += -0.5 * dy * dz * f;
  });
  if (energy) {
    energy[0] = ev;
    energy[1] = ecl;
  }
  if (virial)
    for (int k = 0; k < 6; ++k) virial[k] = vir[k];
}

// ---- k-space --------------------------------------------------------------
// Structure factor, force, energy and virial of a weighted k-set (FP64).
// kspace_calculator_op.hip.cu:238-289 (rho), :98-125 + :678-738 (force),
// :573-674 (energy/virial without the self terms).  For RBE the weight is
// w_k = S/P for every sample; for the exact Ewald sum w_k = exp(-k^2/4alpha)
// (kspace_calculator_op.hip.cu:334-403).
//   F_i = sum_k w_k (-4 pi qqr2e q_i / (V k^2)) [cos(k.r_i) Im rho_k
//                                              - sin(k.r_i) Re rho_k] k
//   E   = qqr2e (2 pi / V) sum_k w_k |rho_k|^2 / k^2
//   W_ab= qqr2e (2 pi / V) sum_k w_k |rho_k|^2 / k^2
//                 (delta_ab - 2 k_a k_b (1/(4 alpha) + 1/k^2))
void orc_kspace_f64(int n, const double* x, const double* y, const double* z,
                    const double* q, const double* L, int K, const int* m,
                    const double* w, double qqr2e, double alpha, double* fx,
                    double* fy, double* fz, double* rho, double* energy,
                    double* virial) {
  const double V = L[0] * L[1] * L[2];
  const double tp = 2.0 * M_PI;
  std::vector<double> re(K, 0.0), im(K, 0.0);
  for (int k = 0; k < K; ++k) {
    const double kx = tp * m[3 * k] / L[0], ky = tp * m[3 * k + 1] / L[1],
                 kz = tp * m[3 * k + 2] / L[2];
    double sr = 0, si = 0;
    for (int i = 0; i < n; ++i) {
      const double th = kx * x[i] + ky * y[i] + kz * z[i];
      sr += q[i] * std::cos(th);
      si += q[i] * std::sin(th);
    }
    re[k] = sr;
    im[k] = si;
    if (rho) {
      rho[2 * k] = sr;
      rho[2 * k + 1] = si;
    }
  }
  if (fx) {
    for (int i = 0; i < n; ++i) {
      double ax = 0, ay = 0, az = 0;
      for (int k = 0; k < K; ++k) {
        const double kx = tp * m[3 * k] / L[0], ky = tp * m[3 * k + 1] / L[1],
                     kz = tp * m[3 * k + 2] / L[2];
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double th = kx * x[i] + ky * y[i] + kz * z[i];
        const double c = w[k] * (-4.0 * M_PI * qqr2e * q[i] / (V * k2)) *
                         (std::cos(th) * im[k] - std::sin(th) * re[k]);
        ax += c * kx;
        ay += c * ky;
        az += c * kz;
      }
      fx[i] = ax;
      fy[i] = ay;
      fz[i] = az;
    }
  }
  double e = 0, vir[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < K; ++k) {
    const double kx = tp * m[3 * k] / L[0], ky = tp * m[3 * k + 1] / L[1],
                 kz = tp * m[3 * k + 2] / L[2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double base = qqr2e * 2.0 * M_PI / V * w[k] *
                        (re[k] * re[k] + im[k] * im[k]) / k2;
    const double c = 2.0 * (1.0 / (4.0 * alpha) + 1.0 / k2);
    e += base;
    vir[0] += base * (1.0 - c * kx * kx);
    vir[1] += base * (1.0 - c * ky * ky);
    vir[2] += base * (1.0 - c * kz * kz);
    vir[3] += base * (-c * kx * ky);
    vir[4] += base * (-c * kx * kz);
    vir[5] += base * (-c * ky * kz);
  }
  if (energy) *energy = e;
  if (virial)
    for (int k = 0; k < 6; ++k) virial[k] = vir[k];
}

// Self and neutralising-background energy (kspace_calculator_op.hip.cu:662):
// -qqr2e [ sqrt(alpha/pi) sum q^2 + (pi/2) (sum q)^2 / (alpha V) ].
double orc_kspace_self(double sum_q2, double sum_q, double alpha, double V,
                       double qqr2e) {
  return -qqr2e * (std::sqrt(alpha / M_PI) * sum_q2 +
                   0.5 * M_PI * sum_q * sum_q / (alpha * V));
}

// RBE normalisation S = prod_d H_d - 1, H_d = sqrt(alpha L_d^2/pi)
// sum_{m=-10}^{10} exp(-alpha L_d^2 m^2) (kspace_calculator.cpp:68-78), in the
// reference's float arithmetic (f32 = 1) or in double.
double orc_rbe_S(int f32, double alpha, const double* L) {
  if (f32) {
    const float a = (float)alpha;
    float h[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < 3; ++d) {
      const float Ld = (float)L[d];
      const float factor = -(a * Ld * Ld);
      for (int mm = -10; mm <= 10; ++mm) h[d] += std::exp(mm * mm * factor);
      h[d] *= std::sqrt((float)((double)(-factor) / M_PI));
    }
    return (double)(h[0] * h[1] * h[2] - 1.0f);
  }
  double h[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    const double factor = -(alpha * L[d] * L[d]);
    for (int mm = -10; mm <= 10; ++mm) h[d] += std::exp(mm * mm * factor);
    h[d] *= std::sqrt(-factor / M_PI);
  }
  return h[0] * h[1] * h[2] - 1.0;
}

// The Poisson dual of S: sum over k != 0 of exp(-k^2/(4 alpha)), |m_d| <= kmax.
double orc_gauss_sum_direct(double alpha, const double* L, int kmax) {
  double s = 0;
  for (int a = -kmax; a <= kmax; ++a)
    for (int b = -kmax; b <= kmax; ++b)
      for (int c = -kmax; c <= kmax; ++c) {
        if (!a && !b && !c) continue;
        const double kx = 2 * M_PI * a / L[0], ky = 2 * M_PI * b / L[1],
                     kz = 2 * M_PI * c / L[2];
        s += std::exp(-(kx * kx + ky * ky + kz * kz) / (4 * alpha));
      }
  return s;
}

// Real-space erfc Ewald energy and forces over all min-image pairs within rc
// (FP64), for the Madelung / alpha-independence checks.
void orc_ewald_real_f64(int n, const double* x, const double* y,
                        const double* z, const double* q, const double* L,
                        double alpha, double rc, double qqr2e, double* fx,
                        double* fy, double* fz, double* energy) {
  double e = 0;
  if (fx) {
    std::fill(fx, fx + n, 0.0);
    std::fill(fy, fy + n, 0.0);
    std::fill(fz, fz + n, 0.0);
  }
  for_each_candidate<double>(n, x, y, z, L, rc * 1.0001 + 1e-9, [&](int i, int j) {
    const double dx = min_image_1<double>(x[j] - x[i], L[0]);
    const double dy = min_image_1<double>(y[j] - y[i], L[1]);
    const double dz = min_image_1<double>(z[j] - z[i], L[2]);
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 >= rc * rc) return;
    double f, ec;
    coul_pair<double>(d2, q[i] * q[j], alpha, qqr2e, f, ec);
    e += ec;
    if (fx) {
      fx[i] += f * dx;
      fy[i] += f * dy;
      fz[i] += f * dz;
    }
  });
  *energy = e;
}

// KSpaceCalculator::DeriveAlphaFromAccuracy, kspace_calculator.cpp:120-157.
// Returns alpha = g_ewald^2; q2 = qqr2e * sum q^2.
// DeriveAlphaFromAccuracy (kspace_calculator.cpp:120-157) with the
// reference's types: accuracy, g and q2 are double, the cutoff is float
// (rbmd::Real), SQRT / LOG are sqrtf / logf (rbmd_define.h:199-209) so their
// double arguments are rounded to float, and alpha = g * g is stored as float.
// qqr2e and q2 = qqr2e * sum q^2 are float products (kspace_calculator.h:103,123;
// kspace_calculator.cpp:740); q2 is taken as given and rounded to float.
// Returns alpha as that float value.
double orc_alpha_from_accuracy(double acc_rel, double qqr2e, double n_global,
                               double rc_in, double V, double q2_in) {
  const float rc = (float)rc_in;
  const double q2 = (double)(float)q2_in;
  const double acc = (double)(float)acc_rel * (double)(float)qqr2e;
  double g = acc * (double)sqrtf((float)(n_global * (double)rc * V)) / (2.0 * q2);
  if (g >= 1.0)
    g = (1.35 - 0.15 * (double)logf((float)acc)) / (double)rc;
  else
    g = (double)(sqrtf(-logf((float)g)) / rc);
  return (double)(float)(g * g);
}

// ---- RBE k-sampler ----------------------------------------------------------
// RbmdParallelUntil::GetRbeRadomM (rbmd_parallel_until.cpp:557-724) with the
// seed chain of :305-325 and :484-501: a base std::mt19937 seeded with
// coulomb.random_seed yields one 32-bit seed per force evaluation; that seed
// starts a fresh std::mt19937 that drives a per-dimension Metropolis-Hastings
// chain.  All arithmetic is float (rbmd::Real), with the double promotions the
// reference source implies; RBMD_PI is the float literal 3.14159265...f.
namespace sampler {
const float PI_F = 3.14159265358979323846f;

struct Chain {
  float alpha;
  float L[3];
  std::mt19937& gen;

  float uniform01() {
    std::uniform_real_distribution<float> dis(0.0f, 1.0f);
    return dis(gen);
  }
  float H(int d) const {  // :609-624
    const float factor = -(alpha * L[d] * L[d]);
    float h = 0.0f;
    for (int mm = -10; mm <= 10; ++mm) {
      const float expx = (float)(mm * mm) * factor;
      h += std::exp(expx);
    }
    h *= std::sqrt(-(factor) / PI_F);
    return h;
  }
  float P(float m, int d) const {  // :645-652
    const float t = 2 * PI_F * m / L[d];
    float pm = std::exp(-(t * t) / (4 * alpha));
    return pm / H(d);
  }
  float Q(float m, int d) const {  // :654-670
    const float s = std::sqrt(alpha * (L[d] * L[d]) / (PI_F * PI_F));
    if (m == 0) return std::erf((float)((1.0 / 2) / s));
    const float a = std::erf((float)(((1.0 / 2) + (double)std::fabs(m)) / s));
    const float b = std::erf((float)(((double)std::fabs(m) - (1.0 / 2)) / s));
    return (a - b) / 2;
  }
  float gauss(float mu, float sigma) {  // :672-692
    float u1;
    do {
      u1 = uniform01();
    } while (u1 < 1e-6f);  // epsilon is rbmd::Real
    const float u2 = uniform01();
    const float r = std::sqrt((float)(-2.0 * (double)std::log(u1)));
    const float c = std::cos(2 * PI_F * u2);
    return sigma * r * c + mu;
  }
  float mh(float m, float mu, const float sigma[3], int d) {  // :627-643
    const float xw = gauss(mu, sigma[d]);
    const float mw = std::round(xw);
    float prob = (P(mw, d) / P(m, d)) * (Q(m, d) / Q(mw, d));
    prob = prob < 1.0f ? prob : 1.0f;
    const float u = uniform01();
    if (u <= prob) m = mw;
    return m;
  }
};
}  // namespace sampler

// Draw the k-sample of force evaluation `call_index` (0 = the first call after
// set-up) for base seed `base_seed`.  Writes P integer m-vectors.
void orc_rbe_sample(uint32_t base_seed, long call_index, int P, double alpha,
                    const double* L, int* m_out) {
  std::mt19937 base(base_seed);
  if (call_index > 0) base.discard((unsigned long long)call_index);
  const uint32_t seed = (uint32_t)base();
  std::mt19937 gen(seed);
  sampler::Chain ch{(float)alpha, {(float)L[0], (float)L[1], (float)L[2]}, gen};
  float sigma[3];
  for (int d = 0; d < 3; ++d)  // :710-720
    sigma[d] = std::sqrt(ch.alpha * ch.L[d] * ch.L[d] /
                         (2 * sampler::PI_F * sampler::PI_F));
  const float tiny = 1e-6f;
  float x0[3];
  do {  // :697-704, braced initialiser: x, y, z in order
    x0[0] = std::round(ch.gauss(0.0f, sigma[0]));
    x0[1] = std::round(ch.gauss(0.0f, sigma[1]));
    x0[2] = std::round(ch.gauss(0.0f, sigma[2]));
  } while (std::fabs(x0[0]) < tiny && std::fabs(x0[1]) < tiny &&
           std::fabs(x0[2]) < tiny);
  for (int d = 0; d < 3; ++d) m_out[d] = (int)x0[d];
  for (int i = 1; i < P; ++i) {  // :709-723
    float x1[3];
    x1[0] = ch.mh(x0[0], 0.0f, sigma, 0);
    x1[1] = ch.mh(x0[1], 0.0f, sigma, 1);
    x1[2] = ch.mh(x0[2], 0.0f, sigma, 2);
    for (int d = 0; d < 3; ++d) m_out[3 * i + d] = (int)x1[d];
    for (int d = 0; d < 3; ++d) x0[d] = x1[d];
    if (std::fabs(x1[0]) < tiny && std::fabs(x1[1]) < tiny &&
        std::fabs(x1[2]) < tiny) {
      i = i - 1;  // re-draw this slot; the chain state stays at (0,0,0)
      continue;
    }
  }
}

// Raw std::mt19937 outputs (to pin the generator with the C++ standard's
// check value: the 10000th output of a default-seeded engine is 4123659995).
void orc_mt19937(uint32_t seed, long skip, int count, uint32_t* out) {
  std::mt19937 g(seed);
  if (skip > 0) g.discard((unsigned long long)skip);
  for (int i = 0; i < count; ++i) out[i] = (uint32_t)g();
}

// Time budget helper for the CPU baseline: one full force evaluation of the
// LJ-electrolyte RBL + RBE path (pair FP32 mirror + k-space FP64 over the
// charged atoms), returning nothing but writing forces.
void orc_force_step(int n, const float* x, const float* y, const float* z,
                    const float* q, const int* type, const int* gid,
                    int ntypes, const double* eps, const double* sig,
                    const double* L, const double* pp, int K, const int* m,
                    double rbe_weight, double* fx, double* fy, double* fz) {
  pair_force_f32pos<float>(n, x, y, z, q, type, gid, ntypes, eps, sig, L, pp,
                           fx, fy, fz, nullptr, nullptr, nullptr);
  std::vector<int> ci;
  for (int i = 0; i < n; ++i)
    if (q[i] != 0.0f) ci.push_back(i);
  const int nc = (int)ci.size();
  std::vector<double> cx(nc), cy(nc), cz(nc), cq(nc), gx(nc), gy(nc), gz(nc);
  for (int c = 0; c < nc; ++c) {
    cx[c] = x[ci[c]];
    cy[c] = y[ci[c]];
    cz[c] = z[ci[c]];
    cq[c] = q[ci[c]];
  }
  std::vector<double> w(K, rbe_weight);
  const PairParams p = unpack(pp);
  orc_kspace_f64(nc, cx.data(), cy.data(), cz.data(), cq.data(), L, K, m,
                 w.data(), p.qqr2e, p.alpha, gx.data(), gy.data(), gz.data(),
                 nullptr, nullptr, nullptr);
  for (int c = 0; c < nc; ++c) {
    fx[ci[c]] += gx[c];
    fy[ci[c]] += gy[c];
    fz[ci[c]] += gz[c];
  }
}

// ---- the CPU baseline: rebuild-amortised force steps over stored lists ------
// The per-step force path as a CPU program, OpenMP over atoms, for timing on
// the host cores: every `rebuild_every` steps a full candidate list within
// rc + skin (the reference's single RBL candidate radius,
// rbl_full_neighbor_list_builder.cpp:90-185), CSR per atom, from a cell grid;
// every step the RBL core / selected-shell classification of the step's epoch
// (rbl_full_neighbor_list_op.hip.cu:237-304), the FP32 pair force of
// lj_cut_coul_kspace_op.hip.cu:569-682 with the mean-force fix
// (lj_cut_coul_kspace.cpp:537-563), and the RBE k-space of the charged atoms
// with FP32 phases and FP64 sums (kspace_calculator_op.hip.cu:238-289,
// 678-738).  Each atom's force is summed by one thread in list order and each
// k's structure factor in atom order, so the result does not depend on the
// thread count.  Positions stay fixed (a timing sample); forces of the last
// step are returned.
void orc_cpu_force_steps(int n, const float* x, const float* y, const float* z,
                         const float* q, const int* type, const int* gid,
                         int ntypes, const double* eps, const double* sig,
                         const double* L64, const double* pp, double skin, int K,
                         const int* m, double rbe_weight, int nsteps,
                         int rebuild_every, int nthreads, double* fx, double* fy,
                         double* fz) {
  PairParams p = unpack(pp);
  const Thresholds t = thresholds(p);
  const float L[3] = {(float)L64[0], (float)L64[1], (float)L64[2]};
  const long long epoch0 = p.epoch;
  const double r_list = p.rc + skin;
  const float rl2 = (float)(r_list * r_list);
  std::vector<float> e_ij((size_t)ntypes * ntypes), s_ij((size_t)ntypes * ntypes);
  for (int a = 0; a < ntypes; ++a)
    for (int b = 0; b < ntypes; ++b) {
      e_ij[a * ntypes + b] = (float)std::sqrt(eps[a] * eps[b]);
      s_ij[a * ntypes + b] = (float)((sig[a] + sig[b]) / 2.0);
    }
  // atoms are processed in cell order: order[a] = caller index of sorted atom
  // a; lists hold sorted indices
  std::vector<int> order(n), cell(n), tys(n), gids(n), cstart;
  std::vector<float> xs(n), ys(n), zs(n), qs(n), ax(n), ay(n), az(n);
  std::vector<long> off(n + 1, 0);
  std::vector<int> nbr, cnt(n), ci;
  std::vector<double> re(K), im(K), kxv(K), kyv(K), kzv(K);
  const double V = L64[0] * L64[1] * L64[2];
  int ncell[3] = {1, 1, 1};
  for (int s = 0; s < nsteps; ++s) {
    const bool rebuild = s % std::max(1, rebuild_every) == 0;
    if (rebuild) {
      for (int d = 0; d < 3; ++d) {
        ncell[d] = (int)std::floor(L64[d] / r_list);
        if (ncell[d] < 3) ncell[d] = 1;
      }
      const int ncells = ncell[0] * ncell[1] * ncell[2];
      cstart.assign(ncells + 1, 0);
      for (int i = 0; i < n; ++i) {
        const float pos[3] = {x[i], y[i], z[i]};
        int c[3];
        for (int d = 0; d < 3; ++d) {
          double f = (double)pos[d] / L64[d];
          f -= std::floor(f);
          c[d] = std::min(ncell[d] - 1, std::max(0, (int)(f * ncell[d])));
        }
        cell[i] = (c[2] * ncell[1] + c[1]) * ncell[0] + c[0];
        ++cstart[cell[i] + 1];
      }
      for (int c = 0; c < ncells; ++c) cstart[c + 1] += cstart[c];
      std::vector<int> fill(cstart.begin(), cstart.end() - 1);
      for (int i = 0; i < n; ++i) order[fill[cell[i]]++] = i;
      ci.clear();
      for (int a = 0; a < n; ++a) {
        const int i = order[a];
        tys[a] = type[i];
        gids[a] = gid[i];
        qs[a] = q[i];
        if (q[i] != 0.0f) ci.push_back(a);
      }
    }
    // this step's positions in cell order
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int a = 0; a < n; ++a) {
      xs[a] = x[order[a]];
      ys[a] = y[order[a]];
      zs[a] = z[order[a]];
    }
    if (rebuild) {
      auto scan = [&](int a, auto&& emit) {
        const int id = cell[order[a]];
        const int cx = id % ncell[0], cy = (id / ncell[0]) % ncell[1],
                  cz = id / (ncell[0] * ncell[1]);
        const int ex = ncell[0] > 1, ey = ncell[1] > 1, ez = ncell[2] > 1;
        for (int oz = -ez; oz <= ez; ++oz)
          for (int oy = -ey; oy <= ey; ++oy)
            for (int ox = -ex; ox <= ex; ++ox) {
              const int jc = (((cz + oz + ncell[2]) % ncell[2]) * ncell[1] +
                              (cy + oy + ncell[1]) % ncell[1]) * ncell[0] +
                             (cx + ox + ncell[0]) % ncell[0];
              for (int b = cstart[jc]; b < cstart[jc + 1]; ++b)
                if (b != a && dist2_f32(xs[a], ys[a], zs[a], xs[b], ys[b], zs[b], L) < rl2)
                  emit(b);
            }
      };
      // one scan: each thread appends its atoms' rows to its own buffer,
      // then the rows are copied into CSR order
      const int nchunk = (n + 255) / 256;
      std::vector<std::vector<int>> rows(nchunk);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
      for (int ch = 0; ch < nchunk; ++ch) {
        std::vector<int>& r = rows[ch];
        for (int a = ch * 256; a < std::min(n, ch * 256 + 256); ++a) {
          const size_t before = r.size();
          scan(a, [&](int b) { r.push_back(b); });
          cnt[a] = (int)(r.size() - before);
        }
      }
      for (int a = 0; a < n; ++a) off[a + 1] = off[a] + cnt[a];
      nbr.resize(off[n]);
#pragma omp parallel for schedule(static) num_threads(nthreads)
      for (int ch = 0; ch < nchunk; ++ch)
        std::copy(rows[ch].begin(), rows[ch].end(), nbr.begin() + off[ch * 256]);
    }
    p.epoch = epoch0 + s;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int a = 0; a < n; ++a) {
      float fxi = 0, fyi = 0, fzi = 0;
      const int ti = tys[a];
      for (long k = off[a]; k < off[a + 1]; ++k) {
        const int b = nbr[k];
        const float dx = min_image_1<float>(xs[b] - xs[a], L[0]);
        const float dy = min_image_1<float>(ys[b] - ys[a], L[1]);
        const float dz = min_image_1<float>(zs[b] - zs[a], L[2]);
        const float d2 = std::fmaf(dz, dz, std::fmaf(dy, dy, dx * dx));
        if (!(d2 < t.rc2)) continue;
        const int cls = p.rbl ? rbl_class(d2, t, p, gids[a], gids[b]) : 1;
        const double w = pair_weight(d2, cls, t, p);
        if (w == 0.0) continue;
        float flj = 0, elj = 0, fcl = 0, ecl = 0;
        if (d2 < t.rclj2)
          lj_pair<float>(d2, e_ij[ti * ntypes + tys[b]], s_ij[ti * ntypes + tys[b]], flj,
                         elj);
        const float qq = qs[a] * qs[b];
        if (p.coulomb && qq != 0.0f)
          coul_pair<float>(d2, qq, (float)p.alpha, (float)p.qqr2e, fcl, ecl);
        const float f = (float)w * (flj + fcl);
        fxi += f * dx;
        fyi += f * dy;
        fzi += f * dz;
      }
      ax[a] = fxi;
      ay[a] = fyi;
      az[a] = fzi;
    }
    double mean[3] = {0, 0, 0};
    if (p.mean_fix) {
      for (int a = 0; a < n; ++a) {
        mean[0] += ax[a];
        mean[1] += ay[a];
        mean[2] += az[a];
      }
      for (int d = 0; d < 3; ++d) mean[d] /= (double)n;
    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int a = 0; a < n; ++a) {
      const int i = order[a];
      fx[i] = (double)ax[a] - mean[0];
      fy[i] = (double)ay[a] - mean[1];
      fz[i] = (double)az[a] - mean[2];
    }
    const int nc = (int)ci.size();
    if (K > 0 && p.coulomb && nc > 0) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
      for (int k = 0; k < K; ++k) {
        const float kx = (float)(2.0 * M_PI * m[3 * k] / L64[0]),
                    ky = (float)(2.0 * M_PI * m[3 * k + 1] / L64[1]),
                    kz = (float)(2.0 * M_PI * m[3 * k + 2] / L64[2]);
        double sr = 0, si = 0;
        for (int c = 0; c < nc; ++c) {
          const int a = ci[c];
          const float th = kx * xs[a] + ky * ys[a] + kz * zs[a];
          sr += (double)(qs[a] * std::cos(th));
          si += (double)(qs[a] * std::sin(th));
        }
        re[k] = sr;
        im[k] = si;
        kxv[k] = kx;
        kyv[k] = ky;
        kzv[k] = kz;
      }
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
      for (int c = 0; c < nc; ++c) {
        const int a = ci[c];
        double gx = 0, gy = 0, gz = 0;
        for (int k = 0; k < K; ++k) {
          const float th =
              (float)kxv[k] * xs[a] + (float)kyv[k] * ys[a] + (float)kzv[k] * zs[a];
          const double k2 = kxv[k] * kxv[k] + kyv[k] * kyv[k] + kzv[k] * kzv[k];
          const double g = rbe_weight * (-4.0 * M_PI * p.qqr2e * qs[a] / (V * k2)) *
                           ((double)std::cos(th) * im[k] - (double)std::sin(th) * re[k]);
          gx += g * kxv[k];
          gy += g * kyv[k];
          gz += g * kzv[k];
        }
        const int i = order[a];
        fx[i] += gx;
        fy[i] += gy;
        fz[i] += gz;
      }
    }
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Force-matching adjoint (FP64).  The force of one evaluation (RBL pair forces
// with the mean-force fix, lj_cut_coul_kspace.cpp:494-573; RBE k-space over
// the atoms of nonzero charge, kspace_calculator_op.hip.cu:98-125,678-738)
// as a function of positions and the per-type parameters Q (charge), eps and
// sigma, with q_i = Q[type_i], and the vector-Jacobian product of a cotangent
// G (N x 3): dL/dtheta for L = sum_i G_i . F_i.  Pair classification (which
// pairs exist, RBL classes and selections) uses the FP32 mirror distance of
// the float-rounded positions, as the product does; arithmetic is FP64.
//
// Per directed pair (i, j), d = r_j - r_i, u = |d|^2, weight w (RBL class):
//   f(u)  = w [f_LJ(u) + f_C(u)],   F_i += f d
//   f_LJ  = -24 eps_ij (2 s6 - 1) s6 / u,   s6 = (sigma_ij^2 / u)^3
//   f_C   = -qqr2e q_i q_j h(u),  h(u) = erfc(g r) / (u r)
//           + (2 g / sqrt(pi)) e^{-alpha u} / u,  g = sqrt(alpha), r = sqrt(u).  The pair
// adds c = f (G'_i . d) to L (G' = G - mean(G) under the mean-force fix), so
//   dc/dr_j = f G'_i + 2 f'(u) (G'_i . d) d = -dc/dr_i,
//   dc/d eps_ij = w (-24 (2 s6 - 1) s6 / u) (G'_i . d),
//   dc/d sigma_ij = w (-24 eps_ij (4 s6 - 1) 6 s6 / (sigma_ij u)) (G'_i . d),
//   dc/d(q_i q_j) = -w qqr2e h(u) (G'_i . d),
// and the mixing rules eps_ij = sqrt(eps_a eps_b), sigma_ij = (sigma_a +
// sigma_b) / 2, q_i q_j = Q_a Q_b carry them to the types a = t_i, b = t_j.
// k-space: A_k = w_k (-4 pi qqr2e) / (V k^2), theta_i = k . r_i,
// T_a(k) = sum_{t_i = a} e^{i theta_i}, U_a(k) = sum_{t_i = a} (G_i . k)
// e^{i theta_i}, rho = sum_a Q_a T_a, beta = sum_a Q_a U_a; then
// L_k = sum_k A_k Im(conj(beta) rho), so
//   dL_k/dQ_a = sum_k A_k [Im(conj(U_a) rho) + Im(conj(beta) T_a)],
//   dL_k/dr_m = q_m sum_k A_k [Re(conj(beta) e^{i theta_m})
//                              - (G_m . k) Re(e^{-i theta_m} rho)] k.
// ---------------------------------------------------------------------------
namespace {

struct Adj {
  std::vector<double> fx, fy, fz;        // forward force
  std::vector<double> dx, dy, dz;        // dL/dr
  std::vector<double> dq, de, ds;        // dL/dQ, dL/deps, dL/dsigma per type
};

void force_adjoint(int n, const double* x, const double* y, const double* z,
                   const int* type, const int* gid, int ntypes,
                   const double* Q, const double* eps, const double* sig,
                   const double* L, const double* pp, int K, const int* m,
                   const double* wk, const double* G, bool want_adj, Adj& out) {
  const PairParams p = unpack(pp);
  const Thresholds t = thresholds(p);
  const float Lf[3] = {(float)L[0], (float)L[1], (float)L[2]};
  std::vector<float> xf(n), yf(n), zf(n);
  for (int i = 0; i < n; ++i) {
    xf[i] = (float)x[i];
    yf[i] = (float)y[i];
    zf[i] = (float)z[i];
  }
  out.fx.assign(n, 0.0);
  out.fy.assign(n, 0.0);
  out.fz.assign(n, 0.0);
  out.dx.assign(n, 0.0);
  out.dy.assign(n, 0.0);
  out.dz.assign(n, 0.0);
  out.dq.assign(ntypes, 0.0);
  out.de.assign(ntypes, 0.0);
  out.ds.assign(ntypes, 0.0);
  // mean-force fix: L = sum (G_i - mean G) . F_pair,i
  double gm[3] = {0, 0, 0};
  if (want_adj && p.mean_fix) {
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) gm[d] += G[3 * i + d];
    for (int d = 0; d < 3; ++d) gm[d] /= (double)n;
  }
  const double g = std::sqrt(p.alpha), tsp = 2.0 / std::sqrt(M_PI);
  const double radius = std::sqrt((double)t.rc2) * 1.0001 + 1e-6;
  std::vector<double> px(n, 0.0), py(n, 0.0), pz(n, 0.0);
  for_each_candidate<float>(n, xf.data(), yf.data(), zf.data(), L, radius,
                            [&](int i, int j) {
    const float d2f = dist2_f32(xf[i], yf[i], zf[i], xf[j], yf[j], zf[j], Lf);
    if (!(d2f < t.rc2)) return;
    const int cls = p.rbl ? rbl_class(d2f, t, p, gid[i], gid[j]) : 1;
    const double w = pair_weight(d2f, cls, t, p);
    if (w == 0.0) return;
    const double ddx = min_image_1<double>(x[j] - x[i], L[0]);
    const double ddy = min_image_1<double>(y[j] - y[i], L[1]);
    const double ddz = min_image_1<double>(z[j] - z[i], L[2]);
    const double u = ddx * ddx + ddy * ddy + ddz * ddz;
    const int a = type[i], b = type[j];
    double flj = 0, dflj = 0, dle = 0, dls = 0;
    const double eij = std::sqrt(eps[a] * eps[b]), sij = 0.5 * (sig[a] + sig[b]);
    if (d2f < t.rclj2) {
      const double s2 = sij * sij / u, s6 = s2 * s2 * s2;
      flj = -24.0 * eij * (2.0 * s6 - 1.0) * s6 / u;
      dflj = 24.0 * eij * (14.0 * s6 * s6 - 4.0 * s6) / (u * u);
      dle = -24.0 * (2.0 * s6 - 1.0) * s6 / u;
      dls = sij > 0 ? -24.0 * eij * (4.0 * s6 - 1.0) * 6.0 * s6 / (sij * u) : 0.0;
    }
    double fc = 0, dfc = 0, hC = 0;
    const double qq = Q[a] * Q[b];
    if (p.coulomb) {
      const double r = std::sqrt(u), er = std::erfc(g * r), ex = std::exp(-p.alpha * u);
      hC = er / (u * r) + g * tsp * ex / u;
      const double dh = -1.5 * er / (u * u * r) - 1.5 * g * tsp * ex / (u * u) -
                        g * g * g * tsp * ex / u;
      fc = -p.qqr2e * qq * hC;
      dfc = -p.qqr2e * qq * dh;
    }
    const double f = w * (flj + fc), fp = w * (dflj + dfc);
    px[i] += f * ddx;
    py[i] += f * ddy;
    pz[i] += f * ddz;
    if (!want_adj) return;
    const double Gx = G[3 * i] - gm[0], Gy = G[3 * i + 1] - gm[1], Gz = G[3 * i + 2] - gm[2];
    const double gd = Gx * ddx + Gy * ddy + Gz * ddz;
    const double vx = f * Gx + 2.0 * fp * gd * ddx, vy = f * Gy + 2.0 * fp * gd * ddy,
                 vz = f * Gz + 2.0 * fp * gd * ddz;
    out.dx[j] += vx;
    out.dy[j] += vy;
    out.dz[j] += vz;
    out.dx[i] -= vx;
    out.dy[i] -= vy;
    out.dz[i] -= vz;
    const double ce = w * dle * gd, cs = w * dls * gd;
    if (eps[a] > 0) out.de[a] += ce * eij / (2.0 * eps[a]);
    if (eps[b] > 0) out.de[b] += ce * eij / (2.0 * eps[b]);
    out.ds[a] += 0.5 * cs;
    out.ds[b] += 0.5 * cs;
    if (p.coulomb) {
      const double cq = -w * p.qqr2e * hC * gd;
      out.dq[a] += cq * Q[b];
      out.dq[b] += cq * Q[a];
    }
  });
  double mean[3] = {0, 0, 0};
  if (p.mean_fix) {
    for (int i = 0; i < n; ++i) {
      mean[0] += px[i];
      mean[1] += py[i];
      mean[2] += pz[i];
    }
    for (int d = 0; d < 3; ++d) mean[d] /= (double)n;
  }
  for (int i = 0; i < n; ++i) {
    out.fx[i] = px[i] - mean[0];
    out.fy[i] = py[i] - mean[1];
    out.fz[i] = pz[i] - mean[2];
  }
  if (!p.coulomb || K <= 0) return;
  // k-space
  const double V = L[0] * L[1] * L[2], tp = 2.0 * M_PI;
  std::vector<double> Tr(ntypes), Ti(ntypes), Ur(ntypes), Ui(ntypes);
  for (int k = 0; k < K; ++k) {
    const double kx = tp * m[3 * k] / L[0], ky = tp * m[3 * k + 1] / L[1],
                 kz = tp * m[3 * k + 2] / L[2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double A = wk[k] * (-4.0 * M_PI * p.qqr2e) / (V * k2);
    std::fill(Tr.begin(), Tr.end(), 0.0);
    std::fill(Ti.begin(), Ti.end(), 0.0);
    std::fill(Ur.begin(), Ur.end(), 0.0);
    std::fill(Ui.begin(), Ui.end(), 0.0);
    for (int i = 0; i < n; ++i) {
      const double th = kx * x[i] + ky * y[i] + kz * z[i];
      const double c = std::cos(th), s = std::sin(th);
      Tr[type[i]] += c;
      Ti[type[i]] += s;
      if (want_adj) {
        const double gk = G[3 * i] * kx + G[3 * i + 1] * ky + G[3 * i + 2] * kz;
        Ur[type[i]] += gk * c;
        Ui[type[i]] += gk * s;
      }
    }
    double rr = 0, ri = 0, br = 0, bi = 0;
    for (int a = 0; a < ntypes; ++a) {
      rr += Q[a] * Tr[a];
      ri += Q[a] * Ti[a];
      br += Q[a] * Ur[a];
      bi += Q[a] * Ui[a];
    }
    for (int i = 0; i < n; ++i) {
      const double qi = Q[type[i]];
      if (qi == 0.0) continue;
      const double th = kx * x[i] + ky * y[i] + kz * z[i];
      const double c = std::cos(th), s = std::sin(th);
      const double fk = A * qi * (c * ri - s * rr);
      out.fx[i] += fk * kx;
      out.fy[i] += fk * ky;
      out.fz[i] += fk * kz;
      if (want_adj) {
        const double gk = G[3 * i] * kx + G[3 * i + 1] * ky + G[3 * i + 2] * kz;
        const double v = A * qi * ((br * c + bi * s) - gk * (c * rr + s * ri));
        out.dx[i] += v * kx;
        out.dy[i] += v * ky;
        out.dz[i] += v * kz;
      }
    }
    if (want_adj)
      for (int a = 0; a < ntypes; ++a)
        // Im(conj(U) rho) + Im(conj(beta) T)
        out.dq[a] += A * ((Ur[a] * ri - Ui[a] * rr) + (br * Ti[a] - bi * Tr[a]));
  }
}

}  // namespace

extern "C" {

// FP64 force of one evaluation for double positions and per-type parameters.
void orc_force_f64(int n, const double* x, const double* y, const double* z,
                   const int* type, const int* gid, int ntypes, const double* Q,
                   const double* eps, const double* sig, const double* L,
                   const double* pp, int K, const int* m, const double* wk,
                   double* fx, double* fy, double* fz) {
  Adj a;
  force_adjoint(n, x, y, z, type, gid, ntypes, Q, eps, sig, L, pp, K, m, wk,
                nullptr, false, a);
  std::copy(a.fx.begin(), a.fx.end(), fx);
  std::copy(a.fy.begin(), a.fy.end(), fy);
  std::copy(a.fz.begin(), a.fz.end(), fz);
}

// Its vector-Jacobian product for the cotangent G (N x 3, row-major).
void orc_force_vjp_f64(int n, const double* x, const double* y, const double* z,
                       const int* type, const int* gid, int ntypes,
                       const double* Q, const double* eps, const double* sig,
                       const double* L, const double* pp, int K, const int* m,
                       const double* wk, const double* G, double* dr,
                       double* dq, double* de, double* ds) {
  Adj a;
  force_adjoint(n, x, y, z, type, gid, ntypes, Q, eps, sig, L, pp, K, m, wk, G,
                true, a);
  for (int i = 0; i < n; ++i) {
    dr[3 * i] = a.dx[i];
    dr[3 * i + 1] = a.dy[i];
    dr[3 * i + 2] = a.dz[i];
  }
  for (int t = 0; t < ntypes; ++t) {
    dq[t] = a.dq[t];
    de[t] = a.de[t];
    ds[t] = a.ds[t];
  }
}

}  // extern "C"
