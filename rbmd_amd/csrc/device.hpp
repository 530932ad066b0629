// Device helpers shared by the force-path kernels (kernels.hip) and the
// replica-batch kernels (batch.hip): vector types, rounding-exact
// arithmetic, minimum image, RBL hashes and selection, cell coordinates,
// wave/block reductions and scans, the pair weight and accumulator, the
// erfc Coulomb scalar, and phases in revolutions.  Header-only; every TU gets
// its own copies (internal linkage).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine.hpp"

namespace rbmdx {
namespace {

template <typename R>
using R4T = typename Dev<R>::R4;

// LDS budget of one brick stage (24 KB of positions).
template <typename R>
struct StageCap {
  static constexpr int value = (24 * 1024) / (int)sizeof(R4T<R>);
};

// List A layout: entries interleaved in groups of 8 so that one 16-byte load
// fetches 8 consecutive entries of an atom: entry k of sorted atom i lives
// at ((k / 8) * n + i) * 8 + k % 8.
__device__ __forceinline__ size_t list_a_index(int k, int i, int n) {
  return ((size_t)(k >> 3) * n + i) * 8 + (k & 7);
}

__device__ __forceinline__ float mul_rn(float a, float b) {
  return __fmul_rn(a, b);
}
__device__ __forceinline__ double mul_rn(double a, double b) {
  return __dmul_rn(a, b);
}
__device__ __forceinline__ float fma_r(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fma_r(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

template <typename R>
__device__ __forceinline__ R4T<R> make4(R x, R y, R z, R w) {
  R4T<R> v;
  v.x = x;
  v.y = y;
  v.z = z;
  v.w = w;
  return v;
}

// Single-shift minimum image (valid while every interaction radius is below
// half the box edge, which set-up enforces).
template <typename R>
__device__ __forceinline__ R min_image(R d, R L, R half) {
  // one shift by L towards zero when |d| > L/2 (d - sign(d) L), as a select
  return d - (fabs(d) > half ? copysign(L, d) : R(0));
}

// The squared pair distance every classification uses:
// d2 = fma(dz, dz, fma(dy, dy, dx*dx)) with d = min_image(r_j - r_i).
// MI = false when |r_j - r_i| < L/2 is known (the minimum image is then the
// identity and the result is the same bit for bit).
template <typename R, bool MI = true>
__device__ __forceinline__ R pair_d2(const Box<R>& b, R xi, R yi, R zi, R xj,
                                     R yj, R zj, R& dx, R& dy, R& dz) {
  dx = xj - xi;
  dy = yj - yi;
  dz = zj - zi;
  if (MI) {
    dx = min_image<R>(dx, b.L[0], b.half[0]);
    dy = min_image<R>(dy, b.L[1], b.half[1]);
    dz = min_image<R>(dz, b.L[2], b.half[2]);
  }
  return fma_r(dz, dz, fma_r(dy, dy, mul_rn(dx, dx)));
}

// pair_d2 for a list entry: the image shift as k = rint(d / L) in {-1, 0, 1}
// and d - k L as one fma.  Every pair that can pass a cutoff (|d| <= cut +
// skin < L/2, which the configuration enforces) gets the same d and d^2 bit
// for bit as pair_d2 (no shift, or one subtraction of L); the two differ only
// for |d| within rounding of L/2, far outside every cutoff under either.
template <typename R>
__device__ __forceinline__ R pair_d2_list(const Box<R>& b, R xi, R yi, R zi, R xj,
                                          R yj, R zj, R& dx, R& dy, R& dz) {
  dx = xj - xi;
  dy = yj - yi;
  dz = zj - zi;
  dx = fma_r(-rint(dx * b.invL[0]), b.L[0], dx);
  dy = fma_r(-rint(dy * b.invL[1]), b.L[1], dy);
  dz = fma_r(-rint(dz * b.invL[2]), b.L[2], dz);
  return fma_r(dz, dz, fma_r(dy, dy, mul_rn(dx, dx)));
}

// Fold a coordinate into [lo, hi).  A no-op for coordinates already inside.
template <typename R>
__device__ __forceinline__ R wrap(R x, R lo, R hi, R L, R invL) {
  if (x < lo) {
    x += L;
  } else if (x >= hi) {
    x -= L;
  }
  if (x < lo || x >= hi) {
    x = x - L * floor((x - lo) * invL);
    if (x >= hi) x -= L;
    if (x < lo) x = lo;
  }
  return x;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long v) {
  v += 0x9e3779b97f4a7c15ULL;
  v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
  v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
  return v ^ (v >> 31);
}

// Directed RBL shell selection for sampling epoch `epoch`: with
// hi = mix64(gid_i) and hj = mix64(gid_j + 0xd1b54a32d192ed03), the pair is
// kept iff mix64(epoch ^ hi ^ hj) mod f == 0.  hj is precomputed per atom at
// rebuild.  The remainder is exact: a mask for powers of two, else
// h mod f = ((h_hi mod f)(2^32 mod f) + h_lo mod f) mod f in 32-bit
// arithmetic (f < 2^16), else the 64-bit remainder.
__device__ __forceinline__ unsigned long long rbl_hi(int gi) {
  return mix64((unsigned long long)(long long)gi);
}
__device__ __forceinline__ unsigned long long rbl_hj(int gj) {
  return mix64((unsigned long long)(long long)gj + 0xd1b54a32d192ed03ULL);
}
__device__ __forceinline__ bool rbl_keep(int f, unsigned fmask, unsigned c32,
                                         long long epoch,
                                         unsigned long long hi,
                                         unsigned long long hj) {
  if (f <= 1) return true;
  const unsigned long long h = mix64((unsigned long long)epoch ^ hi ^ hj);
  if (fmask) return ((unsigned)h & fmask) == 0u;
  const unsigned uf = (unsigned)f;
  if (uf < 65536u) {
    const unsigned r1 = (unsigned)(h >> 32) % uf;
    unsigned r = (r1 * c32) % uf + (unsigned)h % uf;
    if (r >= uf) r -= uf;
    return r == 0u;
  }
  return h % (unsigned long long)f == 0ULL;
}
// rbl_keep with the hash formed ahead of every branch (f = 1 is the power of
// two 2^0, whose mask 0 keeps every pair): the same decision, and no load of
// hj sinks into a branch where it would wait for every load before it.
__device__ __forceinline__ bool rbl_keep_flat(int f, unsigned fmask, unsigned c32,
                                              long long epoch,
                                              unsigned long long hi,
                                              unsigned long long hj) {
  const unsigned long long h = mix64((unsigned long long)epoch ^ hi ^ hj);
  if (fmask != 0u || f <= 1) return ((unsigned)h & fmask) == 0u;
  const unsigned uf = (unsigned)f;
  if (uf < 65536u) {
    const unsigned r1 = (unsigned)(h >> 32) % uf;
    unsigned r = (r1 * c32) % uf + (unsigned)h % uf;
    if (r >= uf) r -= uf;
    return r == 0u;
  }
  return h % (unsigned long long)f == 0ULL;
}
template <typename PC>
__device__ __forceinline__ bool rbl_selected(const PC& pc, int gi, int gj) {
  return rbl_keep(pc.ifreq, pc.fmask, pc.c32, pc.epoch, rbl_hi(gi), rbl_hj(gj));
}

__device__ __forceinline__ bool rebuild_now(const Ctrl* c, unsigned thr) {
  return c->maxdisp_bits > thr;
}

__device__ __forceinline__ int cell_coord(float u, float scale, int n) {
  int c = (int)(u * scale);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

__device__ __forceinline__ int wrapi(int c, int n) {
  c %= n;
  return c < 0 ? c + n : c;
}

// Stencil along one dimension: `ext` cells either side, or the whole
// dimension once when the stencil would wrap onto itself.
struct Span {
  int start, count, n;
  __device__ __forceinline__ int at(int k) const { return wrapi(start + k, n); }
};
__device__ __forceinline__ Span make_span(int c, int ext, int n) {
  Span s;
  s.n = n;
  if (2 * ext + 1 <= n) {
    s.start = c - ext;
    s.count = 2 * ext + 1;
  } else {
    s.start = 0;
    s.count = n;
  }
  return s;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
// Inclusive prefix sum across the wave.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// Block sum of NV doubles (blockDim.x multiple of 64, <= 1024; smem holds
// 16 * NV).  Result valid in thread 0.  Fixed reduction tree: deterministic.
// The cross-wave sums of the NV values run in NV threads side by side (each
// in wave order), not as one thread's chain of NV * waves LDS reads.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* smem) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) smem[wid * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += smem[w * NV + k];
    // column k of wave 0 is read by this thread alone
    smem[k] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = smem[k];
  }
  __syncthreads();
}

// Exclusive prefix over the block (blockDim <= 1024, multiple of 64); the
// block total is returned in `total`.
__device__ __forceinline__ int block_excl_scan(int v, int& total, int* smem) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) smem[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    const int wv = lane < nw ? smem[lane] : 0;
    const int wi = wave_incl_scan(wv);
    __builtin_amdgcn_wave_barrier();
    if (lane < nw) smem[lane] = wi - wv;  // exclusive wave offsets
    if (lane == nw - 1) smem[32] = wi;
  }
  __syncthreads();
  const int r = smem[wid] + inc - v;
  total = smem[32];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// Pair interaction of one directed pair.  Weight applied to the force (RBL:
// 1 core, f selected shell, 0 otherwise; Verlet: 1 inside rc).
// ---------------------------------------------------------------------------
template <typename R, bool RBL, typename HashJ>
__device__ __forceinline__ R pair_weight(const PairConst<R>& pc, R d2,
                                         unsigned long long hi, HashJ hash_j) {
  if (!RBL) return R(1);
  if (d2 < pc.t_core) return (d2 < pc.rs2 && d2 > pc.eps_min) ? R(1) : R(0);
  if (d2 < pc.t_shell && d2 > pc.rs2 &&
      rbl_keep(pc.ifreq, pc.fmask, pc.c32, pc.epoch, hi, hash_j()))
    return pc.freq;
  return R(0);
}

// pair_weight as selects: the selection hash is formed for every candidate,
// which on a wave costs less than the divergent branches around it.
template <typename R, bool RBL>
__device__ __forceinline__ R pair_weight_flat(const PairConst<R>& pc, R d2,
                                              unsigned long long hi,
                                              unsigned long long hj) {
  if (!RBL) return R(1);
  const bool sel = rbl_keep_flat(pc.ifreq, pc.fmask, pc.c32, pc.epoch, hi, hj);
  const R wc = ((d2 < pc.rs2) & (d2 > pc.eps_min)) ? R(1) : R(0);
  const R wsh = ((d2 < pc.t_shell) & (d2 > pc.rs2) & sel) ? pc.freq : R(0);
  return d2 < pc.t_core ? wc : wsh;
}

template <typename R>
struct Acc {
  R fx = 0, fy = 0, fz = 0;
  double e_lj = 0, e_c = 0, v[6] = {0, 0, 0, 0, 0, 0};
};

template <typename R, bool EFLAG>
__device__ __forceinline__ void add_virial(Acc<R>& a, R dx, R dy, R dz, R f) {
  if (EFLAG) {
    const double h = -0.5 * (double)f;
    a.v[0] += h * (double)dx * (double)dx;
    a.v[1] += h * (double)dy * (double)dy;
    a.v[2] += h * (double)dz * (double)dz;
    a.v[3] += h * (double)dx * (double)dy;
    a.v[4] += h * (double)dx * (double)dz;
    a.v[5] += h * (double)dy * (double)dz;
  }
}

// Reciprocal and reciprocal square root of d^2 in the pair terms: the
// hardware instructions (1 ulp) in FP32, IEEE division and square root in FP64.
__device__ __forceinline__ float pair_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double pair_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float pair_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double pair_rsqrt(double x) { return 1.0 / sqrt(x); }

// erfc-screened Coulomb scalar and half entry energy, from d^2, 1/r and
// 1/r^2.
template <typename R>
__device__ __forceinline__ void coul(const PairConst<R>& pc, R d2, R rinv, R inv, R qq,
                                     R& f, R& e) {
  const R ec = erfc(pc.g * (d2 * rinv));
  const R gn = ec * inv + pc.two_g_over_sqrtpi * exp(-pc.alpha * d2) * rinv;
  f = -pc.qqr2e * qq * gn * rinv;
  e = R(0.5) * pc.qqr2e * qq * ec * rinv;
}

// CVFF special-bond row of one atom: cumulative counts n12 <= n13 <= n14 of
// its 1-2, 1-3 and 1-4 partners (clamped on the host as LoadSpecialCounts
// clamps them, cvff_op.hip.cu:51-77) and their global ids.
struct SpecRow {
  int n12 = 0, n13 = 0, n14 = 0;
  const int* ids = nullptr;
};

// Weight of the pair (i, gid) for i's row: the first match decides, 1 when
// gid is no special partner (LookupSpecialWeight, cvff_op.hip.cu:79-110).
template <typename R>
__device__ __forceinline__ R special_w(const SpecRow& r, int gid,
                                       const PairConst<R>& pc) {
  // the weights are selected as values: a select of their addresses would
  // keep a kernel's private copy of pc in scratch memory
  const R w12 = pc.w12, w13 = pc.w13, w14 = pc.w14;
  int hit = -1;
  for (int k = 0; k < r.n14; ++k)
    if (hit < 0 && r.ids[k] == gid) hit = k;
  R w = R(1);
  if (hit >= 0) w = w14;
  if (hit >= 0 && hit < r.n13) w = w13;
  if (hit >= 0 && hit < r.n12) w = w12;
  return w;
}

// A row of up to kSpecReg partners held in registers (one atom per wave:
// scalar registers), so that the weight of a pair is a few compares and
// selects with no load.  Entries past n14 repeat the last partner, which
// leaves the first match unchanged.
constexpr int kSpecReg = 8;
struct SpecRowReg {
  int n12 = 0, n13 = 0, n14 = 0;
  int id[kSpecReg] = {};
};
__device__ __forceinline__ SpecRowReg spec_row_reg(const SpecRow& r) {
  SpecRowReg g;
  g.n12 = r.n12;
  g.n13 = r.n13;
  g.n14 = r.n14;
#pragma unroll
  for (int k = 0; k < kSpecReg; ++k) g.id[k] = r.n14 > 0 ? r.ids[min(k, r.n14 - 1)] : 0;
  return g;
}
// special_w for a row of at most kSpecReg partners
template <typename R>
__device__ __forceinline__ R special_w_reg(const SpecRowReg& r, int gid,
                                           const PairConst<R>& pc) {
  int hit = kSpecReg;
#pragma unroll
  for (int k = kSpecReg - 1; k >= 0; --k) hit = r.id[k] == gid ? k : hit;
  const R w12 = pc.w12, w13 = pc.w13, w14 = pc.w14;
  return hit < r.n12 ? w12 : (hit < r.n13 ? w13 : (hit < r.n14 ? w14 : R(1)));
}

// SOG near-field Coulomb (VERLET-SOG, lj_cut_coul_kspace_op.hip.cu:196-252):
// f = -qqr2e qq (1/r^3 + sum_n T_n r^{2n}) and, when energies are wanted,
// e = 1/2 qqr2e qq (1/r - sum_l omega_l exp(-r^2 / (2 s_l^2))).
template <typename R, bool EFLAG>
__device__ __forceinline__ void coul_sog(const PairConst<R>& pc, R d2, R rinv, R inv,
                                         R qq, R& f, R& e) {
  const R r3inv = rinv * inv;
  R poly = pc.T[5];
#pragma unroll
  for (int n = 4; n >= 0; --n) poly = fma_r(poly, d2, pc.T[n]);
  f = -pc.qqr2e * qq * (r3inv + poly);
  e = R(0);
  if (EFLAG) {
    R sume = R(0);
    for (int l = 0; l < pc.far_m; ++l)
      sume += pc.far_tab[2 * l] * exp(-d2 * pc.far_tab[2 * l + 1]);
    e = R(0.5) * pc.qqr2e * qq * (rinv - sume);
  }
}

// The terms of a charged pair inside r_c once its RBL weight w and (with
// SPECIAL) its special-bond weight ws are known: qq = q_i q_j,
// (dx, dy, dz) = r_j - r_i, d2 its square norm.
template <typename R, bool EFLAG, bool SPECIAL>
__device__ __forceinline__ void eval_b(const PairConst<R>& pc, const R* s_lj,
                                       int ti, int tj, R qq, R dx, R dy, R dz,
                                       R d2, R w, Acc<R>& a, R ws) {
  const R rinv = pair_rsqrt(d2);
  const R inv = pair_rcp(d2);
  R flj = 0, elj = 0;
  if (d2 < pc.rclj2) {
    const int tt = 4 * (ti * pc.ntypes + tj);
    const R m24 = s_lj[tt], sig2 = s_lj[tt + 1], eps = s_lj[tt + 2];
    const R s2 = sig2 * inv;
    const R s6 = s2 * s2 * s2;
    flj = m24 * (R(2) * s6 - R(1)) * s6 * inv;
    elj = R(2) * eps * (s6 - R(1)) * s6;
  }
  R fc, ec;
  if (pc.sog)
    coul_sog<R, EFLAG>(pc, d2, rinv, inv, qq, fc, ec);
  else
    coul<R>(pc, d2, rinv, inv, qq, fc, ec);
  if (!SPECIAL) {
    const R f = w * (flj + fc);
    a.fx += f * dx;
    a.fy += f * dy;
    a.fz += f * dz;
    if (EFLAG) {
      a.e_lj += (double)elj;
      a.e_c += (double)ec;
      add_virial<R, EFLAG>(a, dx, dy, dz, flj + fc);
    }
    return;
  }
  R ffull = R(0), efull = R(0);
  if (ws != R(1)) {
    ffull = -pc.qqr2e * qq * rinv * rinv * rinv;
    efull = R(0.5) * pc.qqr2e * qq * rinv;
  }
  const R corr = (R(1) - ws) * ffull;
  const R f = w * (ws * flj + fc) - (w != R(0) ? corr : R(0));
  a.fx += f * dx;
  a.fy += f * dy;
  a.fz += f * dz;
  if (EFLAG) {
    a.e_lj += (double)(ws * elj);
    a.e_c += (double)(ec - (R(1) - ws) * efull);
    add_virial<R, EFLAG>(a, dx, dy, dz, ws * flj + fc - corr);
  }
}

// ---------------------------------------------------------------------------
// k-space.  Phases are evaluated in revolutions from fractional coordinates,
// t = m . s with s = (r - lo)/L, reduced to [-1/2, 1/2]; e^{ik.r} then differs
// from the reference's e^{ik.r} by the constant phase e^{ik.lo}, which cancels
// in every force, energy and virial term.
// ---------------------------------------------------------------------------
// sin and cos of 2 pi t for t already reduced to [-1/2, 1/2]: the hardware
// transcendental unit takes its argument in revolutions.
__device__ __forceinline__ void sincos_rev(float t, float& s, float& c) {
  t -= rintf(t);
  s = __builtin_amdgcn_sinf(t);
  c = __builtin_amdgcn_cosf(t);
}

template <typename R>
__device__ __forceinline__ void frac_coords(const Box<R>& b, const R4T<R>& p,
                                            float& sx, float& sy, float& sz) {
  sx = (float)((p.x - b.lo[0]) * b.invL[0]);
  sy = (float)((p.y - b.lo[1]) * b.invL[1]);
  sz = (float)((p.z - b.lo[2]) * b.invL[2]);
}

}  // namespace
}  // namespace rbmdx
