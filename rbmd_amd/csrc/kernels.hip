// HIP kernels of the RBMD force path for gfx950 (CDNA4, wave64).
//
// Per force evaluation (one stream, no host synchronisation):
//   k_skin       max displacement since the last rebuild -> rebuild decision
//                (also stages the host k-sample into device memory)
//   k_bin .. k_list_b   (early exit unless the decision says rebuild)
//                cell keys, counting sort with a deterministic in-cell order,
//                all-atom and charged-pair candidate lists (cut + skin)
//   k_gather     caller-order positions -> cell-sorted float4 tiles
//   k_pair_a     LJ over the all-atom list, RBL classification on the fly
//   k_pair_b     LJ + erfc Coulomb over the charged-pair list
//   k_sk         S(k) partial sums over the charged atoms (k tiles x atoms)
//   k_finalize   one block: FP64 reductions of S(k), energy, virial, mean
//                pair force; per-k force coefficients
//   k_rbe_force  per charged atom, k split over the waves of a block
//   k_assemble   F = F_pair - <F_pair> + F_kspace, scattered to caller order
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace rbmdx {
namespace {

template <typename R>
using R4T = typename Dev<R>::R4;

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
  if (fabs(d) > half) d = d > R(0) ? d - L : d + L;
  return d;
}

// The squared pair distance every classification uses:
// d2 = fma(dz, dz, fma(dy, dy, dx*dx)) with d = min_image(r_j - r_i).
template <typename R>
__device__ __forceinline__ R pair_d2(const Box<R>& b, R xi, R yi, R zi, R xj,
                                     R yj, R zj, R& dx, R& dy, R& dz) {
  dx = min_image<R>(xj - xi, b.L[0], b.half[0]);
  dy = min_image<R>(yj - yi, b.L[1], b.half[1]);
  dz = min_image<R>(zj - zi, b.L[2], b.half[2]);
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

// Directed RBL shell selection for sampling epoch `epoch`.
__device__ __forceinline__ bool rbl_selected(int f, long long epoch, int gi,
                                             int gj) {
  if (f <= 1) return true;
  unsigned long long key = (unsigned long long)epoch;
  key ^= mix64((unsigned long long)(long long)gi);
  key ^= mix64((unsigned long long)(long long)gj + 0xd1b54a32d192ed03ULL);
  return mix64(key) % (unsigned long long)f == 0ULL;
}

__device__ __forceinline__ bool rebuild_now(const Ctrl* c, unsigned thr) {
  return c->maxdisp_bits > thr;
}

__device__ __forceinline__ int cell_coord(float u, float scale, int n) {
  int c = (int)(u * scale);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// Stencil along one dimension: `ext` cells either side, or the whole
// dimension once when the stencil would wrap onto itself.
struct Span {
  int start, count, n;
  __device__ __forceinline__ int at(int k) const {
    int c = start + k;
    c %= n;
    return c < 0 ? c + n : c;
  }
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

// Block sum of NV doubles (blockDim.x multiple of 64, <= 1024).  Result valid
// in thread 0.  Fixed reduction tree: deterministic.
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
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += smem[w * NV + k];
      v[k] = s;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Rebuild decision.  Max |r - r_ref|^2 (minimum image), as float bits in
// ctrl->maxdisp_bits; +inf when a rebuild is forced.  Block 0 also copies the
// k-sample from host-mapped staging memory.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_skin(Dev<R> d, Box<R> b, int force,
                                              const int* m_host, int nk3) {
  if (m_host && blockIdx.x == 0)
    for (int t = threadIdx.x; t < nk3; t += blockDim.x) d.m_dev[t] = m_host[t];
  if (force) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      atomicMax(&d.ctrl->maxdisp_bits, 0x7f800000u);
    return;
  }
  float mx = 0.f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
       i += gridDim.x * blockDim.x) {
    const R x = wrap<R>(d.px[i], b.lo[0], b.hi[0], b.L[0], b.invL[0]);
    const R y = wrap<R>(d.py[i], b.lo[1], b.hi[1], b.L[1], b.invL[1]);
    const R z = wrap<R>(d.pz[i], b.lo[2], b.hi[2], b.L[2], b.invL[2]);
    const R4T<R> r = d.ref[i];
    const R dx = min_image<R>(x - r.x, b.L[0], b.half[0]);
    const R dy = min_image<R>(y - r.y, b.L[1], b.half[1]);
    const R dz = min_image<R>(z - r.z, b.L[2], b.half[2]);
    const float d2 = (float)(dx * dx + dy * dy + dz * dz);
    mx = fmaxf(mx, d2 != d2 ? __int_as_float(0x7f800000) : d2);
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0 && mx > 0.f)
    atomicMax(&d.ctrl->maxdisp_bits, __float_as_uint(mx));
}

// Cell key per atom (caller order): key = 2 * cell + neutral, so that inside
// every cell the charged atoms come first.  Arrival slots from atomics; the
// in-cell order is made deterministic by k_segsort.
template <typename R>
__global__ __launch_bounds__(256) void k_bin(Dev<R> d, Box<R> b, Grid g,
                                             int split, unsigned thr) {
  if (!rebuild_now(d.ctrl, thr)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
       i += gridDim.x * blockDim.x) {
    const R x = wrap<R>(d.px[i], b.lo[0], b.hi[0], b.L[0], b.invL[0]);
    const R y = wrap<R>(d.py[i], b.lo[1], b.hi[1], b.L[1], b.invL[1]);
    const R z = wrap<R>(d.pz[i], b.lo[2], b.hi[2], b.L[2], b.invL[2]);
    const int cx = cell_coord((float)(x - b.lo[0]), g.scale[0], g.nc[0]);
    const int cy = cell_coord((float)(y - b.lo[1]), g.scale[1], g.nc[1]);
    const int cz = cell_coord((float)(z - b.lo[2]), g.scale[2], g.nc[2]);
    const int cell = (cz * g.nc[1] + cy) * g.nc[0] + cx;
    const int neutral = (!split || d.charge[i] == R(0)) ? 1 : 0;
    const int key = 2 * cell + neutral;
    d.key[i] = key;
    d.slot[i] = atomicAdd(&d.key_count[key], 1);
  }
}

// Exclusive scans (one block of 1024): key_start over all keys, cstart over
// the charged counts of each cell.  Clears key_count for the next rebuild.
__global__ __launch_bounds__(1024) void k_scan(int* key_count, int* key_start,
                                               int* cstart, int ncell,
                                               Ctrl* ctrl, unsigned thr) {
  if (!rebuild_now(ctrl, thr)) return;
  if (threadIdx.x == 0) {
    ctrl->rebuild_count += 1;
    ctrl->sum_count_a = 0ull;
    ctrl->sum_count_b = 0ull;
  }
  __shared__ int s[1024];
  const int nkeys = 2 * ncell;
  const int t = threadIdx.x;
  for (int pass = 0; pass < 2; ++pass) {
    const int n = pass == 0 ? nkeys : ncell;
    const int per = (n + 1023) / 1024;
    const int a = t * per, e = min(n, a + per);
    int sum = 0;
    for (int k = a; k < e; ++k) sum += pass == 0 ? key_count[k] : key_count[2 * k];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int v = t >= off ? s[t - off] : 0;
      __syncthreads();
      s[t] += v;
      __syncthreads();
    }
    int run = s[t] - sum;  // exclusive prefix of this thread's segment
    for (int k = a; k < e; ++k) {
      const int c = pass == 0 ? key_count[k] : key_count[2 * k];
      if (pass == 0)
        key_start[k] = run;
      else
        cstart[k] = run;
      run += c;
    }
    if (t == 1023) {
      if (pass == 0)
        key_start[nkeys] = s[1023];
      else
        cstart[ncell] = s[1023];
    }
    __syncthreads();
  }
  for (int k = t; k < nkeys; k += 1024) key_count[k] = 0;
}

template <typename R>
__global__ __launch_bounds__(256) void k_scatter(Dev<R> d, unsigned thr) {
  if (!rebuild_now(d.ctrl, thr)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
       i += gridDim.x * blockDim.x)
    d.perm[d.key_start[d.key[i]] + d.slot[i]] = i;
}

// Sort every key segment of perm by caller index: a stable counting sort.
__global__ __launch_bounds__(256) void k_segsort(int* perm,
                                                 const int* key_start,
                                                 int nkeys, const Ctrl* ctrl,
                                                 unsigned thr) {
  if (!rebuild_now(ctrl, thr)) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nkeys;
       k += gridDim.x * blockDim.x) {
    const int a = key_start[k], e = key_start[k + 1];
    for (int p = a + 1; p < e; ++p) {
      const int v = perm[p];
      int q = p - 1;
      while (q >= a && perm[q] > v) {
        perm[q + 1] = perm[q];
        --q;
      }
      perm[q + 1] = v;
    }
  }
}

// Caller-order positions -> sorted tiles (every evaluation).  On rebuild
// also records reference positions, sorted ids/cells and the charged map.
template <typename R>
__global__ __launch_bounds__(256) void k_gather(Dev<R> d, Box<R> b,
                                                unsigned thr) {
  const bool rb = rebuild_now(d.ctrl, thr);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
       i += gridDim.x * blockDim.x) {
    const int src = d.perm[i];
    const R x = wrap<R>(d.px[src], b.lo[0], b.hi[0], b.L[0], b.invL[0]);
    const R y = wrap<R>(d.py[src], b.lo[1], b.hi[1], b.L[1], b.invL[1]);
    const R z = wrap<R>(d.pz[src], b.lo[2], b.hi[2], b.L[2], b.invL[2]);
    const int t = d.type[src];
    d.xs[i] = make4<R>(x, y, z, (R)t);
    int c;
    if (rb) {
      d.ref[src] = make4<R>(x, y, z, R(0));
      d.gid_s[i] = d.gid[src];
      const int key = d.key[src];
      d.cell_s[i] = key >> 1;
      if (!(key & 1)) {
        c = d.cstart[key >> 1] + (i - d.key_start[key]);
        d.c2s[c] = i;
        d.gidc[c] = d.gid[src];
        d.typec[c] = t;
      } else {
        c = -1;
      }
      d.s2c[i] = c;
    } else {
      c = d.s2c[i];
    }
    if (c >= 0) d.xqc[c] = make4<R>(x, y, z, d.charge[src]);
  }
}

// All-atom candidate list: every j != i within r2 in the stencil, except
// charged-charged pairs when the work is split (those go to list B).
// Column-major storage list[k * n + i].  count_only: counts, no writes.
template <typename R>
__global__ __launch_bounds__(256) void k_list_a(Dev<R> d, Box<R> b, Grid g,
                                                R r2, int split, int cap,
                                                unsigned thr) {
  if (!rebuild_now(d.ctrl, thr)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int count = 0;
  if (i < d.n) {
    const R4T<R> xi = d.xs[i];
    const bool qi = split && d.s2c[i] >= 0;
    const int cell = d.cell_s[i];
    const int cx = cell % g.nc[0], cy = (cell / g.nc[0]) % g.nc[1],
              cz = cell / (g.nc[0] * g.nc[1]);
    const Span sx = make_span(cx, g.ext_a[0], g.nc[0]);
    const Span sy = make_span(cy, g.ext_a[1], g.nc[1]);
    const Span sz = make_span(cz, g.ext_a[2], g.nc[2]);
    for (int kz = 0; kz < sz.count; ++kz)
      for (int ky = 0; ky < sy.count; ++ky) {
        const int row = (sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0];
        for (int kx = 0; kx < sx.count; ++kx) {
          const int nb = row + sx.at(kx);
          const int js = qi ? d.key_start[2 * nb + 1] : d.key_start[2 * nb];
          const int je = d.key_start[2 * nb + 2];
          for (int j = js; j < je; ++j) {
            if (j == i) continue;
            const R4T<R> xj = d.xs[j];
            R dx, dy, dz;
            const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z,
                                    dx, dy, dz);
            if (d2 < r2) {
              if (count < cap) d.list_a[(size_t)count * d.n + i] = j;
              ++count;
            }
          }
        }
      }
    d.cnt_a[i] = count;
  }
  const int m = wave_max_i(count);
  const int tot = wave_sum(count);
  if ((threadIdx.x & 63) == 0 && m > 0) {
    atomicMax(&d.ctrl->max_count_a, m);
    atomicAdd(&d.ctrl->sum_count_a, (unsigned long long)tot);
  }
}

// Charged-pair candidate list (charged compact indices), stencil ext_b.
template <typename R>
__global__ __launch_bounds__(256) void k_list_b(Dev<R> d, Box<R> b, Grid g,
                                                R r2, int cap, unsigned thr) {
  if (!rebuild_now(d.ctrl, thr)) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  int count = 0;
  if (c < d.nc) {
    const int i = d.c2s[c];
    const R4T<R> xi = d.xs[i];
    const int cell = d.cell_s[i];
    const int cx = cell % g.nc[0], cy = (cell / g.nc[0]) % g.nc[1],
              cz = cell / (g.nc[0] * g.nc[1]);
    const Span sx = make_span(cx, g.ext_b[0], g.nc[0]);
    const Span sy = make_span(cy, g.ext_b[1], g.nc[1]);
    const Span sz = make_span(cz, g.ext_b[2], g.nc[2]);
    for (int kz = 0; kz < sz.count; ++kz)
      for (int ky = 0; ky < sy.count; ++ky) {
        const int row = (sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0];
        for (int kx = 0; kx < sx.count; ++kx) {
          const int nb = row + sx.at(kx);
          const int js = d.key_start[2 * nb], je = d.key_start[2 * nb + 1];
          for (int j = js; j < je; ++j) {
            if (j == i) continue;
            const R4T<R> xj = d.xs[j];
            R dx, dy, dz;
            const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z,
                                    dx, dy, dz);
            if (d2 < r2) {
              if (count < cap) d.list_b[(size_t)count * d.nc + c] = d.s2c[j];
              ++count;
            }
          }
        }
      }
    d.cnt_b[c] = count;
  }
  const int m = wave_max_i(count);
  const int tot = wave_sum(count);
  if ((threadIdx.x & 63) == 0 && m > 0) {
    atomicMax(&d.ctrl->max_count_b, m);
    atomicAdd(&d.ctrl->sum_count_b, (unsigned long long)tot);
  }
}

// ---------------------------------------------------------------------------
// Pair interaction of one directed pair.  Returns the weight w applied to the
// force (RBL: 1 core, f selected shell, 0 otherwise; Verlet: 1 inside rc).
// ---------------------------------------------------------------------------
template <typename R, bool RBL>
__device__ __forceinline__ R pair_weight(const PairConst<R>& pc, R d2,
                                         bool nonzero, int gi,
                                         const int* gid_j_ptr) {
  if (!RBL) return R(1);
  if (d2 < pc.t_core) return (d2 < pc.rs2 && d2 > pc.eps_min) ? R(1) : R(0);
  if (nonzero && d2 < pc.t_shell && d2 > pc.rs2 &&
      rbl_selected(pc.ifreq, pc.epoch, gi, *gid_j_ptr))
    return pc.freq;
  return R(0);
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

// LJ-only pair of list A.
template <typename R, bool RBL, bool EFLAG>
__device__ __forceinline__ void interact_a(const PairConst<R>& pc,
                                           const Box<R>& b, const R* s_lj,
                                           int ti, const R4T<R>& xi, int gi,
                                           const R4T<R>& xj, const int* gidj,
                                           Acc<R>& a) {
  R dx, dy, dz;
  const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
  if (!(d2 < pc.rclj2) || !(d2 < pc.rc2)) return;
  const int tj = (int)xj.w;
  const int tt = 2 * (ti * pc.ntypes + tj);
  const R eps = s_lj[tt], sig2 = s_lj[tt + 1];
  const R inv = R(1) / d2;
  const R s2 = sig2 * inv;
  const R s6 = s2 * s2 * s2;
  const R flj = R(-24) * eps * (R(2) * s6 - R(1)) * s6 * inv;
  const R w = pair_weight<R, RBL>(pc, d2, true, gi, gidj);
  const R f = w * flj;
  a.fx += f * dx;
  a.fy += f * dy;
  a.fz += f * dz;
  if (EFLAG) {
    a.e_lj += (double)(R(2) * eps * (s6 - R(1)) * s6);
    add_virial<R, EFLAG>(a, dx, dy, dz, flj);
  }
}

// erfc-screened Coulomb scalar and half entry energy.
template <typename R>
__device__ __forceinline__ void coul(const PairConst<R>& pc, R d2, R qq, R& f,
                                     R& e) {
  const R r = sqrt(d2);
  const R ec = erfc(pc.g * r);
  const R gn = ec / d2 + pc.two_g_over_sqrtpi * exp(-pc.alpha * d2) / r;
  f = -pc.qqr2e * qq * gn / r;
  e = R(0.5) * pc.qqr2e * qq * ec / r;
}

// LJ + Coulomb pair of list B (both atoms charged).
template <typename R, bool RBL, bool EFLAG>
__device__ __forceinline__ void interact_b(const PairConst<R>& pc,
                                           const Box<R>& b, const R* s_lj,
                                           int ti, const R4T<R>& xi, int gi,
                                           const R4T<R>& xj, int tj,
                                           const int* gidj, Acc<R>& a) {
  R dx, dy, dz;
  const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
  if (!(d2 < pc.rc2)) return;
  R flj = 0, elj = 0;
  if (d2 < pc.rclj2) {
    const int tt = 2 * (ti * pc.ntypes + tj);
    const R eps = s_lj[tt], sig2 = s_lj[tt + 1];
    const R inv = R(1) / d2;
    const R s2 = sig2 * inv;
    const R s6 = s2 * s2 * s2;
    flj = R(-24) * eps * (R(2) * s6 - R(1)) * s6 * inv;
    elj = R(2) * eps * (s6 - R(1)) * s6;
  }
  R fc, ec;
  coul<R>(pc, d2, xi.w * xj.w, fc, ec);
  const R w = pair_weight<R, RBL>(pc, d2, true, gi, gidj);
  const R f = w * (flj + fc);
  a.fx += f * dx;
  a.fy += f * dy;
  a.fz += f * dz;
  if (EFLAG) {
    a.e_lj += (double)elj;
    a.e_c += (double)ec;
    add_virial<R, EFLAG>(a, dx, dy, dz, flj + fc);
  }
}

template <typename R>
__device__ __forceinline__ void load_lj(const R* tab, int ntypes, R* s) {
  for (int t = threadIdx.x; t < 2 * ntypes * ntypes; t += blockDim.x)
    s[t] = tab[t];
  __syncthreads();
}

template <typename R, bool EFLAG>
__device__ __forceinline__ void write_partials(const Acc<R>& a, double* part,
                                               double* smem) {
  double v[11] = {(double)a.fx, (double)a.fy, (double)a.fz, a.e_lj, a.e_c,
                  a.v[0],       a.v[1],       a.v[2],       a.v[3], a.v[4],
                  a.v[5]};
  if (EFLAG) {
    block_sum<11>(v, smem);
    if (threadIdx.x == 0)
      for (int k = 0; k < 11; ++k) part[11 * blockIdx.x + k] = v[k];
  } else {
    double u[3] = {v[0], v[1], v[2]};
    block_sum<3>(u, smem);
    if (threadIdx.x == 0) {
      for (int k = 0; k < 3; ++k) part[11 * blockIdx.x + k] = u[k];
      for (int k = 3; k < 11; ++k) part[11 * blockIdx.x + k] = 0.0;
    }
  }
}

template <typename R, bool RBL, bool EFLAG>
__global__ __launch_bounds__(256) void k_pair_a(Dev<R> d, Box<R> b, Grid g,
                                                PairConst<R> pc,
                                                const R* lj_tab, int split) {
  __shared__ R s_lj[2 * kMaxTypes * kMaxTypes];
  __shared__ double smem[4 * 11];
  load_lj<R>(lj_tab, pc.ntypes, s_lj);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  Acc<R> a;
  if (i < d.n) {
    const R4T<R> xi = d.xs[i];
    const int ti = (int)xi.w;
    const int gi = d.gid_s[i];
    const int cnt = d.cnt_a[i];
    if (cnt <= d.cap_a) {
      const int* lp = d.list_a + i;
      for (int k = 0; k < cnt; ++k) {
        const int j = lp[(size_t)k * d.n];
        interact_a<R, RBL, EFLAG>(pc, b, s_lj, ti, xi, gi, d.xs[j],
                                  d.gid_s + j, a);
      }
    } else {
      // list overflow since the last rebuild: scan the stencil directly
      const bool qi = split && d.s2c[i] >= 0;
      const int cell = d.cell_s[i];
      const int cx = cell % g.nc[0], cy = (cell / g.nc[0]) % g.nc[1],
                cz = cell / (g.nc[0] * g.nc[1]);
      const Span sx = make_span(cx, g.ext_a[0], g.nc[0]);
      const Span sy = make_span(cy, g.ext_a[1], g.nc[1]);
      const Span sz = make_span(cz, g.ext_a[2], g.nc[2]);
      for (int kz = 0; kz < sz.count; ++kz)
        for (int ky = 0; ky < sy.count; ++ky)
          for (int kx = 0; kx < sx.count; ++kx) {
            const int nb =
                (sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0] + sx.at(kx);
            const int js = qi ? d.key_start[2 * nb + 1] : d.key_start[2 * nb];
            const int je = d.key_start[2 * nb + 2];
            for (int j = js; j < je; ++j)
              if (j != i)
                interact_a<R, RBL, EFLAG>(pc, b, s_lj, ti, xi, gi, d.xs[j],
                                          d.gid_s + j, a);
          }
    }
    d.fa[i] = make4<R>(a.fx, a.fy, a.fz, R(0));
  }
  write_partials<R, EFLAG>(a, d.pair_part, smem);
}

template <typename R, bool RBL, bool EFLAG>
__global__ __launch_bounds__(256) void k_pair_b(Dev<R> d, Box<R> b, Grid g,
                                                PairConst<R> pc,
                                                const R* lj_tab,
                                                int part_offset) {
  __shared__ R s_lj[2 * kMaxTypes * kMaxTypes];
  __shared__ double smem[4 * 11];
  load_lj<R>(lj_tab, pc.ntypes, s_lj);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  Acc<R> a;
  if (c < d.nc) {
    const R4T<R> xi = d.xqc[c];
    const int ti = d.typec[c];
    const int gi = d.gidc[c];
    const int cnt = d.cnt_b[c];
    if (cnt <= d.cap_b) {
      const int* lp = d.list_b + c;
      for (int k = 0; k < cnt; ++k) {
        const int e = lp[(size_t)k * d.nc];
        interact_b<R, RBL, EFLAG>(pc, b, s_lj, ti, xi, gi, d.xqc[e],
                                  d.typec[e], d.gidc + e, a);
      }
    } else {
      const int i = d.c2s[c];
      const int cell = d.cell_s[i];
      const int cx = cell % g.nc[0], cy = (cell / g.nc[0]) % g.nc[1],
                cz = cell / (g.nc[0] * g.nc[1]);
      const Span sx = make_span(cx, g.ext_b[0], g.nc[0]);
      const Span sy = make_span(cy, g.ext_b[1], g.nc[1]);
      const Span sz = make_span(cz, g.ext_b[2], g.nc[2]);
      for (int kz = 0; kz < sz.count; ++kz)
        for (int ky = 0; ky < sy.count; ++ky)
          for (int kx = 0; kx < sx.count; ++kx) {
            const int nb =
                (sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0] + sx.at(kx);
            const int js = d.key_start[2 * nb], je = d.key_start[2 * nb + 1];
            for (int j = js; j < je; ++j) {
              if (j == i) continue;
              const int e = d.s2c[j];
              interact_b<R, RBL, EFLAG>(pc, b, s_lj, ti, xi, gi, d.xqc[e],
                                        d.typec[e], d.gidc + e, a);
            }
          }
    }
    d.fb[c] = make4<R>(a.fx, a.fy, a.fz, R(0));
  }
  write_partials<R, EFLAG>(a, d.pair_part + 11 * part_offset, smem);
}

// ---------------------------------------------------------------------------
// k-space.  Phases are evaluated in revolutions from fractional coordinates,
// t = m . s with s = (r - lo)/L, reduced to [-1/2, 1/2]; e^{ik.r} then differs
// from the reference's e^{ik.r} by the constant phase e^{ik.lo}, which cancels
// in every force, energy and virial term.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sincos_rev(float t, float& s, float& c) {
  t -= rintf(t);
  sincospif(2.0f * t, &s, &c);
}

template <typename R>
__device__ __forceinline__ void frac_coords(const Box<R>& b, const R4T<R>& p,
                                            float& sx, float& sy, float& sz) {
  sx = (float)((p.x - b.lo[0]) * b.invL[0]);
  sy = (float)((p.y - b.lo[1]) * b.invL[1]);
  sz = (float)((p.z - b.lo[2]) * b.invL[2]);
}

// S(k) partial sums: block (chunk of kSkChunk charged atoms) x (kSkTile k's).
template <typename R>
__global__ __launch_bounds__(256) void k_sk(Dev<R> d, Box<R> b,
                                            const int* __restrict__ m, int K) {
  __shared__ double smem[4 * 2 * kSkTile];
  const int chunk = blockIdx.x, k0 = blockIdx.y * kSkTile;
  const int kn = min(kSkTile, K - k0);
  float mx[kSkTile], my[kSkTile], mz[kSkTile];
#pragma unroll
  for (int kk = 0; kk < kSkTile; ++kk) {
    const int k = k0 + (kk < kn ? kk : 0);
    mx[kk] = (float)m[3 * k];
    my[kk] = (float)m[3 * k + 1];
    mz[kk] = (float)m[3 * k + 2];
  }
  float re[kSkTile], im[kSkTile];
#pragma unroll
  for (int kk = 0; kk < kSkTile; ++kk) re[kk] = im[kk] = 0.f;
  const int a0 = chunk * kSkChunk, a1 = min(d.nc, a0 + kSkChunk);
  for (int a = a0 + threadIdx.x; a < a1; a += blockDim.x) {
    const R4T<R> p = d.xqc[a];
    float sx, sy, sz;
    frac_coords<R>(b, p, sx, sy, sz);
    const float q = (float)p.w;
#pragma unroll
    for (int kk = 0; kk < kSkTile; ++kk) {
      const float t = fma_r(mx[kk], sx, fma_r(my[kk], sy, mz[kk] * sz));
      float sn, cs;
      sincos_rev(t, sn, cs);
      re[kk] = fma_r(q, cs, re[kk]);
      im[kk] = fma_r(q, sn, im[kk]);
    }
  }
  double v[2 * kSkTile];
#pragma unroll
  for (int kk = 0; kk < kSkTile; ++kk) {
    v[2 * kk] = (double)re[kk];
    v[2 * kk + 1] = (double)im[kk];
  }
  block_sum<2 * kSkTile>(v, smem);
  if (threadIdx.x == 0)
    for (int kk = 0; kk < kn; ++kk)
      d.sk_part[(size_t)chunk * d.kcap + k0 + kk] =
          make_double2(v[2 * kk], v[2 * kk + 1]);
}

// One block of 1024: pair partial sums -> mean force / energies / virial;
// S(k) partials -> rho, k-space energy and virial, per-k force coefficients.
template <typename R>
__global__ __launch_bounds__(1024) void k_finalize(Dev<R> d, Box<R> b,
                                                   const int* __restrict__ m,
                                                   int K, int nchunks,
                                                   int npart, Params p,
                                                   double e_self) {
  __shared__ double smem[16 * 11];
  double v[11];
  for (int k = 0; k < 11; ++k) v[k] = 0.0;
  for (int blk = threadIdx.x; blk < npart; blk += blockDim.x)
    for (int k = 0; k < 11; ++k) v[k] += d.pair_part[11 * blk + k];
  block_sum<11>(v, smem);
  double pair[11];
  if (threadIdx.x == 0)
    for (int k = 0; k < 11; ++k) pair[k] = v[k];

  double kv[7] = {0, 0, 0, 0, 0, 0, 0};
  if (p.kspace) {
    const double V = p.volume;
    const double two_pi = 6.283185307179586476925;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
      double re = 0.0, im = 0.0;
      for (int c = 0; c < nchunks; ++c) {
        const double2 s = d.sk_part[(size_t)c * d.kcap + k];
        re += s.x;
        im += s.y;
      }
      d.rho[k] = make_double2(re, im);
      const double kx = two_pi * m[3 * k] / (double)b.L[0];
      const double ky = two_pi * m[3 * k + 1] / (double)b.L[1];
      const double kz = two_pi * m[3 * k + 2] / (double)b.L[2];
      const double k2 = kx * kx + ky * ky + kz * kz;
      const double w = d.kw[k];
      const double A = w * (-4.0 * 3.141592653589793238463 * p.qqr2e) / (V * k2);
      d.kf0[k] = make_float4((float)m[3 * k], (float)m[3 * k + 1],
                             (float)m[3 * k + 2], (float)(A * re));
      d.kf1[k] = make_float4((float)kx, (float)ky, (float)kz, (float)(A * im));
      const double base = p.qqr2e * two_pi / V * w * (re * re + im * im) / k2;
      const double cc = 2.0 * (1.0 / (4.0 * p.alpha) + 1.0 / k2);
      kv[0] += base;
      kv[1] += base * (1.0 - cc * kx * kx);
      kv[2] += base * (1.0 - cc * ky * ky);
      kv[3] += base * (1.0 - cc * kz * kz);
      kv[4] += base * (-cc * kx * ky);
      kv[5] += base * (-cc * kx * kz);
      kv[6] += base * (-cc * ky * kz);
    }
    block_sum<7>(kv, smem);
  }
  if (threadIdx.x == 0) {
    double mean[3] = {0, 0, 0};
    if (p.mean_fix)
      for (int k = 0; k < 3; ++k) mean[k] = pair[k] / p.n_global;
    for (int k = 0; k < 3; ++k) d.mean[k] = mean[k];
    double* t = d.thermo;
    t[0] = pair[3];
    t[1] = pair[4];
    t[2] = p.kspace ? kv[0] + e_self : 0.0;
    for (int k = 0; k < 6; ++k) {
      t[3 + k] = pair[5 + k] + (p.kspace ? kv[1 + k] : 0.0);
      t[9 + k] = p.kspace ? kv[1 + k] : 0.0;
    }
    for (int k = 0; k < 3; ++k) t[15 + k] = mean[k];
  }
}

// RBE / Ewald force on the charged atoms: 64 atoms per block, the k range
// split over kRbeWaves waves, reduced through LDS.
template <typename R>
__global__ __launch_bounds__(64 * kRbeWaves) void k_rbe_force(Dev<R> d,
                                                              Box<R> b,
                                                              int K) {
  __shared__ float s_f[kRbeWaves][3][64];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = blockIdx.x * 64 + lane;
  const int per = (K + kRbeWaves - 1) / kRbeWaves;
  const int k0 = wv * per, k1 = min(K, k0 + per);
  float fx = 0.f, fy = 0.f, fz = 0.f;
  float q = 0.f;
  if (c < d.nc) {
    const R4T<R> p = d.xqc[c];
    float sx, sy, sz;
    frac_coords<R>(b, p, sx, sy, sz);
    q = (float)p.w;
    for (int k = k0; k < k1; ++k) {
      const float4 a = d.kf0[k];
      const float4 e = d.kf1[k];
      const float t = fma_r(a.x, sx, fma_r(a.y, sy, a.z * sz));
      float sn, cs;
      sincos_rev(t, sn, cs);
      const float gg = fma_r(cs, e.w, -sn * a.w);
      fx = fma_r(gg, e.x, fx);
      fy = fma_r(gg, e.y, fy);
      fz = fma_r(gg, e.z, fz);
    }
  }
  s_f[wv][0][lane] = fx;
  s_f[wv][1][lane] = fy;
  s_f[wv][2][lane] = fz;
  __syncthreads();
  if (wv == 0 && c < d.nc) {
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int w = 0; w < kRbeWaves; ++w) {
      sx += s_f[w][0][lane];
      sy += s_f[w][1][lane];
      sz += s_f[w][2][lane];
    }
    d.fk[c] = make4<R>((R)(q * sx), (R)(q * sy), (R)(q * sz), R(0));
  }
}

// Total force in caller order; resets the skin accumulator for the next call.
template <typename R>
__global__ __launch_bounds__(256) void k_assemble(Dev<R> d, int mean_fix,
                                                  int has_b, int has_k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    d.ctrl->maxdisp_bits = 0u;
    d.ctrl->step_count += 1;
    if (d.ctrl_host) {
      d.ctrl_host->max_count_a = d.ctrl->max_count_a;
      d.ctrl_host->max_count_b = d.ctrl->max_count_b;
      d.ctrl_host->step_count = d.ctrl->step_count;
    }
  }
  if (i >= d.n) return;
  R4T<R> f = d.fa[i];
  if (mean_fix) {
    f.x -= (R)d.mean[0];
    f.y -= (R)d.mean[1];
    f.z -= (R)d.mean[2];
  }
  const int c = d.s2c[i];
  if (c >= 0) {
    if (has_b) {
      const R4T<R> g = d.fb[c];
      f.x += g.x;
      f.y += g.y;
      f.z += g.z;
    }
    if (has_k) {
      const R4T<R> g = d.fk[c];
      f.x += g.x;
      f.y += g.y;
      f.z += g.z;
    }
  }
  const int o = d.perm[i];
  d.fx[o] = f.x;
  d.fy[o] = f.y;
  d.fz[o] = f.z;
}

// Parity export: classified pairs (or raw candidates) as (gid_i, gid_j, cls).
template <typename R, bool RBL>
__global__ __launch_bounds__(256) void k_export(Dev<R> d, Box<R> b,
                                                PairConst<R> pc,
                                                int candidates, int cap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  auto emit = [&](int gi, int gj, int cls) {
    const int pos = atomicAdd(d.export_count, 1);
    if (pos < cap) {
      d.export_buf[3 * pos] = gi;
      d.export_buf[3 * pos + 1] = gj;
      d.export_buf[3 * pos + 2] = cls;
    }
  };
  auto classify = [&](R d2, int gi, int gj) -> int {
    if (!RBL) return d2 < pc.rc2 ? 1 : 0;
    if (d2 < pc.t_core) return 1;
    if (d2 < pc.t_shell && rbl_selected(pc.ifreq, pc.epoch, gi, gj)) return 2;
    return 0;
  };
  if (t < d.n) {
    const int i = t;
    const R4T<R> xi = d.xs[i];
    const int gi = d.gid_s[i];
    const int cnt = min(d.cnt_a[i], d.cap_a);
    for (int k = 0; k < cnt; ++k) {
      const int j = d.list_a[(size_t)k * d.n + i];
      const int gj = d.gid_s[j];
      if (candidates) {
        emit(gi, gj, 1);
        continue;
      }
      const R4T<R> xj = d.xs[j];
      R dx, dy, dz;
      const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
      const int cls = classify(d2, gi, gj);
      if (cls) emit(gi, gj, cls);
    }
  }
  if (t < d.nc) {
    const int c = t;
    const R4T<R> xi = d.xqc[c];
    const int gi = d.gidc[c];
    const int cnt = min(d.cnt_b[c], d.cap_b);
    for (int k = 0; k < cnt; ++k) {
      const int e = d.list_b[(size_t)k * d.nc + c];
      const int gj = d.gidc[e];
      if (candidates) {
        emit(gi, gj, 2);
        continue;
      }
      const R4T<R> xj = d.xqc[e];
      R dx, dy, dz;
      const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
      const int cls = classify(d2, gi, gj);
      if (cls) emit(gi, gj, cls);
    }
  }
}

inline int blocks_for(int n, int per = kBlock, int cap = 1 << 20) {
  int b = (n + per - 1) / per;
  if (b < 1) b = 1;
  return b > cap ? cap : b;
}

template <typename R>
hipError_t build_part(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s) {
  const Grid& g = a.grid;
  const unsigned thr = a.rebuild_thr;
  const int nb = blocks_for(d.n);
  const int gb = blocks_for(d.n, kBlock, 2048);
  const int cap_a = a.count_only ? 0 : d.cap_a;
  const int cap_b = a.count_only ? 0 : d.cap_b;
  hipLaunchKernelGGL((k_bin<R>), dim3(gb), dim3(kBlock), 0, s, d, a.box, g,
                     a.p.split, thr);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, d.key_count,
                     d.key_start, d.cstart, g.ncell, d.ctrl, thr);
  hipLaunchKernelGGL((k_scatter<R>), dim3(gb), dim3(kBlock), 0, s, d, thr);
  hipLaunchKernelGGL(k_segsort, dim3(blocks_for(2 * g.ncell, kBlock, 2048)),
                     dim3(kBlock), 0, s, d.perm, d.key_start, 2 * g.ncell,
                     d.ctrl, thr);
  hipLaunchKernelGGL((k_gather<R>), dim3(gb), dim3(kBlock), 0, s, d, a.box,
                     thr);
  hipLaunchKernelGGL((k_list_a<R>), dim3(nb), dim3(kBlock), 0, s, d, a.box, g,
                     a.pc.r2_list_a, a.p.split, cap_a, thr);
  if (a.p.split && d.nc > 0)
    hipLaunchKernelGGL((k_list_b<R>), dim3(blocks_for(d.nc)), dim3(kBlock), 0,
                       s, d, a.box, g, a.pc.r2_list_b, cap_b, thr);
  return hipGetLastError();
}

}  // namespace

template <typename R>
hipError_t launch_build(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s) {
  hipLaunchKernelGGL((k_skin<R>), dim3(1), dim3(kBlock), 0, s, d, a.box, 1,
                     nullptr, 0);
  return build_part<R>(d, a, s);
}

template <typename R>
hipError_t launch_force(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s) {
  const Params& p = a.p;
  const int gb = blocks_for(d.n, kBlock, 2048);
  if (a.ev[0]) hipEventRecord(a.ev[0], s);
  hipLaunchKernelGGL((k_skin<R>), dim3(gb), dim3(kBlock), 0, s, d, a.box,
                     a.force_rebuild, a.m_host, a.m_host ? 3 * p.num_k : 0);
  hipError_t e = build_part<R>(d, a, s);
  if (e != hipSuccess) return e;
  if (a.ev[1]) hipEventRecord(a.ev[1], s);
  const int na = blocks_for(d.n);
  const int nbb = (p.split && d.nc > 0) ? blocks_for(d.nc) : 0;
#define RBMDX_PAIR(RBLv, EFv)                                                  \
  do {                                                                         \
    hipLaunchKernelGGL((k_pair_a<R, RBLv, EFv>), dim3(na), dim3(kBlock), 0, s, \
                       d, a.box, a.grid, a.pc, a.lj_tab, p.split);             \
    if (a.ev[2]) hipEventRecord(a.ev[2], s);                                   \
    if (nbb)                                                                   \
      hipLaunchKernelGGL((k_pair_b<R, RBLv, EFv>), dim3(nbb), dim3(kBlock), 0, \
                         s, d, a.box, a.grid, a.pc, a.lj_tab, na);             \
  } while (0)
  if (p.rbl) {
    if (p.eflag)
      RBMDX_PAIR(true, true);
    else
      RBMDX_PAIR(true, false);
  } else {
    if (p.eflag)
      RBMDX_PAIR(false, true);
    else
      RBMDX_PAIR(false, false);
  }
#undef RBMDX_PAIR
  if (a.ev[3]) hipEventRecord(a.ev[3], s);
  const int K = p.kspace ? p.num_k : 0;
  const int nchunks = (d.nc + kSkChunk - 1) / kSkChunk;
  if (K > 0 && d.nc > 0)
    hipLaunchKernelGGL((k_sk<R>), dim3(nchunks, (K + kSkTile - 1) / kSkTile),
                       dim3(kBlock), 0, s, d, a.box, a.m_use, K);
  Params pf = p;
  if (d.nc == 0) pf.kspace = 0;
  hipLaunchKernelGGL((k_finalize<R>), dim3(1), dim3(1024), 0, s, d, a.box,
                     a.m_use, K, nchunks, na + nbb, pf, a.e_self);
  if (K > 0 && d.nc > 0)
    hipLaunchKernelGGL((k_rbe_force<R>), dim3(blocks_for(d.nc, 64)),
                       dim3(64 * kRbeWaves), 0, s, d, a.box, K);
  if (a.ev[4]) hipEventRecord(a.ev[4], s);
  hipLaunchKernelGGL((k_assemble<R>), dim3(na), dim3(kBlock), 0, s, d,
                     p.mean_fix, nbb > 0 ? 1 : 0, (K > 0 && d.nc > 0) ? 1 : 0);
  if (a.ev[5]) hipEventRecord(a.ev[5], s);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_export(const Dev<R>& d, const StepArgs<R>& a, int candidates,
                         int cap, hipStream_t s) {
  const int nmax = d.n > d.nc ? d.n : d.nc;
  if (a.p.rbl)
    hipLaunchKernelGGL((k_export<R, true>), dim3(blocks_for(nmax)),
                       dim3(kBlock), 0, s, d, a.box, a.pc, candidates, cap);
  else
    hipLaunchKernelGGL((k_export<R, false>), dim3(blocks_for(nmax)),
                       dim3(kBlock), 0, s, d, a.box, a.pc, candidates, cap);
  return hipGetLastError();
}

template hipError_t launch_force<float>(const Dev<float>&,
                                        const StepArgs<float>&, hipStream_t);
template hipError_t launch_force<double>(const Dev<double>&,
                                         const StepArgs<double>&, hipStream_t);
template hipError_t launch_build<float>(const Dev<float>&,
                                        const StepArgs<float>&, hipStream_t);
template hipError_t launch_build<double>(const Dev<double>&,
                                         const StepArgs<double>&, hipStream_t);
template hipError_t launch_export<float>(const Dev<float>&,
                                         const StepArgs<float>&, int, int,
                                         hipStream_t);
template hipError_t launch_export<double>(const Dev<double>&,
                                          const StepArgs<double>&, int, int,
                                          hipStream_t);

}  // namespace rbmdx
