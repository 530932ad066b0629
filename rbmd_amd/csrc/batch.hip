// Replica-batch kernels: one force evaluation of many independent periodic
// boxes of one force field (LJ + erfc Coulomb pairs with RBL or a full pair
// set, RBE k-space), every stage one launch over a range of replicas:
//
//   kb_bin       wrap, cell of every atom, per-cell counts (atomics)
//   kb_scan      per-cell offsets of atoms and charged atoms (one workgroup)
//   kb_scatter   cell buckets in arrival order
//   kb_cellsort  per cell: charged atoms first, then by caller index (a
//                deterministic order); sorted positions, types, ids, hashes
//   kb_pair      persistent over the (replica, brick) work list: the brick's
//                neighbourhood staged in LDS; LJ for every pair with a
//                neutral atom, LJ + Coulomb for charged pairs (stencil of
//                ext_c cells); FP64 energy, virial and force sum per replica
//   kb_kspace    persistent over replicas: S(k) over the replica's charged
//                atoms (FP64 across tiles), RBE energy and virial, and the
//                k-space force of every charged atom
//   kb_finish    mean pair force removed (RBL), forces to caller order
//
// Cells are rebuilt every evaluation (no skin, no stored lists), so an
// evaluation has no host-side decisions and the whole sequence replays as
// one captured graph.  Small boxes (fewer than three cells in a dimension)
// are handled exactly: such a dimension is staged once and pairs take the
// minimum image; elsewhere stage coordinates are shifted across the box
// edges and the minimum image is the identity.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "device.hpp"

namespace rbmdx {
namespace {

__device__ __forceinline__ float wrap_coord(float x, float lo, float L, float invL,
                                            float& s) {
  s = (x - lo) * invL;
  const float f = floorf(s);
  s -= f;
  return f != 0.f ? x - f * L : x;
}

__device__ __forceinline__ int batch_cell(const RepInfo& p, float x, float y, float z,
                                          float w[3]) {
  float s[3];
  w[0] = wrap_coord(x, p.box.lo[0], p.box.L[0], p.box.invL[0], s[0]);
  w[1] = wrap_coord(y, p.box.lo[1], p.box.L[1], p.box.invL[1], s[1]);
  w[2] = wrap_coord(z, p.box.lo[2], p.box.L[2], p.box.invL[2], s[2]);
  int c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    c[d] = max(0, min((int)(s[d] * (float)p.nc[d]), p.nc[d] - 1));
  return p.c0 + (c[2] * p.nc[1] + c[1]) * p.nc[0] + c[0];
}

__global__ __launch_bounds__(256) void kb_bin(BatchDev d, BatchRange rg) {
  const int i = rg.a0 + blockIdx.x * blockDim.x + threadIdx.x;
  // per-replica sums are accumulated with atomics by later stages
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < (rg.r1 - rg.r0) * kBatchAcc;
       q += gridDim.x * blockDim.x)
    d.racc[(size_t)rg.r0 * kBatchAcc + q] = 0.0;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < (rg.r1 - rg.r0) * kBatchKAcc;
       q += gridDim.x * blockDim.x)
    d.kacc[(size_t)rg.r0 * kBatchKAcc + q] = 0.0;
  if (i >= rg.a1) return;
  const RepInfo& p = d.rep[d.rep_of[i]];
  float w[3];
  const int cell = batch_cell(p, d.px[i], d.py[i], d.pz[i], w);
  d.cell_of[i] = cell;
  d.slot[i] = atomicAdd(&d.cnt[cell], 1);
  if (d.charge[i] != 0.f) atomicAdd(&d.cntq[cell], 1);
}

// Exclusive offsets of the range's cells (atoms from rg.a0, charged atoms from
// rg.q0); the counters are cleared for the next evaluation.  One workgroup.
__global__ __launch_bounds__(1024) void kb_scan(BatchDev d, BatchRange rg) {
  __shared__ int sm[33];
  const int nc = rg.c1 - rg.c0;
  const int per = (nc + blockDim.x - 1) / blockDim.x;
  const int s0 = rg.c0 + threadIdx.x * per, s1 = min(rg.c1, s0 + per);
  int a = 0, q = 0;
  for (int c = s0; c < s1; ++c) {
    a += d.cnt[c];
    q += d.cntq[c];
  }
  int ta, tq;
  int ra = block_excl_scan(a, ta, sm) + rg.a0;
  int rq = block_excl_scan(q, tq, sm) + rg.q0;
  for (int c = s0; c < s1; ++c) {
    d.start[c] = ra;
    d.qstart[c] = rq;
    ra += d.cnt[c];
    rq += d.cntq[c];
    d.cnt[c] = 0;
    d.cntq[c] = 0;
  }
  if (threadIdx.x == 0) {
    d.start[rg.c1] = rg.a0 + ta;
    d.qstart[rg.c1] = rg.q0 + tq;
  }
}

__global__ __launch_bounds__(256) void kb_scatter(BatchDev d, BatchRange rg) {
  const int i = rg.a0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rg.a1) return;
  d.tmp[d.start[d.cell_of[i]] + d.slot[i]] = i;
}

// 32 lanes per cell: order the bucket (charged first, then caller index) by
// ranking every atom's key within the cell, and write the sorted arrays in
// parallel; a cell of more than 32 atoms is ordered by its first lane.
__device__ __forceinline__ void batch_sorted_write(const BatchDev& d, int i, int k,
                                                   int qslot) {
  const RepInfo& p = d.rep[d.rep_of[i]];
  float w[3];
  batch_cell(p, d.px[i], d.py[i], d.pz[i], w);
  const float qi = d.charge[i];
  d.xs[k] = make_float4(w[0], w[1], w[2], qi);
  const int g = d.gid[i];
  d.tg[k] = make_int2(d.type[i], g);
  d.hj[k] = rbl_hj(g);
  d.perm[k] = i;
  if (qslot >= 0) d.qidx[qslot] = k;
}

__global__ __launch_bounds__(256) void kb_cellsort(BatchDev d, BatchRange rg) {
  const int lane = threadIdx.x & 31;
  const int c = rg.c0 + (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 5);
  if (c >= rg.c1) return;  // whole 32-lane groups leave together
  const int s0 = d.start[c], s1 = d.start[c + 1], len = s1 - s0;
  auto key = [&](int i) { return (d.charge[i] != 0.f ? 0 : (1 << 30)) + i; };
  if (len <= 32) {
    const int i = lane < len ? d.tmp[s0 + lane] : 0;
    const int kv = lane < len ? key(i) : 0x7fffffff;
    int r = 0;
    for (int q = 0; q < len; ++q) r += __shfl(kv, q, 32) < kv ? 1 : 0;
    if (lane < len) {
      const int q0 = d.qstart[c];
      batch_sorted_write(d, i, s0 + r, kv < (1 << 30) ? q0 + r : -1);
    }
    return;
  }
  if (lane != 0) return;
  for (int k = s0 + 1; k < s1; ++k) {
    const int v = d.tmp[k], kv = key(v);
    int j = k - 1;
    while (j >= s0 && key(d.tmp[j]) > kv) {
      d.tmp[j + 1] = d.tmp[j];
      --j;
    }
    d.tmp[j + 1] = v;
  }
  int q = d.qstart[c];
  for (int k = s0; k < s1; ++k) {
    const int i = d.tmp[k];
    batch_sorted_write(d, i, k, d.charge[i] != 0.f ? q++ : -1);
  }
}

// Slot tables of one brick's neighbourhood (the brick widened by one cell,
// a dimension that would wrap onto itself taken whole).
struct BrickNb {
  int o[3], w[3];      // brick cells [o, o + w)
  int rs[3], rn[3];    // neighbourhood start (may be -1) and count
  int full[3];
  int nslots;
  int mi;
};

__device__ __forceinline__ BrickNb batch_brick(const RepInfo& p, int b, int bw) {
  BrickNb h;
  const int bc[3] = {b % p.nb[0], (b / p.nb[0]) % p.nb[1], b / (p.nb[0] * p.nb[1])};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    h.o[d] = bc[d] * bw;
    h.w[d] = min(bw, p.nc[d] - h.o[d]);
    h.full[d] = h.w[d] + 2 >= p.nc[d];
    h.rs[d] = h.full[d] ? 0 : h.o[d] - 1;
    h.rn[d] = h.full[d] ? p.nc[d] : h.w[d] + 2;
  }
  h.nslots = h.rn[0] * h.rn[1] * h.rn[2];
  h.mi = h.full[0] | h.full[1] | h.full[2];
  return h;
}

// LJ interaction of a pair that involves a neutral atom (interact_a's
// scalars and energy convention).
template <bool RBL, bool EFLAG, bool MI, bool SPECIAL>
__device__ __forceinline__ void batch_lj(const PairConst<float>& pc, const Box<float>& b,
                                         const float* s_lj, int ti, const float4& xi,
                                         unsigned long long hi, const float4& xj, int tj,
                                         int gj, float& fx, float& fy, float& fz,
                                         Acc<float>& a, const SpecRow& sr) {
  float dx, dy, dz;
  const float d2 = pair_d2<float, MI>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
  if (!(d2 < pc.rclj2) || !(d2 < pc.rc2)) return;
  const int tt = 4 * (ti * pc.ntypes + tj);
  const float m24 = s_lj[tt], sig2 = s_lj[tt + 1], eps = s_lj[tt + 2];
  const float inv = pair_rcp(d2);
  const float s2 = sig2 * inv;
  const float s6 = s2 * s2 * s2;
  // CVFF: the owner's special-bond weight scales the LJ term
  const float ws = SPECIAL ? special_w<float>(sr, gj, pc) : 1.f;
  const float flj = ws * (m24 * (2.f * s6 - 1.f) * s6 * inv);
  // the neighbour's hash enters only for shell pairs
  const float w = pair_weight<float, RBL>(pc, d2, hi, [&] { return rbl_hj(gj); });
  const float f = w * flj;
  fx += f * dx;
  fy += f * dy;
  fz += f * dz;
  if (EFLAG) {
    a.e_lj += (double)(ws * (2.f * eps * (s6 - 1.f) * s6));
    add_virial<float, EFLAG>(a, dx, dy, dz, flj);
  }
}

// batch_lj over the staged candidates [l0, l1) of one stencil row without
// energies or special bonds, with no branch per candidate: the owner itself
// and (for a charged owner) charged candidates are masked, the cutoffs and
// the RBL weight are selects, and only a shell candidate (rare at an LJ cutoff
// at or below r_core) forms its selection hash.  Per pair the same terms in
// the same order as batch_lj.
template <bool RBL, bool MI>
__device__ __forceinline__ void batch_lj_row(const PairConst<float>& pc, const Box<float>& b,
                                             const float* s_lj, int ti, const float4& xi,
                                             unsigned long long hi, bool qi, int li,
                                             const float4* sx, const int2* stg, int l0,
                                             int l1, float& fx, float& fy, float& fz) {
  const int row = 4 * ti * pc.ntypes;
  for (int l = l0; l < l1; ++l) {
    const float4 xj = sx[l];
    const int2 tgj = stg[l];
    float dx, dy, dz;
    const float d2 = pair_d2<float, MI>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
    const bool in = (l != li) & !(qi & (xj.w != 0.f)) & (d2 < pc.rclj2) & (d2 < pc.rc2);
    float w = 1.f;
    if (RBL) {
      const bool core = d2 < pc.t_core;
      const float wc = ((d2 < pc.rs2) & (d2 > pc.eps_min)) ? 1.f : 0.f;
      float wsh = 0.f;
      if (in & !core & (d2 < pc.t_shell) & (d2 > pc.rs2))
        wsh = rbl_keep(pc.ifreq, pc.fmask, pc.c32, pc.epoch, hi, rbl_hj(tgj.y)) ? pc.freq
                                                                               : 0.f;
      w = core ? wc : wsh;
    }
    const int tt = row + 4 * tgj.x;
    const float m24 = s_lj[tt], sig2 = s_lj[tt + 1];
    const float inv = pair_rcp(d2);
    const float s2 = sig2 * inv;
    const float s6 = s2 * s2 * s2;
    const float flj = m24 * (2.f * s6 - 1.f) * s6 * inv;
    const float f = (in & (w != 0.f)) ? w * flj : 0.f;
    fx += f * dx;
    fy += f * dy;
    fz += f * dz;
  }
}

// LJ + erfc Coulomb of a charged pair (interact_b's terms: eval_b, with the
// owner's special-bond weight under CVFF; minimum image).
template <bool RBL, bool EFLAG, bool MI, bool SPECIAL>
__device__ __forceinline__ void batch_charged(const PairConst<float>& pc,
                                              const Box<float>& b, const float* s_lj,
                                              int ti, const float4& xi,
                                              unsigned long long hi, const float4& xj,
                                              int tj, int gj, Acc<float>& a,
                                              const SpecRow& sr) {
  float dx, dy, dz;
  const float d2 = pair_d2<float, MI>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
  if (!(d2 < pc.rc2)) return;
  const float w = pair_weight<float, RBL>(pc, d2, hi, [&] { return rbl_hj(gj); });
  const float ws = SPECIAL ? special_w<float>(sr, gj, pc) : 1.f;
  eval_b<float, EFLAG, SPECIAL>(pc, s_lj, ti, tj, xi.w * xj.w, dx, dy, dz, d2, w, a, ws);
}

// The CVFF special-bond row of sorted atom s (counts and partner ids in
// caller order).
template <bool SPECIAL>
__device__ __forceinline__ SpecRow batch_spec_row(const BatchDev& d, int ms, int s) {
  SpecRow r;
  if (SPECIAL) {
    const int i = d.perm[s];
    const int4 c = d.spn[i];
    r.n12 = c.x;
    r.n13 = c.y;
    r.n14 = c.z;
    r.ids = d.spid + (size_t)i * ms;
  }
  return r;
}

template <bool RBL, bool EFLAG, bool SPECIAL>
__global__ __launch_bounds__(256) void kb_pair(BatchDev d, PairConst<float> pc_arg,
                                               BatchRange rg, int bw, int qlanes) {
  __shared__ float4 sx[kBatchStageCap];
  __shared__ int2 stg[kBatchStageCap];
  // slot tables of the staged neighbourhood: up to kBatchMaxSlots cells for
  // the LJ pass, up to kBatchQSlots for the charged pass
  __shared__ int soff[kBatchQSlots + 1];
  __shared__ int sfirst[kBatchQSlots];
  __shared__ int sedge[kBatchQSlots];
  __shared__ int sscan[33];
  __shared__ float s_lj[4 * kBatchMaxTypes * kBatchMaxTypes];
  __shared__ double smem[16 * kBatchAcc];
  const int nt = pc_arg.ntypes;
  for (int q = threadIdx.x; q < 4 * nt * nt; q += blockDim.x) s_lj[q] = d.lj[q];
  const long long epoch = *d.epoch;
  for (int wi = rg.w0 + blockIdx.x; wi < rg.w1; wi += gridDim.x) {
    const BatchWork wk = d.work[wi];
    const RepInfo& p = d.rep[wk.rep];
    PairConst<float> pc = pc_arg;
    if (RBL) {
      pc.epoch = epoch;
      pc.freq = p.freq;
      pc.ifreq = p.ifreq;
      pc.fmask = p.fmask;
      pc.c32 = p.c32;
    }
    const Box<float> b = p.box;
    const BrickNb h = batch_brick(p, wk.brick, bw);
    __syncthreads();  // the previous item's readers of the tables are done
    if (threadIdx.x < 64) {
      int cnt = 0;
      const int s = threadIdx.x;
      if (s < h.nslots) {
        const int k[3] = {s % h.rn[0], (s / h.rn[0]) % h.rn[1], s / (h.rn[0] * h.rn[1])};
        int cw[3], e = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int u = h.rs[q] + k[q];
          cw[q] = wrapi(u, p.nc[q]);
          e |= (u < 0 ? 0 : (u >= p.nc[q] ? 2 : 1)) << (2 * q);
        }
        const int cell = p.c0 + (cw[2] * p.nc[1] + cw[1]) * p.nc[0] + cw[0];
        sfirst[s] = d.start[cell];
        cnt = d.start[cell + 1] - d.start[cell];
        sedge[s] = e;
      }
      const int inc = wave_incl_scan(cnt);
      if (s < h.nslots) soff[s] = inc - cnt;
      if (s == 63) soff[h.nslots] = inc;
    }
    __syncthreads();
    const int total = soff[h.nslots];
    const bool staged = total <= kBatchStageCap;
    if (staged) {
      const int lane = threadIdx.x & 15;
      for (int s = threadIdx.x >> 4; s < h.nslots; s += blockDim.x >> 4) {
        const int o = soff[s], len = soff[s + 1] - o, f = sfirst[s], e = sedge[s];
        float sh[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int c = (e >> (2 * q)) & 3;
          sh[q] = c == 0 ? -b.L[q] : (c == 2 ? b.L[q] : 0.f);
        }
        for (int j = lane; j < len; j += 16) {
          float4 v = d.xs[f + j];
          v.x += sh[0];
          v.y += sh[1];
          v.z += sh[2];
          sx[o + j] = v;
          stg[o + j] = d.tg[f + j];
        }
      }
    }
    __syncthreads();
    Acc<float> a;
    float fsx = 0.f, fsy = 0.f, fsz = 0.f;  // this thread's share of the force sum
    // owners: the atoms of the brick's own cells, one per thread
    // own cells u = 0..own_n-1 (x-fastest within the brick) and their
    // atoms' running offsets (at most 8 cells)
    const int own_n = h.w[0] * h.w[1] * h.w[2];
    int own_end[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      int len = 0;
      if (u < own_n) {
        const int k[3] = {u % h.w[0], (u / h.w[0]) % h.w[1], u / (h.w[0] * h.w[1])};
        const int s = ((h.o[2] - h.rs[2] + k[2]) * h.rn[1] + (h.o[1] - h.rs[1] + k[1])) *
                          h.rn[0] +
                      (h.o[0] - h.rs[0] + k[0]);
        len = soff[s + 1] - soff[s];
      }
      own_end[u] = (u > 0 ? own_end[u - 1] : 0) + len;
    }
    const int nown = own_end[7];
    // lpo adjacent lanes per owner (up to 4, as many as the block leaves
    // idle); lane `sub` takes every lpo-th row of the owner's stencil
    int lpo = 1;
    while (lpo < 4 && nown * lpo * 2 <= (int)blockDim.x) lpo *= 2;
    for (int tt = threadIdx.x; tt < nown * lpo; tt += blockDim.x) {
      const int t = tt / lpo, sub = tt % lpo;
      int u = 0, ub = 0;
#pragma unroll
      for (int q = 0; q < 7; ++q)
        if (own_end[q] <= t) {
          u = q + 1;
          ub = own_end[q];
        }
      const int j0 = t - ub;
      const int k[3] = {u % h.w[0], (u / h.w[0]) % h.w[1], u / (h.w[0] * h.w[1])};
      int ko[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) ko[q] = h.o[q] - h.rs[q] + k[q];
      const int s_own = (ko[2] * h.rn[1] + ko[1]) * h.rn[0] + ko[0];
      const int i = sfirst[s_own] + j0;
      const int li = soff[s_own] + j0;
      const float4 xi = d.xs[i];
      const int2 tgi = d.tg[i];
      const unsigned long long hi = RBL ? rbl_hi(tgi.y) : 0ull;
      const bool qi = xi.w != 0.f;
      const SpecRow sr = batch_spec_row<SPECIAL>(d, pc.ms, i);
      float fx = 0.f, fy = 0.f, fz = 0.f;
      int lo3[3], hi3[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        lo3[q] = h.full[q] ? 0 : ko[q] - 1;
        hi3[q] = h.full[q] ? h.rn[q] - 1 : ko[q] + 1;
      }
      const int nyr = hi3[1] - lo3[1] + 1;
      const int nrow = (hi3[2] - lo3[2] + 1) * nyr;
      for (int row = sub; row < nrow; row += lpo) {
        const int kz = lo3[2] + row / nyr, ky = lo3[1] + row % nyr;
        const int srow = (kz * h.rn[1] + ky) * h.rn[0];
        if (staged) {
          const int l0 = soff[srow + lo3[0]], l1 = soff[srow + hi3[0] + 1];
          if (!EFLAG && !SPECIAL) {
            // the force pass: every candidate's terms as selects (the same
            // pairs, weights and arithmetic as batch_lj)
            if (h.mi)
              batch_lj_row<RBL, true>(pc, b, s_lj, tgi.x, xi, hi, qi, li, sx, stg, l0, l1, fx,
                                      fy, fz);
            else
              batch_lj_row<RBL, false>(pc, b, s_lj, tgi.x, xi, hi, qi, li, sx, stg, l0, l1,
                                       fx, fy, fz);
          } else {
            for (int l = l0; l < l1; ++l) {
              const float4 xj = sx[l];
              if (l == li || (qi && xj.w != 0.f)) continue;
              const int2 tgj = stg[l];
              if (h.mi)
                batch_lj<RBL, EFLAG, true, SPECIAL>(pc, b, s_lj, tgi.x, xi, hi, xj, tgj.x,
                                                    tgj.y, fx, fy, fz, a, sr);
              else
                batch_lj<RBL, EFLAG, false, SPECIAL>(pc, b, s_lj, tgi.x, xi, hi, xj, tgj.x,
                                                     tgj.y, fx, fy, fz, a, sr);
            }
          }
        } else {
          // beyond the stage: the same cells from global memory, minimum image
          for (int kx = lo3[0]; kx <= hi3[0]; ++kx) {
            const int s = srow + kx;
            const int f = sfirst[s], len = soff[s + 1] - soff[s];
            for (int j = 0; j < len; ++j) {
              const float4 xj = d.xs[f + j];
              if (f + j == i || (qi && xj.w != 0.f)) continue;
              const int2 tgj = d.tg[f + j];
              batch_lj<RBL, EFLAG, true, SPECIAL>(pc, b, s_lj, tgi.x, xi, hi, xj, tgj.x,
                                                  tgj.y, fx, fy, fz, a, sr);
            }
          }
        }
      }
      fsx += fx;
      fsy += fy;
      fsz += fz;
      for (int o = lpo / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o, lpo);
        fy += __shfl_xor(fy, o, lpo);
        fz += __shfl_xor(fz, o, lpo);
      }
      if (sub == 0) d.fs[i] = make_float4(fx, fy, fz, 0.f);
    }
    // charged owners: LJ + Coulomb with the charged atoms of the ext_c
    // stencil, qlanes lanes per owner.  The charged atoms of the brick
    // widened by ext_c (a dimension that would wrap onto itself taken whole)
    // are restaged into the LJ stage buffers, edge-shifted; a neighbourhood
    // beyond the tables or the stage is read from global memory.
    __syncthreads();
    {
      int qs[3], qn[3], qfull[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        qfull[q] = h.w[q] + 2 * p.ext_c[q] >= p.nc[q];
        qs[q] = qfull[q] ? 0 : h.o[q] - p.ext_c[q];
        qn[q] = qfull[q] ? p.nc[q] : h.w[q] + 2 * p.ext_c[q];
      }
      const int qslots = qn[0] * qn[1] * qn[2];
      const bool qmi = qfull[0] | qfull[1] | qfull[2];
      bool qstaged = false;
      if (qslots <= kBatchQSlots) {
        const int per = (qslots + blockDim.x - 1) / blockDim.x;
        const int s0 = threadIdx.x * per, s1 = min(qslots, s0 + per);
        int sum = 0;
        for (int s = s0; s < s1; ++s) {
          const int kk[3] = {s % qn[0], (s / qn[0]) % qn[1], s / (qn[0] * qn[1])};
          int cw[3], e = 0;
#pragma unroll
          for (int q = 0; q < 3; ++q) {
            const int u = qs[q] + kk[q];
            cw[q] = wrapi(u, p.nc[q]);
            e |= (u < 0 ? 0 : (u >= p.nc[q] ? 2 : 1)) << (2 * q);
          }
          const int cell = p.c0 + (cw[2] * p.nc[1] + cw[1]) * p.nc[0] + cw[0];
          sfirst[s] = d.start[cell];
          sedge[s] = e;
          const int cnt = d.qstart[cell + 1] - d.qstart[cell];
          soff[s] = cnt;
          sum += cnt;
        }
        __syncthreads();
        int tot;
        int run = block_excl_scan(sum, tot, sscan);
        for (int s = s0; s < s1; ++s) {
          const int c = soff[s];
          soff[s] = run;
          run += c;
        }
        if (threadIdx.x == 0) soff[qslots] = tot;
        __syncthreads();
        qstaged = tot <= kBatchStageCap;
        if (qstaged) {
          const int lane = threadIdx.x & 15;
          for (int s = threadIdx.x >> 4; s < qslots; s += blockDim.x >> 4) {
            const int o = soff[s], len = soff[s + 1] - o, f = sfirst[s], ec = sedge[s];
            float sh[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              const int cc = (ec >> (2 * q)) & 3;
              sh[q] = cc == 0 ? -b.L[q] : (cc == 2 ? b.L[q] : 0.f);
            }
            for (int j = lane; j < len; j += 16) {
              float4 v = d.xs[f + j];
              v.x += sh[0];
              v.y += sh[1];
              v.z += sh[2];
              sx[o + j] = v;
              stg[o + j] = d.tg[f + j];
            }
          }
        }
        __syncthreads();
      }
      const int grp = threadIdx.x / qlanes, gl = threadIdx.x % qlanes;
      const int ngrp = blockDim.x / qlanes;
      int qend[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        int len = 0;
        if (u < own_n) {
          const int k[3] = {u % h.w[0], (u / h.w[0]) % h.w[1], u / (h.w[0] * h.w[1])};
          const int cell = p.c0 + ((h.o[2] + k[2]) * p.nc[1] + h.o[1] + k[1]) * p.nc[0] +
                           h.o[0] + k[0];
          len = d.qstart[cell + 1] - d.qstart[cell];
        }
        qend[u] = (u > 0 ? qend[u - 1] : 0) + len;
      }
      const int nqo = qend[7];
      for (int t0 = 0; t0 < nqo; t0 += ngrp) {
        const int t = t0 + grp;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        int i = -1;
        if (t < nqo) {
          int u = 0, ub = 0;
#pragma unroll
          for (int q = 0; q < 7; ++q)
            if (qend[q] <= t) {
              u = q + 1;
              ub = qend[q];
            }
          const int k[3] = {u % h.w[0], (u / h.w[0]) % h.w[1], u / (h.w[0] * h.w[1])};
          const int c[3] = {h.o[0] + k[0], h.o[1] + k[1], h.o[2] + k[2]};
          const int cell = p.c0 + (c[2] * p.nc[1] + c[1]) * p.nc[0] + c[0];
          i = d.start[cell] + (t - ub);
          const float4 xi = d.xs[i];
          const int ti = d.tg[i].x;
          const unsigned long long hi = RBL ? rbl_hi(d.tg[i].y) : 0ull;
          const SpecRow sr = batch_spec_row<SPECIAL>(d, pc.ms, i);
          Acc<float> ac;
          if (qstaged) {
            // rows of the stencil in the restaged neighbourhood
            int lo3[3], nn[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
              lo3[q] = qfull[q] ? 0 : c[q] - h.o[q];
              nn[q] = qfull[q] ? qn[q] : 2 * p.ext_c[q] + 1;
            }
            const int li = soff[((c[2] - qs[2]) * qn[1] + (c[1] - qs[1])) * qn[0] +
                                (c[0] - qs[0])] + (t - ub);
            for (int row = gl; row < nn[1] * nn[2]; row += qlanes) {
              const int ky = lo3[1] + row % nn[1], kz = lo3[2] + row / nn[1];
              const int srow = (kz * qn[1] + ky) * qn[0];
              const int l1 = soff[srow + lo3[0] + nn[0]];
              for (int l = soff[srow + lo3[0]]; l < l1; ++l) {
                if (l == li) continue;
                const int2 tgj = stg[l];
                if (qmi)
                  batch_charged<RBL, EFLAG, true, SPECIAL>(pc, b, s_lj, ti, xi, hi, sx[l],
                                                           tgj.x, tgj.y, ac, sr);
                else
                  batch_charged<RBL, EFLAG, false, SPECIAL>(pc, b, s_lj, ti, xi, hi, sx[l],
                                                            tgj.x, tgj.y, ac, sr);
              }
            }
          } else {
            const Span spx = make_span(c[0], p.ext_c[0], p.nc[0]);
            const Span spy = make_span(c[1], p.ext_c[1], p.nc[1]);
            const Span spz = make_span(c[2], p.ext_c[2], p.nc[2]);
            const int nst = spx.count * spy.count * spz.count;
            for (int s = gl; s < nst; s += qlanes) {
              const int kx = s % spx.count, ky = (s / spx.count) % spy.count,
                        kz = s / (spx.count * spy.count);
              const int cj =
                  p.c0 + (spz.at(kz) * p.nc[1] + spy.at(ky)) * p.nc[0] + spx.at(kx);
              const int j0 = d.start[cj], j1 = j0 + (d.qstart[cj + 1] - d.qstart[cj]);
              for (int j = j0; j < j1; ++j) {
                if (j == i) continue;
                batch_charged<RBL, EFLAG, true, SPECIAL>(pc, b, s_lj, ti, xi, hi, d.xs[j],
                                                         d.tg[j].x, d.tg[j].y, ac, sr);
              }
            }
          }
          fx = ac.fx;
          fy = ac.fy;
          fz = ac.fz;
          a.e_lj += ac.e_lj;
          a.e_c += ac.e_c;
#pragma unroll
          for (int q = 0; q < 6; ++q) a.v[q] += ac.v[q];
        }
        for (int o = qlanes / 2; o > 0; o >>= 1) {
          fx += __shfl_xor(fx, o, qlanes);
          fy += __shfl_xor(fy, o, qlanes);
          fz += __shfl_xor(fz, o, qlanes);
        }
        if (i >= 0 && gl == 0) {
          float4 f = d.fs[i];
          f.x += fx;
          f.y += fy;
          f.z += fz;
          d.fs[i] = f;
          fsx += fx;
          fsy += fy;
          fsz += fz;
        }
      }
    }
    double v[kBatchAcc] = {(double)fsx, (double)fsy, (double)fsz, a.e_lj, a.e_c,
                           a.v[0],      a.v[1],      a.v[2],      a.v[3], a.v[4],
                           a.v[5]};
    block_sum<kBatchAcc>(v, smem);
    if (threadIdx.x == 0)
      for (int q = 0; q < (EFLAG ? kBatchAcc : 3); ++q)
        if (v[q] != 0.0) atomicAdd(&d.racc[(size_t)wk.rep * kBatchAcc + q], v[q]);
  }
}

// k-space of one replica per workgroup pass: S(k) (one k per thread,
// charged atoms staged in tiles, FP32 within a tile and FP64 across), the
// energy and virial (FP64), then the force on every charged atom.  RBE and
// EWALD: the k's of m, weights kw or exp(-k^2 / 4 alpha), virial
// E_k (delta_ab - 2 k_a k_b (1/(4 alpha) + 1/k^2)).  RBSOG: the P sampled k's
// (weight kw) and the nd direct vectors (per-replica weights kwd), every k in
// the RBE form E_k = qqr2e (2 pi / V) w |rho|^2 / k^2; virial W_aa += E_k and
// W_ab += tau_k |rho_s|^2 k_s,a k_s,b with s the virial-chain source of a
// sampled k (tau = ktau, over |k_s|^4) or the direct vector itself (tau = kvd)
// (kspace_calculator_op.hip.cu:1133-1222, :1308-1352).
constexpr int kKsTile = 256;
template <bool EFLAG>
__global__ __launch_bounds__(256) void kb_kspace(BatchDev d, PairConst<float> pc,
                                                 BatchRange rg) {
  __shared__ float4 sa[kKsTile];
  __shared__ float4 kf0[kBatchMaxK];
  __shared__ float4 kf1[kBatchMaxK];
  __shared__ double2 srho[kBatchMaxK];
  __shared__ double smem[16 * kBatchKAcc];
  const double two_pi = 6.283185307179586476925;
  const int P = d.P;
  const bool sog = d.kmode == 2;
  const int K = P + (sog ? d.nd : 0);
  for (int r = rg.r0 + blockIdx.x; r < rg.r1; r += gridDim.x) {
    const RepInfo& p = d.rep[r];
    const int* m = d.m + (size_t)r * d.mstride;
    // integer vector of k: sampled (replica's m) or direct (shared table)
    auto mvec = [&](int k) { return k < P ? m + 3 * k : d.dm + 3 * (k - P); };
    double kv[kBatchKAcc] = {0, 0, 0, 0, 0, 0, 0};
    for (int k0 = 0; k0 < K; k0 += blockDim.x) {
      const int k = k0 + threadIdx.x;
      const int* mk = mvec(k < K ? k : 0);
      const float mx = (float)mk[0], my = (float)mk[1], mz = (float)mk[2];
      double re = 0.0, im = 0.0;
      for (int t0 = 0; t0 < p.nq; t0 += kKsTile) {
        const int tn = min(kKsTile, p.nq - t0);
        __syncthreads();
        for (int j = threadIdx.x; j < tn; j += blockDim.x) {
          const float4 x = d.xs[d.qidx[p.q0 + t0 + j]];
          sa[j] = make_float4((x.x - p.box.lo[0]) * p.box.invL[0],
                              (x.y - p.box.lo[1]) * p.box.invL[1],
                              (x.z - p.box.lo[2]) * p.box.invL[2], x.w);
        }
        __syncthreads();
        float fr = 0.f, fi = 0.f;
        for (int j = 0; j < tn; ++j) {
          const float4 a = sa[j];
          float sn, cs;
          sincos_rev(fma_r(mx, a.x, fma_r(my, a.y, mz * a.z)), sn, cs);
          fr = fma_r(a.w, cs, fr);
          fi = fma_r(a.w, sn, fi);
        }
        re += (double)fr;
        im += (double)fi;
      }
      if (k < K) {
        const double kx = two_pi * mk[0] / (double)p.box.L[0];
        const double ky = two_pi * mk[1] / (double)p.box.L[1];
        const double kz = two_pi * mk[2] / (double)p.box.L[2];
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double wk = d.kmode == 1 ? exp(-k2 / (4.0 * (double)pc.alpha))
                          : (k < P ? p.kw : d.kwd[(size_t)r * d.nd + (k - P)]);
        const double A = wk * (-4.0 * 3.141592653589793238463 * (double)pc.qqr2e) /
                         (p.volume * k2);
        kf0[k] = make_float4((float)mk[0], (float)mk[1], (float)mk[2], (float)(A * re));
        kf1[k] = make_float4((float)kx, (float)ky, (float)kz, (float)(A * im));
        srho[k] = make_double2(re, im);
        if (EFLAG) {
          const double base =
              (double)pc.qqr2e * two_pi / p.volume * wk * (re * re + im * im) / k2;
          kv[0] += base;
          if (sog) {
            kv[1] += base;
            kv[2] += base;
            kv[3] += base;
          } else {
            const double cc = 2.0 * (1.0 / (4.0 * (double)pc.alpha) + 1.0 / k2);
            kv[1] += base * (1.0 - cc * kx * kx);
            kv[2] += base * (1.0 - cc * ky * ky);
            kv[3] += base * (1.0 - cc * kz * kz);
            kv[4] += base * (-cc * kx * ky);
            kv[5] += base * (-cc * kx * kz);
            kv[6] += base * (-cc * ky * kz);
          }
        }
      }
    }
    __syncthreads();  // srho of every k is written
    if (EFLAG) {
      if (sog)
        for (int k = threadIdx.x; k < K; k += blockDim.x) {
          const int src = k < P ? m[3 * P + k] : k;
          const int* ms = mvec(src);
          const double kx = two_pi * ms[0] / (double)p.box.L[0];
          const double ky = two_pi * ms[1] / (double)p.box.L[1];
          const double kz = two_pi * ms[2] / (double)p.box.L[2];
          const double2 rs = srho[src];
          double t = (k < P ? p.ktau : d.kvd[(size_t)r * d.nd + (k - P)]) *
                     (rs.x * rs.x + rs.y * rs.y);
          if (k < P) {
            const double k2 = kx * kx + ky * ky + kz * kz;
            t /= k2 * k2;
          }
          kv[1] += t * kx * kx;
          kv[2] += t * ky * ky;
          kv[3] += t * kz * kz;
          kv[4] += t * kx * ky;
          kv[5] += t * kx * kz;
          kv[6] += t * ky * kz;
        }
      block_sum<kBatchKAcc>(kv, smem);
      if (threadIdx.x == 0)
        for (int q = 0; q < kBatchKAcc; ++q) d.kacc[(size_t)r * kBatchKAcc + q] = kv[q];
    }
    __syncthreads();
    for (int c = threadIdx.x; c < p.nq; c += blockDim.x) {
      const int s = d.qidx[p.q0 + c];
      const float4 x = d.xs[s];
      const float sx = (x.x - p.box.lo[0]) * p.box.invL[0];
      const float sy = (x.y - p.box.lo[1]) * p.box.invL[1];
      const float sz = (x.z - p.box.lo[2]) * p.box.invL[2];
      float fx = 0.f, fy = 0.f, fz = 0.f;
      for (int k = 0; k < K; ++k) {
        const float4 a = kf0[k];
        const float4 e = kf1[k];
        float sn, cs;
        sincos_rev(fma_r(a.x, sx, fma_r(a.y, sy, a.z * sz)), sn, cs);
        const float gg = fma_r(cs, e.w, -sn * a.w);
        fx = fma_r(gg, e.x, fx);
        fy = fma_r(gg, e.y, fy);
        fz = fma_r(gg, e.z, fz);
      }
      float4 f = d.fs[s];
      f.x += x.w * fx;
      f.y += x.w * fy;
      f.z += x.w * fz;
      d.fs[s] = f;
    }
    __syncthreads();
  }
}

// Forces to caller order; the RBL mean pair force of each replica removed.
__global__ __launch_bounds__(256) void kb_finish(BatchDev d, BatchRange rg, int mean_fix) {
  const int s = rg.a0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= rg.a1) return;
  const int i = d.perm[s];
  float4 f = d.fs[s];
  if (mean_fix) {
    const int r = d.rep_of[i];
    const double* acc = d.racc + (size_t)r * kBatchAcc;
    const double inv = 1.0 / (double)d.rep[r].n;
    f.x -= (float)(acc[0] * inv);
    f.y -= (float)(acc[1] * inv);
    f.z -= (float)(acc[2] * inv);
  }
  d.fx[i] = f.x;
  d.fy[i] = f.y;
  d.fz[i] = f.z;
}

__global__ void kb_epoch(long long* e, long long step) { *e = step; }

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

hipError_t batch_set_epoch(long long* epoch, long long step, hipStream_t s) {
  hipLaunchKernelGGL(kb_epoch, dim3(1), dim3(1), 0, s, epoch, step);
  return hipGetLastError();
}

hipError_t batch_launch(const BatchDev& d, const PairConst<float>& pc, int rbl,
                        int coulomb, int eflag, int special, const BatchRange& rg,
                        const BatchTune& t, int num_cu, hipStream_t s) {
  const int na = rg.a1 - rg.a0, ncl = rg.c1 - rg.c0, nw = rg.w1 - rg.w0;
  if (na <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_bin, dim3(cdiv(na, 256)), dim3(256), 0, s, d, rg);
  hipLaunchKernelGGL(kb_scan, dim3(1), dim3(1024), 0, s, d, rg);
  hipLaunchKernelGGL(kb_scatter, dim3(cdiv(na, 256)), dim3(256), 0, s, d, rg);
  hipLaunchKernelGGL(kb_cellsort, dim3(cdiv(ncl, 8)), dim3(256), 0, s, d, rg);
  const int pg = std::max(1, std::min(nw, t.grid_per_cu * num_cu));
#define RBMDX_KB_PAIR(RB, EF, SP)                                                      \
  hipLaunchKernelGGL((kb_pair<RB, EF, SP>), dim3(pg), dim3(t.pair_block), 0, s, d, pc, rg, \
                     t.bw, t.qlanes)
#define RBMDX_KB_PAIR2(RB, EF)   \
  if (special)                   \
    RBMDX_KB_PAIR(RB, EF, true); \
  else                           \
    RBMDX_KB_PAIR(RB, EF, false)
  if (rbl && eflag) {
    RBMDX_KB_PAIR2(true, true);
  } else if (rbl) {
    RBMDX_KB_PAIR2(true, false);
  } else if (eflag) {
    RBMDX_KB_PAIR2(false, true);
  } else {
    RBMDX_KB_PAIR2(false, false);
  }
#undef RBMDX_KB_PAIR2
#undef RBMDX_KB_PAIR
  if (coulomb && d.P > 0) {
    const int kg = std::min(rg.r1 - rg.r0, 4 * num_cu);
    if (eflag)
      hipLaunchKernelGGL((kb_kspace<true>), dim3(kg), dim3(256), 0, s, d, pc, rg);
    else
      hipLaunchKernelGGL((kb_kspace<false>), dim3(kg), dim3(256), 0, s, d, pc, rg);
  }
  hipLaunchKernelGGL(kb_finish, dim3(cdiv(na, 256)), dim3(256), 0, s, d, rg, rbl);
  return hipGetLastError();
}

}  // namespace rbmdx
