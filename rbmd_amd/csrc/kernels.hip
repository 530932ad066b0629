// HIP kernels of the RBMD force path for gfx950 (CDNA4, wave64).
//
// Per force evaluation, one stream, no host synchronisation:
//   k_skin        rebuild decision, part 1 (max displacement since the last
//                 rebuild, per block) and the gather of caller-order
//                 positions into the cell-sorted tiles
//   k_bin_sk      folds the block maxima into the decision; on a rebuild the
//                 cell keys.  Beside them, the S(k) partials (atom chunks x
//                 k tiles) of sampled k-space
//   k_scan_fin    on a rebuild the scan of the cell counts; beside it, the
//                 S(k) reduction per k tile (rho_k, per-k energy terms and
//                 force coefficients)
//   k_scatter_rbe on a rebuild the scatter into cell order; beside it, the
//                 RBE / RBSOG force per charged atom
//   k_cellsort,   on a rebuild only (early exit otherwise): deterministic
//   k_lists,      in-cell order and the gather in it; brick-local all-atom
//   k_lists_dense candidate lists from LDS-staged neighbourhoods and the
//                 charged-pair lists (per cell in mostly charged systems)
//   k_pair        one workgroup per brick: the brick's neighbourhood is
//                 staged in LDS; LJ pairs gather from LDS through 16-bit
//                 brick-local lists with the RBL classification applied on
//                 the fly; in mostly neutral systems the charged atoms then
//                 add LJ + Coulomb over their charged-pair lists
//   k_pair_b      mostly charged systems: the charged pairs, one wave per
//                 atom (k_pair is not launched when list A is empty)
//   k_reduce      one block: the per-block partials in a fixed order (mean
//                 pair force, energies, virial, k-space sums; list
//                 statistics)
//   k_assemble    F = F_pair - <F_pair> + F_kspace in caller order (and the
//                 leapfrog update when asked)
// Cross-block reductions happen at kernel boundaries only: no same-address
// atomics in the hot kernels and no last-block election (its device-scope
// fences write back the L2 of every XCD).
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "device.hpp"
#include "engine.hpp"

namespace rbmdx {
namespace {

// ---------------------------------------------------------------------------
// Brick neighbourhood.  A brick is a bw^3 block of cells; its neighbourhood
// with extension e adds e cells on every side (the whole dimension when that
// would wrap onto itself).  Slots are enumerated x-fastest; every slot holds
// one cell whose atoms are contiguous in sorted order (charged first).
// ---------------------------------------------------------------------------
struct Neigh {
  int start[3], cnt[3], full[3], e[3];
  int nslots;
  int cell0, cell1;  // cell id range of the brick
  // 1 when the neighbourhood covers a whole dimension: its stage then holds
  // that dimension once, in box coordinates, and pairs need the minimum
  // image.  Otherwise stage_atoms shifts the cells that lie across the box
  // edge by +-L, and every stencil pair (cells at most 2 apart, while a
  // dimension that is not whole has at least 5 cells) has |r_j - r_i| < L/2
  // in stage coordinates: the minimum image is the identity.
  int mi;
};

// Brick of workgroup b of a one-brick-per-workgroup launch.  Workgroups are
// dispatched round-robin over the 8 XCDs (workgroup b on XCD b % 8); each XCD
// gets a contiguous range of the brick-major numbering, a compact slab of the
// box, so the neighbourhoods it stages overlap in its L2.
__device__ __forceinline__ int xcd_brick(int b, int nb) {
  const int x = b & 7, k = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + k;
}

__device__ __forceinline__ Neigh brick_neigh(const Grid& g, int b,
                                             const int e[3]) {
  Neigh h;
  const int bx = b % g.nb[0], by = (b / g.nb[0]) % g.nb[1],
            bz = b / (g.nb[0] * g.nb[1]);
  const int bc[3] = {bx, by, bz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int o = bc[d] * g.bw;
    const int w = min(g.bw, g.nc[d] - o);
    h.e[d] = e[d];
    if (w + 2 * e[d] >= g.nc[d]) {
      h.start[d] = 0;
      h.cnt[d] = g.nc[d];
      h.full[d] = 1;
    } else {
      h.start[d] = o - e[d];
      h.cnt[d] = w + 2 * e[d];
      h.full[d] = 0;
    }
  }
  h.nslots = h.cnt[0] * h.cnt[1] * h.cnt[2];
  h.mi = (h.full[0] | h.full[1] | h.full[2]) ? 1 : 0;
  h.cell0 = g.brick_cell0[b];
  h.cell1 = g.brick_cell0[b + 1];
  return h;
}

__device__ __forceinline__ int slot_cell(const Grid& g, const Neigh& h, int s) {
  const int kx = s % h.cnt[0], ky = (s / h.cnt[0]) % h.cnt[1],
            kz = s / (h.cnt[0] * h.cnt[1]);
  const int cx = wrapi(h.start[0] + kx, g.nc[0]);
  const int cy = wrapi(h.start[1] + ky, g.nc[1]);
  const int cz = wrapi(h.start[2] + kz, g.nc[2]);
  return g.cell_of_coord[(cz * g.nc[1] + cy) * g.nc[0] + cx];
}

// Slot tables in LDS: off (prefix of staged atom counts), first (sorted index
// of the slot cell's first atom), nq (charged atoms of the slot cell), cfirst
// (charged compact index of the slot cell's first charged atom), edge (the
// slot's box-edge shift: 2 bits per dimension holding 0, 1 or 2 for -L, 0,
// +L; see stage_atoms).
template <int NS>
struct SlotTables {
  int off[NS + 1];
  int first[NS];
  int nq[NS];
  int cfirst[NS];
  int edge[NS];
  int scan[33];
};

// charged_only: stage only the charged atoms of every slot cell.
template <int NS>
__device__ __forceinline__ void build_slots(const Grid& g, const Neigh& h,
                                            const int* key_start,
                                            const int* cstart, bool charged_only,
                                            SlotTables<NS>& t) {
  const int per = (h.nslots + blockDim.x - 1) / blockDim.x;
  const int s0 = threadIdx.x * per;
  int sum = 0;
  for (int s = s0; s < min(h.nslots, s0 + per); ++s) {
    const int cell = slot_cell(g, h, s);
    {
      const int kk[3] = {s % h.cnt[0], (s / h.cnt[0]) % h.cnt[1],
                         s / (h.cnt[0] * h.cnt[1])};
      int e = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int u = h.start[d] + kk[d];
        e |= (u < 0 ? 0 : (u >= g.nc[d] ? 2 : 1)) << (2 * d);
      }
      t.edge[s] = e;
    }
    const int a = key_start[2 * cell];
    const int nq = key_start[2 * cell + 1] - a;
    t.first[s] = a;
    t.nq[s] = nq;
    t.cfirst[s] = cstart ? cstart[cell] : 0;
    const int cnt = charged_only ? nq : key_start[2 * cell + 2] - a;
    t.off[s] = cnt;  // count, turned into an offset below
    sum += cnt;
  }
  __syncthreads();
  int total;
  int run = block_excl_scan(sum, total, t.scan);
  for (int s = s0; s < min(h.nslots, s0 + per); ++s) {
    const int c = t.off[s];
    t.off[s] = run;
    run += c;
  }
  if (threadIdx.x == 0) t.off[h.nslots] = total;
  __syncthreads();
}

// Slot holding stage position l (binary search over the prefix table).
template <int NS>
__device__ __forceinline__ int slot_of_pos(const SlotTables<NS>& t, int nslots,
                                           int l) {
  int lo = 0, hi = nslots - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t.off[mid] <= l)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Byte offsets of stage entries fit the 16-bit list entries: entry l of the
// stage is stored as l << kStageShift<R> (sizeof of one entry).
template <typename R>
struct StageShift {
  static constexpr int value = sizeof(R) == 4 ? 4 : 5;
};

// .w of a stage entry.  Pair kernel: the byte offset of the atom type's
// entry in one row of the LDS LJ table (4 values per type pair).  List
// build: 0, or 1e30 for a charged atom (added to d^2 for a charged i, so
// that charged-charged pairs fail the cutoff test: they belong to list B).
__device__ __forceinline__ float pair_w(float w) {
  return __int_as_float(((int)w) << 4);
}
__device__ __forceinline__ double pair_w(double w) {
  return (double)(((int)w) << 5);
}
__device__ __forceinline__ int pair_woff(float w) { return __float_as_int(w); }
__device__ __forceinline__ int pair_woff(double w) { return (int)w; }

// Stage the neighbourhood's atoms in slot order.  Cells that lie across a
// box edge (neighbourhood coordinate outside [0, nc) in a dimension that is
// not whole) are shifted by +-L, so that stage coordinates are continuous.
// The pair kernel also gets a sentinel entry at StageCap - 1, far from every
// atom: list A pads its last group of 8 with it.
// With sgid, the atoms' global ids are staged beside them (gid_s in sorted
// order; the sentinel's id is -1).
template <typename R, bool PAIR, int NS>
__device__ __forceinline__ void stage_atoms(const Grid& g, const Box<R>& b,
                                            const Neigh& h,
                                            const SlotTables<NS>& t,
                                            const R4T<R>* xs, R4T<R>* stage,
                                            const int* gid_s = nullptr,
                                            int* sgid = nullptr) {
  // every thread takes stage entries tid, tid + T, ... in batches of
  // kStageBatch: the entries' slots (binary search of the offsets) and
  // source rows first, then all their loads, then the shifted stores, so
  // that a thread waits for memory once per batch
  constexpr int kStageBatch = 8;
  const int total = t.off[h.nslots];
  const int T = (int)blockDim.x;
  for (int l0 = threadIdx.x; l0 < total; l0 += kStageBatch * T) {
    int src[kStageBatch], sl[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int l = min(l0 + u * T, total - 1);
      sl[u] = slot_of_pos(t, h.nslots, l);
      src[u] = t.first[sl[u]] + (l - t.off[sl[u]]);
    }
    R4T<R> p[kStageBatch];
    int gs[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      p[u] = xs[src[u]];
      gs[u] = sgid ? gid_s[src[u]] : 0;
    }
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int l = l0 + u * T;
      if (l >= total) break;
      const int e = t.edge[sl[u]];
      R sh[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int q = (e >> (2 * d)) & 3;
        sh[d] = q == 0 ? -b.L[d] : (q == 2 ? b.L[d] : R(0));
      }
      R4T<R> v = p[u];
      v.x += sh[0];
      v.y += sh[1];
      v.z += sh[2];
      if (PAIR)
        v.w = pair_w(v.w);
      else
        v.w = v.w != (R)(int)v.w ? R(1e30) : R(0);
      stage[l] = v;
      if (sgid) sgid[l] = gs[u];
    }
  }
  if (PAIR && threadIdx.x == 0) {
    stage[StageCap<R>::value - 1] = make4<R>(R(1e18), R(1e18), R(1e18), pair_w(R(0)));
    if (sgid) sgid[StageCap<R>::value - 1] = -1;
  }
  __syncthreads();
}

// The (2e+1)^3 stencil of an atom in neighbourhood slot coordinates: per
// dimension n[d] distinct cells, the k-th at slot coordinate at(d, k).
struct Stencil {
  int n[3], base[3], mode[3], nc[3], e[3];
  __device__ __forceinline__ int at(int d, int k) const {
    // mode 0: window inside the neighbourhood; 1: wrapped window over a full
    // dimension; 2: the whole dimension once
    return mode[d] == 0 ? base[d] - e[d] + k
                        : (mode[d] == 1 ? wrapi(base[d] - e[d] + k, nc[d]) : k);
  }
  __device__ __forceinline__ int count() const { return n[0] * n[1] * n[2]; }
};
__device__ __forceinline__ Stencil atom_stencil(const Grid& g, const Neigh& h,
                                                const int c[3]) {
  Stencil st;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    st.nc[d] = g.nc[d];
    st.e[d] = h.e[d];
    const int w = 2 * h.e[d] + 1;
    if (h.full[d] && w > g.nc[d]) {
      st.n[d] = g.nc[d];
      st.mode[d] = 2;
      st.base[d] = 0;
    } else if (h.full[d]) {
      st.n[d] = w;
      st.mode[d] = 1;
      st.base[d] = c[d];
    } else {
      st.n[d] = w;
      st.mode[d] = 0;
      st.base[d] = c[d] - h.start[d];
    }
  }
  return st;
}

__device__ __forceinline__ int own_slot(const Neigh& h, const int c[3]) {
  int k[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) k[d] = h.full[d] ? c[d] : c[d] - h.start[d];
  return (k[2] * h.cnt[1] + k[1]) * h.cnt[0] + k[0];
}

__device__ __forceinline__ void unpack_coord(int packed, int c[3]) {
  c[0] = packed & 1023;
  c[1] = (packed >> 10) & 1023;
  c[2] = (packed >> 20) & 1023;
}

// ---------------------------------------------------------------------------
// Every evaluation, one thread per caller atom (coalesced reads, scattered
// writes): the rebuild decision, part 1, and the gather of positions into
// the cell-sorted tiles in the order of the last rebuild.
//   * per-block max |r - r_ref|^2 (minimum image), +inf when a rebuild is
//     forced; k_bin folds the block maxima (the kernel boundary makes them
//     visible to every block) into ctrl->maxdisp_bits;
//   * xs[inv_perm[i]] = (r, type + 1/2 if charged), and the charged copy;
//   * xqs[qpos[i]] = (r, q): the charged atoms in a fixed order (ascending
//     caller index) for the k-space kernels.
// When k_bin then decides to rebuild, k_cellsort gathers again in the new
// order.  The gather runs also when the step ring asks for a rebuild: the
// order maps always hold a valid order (the identity, without charged
// copies, until the first rebuild), and k_cellsort overwrites its result.
// With ctl, the last block is the control block and reads no atoms: it
// takes the step ring's header (host memory), publishes the sampling epoch,
// reports a requested rebuild as an infinite block maximum and copies the
// step's k-sample (kcopy vectors, then with_src their virial-chain sources)
// from the host-mapped ring slot to HBM for the k-space kernels, all beside
// the gather.  force (an explicit rebuild of the list builder alone) skips
// the sorted gather.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_skin(Dev<R> d, Box<R> b, int force, int ctl,
                                              int kcopy, int with_src) {
  __shared__ float s_max[4];
  const int nab = (int)gridDim.x - ctl;  // blocks over the atoms
  if ((int)blockIdx.x == nab) {
    const int slot = (int)(d.ctrl->step_count % kSlots);
    // every load of the slot is issued before the stores
    const int* ms = d.stage_ring + (size_t)slot * d.stage_stride;
    const int n3 = 3 * kcopy, nt = n3 + (with_src ? kcopy : 0);
    for (int q0 = 0; q0 < nt; q0 += 8 * (int)blockDim.x) {
      int v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = q0 + u * (int)blockDim.x + (int)threadIdx.x;
        v[u] = q < nt ? ms[q] : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = q0 + u * (int)blockDim.x + (int)threadIdx.x;
        if (q < n3)
          d.m_dev[q] = v[u];
        else if (q < nt)
          d.src_dev[q - n3] = v[u];
      }
    }
    if (threadIdx.x == 0) {
      const StepHdr h = d.hdr_ring[slot];
      d.ctrl->epoch = h.epoch;
      d.skin_part[blockIdx.x] = h.force ? __int_as_float(0x7f800000) : 0.f;
    }
    return;
  }
  float mx = 0.f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n; i += nab * blockDim.x) {
    const R x = wrap<R>(d.px[i], b.lo[0], b.hi[0], b.L[0], b.invL[0]);
    const R y = wrap<R>(d.py[i], b.lo[1], b.hi[1], b.L[1], b.invL[1]);
    const R z = wrap<R>(d.pz[i], b.lo[2], b.hi[2], b.L[2], b.invL[2]);
    // the charged atoms in their fixed order for the k-space kernels
    // (independent of the sorted order, so also on a forced rebuild)
    if (d.qpos) {
      const int qc = d.qpos[i];
      if (qc >= 0) d.xqs[qc] = make4<R>(x, y, z, d.charge[i]);
    }
    if (force) continue;
    const R4T<R> r = d.ref[i];
    // minimum image of r - r_ref as r_ref + (r - shift) - r_ref: the
    // shift also undoes a wrap across the box edge since the rebuild
    const R ex = x - r.x, ey = y - r.y, ez = z - r.z;
    const R cx = fabs(ex) > b.half[0] ? copysign(b.L[0], ex) : R(0);
    const R cy = fabs(ey) > b.half[1] ? copysign(b.L[1], ey) : R(0);
    const R cz = fabs(ez) > b.half[2] ? copysign(b.L[2], ez) : R(0);
    const R dx = ex - cx, dy = ey - cy, dz = ez - cz;
    const float d2 = (float)(dx * dx + dy * dy + dz * dz);
    mx = fmaxf(mx, d2 != d2 ? __int_as_float(0x7f800000) : d2);
    const int s = d.inv_perm[i], c = d.c_of[i];
    // Sorted tiles keep the frame of the last rebuild: an atom that has
    // wrapped across the box edge since stays beside its cell (x - shift),
    // so that stage coordinates remain continuous while lists are reused.
    // .w: the type, plus one half for a charged atom (truncation gives the
    // type back; the half marks the atom for the list builds)
    const R xu = x - cx, yu = y - cy, zu = z - cz;
    d.xs[s] = make4<R>(xu, yu, zu, (R)d.type[i] + (c >= 0 ? R(0.5) : R(0)));
    if (c >= 0) d.xqc[c] = make4<R>(xu, yu, zu, d.charge[i]);
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = s_max[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, s_max[w]);
    d.skin_part[blockIdx.x] = force ? __int_as_float(0x7f800000) : m;
  }
}

// Cell key per atom (caller order): key = 2 * cell + neutral, so that inside
// every cell the charged atoms come first.  Arrival slots from atomics; the
// in-cell order is made deterministic by k_cellsort.
// Block `bid` of `nblk` binning blocks.
template <typename R>
__device__ __forceinline__ void bin_block(const Dev<R>& d, const Box<R>& b,
                                          const Grid& g, int split, unsigned thr,
                                          int nskin, int bid, int nblk) {
  // rebuild decision, part 2: every block folds the skin-check maxima
  __shared__ float s_max[4];
  float m = 0.f;
  for (int k = threadIdx.x; k < nskin; k += blockDim.x)
    m = fmaxf(m, d.skin_part[k]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  m = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
  if (bid == 0 && threadIdx.x == 0)
    d.ctrl->maxdisp_bits = __float_as_uint(m);
  if (!(__float_as_uint(m) > thr)) return;
  for (int i = bid * blockDim.x + threadIdx.x; i < d.n; i += nblk * blockDim.x) {
    const R x = wrap<R>(d.px[i], b.lo[0], b.hi[0], b.L[0], b.invL[0]);
    const R y = wrap<R>(d.py[i], b.lo[1], b.hi[1], b.L[1], b.invL[1]);
    const R z = wrap<R>(d.pz[i], b.lo[2], b.hi[2], b.L[2], b.invL[2]);
    const int cx = cell_coord((float)(x - b.lo[0]), g.scale[0], g.nc[0]);
    const int cy = cell_coord((float)(y - b.lo[1]), g.scale[1], g.nc[1]);
    const int cz = cell_coord((float)(z - b.lo[2]), g.scale[2], g.nc[2]);
    const int cell = g.cell_of_coord[(cz * g.nc[1] + cy) * g.nc[0] + cx];
    const int neutral = (!split || d.charge[i] == R(0)) ? 1 : 0;
    const int key = 2 * cell + neutral;
    d.key[i] = key;
    d.slot[i] = atomicAdd(&d.key_count[key], 1);
  }
}

// Exclusive scans in one block of 1024 threads: key_start over all 2 ncell
// keys and cstart over the charged key (2 cell) of every cell.  The cells go
// in tiles of kScanPer * blockDim: the tile's counts (one int2 per cell) are
// loaded coalesced into LDS, each thread scans its kScanPer consecutive cells
// after one block scan of the thread totals (plus the earlier tiles' carry),
// the starts go back through LDS and out coalesced, and key_count is cleared
// for the next rebuild.  buf: LDS of kScanPer * blockDim * 12 bytes.
constexpr int kScanPer = 4;
__device__ __forceinline__ void scan_block(int* key_count, int* key_start,
                                           int* cstart, int ncell, Ctrl* ctrl,
                                           unsigned thr, char* buf) {
  if (!rebuild_now(ctrl, thr)) return;
  __shared__ int smem[33];
  const int t = threadIdx.x, nt = blockDim.x, tile = nt * kScanPer;
  int2* s_io = reinterpret_cast<int2*>(buf);                 // counts in, starts out
  int* s_cs = reinterpret_cast<int*>(buf + sizeof(int2) * tile);  // cstart out
  if (t == 0) ctrl->rebuild_count += 1;
  int2* kc = reinterpret_cast<int2*>(key_count);
  int2* ks2 = reinterpret_cast<int2*>(key_start);
  int carry_a = 0, carry_q = 0;
  for (int base = 0; base < ncell; base += tile) {
    const int n = min(tile, ncell - base);
    for (int k = t; k < tile; k += nt) s_io[k] = k < n ? kc[base + k] : make_int2(0, 0);
    __syncthreads();
    int2 v[kScanPer];
    int sa = 0, sq = 0;
#pragma unroll
    for (int u = 0; u < kScanPer; ++u) {
      v[u] = s_io[kScanPer * t + u];
      sa += v[u].x + v[u].y;
      sq += v[u].x;
    }
    int tot_a, tot_q;
    int ra = block_excl_scan(sa, tot_a, smem) + carry_a;
    int rq = block_excl_scan(sq, tot_q, smem) + carry_q;
#pragma unroll
    for (int u = 0; u < kScanPer; ++u) {
      s_io[kScanPer * t + u] = make_int2(ra, ra + v[u].x);
      ra += v[u].x + v[u].y;
      s_cs[kScanPer * t + u] = rq;
      rq += v[u].x;
    }
    __syncthreads();
    for (int k = t; k < n; k += nt) {
      ks2[base + k] = s_io[k];
      cstart[base + k] = s_cs[k];
      kc[base + k] = make_int2(0, 0);
    }
    carry_a += tot_a;
    carry_q += tot_q;
    __syncthreads();  // the tile's LDS is read before the next tile's loads
  }
  if (t == 0) {
    key_start[2 * ncell] = carry_a;
    cstart[ncell] = carry_q;
  }
}

template <typename R>
__device__ __forceinline__ void scatter_block(const Dev<R>& d, unsigned thr, int bid,
                                              int nblk) {
  if (!rebuild_now(d.ctrl, thr)) return;
  for (int i = bid * blockDim.x + threadIdx.x; i < d.n; i += nblk * blockDim.x)
    d.perm[d.key_start[d.key[i]] + d.slot[i]] = i;
}

// One wave per cell: both key segments are sorted by caller index (bitonic
// across the wave up to 64 atoms, a rank sort up to 512, serial by lane 0
// beyond), then the cell's atoms are gathered into the sorted tiles:
// positions (and the charged copy), reference positions for the skin check,
// sorted ids, RBL hashes and cells, the charged map, inv_perm, and the CVFF
// special-bond rows (ms per atom).  A cell of at most 64 atoms is gathered in
// one pass with the caller indices taken from the sort's registers.
template <typename R>
__global__ __launch_bounds__(256) void k_cellsort(Dev<R> d, Box<R> b,
                                                  int ncell, unsigned thr,
                                                  int ms) {
  if (!rebuild_now(d.ctrl, thr)) return;
  int* perm = d.perm;
  const int* key_start = d.key_start;
  const int lane = threadIdx.x & 63;
  const int cell = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (cell >= ncell) return;
  // the key starts (charged segment, neutral segment, end) and, for segments
  // of at most 64 atoms, the sorted caller indices in lane order
  int bnd[3];
  int sv[2] = {0x7fffffff, 0x7fffffff};
  bool small = true;
  for (int seg = 0; seg < 2; ++seg) {
    const int a = key_start[2 * cell + seg], e = key_start[2 * cell + seg + 1];
    bnd[seg] = a;
    bnd[seg + 1] = e;
    const int len = e - a;
    if (len <= 64) {
      int v = lane < len ? perm[a + lane] : 0x7fffffff;
      if (len > 1) {
        for (int k = 2; k <= 64; k <<= 1)
          for (int j = k >> 1; j > 0; j >>= 1) {
            const int p = __shfl_xor(v, j, 64);
            const bool up = (lane & k) == 0;
            const bool lower = (lane & j) == 0;
            v = (lower == up) ? min(v, p) : max(v, p);
          }
        if (lane < len) perm[a + lane] = v;
      }
      sv[seg] = v;
    } else if (len <= 512) {
      small = false;
      // rank sort by the whole wave (the values are distinct caller
      // indices): every lane ranks its up to 8 values against the segment,
      // whose values come from the lanes' registers (no load per value),
      // and all ranks are known before the first store
      int v[8], r[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int p = a + lane + 64 * u;
        v[u] = p < e ? perm[p] : 0x7fffffff;
        r[u] = 0;
      }
      auto rank = [&](auto nu) {
        constexpr int NU = decltype(nu)::value;  // 64-value groups to rank
#pragma unroll
        for (int u2 = 0; u2 < NU; ++u2) {
          const int nl = min(64, len - 64 * u2);
          for (int l = 0; l < nl; ++l) {
            const int w = __builtin_amdgcn_readlane(v[u2], l);
#pragma unroll
            for (int u = 0; u < NU; ++u) r[u] += w < v[u] ? 1 : 0;
          }
        }
      };
      if (len <= 256)
        rank(std::integral_constant<int, 4>());
      else
        rank(std::integral_constant<int, 8>());
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (a + lane + 64 * u < e) perm[a + r[u]] = v[u];
    } else {
      small = false;
      if (lane == 0) {
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
  }
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
  // key 2 cell (charged when the work is split) precedes key 2 cell + 1
  const int a = bnd[0], aq = bnd[1], e = bnd[2];
  const int cq = d.cstart[cell];
  auto gather = [&](int s, int i) {
    d.inv_perm[i] = s;
    const R x = wrap<R>(d.px[i], b.lo[0], b.hi[0], b.L[0], b.invL[0]);
    const R y = wrap<R>(d.py[i], b.lo[1], b.hi[1], b.L[1], b.invL[1]);
    const R z = wrap<R>(d.pz[i], b.lo[2], b.hi[2], b.L[2], b.invL[2]);
    const int t = d.type[i];
    const int gid = d.gid[i];
    const unsigned long long hj = rbl_hj(gid);
    d.ref[i] = make4<R>(x, y, z, R(0));
    d.gid_s[s] = gid;
    d.hj_s[s] = hj;
    d.cell_s[s] = cell;
    int c = -1;
    if (s < aq) {
      c = cq + (s - a);
      d.c2s[c] = s;
      d.c2i[c] = i;
      d.gidc[c] = gid;
      d.hjc[c] = hj;
      d.typec[c] = t;
      d.xqc[c] = make4<R>(x, y, z, d.charge[i]);
    }
    d.s2c[s] = c;
    d.c_of[i] = c;
    if (d.spn) {
      d.spn_s[s] = d.spn[i];
      for (int q = 0; q < ms; ++q)
        d.spid_s[(size_t)s * ms + q] = d.spid[(size_t)i * ms + q];
    }
    // .w: the type, plus one half for a charged atom (as in k_skin)
    d.xs[s] = make4<R>(x, y, z, (R)t + (c >= 0 ? R(0.5) : R(0)));
  };
  if (small && e - a <= 64) {
    // one pass, the caller indices from the sort's registers: lane k takes
    // slot a + k, the charged segment's k-th atom or the neutral one's
    // (k - len0)-th
    const int len0 = aq - a;
    const int i1 = __shfl(sv[1], max(lane - len0, 0), 64);
    if (lane < e - a) gather(a + lane, lane < len0 ? sv[0] : i1);
  } else {
    for (int s = a + lane; s < e; s += 64) gather(s, perm[s]);
  }
}

// Charged flag of a sorted position tile (see k_skin).
template <typename R>
__device__ __forceinline__ bool charged_w(R w) {
  return w != (R)(int)w;
}

// All-atom candidate list of one brick: every j != i within r2 among the 27
// cells around i, except charged-charged pairs when the work is split (those
// go to list B).  Entries are 16-bit positions in the brick's staged
// neighbourhood in the interleaved layout of list_a_index; each lane stores
// its atom's entries as it accepts them (16-bit stores).  count_only
// (cap 0): no writes.  Neighbourhood positions come from the LDS stage when
// STAGED, else from the sorted arrays.
template <typename R, bool STAGED, bool MI>
__device__ __forceinline__ void list_a_atoms(const Dev<R>& d, const Box<R>& b,
                                             const Grid& g, const Neigh& h,
                                             const SlotTables<kMaxSlots>& t,
                                             const R4T<R>* stage, R r2,
                                             int split, int lcap, int cap,
                                             int& mx, int& sum) {
  const int i0 = d.key_start[2 * h.cell0], i1 = d.key_start[2 * h.cell1];
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    int c[3];
    unpack_coord(g.coord_of_cell[d.cell_s[i]], c);
    const Stencil st = atom_stencil(g, h, c);
    const int os = own_slot(h, c);
    const int self = t.off[os] + (i - t.first[os]);
    const bool qi = split && (i - t.first[os]) < t.nq[os];
    const R qf = qi ? R(1) : R(0);
    const R4T<R> xi = d.xs[i];
    int count = 0;
    // entry e of atom i: 16-bit slot e & 7 of group e >> 3 (groups of all
    // atoms interleaved, n8 entries apart), stored as it is accepted
    uint16_t* row = d.list_a + (size_t)i * 8;
    const size_t n8 = (size_t)d.n * 8;
    constexpr int sh = StageShift<R>::value;
    size_t wo = 0;  // (count >> 3) * n8 + (count & 7), kept as entries are added
    auto accept = [&](int l) {
      if (count < lcap) row[wo] = (uint16_t)(l << sh);
      ++count;
      wo += (count & 7) ? 1 : n8 - 7;
    };
    // Staged neighbourhoods scan a stencil row (its x cells are consecutive
    // slots, hence one contiguous stage range) as one run, 4 candidates at a
    // time; the charged atoms of each cell are skipped for a charged i.
    const bool rows = STAGED && st.mode[0] != 1;
    const int runlen = rows ? st.n[0] : 1;
    for (int kz = 0; kz < st.n[2]; ++kz)
      for (int ky = 0; ky < st.n[1]; ++ky) {
        const int srow = (st.at(2, kz) * h.cnt[1] + st.at(1, ky)) * h.cnt[0];
        for (int kx = 0; kx < st.n[0]; kx += runlen) {
          const int s0 = srow + st.at(0, kx);
          const int l0 = t.off[s0], l1 = t.off[s0 + runlen];
          if (STAGED) {
            // chunks of up to 64 candidates: first the accept mask (no
            // branches; candidate cb + 4k + u sets bit 4k + u), then the
            // accepted entries in candidate order
            for (int cb = l0; cb < l1; cb += 64) {
              const int ce = min(l1, cb + 64), nk = ce - cb;
              unsigned long long ok = 0;
#pragma unroll
              for (int k = 0; k < 16; ++k) {
                const int l = cb + 4 * k;
                if (l >= ce) continue;
                R4T<R> xj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) xj[u] = stage[min(l + u, ce - 1)];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  R dx, dy, dz;
                  const R d2 = pair_d2<R, MI>(b, xi.x, xi.y, xi.z, xj[u].x, xj[u].y,
                                              xj[u].z, dx, dy, dz);
                  // charged-charged pairs belong to list B: the stage's .w
                  // (1e30 for a charged j) pushes them past the cutoff
                  ok |= (unsigned long long)(fma_r(qf, xj[u].w, d2) < r2) << (4 * k + u);
                }
              }
              // the bits past the chunk's end (their loads repeat its last
              // candidate) and i itself are not entries
              if (nk < 64) ok &= (1ull << nk) - 1ull;
              const int so = self - cb;
              if (so >= 0 && so < nk) ok &= ~(1ull << so);
              while (ok) {
                const int u = __builtin_ctzll(ok);
                ok &= ok - 1;
                accept(cb + u);
              }
            }
          } else {
            const int jf = t.first[s0];
            const int ls = qi ? l0 + t.nq[s0] : l0;  // charged i skips charged j
            for (int l = ls; l < l1; ++l) {
              const R4T<R> xj = d.xs[jf + (l - l0)];
              R dx, dy, dz;
              const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z,
                                      dx, dy, dz);
              if (d2 < r2 && l != self) accept(l);
            }
          }
        }
      }
    if ((count & 7) && count <= lcap) {
      // the rest of the last group: the pair kernel's sentinel entry
      const uint16_t sent = (uint16_t)((StageCap<R>::value - 1) << sh);
      uint16_t* gp = row + (size_t)(count >> 3) * n8;
      for (int s8 = count & 7; s8 < 8; ++s8) gp[s8] = sent;
    }
    d.cnt_a[i] = (lcap == 0 && cap > 0) ? cap + 1 + count : count;
    mx = max(mx, count);
    sum += count;
  }
}

// List A of brick `brick`, one block.
template <typename R>
__device__ __forceinline__ void list_a_brick(const Dev<R>& d, const Box<R>& b,
                                             const Grid& g, int brick, R r2,
                                             int split, int cap) {
  __shared__ R4T<R> stage[StageCap<R>::value];
  __shared__ SlotTables<kMaxSlots> t;
  const int e1[3] = {1, 1, 1};
  const Neigh h = brick_neigh(g, brick, e1);
  build_slots(g, h, d.key_start, nullptr, false, t);
  // the pair kernel reads this brick's slot table until the next rebuild
  {
    int* bt = d.brick_tab + (size_t)brick * kBrickTabInts;
    for (int k = threadIdx.x; k < kMaxSlots + 1; k += blockDim.x) bt[k] = t.off[k];
    for (int k = threadIdx.x; k < kMaxSlots; k += blockDim.x) {
      bt[kMaxSlots + 1 + k] = t.first[k];
      bt[2 * kMaxSlots + 1 + k] = t.nq[k];
      bt[3 * kMaxSlots + 1 + k] = t.edge[k];
    }
  }
  const int total = t.off[h.nslots];
  // (the pair kernel keeps the last stage entry for its sentinel)
  const bool staged = total < StageCap<R>::value;
  if (staged) stage_atoms<R, false>(g, b, h, t, d.xs, stage);
  // neighbourhoods beyond 16-bit byte offsets are left to the direct scan
  const int lcap = total > (65536 >> StageShift<R>::value) ? 0 : cap;
  int mx = 0, sum = 0;
  if (staged && !h.mi)
    list_a_atoms<R, true, false>(d, b, g, h, t, stage, r2, split, lcap, cap, mx, sum);
  else if (staged)
    list_a_atoms<R, true, true>(d, b, g, h, t, stage, r2, split, lcap, cap, mx, sum);
  else
    list_a_atoms<R, false, true>(d, b, g, h, t, stage, r2, split, lcap, cap, mx, sum);
  // per-block statistics (folded by k_reduce)
  __shared__ int s_red[2 * 8];
  mx = wave_max_i(mx);
  sum = wave_sum(sum);
  const int nw = (int)(blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_red[2 * (threadIdx.x >> 6)] = mx;
    s_red[2 * (threadIdx.x >> 6) + 1] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0, su = 0;
    for (int w = 0; w < nw; ++w) {
      m = max(m, s_red[2 * w]);
      su += s_red[2 * w + 1];
    }
    d.lista_part[brick] = make_int4(m, su, total, staged ? 0 : 1);
  }
  // The pair kernel's order of the brick's atoms: by list length, longest
  // first (ties by position), so that each of its waves takes lists of about
  // equal length and no wave idles most of its lanes behind one long list.
  const int i0 = d.key_start[2 * h.cell0], na = d.key_start[2 * h.cell1] - i0;
  // the keys reuse the stage's LDS (the build is done with it: the block
  // statistics above passed a barrier)
  unsigned* s_key = reinterpret_cast<unsigned*>(stage);
  static_assert(sizeof(stage) >= kOrderMax * sizeof(unsigned), "order keys");
  if (na <= kOrderMax) {
    __threadfence_block();  // this block's counts are read back below
    __syncthreads();
    // keys (longer list first, then position) in a bitonic sort over the next
    // power of two, padded with the largest key
    int np = 1;
    while (np < na) np <<= 1;
    for (int k = threadIdx.x; k < np; k += blockDim.x)
      s_key[k] = k < na ? ((unsigned)(0xffff - min(d.cnt_a[i0 + k], 0xffff)) << 16) |
                              (unsigned)k
                        : 0xffffffffu;
    __syncthreads();
    for (int kk = 2; kk <= np; kk <<= 1)
      for (int j = kk >> 1; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < np; t += blockDim.x) {
          const int u = t ^ j;
          if (u > t) {
            const unsigned x = s_key[t], y = s_key[u];
            if ((x > y) == ((t & kk) == 0)) {
              s_key[t] = y;
              s_key[u] = x;
            }
          }
        }
        __syncthreads();
      }
    for (int k = threadIdx.x; k < na; k += blockDim.x)
      d.order_a[i0 + k] = i0 + (int)(s_key[k] & 0xffffu);
  } else {
    for (int k = threadIdx.x; k < na; k += blockDim.x) d.order_a[i0 + k] = i0 + k;
  }
}

// Charged-pair candidate list: kChargedLanes lanes per charged atom scan the
// charged atoms of the cells within ext_b around it, read from the compact
// charged array (20 B per charged atom, cache-resident), count, prefix-sum
// over the group and write in a fixed order (lane-major over the lane's
// cells).  Row-major list_b[c * cap + k] of charged compact indices.
// Block `blk` of list B: blockDim / kChargedLanes charged atoms.
template <typename R>
__device__ __forceinline__ void list_b_block(const Dev<R>& d, const Box<R>& b,
                                             const Grid& g, int blk, R r2,
                                             int cap) {
  const int grp = threadIdx.x / kChargedLanes, gl = threadIdx.x % kChargedLanes;
  const int c = blk * (blockDim.x / kChargedLanes) + grp;
  int tot = 0;
  if (c < d.nc) {
    const int i = d.c2s[c];
    const R4T<R> xi = d.xqc[c];
    int cc[3];
    unpack_coord(g.coord_of_cell[d.cell_s[i]], cc);
    const Span sx = make_span(cc[0], g.ext_b[0], g.nc[0]);
    const Span sy = make_span(cc[1], g.ext_b[1], g.nc[1]);
    const Span sz = make_span(cc[2], g.ext_b[2], g.nc[2]);
    const int nst = sx.count * sy.count * sz.count;
    int count = 0, pos = 0;
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 1) {
        int inc = count;
        for (int o = 1; o < kChargedLanes; o <<= 1) {
          const int u = __shfl_up(inc, o, kChargedLanes);
          if (gl >= o) inc += u;
        }
        pos = inc - count;
      }
      for (int s = gl; s < nst; s += kChargedLanes) {
        const int kx = s % sx.count, ky = (s / sx.count) % sy.count,
                  kz = s / (sx.count * sy.count);
        const int cell =
            g.cell_of_coord[(sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0] + sx.at(kx)];
        const int j0 = d.cstart[cell], j1 = d.cstart[cell + 1];
        for (int j = j0; j < j1; ++j) {
          const R4T<R> xj = d.xqc[j];
          R dx, dy, dz;
          const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
          if (d2 < r2 && j != c) {
            if (pass == 0) {
              ++count;
            } else {
              if (pos < cap) d.list_b[(size_t)c * cap + pos] = j;
              ++pos;
            }
          }
        }
      }
    }
    tot = count;
    for (int o = kChargedLanes / 2; o > 0; o >>= 1)
      tot += __shfl_xor(tot, o, kChargedLanes);
    if (gl == 0) d.cnt_b[c] = tot;
  }
  // per-block statistics (folded by k_reduce)
  __shared__ int s_red[2 * 8];
  int mx = wave_max_i(tot);
  int sum = wave_sum((gl == 0) ? tot : 0);
  if ((threadIdx.x & 63) == 0) {
    s_red[2 * (threadIdx.x >> 6)] = mx;
    s_red[2 * (threadIdx.x >> 6) + 1] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      mx = max(mx, s_red[2 * w]);
      sum += s_red[2 * w + 1];
    }
    d.listb_part[blk] = make_int2(mx, sum);
  }
}

// Charged-pair candidate list of a mostly charged system.  A cell's charged
// atoms are taken 64 at a time by blocks of 256 threads, four lanes per atom
// (g.dsplit blocks per cell, the cell's groups of 64 atoms dealt over them).
// The charged atoms of the stencil cells are staged in LDS, 256 at a time,
// double-buffered (the next chunk's loads are in flight while the current
// one is scanned) and padded to whole groups of 64 with far-away sentinels.
// In each group of 64 staged atoms, lane q of an atom tests the staged atoms
// 4v + q (sixteen; a wave reads four consecutive entries at a time) and the
// atom's four lanes OR their bits into its 64-candidate acceptance mask.
// The accepted atoms are appended in order (stencil-cell order, ascending
// within a cell) with their types in list_t: by the wave, one atom at a time
// with coalesced stores, when some atom of the wave accepted more than
// kDenseCoop of the 64 (near cells), else by each lane for its own
// candidates at their ranks in the mask.  Across a periodic edge the pair
// vector takes the stencil cell's shift by L, which is exactly the
// minimum-image shift of each of its pairs when the dimension has at least 4
// cells (both are one subtraction of L from r_j - r_i; a stencil cell
// without a shift subtracts nothing); with fewer the minimum image is
// applied per pair.
constexpr int kDenseCoop = 40;
constexpr int kDenseBlock = 256;

// Bits 4v (v < NV) of the acceptance mask of one lane: candidate st[4v].
template <typename R, bool MI, bool SHIFT, int NV>
__device__ __forceinline__ unsigned long long dense_mask(const Box<R>& b,
                                                         const R4T<R>& xi,
                                                         const R4T<R>* st, R r2,
                                                         R sx, R sy, R sz) {
  unsigned long long m = 0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const R4T<R> xj = st[4 * v];
    R dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
    if (MI) {
      dx = min_image<R>(dx, b.L[0], b.half[0]);
      dy = min_image<R>(dy, b.L[1], b.half[1]);
      dz = min_image<R>(dz, b.L[2], b.half[2]);
    } else if (SHIFT) {
      dx -= sx;
      dy -= sy;
      dz -= sz;
    }
    const R d2 = fma_r(dz, dz, fma_r(dy, dy, mul_rn(dx, dx)));
    m |= (d2 < r2 ? 1ull : 0ull) << (4 * v);
  }
  return m;
}

// OR over the four lanes of each quad (DPP quad permutations [1,0,3,2] and
// [2,3,0,1]).
__device__ __forceinline__ unsigned quad_or(unsigned x) {
  x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xf, 0xf, false);
  x |= (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xf, 0xf, false);
  return x;
}

// The stencil cell `st` of a dense list-B block: its charged range, its
// periodic shift (0 or +-L per dimension) and whether its pairs need the
// minimum image per pair (a dimension of fewer than 4 cells).
template <typename R>
struct DenseCell {
  int j0, j1;
  bool mi;
  R s[3];
};
template <typename R>
__device__ __forceinline__ DenseCell<R> dense_cell(const Grid& g, const Box<R>& b,
                                                   const int* cstart, const Span* sp,
                                                   int st) {
  const int k3[3] = {st % sp[0].count, (st / sp[0].count) % sp[1].count,
                     st / (sp[0].count * sp[1].count)};
  DenseCell<R> dc;
  dc.mi = false;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int u = sp[q].start + k3[q];
    dc.mi = dc.mi || sp[q].n < 4 || sp[q].count == sp[q].n;
    // a cell below 0 is the top cell of the box: its atoms sit about L above
    // the atom's image (r_j - r_i - L), and one above n - 1 about L below
    dc.s[q] = u < 0 ? b.L[q] : (u >= sp[q].n ? -b.L[q] : R(0));
  }
  const int cl = g.cell_of_coord[(sp[2].at(k3[2]) * g.nc[1] + sp[1].at(k3[1])) * g.nc[0] +
                                 sp[0].at(k3[0])];
  dc.j0 = cstart[cl];
  dc.j1 = cstart[cl + 1];
  return dc;
}

template <typename R>
__device__ __forceinline__ void list_b_dense(const Dev<R>& d, const Box<R>& b,
                                             const Grid& g, int cell, int part, int blk,
                                             R r2, int cap) {
  __shared__ R4T<R> st_x[2][kDenseBlock];
  __shared__ unsigned char st_t[2][kDenseBlock];
  const int cs = d.cstart[cell], ce = d.cstart[cell + 1];
  int cc[3];
  unpack_coord(g.coord_of_cell[cell], cc);
  const Span sp[3] = {make_span(cc[0], g.ext_b[0], g.nc[0]),
                      make_span(cc[1], g.ext_b[1], g.nc[1]),
                      make_span(cc[2], g.ext_b[2], g.nc[2])};
  const int nst = sp[0].count * sp[1].count * sp[2].count;
  const R far = sizeof(R) == 4 ? R(1e30) : R(1e200);
  const int t = (int)threadIdx.x, lane = t & 63, q = t & 3;
  int mx = 0, sum = 0;
  for (int i0 = cs + 64 * part; i0 < ce; i0 += 64 * g.dsplit) {
    const int c = i0 + (t >> 2);
    const bool act = c < ce;
    const R4T<R> xi = d.xqc[act ? c : cs];
    int* row = d.list_b + (size_t)(act ? c : cs) * cap;
    unsigned char* trow = d.list_t + (size_t)(act ? c : cs) * cap;
    int pos = 0;
    // the staged chunks in order: (stencil cell, first atom); (st, jb) is
    // the chunk whose loads are in flight
    int st = 0, jb = 0;
    DenseCell<R> dc;
    auto seek = [&]() {
      // from (st, jb) to the first chunk with atoms, st == nst at the end
      while (st < nst) {
        dc = dense_cell<R>(g, b, d.cstart, sp, st);
        if (jb == 0) jb = dc.j0;
        if (jb < dc.j1) return;
        ++st;
        jb = 0;
      }
    };
    seek();
    R4T<R> px = make4<R>(far, far, far, R(0));
    int pt = 0;
    if (st < nst && t < dc.j1 - jb) {
      px = d.xqc[jb + t];
      pt = d.typec[jb + t];
    }
    int buf = 0;
    while (st < nst) {
      const int nj = min(dc.j1 - jb, kDenseBlock);
      const int npad = (nj + 63) & ~63;
      const int cjb = jb;
      const DenseCell<R> cdc = dc;
      // buffer `buf` was last read two chunks ago, before the barrier of the
      // previous chunk
      if (t < npad) {
        st_x[buf][t] = t < nj ? px : make4<R>(far, far, far, R(0));
        st_t[buf][t] = (unsigned char)(t < nj ? pt : 0);
      }
      __syncthreads();
      // the next chunk's loads, in flight during the scan
      if (jb + kDenseBlock < dc.j1) {
        jb += kDenseBlock;
      } else {
        ++st;
        jb = 0;
        seek();
      }
      if (st < nst && t < dc.j1 - jb) {
        px = d.xqc[jb + t];
        pt = d.typec[jb + t];
      } else {
        px = make4<R>(far, far, far, R(0));
      }
      const bool shift = cdc.s[0] != R(0) || cdc.s[1] != R(0) || cdc.s[2] != R(0);
      for (int k0 = 0; k0 < npad; k0 += 64) {
        const int tyl = st_t[buf][k0 + lane];  // the type of staged atom k0 + lane
        unsigned long long m = 0;
        if (act) {
          const R4T<R>* sq = st_x[buf] + k0 + q;
          auto test = [&](auto nv) {
            constexpr int NV = decltype(nv)::value;
            return cdc.mi ? dense_mask<R, true, false, NV>(b, xi, sq, r2, R(0), R(0), R(0))
                   : shift ? dense_mask<R, false, true, NV>(b, xi, sq, r2, cdc.s[0],
                                                            cdc.s[1], cdc.s[2])
                           : dense_mask<R, false, false, NV>(b, xi, sq, r2, R(0), R(0),
                                                             R(0));
          };
          // a last group of at most 32 atoms: the sentinels after them fail
          // every test, so only the first 32 are tested
          m = nj - k0 <= 32 ? test(std::integral_constant<int, 8>())
                            : test(std::integral_constant<int, 16>());
          m <<= q;
        }
        m = ((unsigned long long)quad_or((unsigned)(m >> 32)) << 32) |
            quad_or((unsigned)m);
        const int self = c - (cjb + k0);
        if (self >= 0 && self < 64) m &= ~(1ull << self);
        const int na = __popcll(m);
        if (__ballot(na > kDenseCoop) != 0ull) {
          // many accepted: the wave writes each atom's entries of the group
          // with one coalesced store, lane k taking staged atom k0 + k
          const int abase = __builtin_amdgcn_readfirstlane(i0 + ((t & ~63) >> 2));
          for (int l = 0; l < 16; ++l) {
            const unsigned lo = __builtin_amdgcn_readlane((unsigned)m, 4 * l);
            const unsigned hi = __builtin_amdgcn_readlane((unsigned)(m >> 32), 4 * l);
            if ((lo | hi) == 0u) continue;
            const unsigned long long ml = ((unsigned long long)hi << 32) | lo;
            if ((ml >> lane) & 1ull) {
              const unsigned r =
                  (unsigned)__builtin_amdgcn_readlane(pos, 4 * l) +
                  __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
              if (r < (unsigned)cap) {
                const size_t o = (size_t)(abase + l) * cap;
                d.list_b[o + r] = cjb + k0 + lane;
                d.list_t[o + r] = (unsigned char)tyl;
              }
            }
          }
        } else {
          // few: lane q of an atom appends its candidates 4v + q at their
          // ranks in the mask, with their staged types
          unsigned long long mq = m & (0x1111111111111111ull << q);
          while (mq) {
            const int u = __builtin_ctzll(mq);
            mq &= mq - 1;
            const int r = pos + __popcll(m & ((1ull << u) - 1ull));
            if (r < cap) {
              row[r] = cjb + k0 + u;
              trow[r] = st_t[buf][k0 + u];
            }
          }
        }
        pos += na;
      }
      buf ^= 1;
    }
    if (act && q == 0) {
      d.cnt_b[c] = pos;
      mx = max(mx, pos);
      sum += pos;
    }
  }
  // per-block statistics (folded by k_reduce)
  __shared__ int s_red[2 * (kDenseBlock / 64)];
  mx = wave_max_i(mx);
  sum = wave_sum(sum);
  if (lane == 0) {
    s_red[2 * (t >> 6)] = mx;
    s_red[2 * (t >> 6) + 1] = sum;
  }
  __syncthreads();
  if (t == 0) {
    for (int w = 1; w < kDenseBlock / 64; ++w) {
      mx = max(mx, s_red[2 * w]);
      sum += s_red[2 * w + 1];
    }
    d.listb_part[blk] = make_int2(mx, sum);
  }
}

// Both candidate lists in one launch (they are independent): blocks
// [0, nbrick) build list A of one brick each, the blocks after them list B,
// so the charged-pair build fills the machine beside the brick builds.
// In mostly charged systems list B has a kernel of its own (k_lists_dense,
// one block per cell) and this launch holds the list-A blocks alone.
template <typename R>
__global__ __launch_bounds__(512) void k_lists(Dev<R> d, Box<R> b, Grid g,
                                               R r2a, R r2b, int split,
                                               int cap_a, int cap_b,
                                               unsigned thr) {
  if (!rebuild_now(d.ctrl, thr)) return;
  if ((int)blockIdx.x < g.nbrick)
    list_a_brick<R>(d, b, g, xcd_brick(blockIdx.x, g.nbrick), r2a, split, cap_a);
  else
    list_b_block<R>(d, b, g, blockIdx.x - g.nbrick, r2b, cap_b);
}

template <typename R>
__global__ __launch_bounds__(kDenseBlock) void k_lists_dense(Dev<R> d, Box<R> b, Grid g,
                                                             R r2b, int cap_b, unsigned thr) {
  if (!rebuild_now(d.ctrl, thr)) return;
  // g.dsplit consecutive blocks per cell, each XCD a contiguous range of cells
  const int blk = xcd_brick(blockIdx.x, gridDim.x);
  list_b_dense<R>(d, b, g, blk / g.dsplit, blk % g.dsplit, blk, r2b, cap_b);
}

template <typename R>
__device__ __forceinline__ SpecRow spec_row(const Dev<R>& d, int ms, int s) {
  SpecRow r;
  const int4 c = d.spn_s[s];
  r.n12 = c.x;
  r.n13 = c.y;
  r.n14 = c.z;
  r.ids = d.spid_s + (size_t)s * ms;
  return r;
}

// LJ-only pair of list A.  hj() yields j's selection hash (evaluated only
// for RBL shell pairs); with SPECIAL, gj is j's global id and the LJ force
// carries i's special-bond weight for it.
template <typename R, bool RBL, bool EFLAG, bool SPECIAL, typename GidJ>
__device__ __forceinline__ void interact_a(const PairConst<R>& pc,
                                           const Box<R>& b, const R* s_lj,
                                           int ti, const R4T<R>& xi,
                                           unsigned long long hi,
                                           const R4T<R>& xj, GidJ hj,
                                           const SpecRow& sr, int gj,
                                           Acc<R>& a) {
  R dx, dy, dz;
  const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
  if (!(d2 < pc.rclj2) || !(d2 < pc.rc2)) return;
  const int tj = (int)xj.w;
  const int tt = 4 * (ti * pc.ntypes + tj);
  const R m24 = s_lj[tt], sig2 = s_lj[tt + 1], eps = s_lj[tt + 2];
  const R inv = R(1) / d2;
  const R s2 = sig2 * inv;
  const R s6 = s2 * s2 * s2;
  const R ws = SPECIAL ? special_w<R>(sr, gj, pc) : R(1);
  const R flj = ws * (m24 * (R(2) * s6 - R(1)) * s6 * inv);
  const R w = pair_weight<R, RBL>(pc, d2, hi, hj);
  const R f = w * flj;
  a.fx += f * dx;
  a.fy += f * dy;
  a.fz += f * dz;
  if (EFLAG) {
    a.e_lj += (double)(ws * (R(2) * eps * (s6 - R(1)) * s6));
    add_virial<R, EFLAG>(a, dx, dy, dz, flj);
  }
}

// LJ + Coulomb pair of list B (both atoms charged).  With SPECIAL (CVFF),
// gj is j's global id; a special-bond weight ws < 1 scales the LJ term and
// removes (1 - ws) of the full Coulomb force -qqr2e q_i q_j / r^3 from every
// pair the lists keep, without the RBL multiplier
// (cvff_op.hip.cu:572-599; energies :401-520).
template <typename R, bool RBL, bool EFLAG, bool SPECIAL = false>
__device__ __forceinline__ void interact_b(const PairConst<R>& pc,
                                           const Box<R>& b, const R* s_lj,
                                           int ti, const R4T<R>& xi,
                                           unsigned long long hi,
                                           const R4T<R>& xj, int tj,
                                           const unsigned long long* hjp,
                                           Acc<R>& a, const SpecRow& sr = SpecRow(),
                                           int gj = 0) {
  R dx, dy, dz;
  const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
  if (!(d2 < pc.rc2)) return;
  const R w = pair_weight<R, RBL>(pc, d2, hi, [&] { return *hjp; });
  const R ws = SPECIAL ? special_w<R>(sr, gj, pc) : R(1);
  eval_b<R, EFLAG, SPECIAL>(pc, s_lj, ti, tj, xi.w * xj.w, dx, dy, dz, d2, w, a, ws);
}

template <typename R>
struct R2Of {
  using type = float2;
};
template <>
struct R2Of<double> {
  using type = double2;
};

// All list-A entries of sorted atom i (LJ only).  Entries are byte offsets
// of stage entries, read 8 at a time (one 16-byte load per group, the next
// group prefetched).  STAGED: neighbours come from the LDS stage, and the
// last group's unused entries name the sentinel entry, which fails every
// cutoff test; otherwise from the sorted arrays, entries beyond cnt masked.
// The LJ table row of type pair (ti, tj) holds (-24 eps, sigma^2, eps, 0).
// Without list-A shell pairs and energies the force window is one interval
// (a_lo, a_hi) of d^2 (the cutoffs, r_core and the coincident-pair guard
// folded on the host).  With SPECIAL (CVFF), the LJ force carries i's
// special-bond weight for j (global ids staged beside the positions).
template <typename R, bool RBL, bool EFLAG, bool SHELL, bool STAGED, bool MI,
          bool SPECIAL>
__device__ __forceinline__ void list_a_atom(
    const Dev<R>& d, const Box<R>& b, const PairConst<R>& pc, const R* s_lj,
    const SlotTables<kMaxSlots>& t, int nslots, const R4T<R>* stage, int i,
    int cnt, const R4T<R>& xi, int ti, unsigned long long hi, Acc<R>& a,
    const SpecRow& sr, const int* sgid) {
  using R2 = typename R2Of<R>::type;
  using R4 = R4T<R>;
  constexpr int sh = StageShift<R>::value;
  const uint4* lp = reinterpret_cast<const uint4*>(d.list_a) + i;
  const char* lj_row =
      reinterpret_cast<const char*>(s_lj) + (size_t)ti * pc.ntypes * 4 * sizeof(R);
  const char* st = reinterpret_cast<const char*>(stage);
  auto global_of = [&](int l) {
    const int s = slot_of_pos(t, nslots, l);
    return t.first[s] + (l - t.off[s]);
  };
  // the next 16-byte group is loaded while the current one is evaluated
  uint4 q_next = cnt > 0 ? lp[0] : make_uint4(0u, 0u, 0u, 0u);
  for (int k0 = 0; k0 < cnt; k0 += 8) {
    const uint4 q8 = q_next;
    if (k0 + 8 < cnt) q_next = lp[(size_t)((k0 >> 3) + 1) * d.n];
    const int kn = cnt - k0;
    const unsigned wv[4] = {q8.x, q8.y, q8.z, q8.w};
#pragma unroll
    for (int half = 0; half < 8; half += 4) {
    int lj_[8];
    R4 xj[8];
#pragma unroll
    for (int kk = half; kk < half + 4; ++kk) {
      const int off = (int)((wv[kk >> 1] >> (16 * (kk & 1))) & 0xffffu);
      if (STAGED) {
        lj_[kk] = off >> sh;
        xj[kk] = *reinterpret_cast<const R4*>(st + off);
      } else {
        lj_[kk] = kk < kn ? global_of(off >> sh) : i;
        xj[kk] = d.xs[lj_[kk]];
      }
    }
#pragma unroll
    for (int kk = half; kk < half + 4; ++kk) {
      R dx, dy, dz;
      const R d2 = pair_d2<R, MI>(b, xi.x, xi.y, xi.z, xj[kk].x, xj[kk].y,
                                  xj[kk].z, dx, dy, dz);
      const int woff = STAGED ? pair_woff(xj[kk].w)
                              : ((int)xj[kk].w) * 4 * (int)sizeof(R);
      const R2 lj = *reinterpret_cast<const R2*>(lj_row + woff);
      const R inv = pair_rcp(d2);
      const R s2 = lj.y * inv;
      const R s6 = s2 * s2 * s2;
      R ws = R(1);
      if (SPECIAL) {
        const int gj = STAGED ? sgid[lj_[kk]] : d.gid_s[lj_[kk]];
        ws = special_w<R>(sr, gj, pc);
      }
      const R flj = ws * (lj.x * (R(2) * s6 - R(1)) * s6 * inv);
      if (!SHELL && !EFLAG) {
        const bool ok = (STAGED || kk < kn) && d2 < pc.a_hi && d2 > pc.a_lo;
        const R f = ok ? flj : R(0);
        a.fx = fma_r(f, dx, a.fx);
        a.fy = fma_r(f, dy, a.fy);
        a.fz = fma_r(f, dz, a.fz);
        continue;
      }
      const bool in = (STAGED || kk < kn) && d2 < pc.rclj2 && d2 < pc.rc2;
      R w = R(1);
      if (RBL) {
        const bool core = d2 < pc.t_core;
        w = (core && d2 < pc.rs2 && d2 > pc.eps_min) ? R(1) : R(0);
        if (SHELL && in && !core && d2 < pc.t_shell && d2 > pc.rs2) {
          const int jg = STAGED ? global_of(lj_[kk]) : lj_[kk];
          if (rbl_keep(pc.ifreq, pc.fmask, pc.c32, pc.epoch, hi, d.hj_s[jg]))
            w = pc.freq;
        }
      }
      const R f = in ? w * flj : R(0);
      a.fx = fma_r(f, dx, a.fx);
      a.fy = fma_r(f, dy, a.fy);
      a.fz = fma_r(f, dz, a.fz);
      if (EFLAG) {
        const R eps = *reinterpret_cast<const R*>(lj_row + woff + 2 * sizeof(R));
        a.e_lj += in ? (double)(ws * (R(2) * eps * (s6 - R(1)) * s6)) : 0.0;
        add_virial<R, EFLAG>(a, dx, dy, dz, in ? flj : R(0));
      }
    }
    }
  }
}

template <typename R>
__device__ __forceinline__ void load_lj(const R* tab, int ntypes, R* s) {
  for (int t = threadIdx.x; t < 4 * ntypes * ntypes; t += blockDim.x)
    s[t] = tab[t];
}

// Per-brick FP64 partials, component-major (part[k * nbrick + brick]) so
// that k_reduce reads each component with coalesced loads: force sum, then
// with energies E_lj, E_coul and the six virial components.
template <typename R, bool EFLAG>
__device__ __forceinline__ void write_partials(const Acc<R>& a, double* part,
                                               int brick, double* smem) {
  double v[11] = {(double)a.fx, (double)a.fy, (double)a.fz, a.e_lj, a.e_c,
                  a.v[0],       a.v[1],       a.v[2],       a.v[3], a.v[4],
                  a.v[5]};
  const size_t nb = gridDim.x;
  if (EFLAG) {
    block_sum<11>(v, smem);
    if (threadIdx.x == 0)
      for (int k = 0; k < 11; ++k) part[k * nb + brick] = v[k];
  } else {
    double u[3] = {v[0], v[1], v[2]};
    block_sum<3>(u, smem);
    if (threadIdx.x == 0)
      for (int k = 0; k < 3; ++k) part[k * nb + brick] = u[k];
  }
}

// Occupancy target of the pair kernel: 6 waves per SIMD for the FP32 force
// pass without energies or list-A shell pairs (the per-step hot case), so
// that every brick of a 200k-atom system is resident at once.
template <typename R, bool EFLAG, bool SHELL>
struct kPairWaves {
  static constexpr int value = (sizeof(R) == 4 && !EFLAG && !SHELL) ? 6 : 1;
};

// Pair forces of one brick.  Phase 1: LJ over list A for every atom of the
// brick (one lane per atom, neighbours from the LDS stage).  Phase 2: the
// brick's charged atoms add LJ + Coulomb over list B, kChargedLanes lanes per
// atom.  fa[i] receives the total pair force of atom i.
// SPECIAL: CVFF special-bond weights (partners looked up by global id).
template <typename R, bool RBL, bool EFLAG, bool SHELL, bool SPECIAL>
__global__ __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(kPairWaves<R, EFLAG, SHELL || SPECIAL>::value))) void k_pair(Dev<R> d, Box<R> b, Grid g,
                                              PairConst<R> pc_arg,
                                              const R* lj_tab, int split) {
  // the sampling epoch of this evaluation comes from the step ring (k_skin)
  PairConst<R> pc = pc_arg;
  if (RBL) pc.epoch = d.ctrl->epoch;
  __shared__ R4T<R> stage[StageCap<R>::value];
  __shared__ int sgid[SPECIAL ? StageCap<R>::value : 1];
  __shared__ SlotTables<kMaxSlots> t;
  __shared__ double smem[8 * 11];
  extern __shared__ char dyn_lds[];  // LJ table: 4 * ntypes^2 values of R
  R* s_lj = reinterpret_cast<R*>(dyn_lds);
  load_lj<R>(lj_tab, pc.ntypes, s_lj);
  const int e1[3] = {1, 1, 1};
  // workgroup b serves brick b % nbrick: the first of the brick's g.psub
  // workgroups stages it and runs list A; all of them share the brick's
  // charged atoms over list B
  const int sub = blockIdx.x / g.nbrick;
  const int brick = xcd_brick(blockIdx.x - sub * g.nbrick, g.nbrick);
  const Neigh h = brick_neigh(g, brick, e1);
  Acc<R> acc;
  // threads [0, TA) take list A; with a list-B wave (g.qwave), the threads
  // after them take the charged pairs from the start
  const int TA = (int)blockDim.x - g.qwave;
  if (sub == 0) {
    // the brick's slot table of the last rebuild (one coalesced read instead
    // of the cell lookups and the block scan)
    {
      const int* bt = d.brick_tab + (size_t)brick * kBrickTabInts;
      for (int k = threadIdx.x; k < kMaxSlots + 1; k += blockDim.x) t.off[k] = bt[k];
      for (int k = threadIdx.x; k < kMaxSlots; k += blockDim.x) {
        t.first[k] = bt[kMaxSlots + 1 + k];
        t.nq[k] = bt[2 * kMaxSlots + 1 + k];
        t.edge[k] = bt[3 * kMaxSlots + 1 + k];
      }
      __syncthreads();
    }
    const int total = t.off[h.nslots];
    // the last stage entry is the sentinel
    const bool staged = total < StageCap<R>::value;
    if (staged)
      stage_atoms<R, true>(g, b, h, t, d.xs, stage, SPECIAL ? d.gid_s : nullptr,
                           SPECIAL ? sgid : nullptr);
    const int i0 = d.key_start[2 * h.cell0], i1 = d.key_start[2 * h.cell1];
    // phase 1: LJ over list A, the atoms in the order of the last rebuild
    // (longest list first)
    for (int k = (int)threadIdx.x; k < i1 - i0 && (int)threadIdx.x < TA; k += TA) {
      const int i = d.order_a[i0 + k];
      Acc<R> a;
      const R4T<R> xi = d.xs[i];
      const int ti = (int)xi.w;
      const unsigned long long hi = RBL ? rbl_hi(d.gid_s[i]) : 0ull;
      const int cnt = d.cnt_a[i];
      const SpecRow sr = SPECIAL ? spec_row<R>(d, pc.ms, i) : SpecRow();
      if (cnt <= d.cap_a) {
        if (staged && !h.mi)
          list_a_atom<R, RBL, EFLAG, SHELL, true, false, SPECIAL>(
              d, b, pc, s_lj, t, h.nslots, stage, i, cnt, xi, ti, hi, a, sr, sgid);
        else if (staged)
          list_a_atom<R, RBL, EFLAG, SHELL, true, true, SPECIAL>(
              d, b, pc, s_lj, t, h.nslots, stage, i, cnt, xi, ti, hi, a, sr, sgid);
        else
          list_a_atom<R, RBL, EFLAG, SHELL, false, true, SPECIAL>(
              d, b, pc, s_lj, t, h.nslots, stage, i, cnt, xi, ti, hi, a, sr, sgid);
      } else {
        // list overflow since the last rebuild: scan the 27 cells directly
        int c[3];
        unpack_coord(g.coord_of_cell[d.cell_s[i]], c);
        const Stencil st = atom_stencil(g, h, c);
        const int os = own_slot(h, c);
        const bool qi = split && (i - t.first[os]) < t.nq[os];
        for (int kz = 0; kz < st.n[2]; ++kz)
          for (int ky = 0; ky < st.n[1]; ++ky)
            for (int kx = 0; kx < st.n[0]; ++kx) {
              const int sl = (st.at(2, kz) * h.cnt[1] + st.at(1, ky)) * h.cnt[0] +
                             st.at(0, kx);
              const int len = t.off[sl + 1] - t.off[sl];
              for (int k = qi ? t.nq[sl] : 0; k < len; ++k) {
                const int jg = t.first[sl] + k;
                if (jg == i) continue;
                interact_a<R, RBL, EFLAG, SPECIAL>(pc, b, s_lj, ti, xi, hi, d.xs[jg],
                                                   [&] { return d.hj_s[jg]; }, sr,
                                                   SPECIAL ? d.gid_s[jg] : 0, a);
              }
            }
      }
      d.fa[d.perm[i]] = make4<R>(a.fx, a.fy, a.fz, R(0));
      acc.fx += a.fx;
      acc.fy += a.fy;
      acc.fz += a.fz;
      if (EFLAG) {
        acc.e_lj += a.e_lj;
        for (int k = 0; k < 6; ++k) acc.v[k] += a.v[k];
      }
    }
  }
  // phase 2: charged atoms of the brick over list B, QL lanes per atom (8 in
  // the mostly neutral hot variant, so that a brick's charged atoms take one
  // pass; 16 where list B carries most of the work).  No LDS is read here but
  // the LJ table, so a wave starts as soon as its list-A atoms are done.
  const bool in_b = g.qwave ? (int)threadIdx.x >= TA : true;
  if (split && in_b && g.qblocks == 0) {
    constexpr int QL = (SHELL || SPECIAL) ? kChargedLanes : 8;
    // the list-B wave takes an equal share of all charged atoms (list B reads
    // no brick stage, and per-brick charged counts vary several-fold); without
    // one, the brick's own charged atoms, shared by its g.psub workgroups
    const int c0 = g.qwave ? (int)((long long)blockIdx.x * d.nc / gridDim.x)
                           : d.cstart[h.cell0];
    const int c1 = g.qwave ? (int)((long long)(blockIdx.x + 1) * d.nc / gridDim.x)
                           : d.cstart[h.cell1];
    // the list-B wave alone, or (without one) every thread after list A
    const int tb = g.qwave ? (int)threadIdx.x - TA : (int)threadIdx.x;
    const int grp = tb / QL, gl = tb % QL;
    const int ngrp = (g.qwave ? g.qwave : (int)blockDim.x) / QL;
    for (int c = c0 + sub * ngrp + grp; c < c1; c += ngrp * g.psub) {
      Acc<R> a;
      const R4T<R> xi = d.xqc[c];
      const int ti = d.typec[c];
      const unsigned long long hi = RBL ? rbl_hi(d.gidc[c]) : 0ull;
      const int cnt = d.cnt_b[c];
      const SpecRow sr = SPECIAL ? spec_row<R>(d, pc.ms, d.c2s[c]) : SpecRow();
      if (cnt <= d.cap_b) {
        const int* lp = d.list_b + (size_t)c * d.cap_b;
        // four entries per lane at a time: their indices, then their
        // gathers, are all in flight before any of them is evaluated
        // (indices past the count are clamped, so every load issues)
        for (int k0 = gl; k0 < cnt; k0 += 4 * QL) {
          int e[4], tj[4], gj[4];
          R4T<R> xj[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) e[u] = lp[min(k0 + u * QL, cnt - 1)];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            xj[u] = d.xqc[e[u]];
            tj[u] = d.typec[e[u]];
            gj[u] = SPECIAL ? d.gidc[e[u]] : 0;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (k0 + u * QL < cnt)
              interact_b<R, RBL, EFLAG, SPECIAL>(pc, b, s_lj, ti, xi, hi, xj[u], tj[u],
                                                 d.hjc + e[u], a, sr, gj[u]);
        }
      } else {
        const int i = d.c2s[c];
        int cc[3];
        unpack_coord(g.coord_of_cell[d.cell_s[i]], cc);
        const Span sx = make_span(cc[0], g.ext_b[0], g.nc[0]);
        const Span sy = make_span(cc[1], g.ext_b[1], g.nc[1]);
        const Span sz = make_span(cc[2], g.ext_b[2], g.nc[2]);
        const int nst = sx.count * sy.count * sz.count;
        for (int s = gl; s < nst; s += QL) {
          const int kx = s % sx.count, ky = (s / sx.count) % sy.count,
                    kz = s / (sx.count * sy.count);
          const int cell =
              g.cell_of_coord[(sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0] + sx.at(kx)];
          const int js = d.key_start[2 * cell], je = d.key_start[2 * cell + 1];
          for (int j = js; j < je; ++j) {
            if (j == i) continue;
            const int e = d.s2c[j];
            interact_b<R, RBL, EFLAG, SPECIAL>(pc, b, s_lj, ti, xi, hi, d.xqc[e],
                                               d.typec[e], d.hjc + e, a, sr,
                                               SPECIAL ? d.gidc[e] : 0);
          }
        }
      }
      // reduce over the group's lanes (fixed tree)
      R fx = a.fx, fy = a.fy, fz = a.fz;
      for (int o = QL / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o, QL);
        fy += __shfl_xor(fy, o, QL);
        fz += __shfl_xor(fz, o, QL);
      }
      if (gl == 0) d.fb[d.c2i[c]] = make4<R>(fx, fy, fz, R(0));
      acc.fx += a.fx;
      acc.fy += a.fy;
      acc.fz += a.fz;
      if (EFLAG) {
        acc.e_lj += a.e_lj;
        acc.e_c += a.e_c;
        for (int k = 0; k < 6; ++k) acc.v[k] += a.v[k];
      }
    }
  }
  write_partials<R, EFLAG>(acc, d.pair_part, blockIdx.x, smem);
}

// ---------------------------------------------------------------------------
// Charged pairs of mostly charged systems (water, molten salts): one wave per
// charged atom.  The wave reads the atom's list-B row and its type bytes 64
// entries per lane step, kQBatch steps' loads in flight (position and
// selection hash of every candidate), applies the distance test and the RBL
// class as selects, without branches, and packs the pairs that add a force
// -- with energies, every pair inside r_c -- into a per-wave LDS ring in list
// order: the pair vector and d^2, q_j, j with a weight code and the type of
// j.  Every 64 packed pairs then take LJ, the Coulomb term and the
// special-bond weight on a full wave without further gathers (but j's global
// id for special bonds), so with RBL at r_c + skin only the core and the
// selected shell (about a quarter of a row) reach that part.  Per-atom sums: each lane's terms in ring order, then
// a fixed tree over the wave; block partials of the thermo sums in `part`
// (k_reduce's second segment).  Each wave takes a contiguous run of atoms,
// and the runs of one XCD are contiguous, so an XCD's L2 holds one region's
// neighbours.
// ---------------------------------------------------------------------------
constexpr int kQBatch = 3;  // list steps of 64 entries per load batch
constexpr int kRing = 256;  // ring entries per wave (>= 63 + 64 * kQBatch)
static_assert(kRing >= 63 + 64 * kQBatch && (kRing & (kRing - 1)) == 0, "ring size");

template <typename R, bool RBL, bool EFLAG, bool SPECIAL>
__global__ __launch_bounds__(64 * kQWaves)
__attribute__((amdgpu_waves_per_eu(4))) void k_pair_b(Dev<R> d, Box<R> b, Grid g,
                                                      PairConst<R> pc_arg,
                                                      const R* lj_tab, double* part) {
  PairConst<R> pc = pc_arg;
  if (RBL) pc.epoch = d.ctrl->epoch;
  __shared__ R4T<R> r_d[kQWaves][kRing];         // (dx, dy, dz, d^2)
  __shared__ R r_q[kQWaves][kRing];              // q_j
  __shared__ int r_e[kQWaves][kRing];            // j | weight code << 30
  __shared__ unsigned char r_t[kQWaves][kRing];  // type of j
  __shared__ double smem[kQWaves * 11];
  extern __shared__ char dyn_lds[];  // LJ table: 4 * ntypes^2 values of R
  R* s_lj = reinterpret_cast<R*>(dyn_lds);
  load_lj<R>(lj_tab, pc.ntypes, s_lj);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  R4T<R>* rd = r_d[wv];
  R* rq = r_q[wv];
  int* re = r_e[wv];
  unsigned char* rt = r_t[wv];
  const int nwav = gridDim.x * kQWaves;
  const int wid = xcd_brick(blockIdx.x, gridDim.x) * kQWaves + wv;
  const int c0 = (int)((long long)wid * d.nc / nwav);
  const int c1 = (int)((long long)(wid + 1) * d.nc / nwav);
  Acc<R> acc;
  for (int c = c0; c < c1; ++c) {
    const R4T<R> xi = d.xqc[c];
    const int ti = d.typec[c];
    const unsigned long long hi = RBL ? rbl_hi(d.gidc[c]) : 0ull;
    const int cnt = d.cnt_b[c];
    const SpecRow sr = SPECIAL ? spec_row<R>(d, pc.ms, d.c2s[c]) : SpecRow();
    // rows of up to kSpecReg partners (water: 2) from registers
    const bool sreg = SPECIAL && sr.n14 <= kSpecReg;
    const SpecRowReg srr = sreg ? spec_row_reg(sr) : SpecRowReg();
    Acc<R> a;
    unsigned head = 0, tail = 0;  // ring positions (wave-uniform)
    // one candidate per lane, appended in lane order when kept; code 1:
    // weight 1, 2: the RBL shell multiplier, 0: weight 0 (energies only)
    auto offer = [&](bool keep, int e, int tj, R dx, R dy, R dz, R d2, R w, R qj) {
      const unsigned long long m = __ballot(keep);
      if (keep) {
        const unsigned r = __builtin_amdgcn_mbcnt_hi(
            (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        const int s = (int)((tail + r) & (kRing - 1));
        const int code = w == R(0) ? 0 : ((RBL && w != R(1)) ? 2 : 1);
        rd[s] = make4<R>(dx, dy, dz, d2);
        rq[s] = qj;
        re[s] = e | (code << 30);
        rt[s] = (unsigned char)tj;
      }
      tail += (unsigned)__popcll(m);
    };
    // the terms of the ring's entries from head, 64 at a time, while at
    // least `keep` remain (all of them for keep = 1)
    auto drain = [&](unsigned keep) {
      while (tail - head >= keep && tail != head) {
        __builtin_amdgcn_wave_barrier();
        const unsigned m = min(tail - head, 64u);
        if ((unsigned)lane < m) {
          const int s = (int)((head + lane) & (kRing - 1));
          const R4T<R> v = rd[s];
          const int ev = re[s];
          const int code = (ev >> 30) & 3;
          const int gj = SPECIAL ? d.gidc[ev & 0x3fffffff] : 0;
          const R w = code == 0 ? R(0) : (code == 1 ? R(1) : pc.freq);
          const R ws = !SPECIAL ? R(1)
                       : (sreg ? special_w_reg<R>(srr, gj, pc) : special_w<R>(sr, gj, pc));
          eval_b<R, EFLAG, SPECIAL>(pc, s_lj, ti, (int)rt[s], xi.w * rq[s], v.x, v.y, v.z,
                                    v.w, w, a, ws);
        }
        head += m;
        __builtin_amdgcn_wave_barrier();
      }
    };
    if (cnt <= d.cap_b) {
      const int* lp = d.list_b + (size_t)c * d.cap_b;
      const unsigned char* tp = d.list_t + (size_t)c * d.cap_b;
      for (int k0 = 0; k0 < cnt; k0 += 64 * kQBatch) {
        int e[kQBatch], tj[kQBatch];
        R4T<R> xj[kQBatch];
        unsigned long long hj[kQBatch];
        // indices past the count are clamped, so every load issues; the
        // selection hash of every candidate is gathered beside its position,
        // so that no load waits for the distance and the filter below has no
        // branches
#pragma unroll
        for (int u = 0; u < kQBatch; ++u) {
          const int k = min(k0 + 64 * u + lane, cnt - 1);
          e[u] = lp[k];
          tj[u] = tp[k];
        }
#pragma unroll
        for (int u = 0; u < kQBatch; ++u) {
          xj[u] = d.xqc[e[u]];
          hj[u] = RBL ? d.hjc[e[u]] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kQBatch; ++u) {
          R dx, dy, dz;
          const R d2 = pair_d2_list<R>(b, xi.x, xi.y, xi.z, xj[u].x, xj[u].y, xj[u].z, dx,
                                       dy, dz);
          const R wf = pair_weight_flat<R, RBL>(pc, d2, hi, hj[u]);
          const bool in = (k0 + 64 * u + lane < cnt) & (d2 < pc.rc2);
          const R w = in ? wf : R(0);
          offer(EFLAG ? in : w != R(0), e[u], tj[u], dx, dy, dz, d2, w, xj[u].w);
        }
        drain(64);
      }
    } else {
      // list overflow since the last rebuild: the charged atoms of the cells
      // within ext_b, 64 per lane step
      int cc[3];
      unpack_coord(g.coord_of_cell[d.cell_s[d.c2s[c]]], cc);
      const Span sx = make_span(cc[0], g.ext_b[0], g.nc[0]);
      const Span sy = make_span(cc[1], g.ext_b[1], g.nc[1]);
      const Span sz = make_span(cc[2], g.ext_b[2], g.nc[2]);
      const int nst = sx.count * sy.count * sz.count;
      for (int st = 0; st < nst; ++st) {
        const int kx = st % sx.count, ky = (st / sx.count) % sy.count,
                  kz = st / (sx.count * sy.count);
        const int cell =
            g.cell_of_coord[(sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0] + sx.at(kx)];
        const int j0 = d.cstart[cell], j1 = d.cstart[cell + 1];
        for (int jb = j0; jb < j1; jb += 64) {
          const int j = min(jb + lane, j1 - 1);
          const R4T<R> xj = d.xqc[j];
          R dx, dy, dz;
          const R d2 = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
          const bool in = jb + lane < j1 && j != c && d2 < pc.rc2;
          const R w = in ? pair_weight<R, RBL>(pc, d2, hi, [&] { return d.hjc[j]; }) : R(0);
          offer(EFLAG ? in : w != R(0), j, d.typec[j], dx, dy, dz, d2, w, xj.w);
          drain(64);
        }
      }
    }
    drain(1);
    R fx = a.fx, fy = a.fy, fz = a.fz;
    for (int o = 32; o > 0; o >>= 1) {
      fx += __shfl_xor(fx, o, 64);
      fy += __shfl_xor(fy, o, 64);
      fz += __shfl_xor(fz, o, 64);
    }
    if (lane == 0) d.fb[d.c2i[c]] = make4<R>(fx, fy, fz, R(0));
    acc.fx += a.fx;
    acc.fy += a.fy;
    acc.fz += a.fz;
    if (EFLAG) {
      acc.e_lj += a.e_lj;
      acc.e_c += a.e_c;
      for (int k = 0; k < 6; ++k) acc.v[k] += a.v[k];
    }
  }
  write_partials<R, EFLAG>(acc, part, blockIdx.x, smem);
}

// The k-space sums k_reduce adds up after the force kernels: the per-k
// energy and virial terms and, for RBSOG, the virial tensor part
// (kspace_calculator_op.hip.cu:1308-1352): W_ab += tau_k |rho_s|^2 k_s,a k_s,b
// with s = src[k] and tau_k = kvir[k].x, divided by |k_s|^4 when kvir[k].y
// != 0.
struct KSum {
  const double* part = nullptr;  // [K * 8]
  int K = 0;                     // 0: no k-space this evaluation
  int tensor = 0;                // RBSOG tensor part
  const int* m = nullptr;
  const int* src = nullptr;
  const double2* kvir = nullptr;
  const double2* rho = nullptr;
  double L[3] = {1.0, 1.0, 1.0};
};

// Per-thread sums in a fixed order, then one block sum of NP pair values
// (force sum; with energies E_lj, E_coul and the virial) and, with k-space,
// the seven k-space values.
template <int NP, bool KS>
__device__ __forceinline__ void reduce_sums(const double* __restrict__ pair_part,
                                            int npair, const double* __restrict__ part2,
                                            int npair2, const KSum& ks, double* thermo,
                                            double* mean, int mean_fix, double n_global,
                                            double* smem) {
  constexpr int NV = NP + (KS ? 7 : 0);
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  const int ntot = npair + npair2;
  for (int blk = threadIdx.x; blk < ntot; blk += blockDim.x)
#pragma unroll
    for (int q = 0; q < NP; ++q)
      v[q] += blk < npair ? pair_part[(size_t)q * npair + blk]
                          : part2[(size_t)q * npair2 + (blk - npair)];
  if (KS) {
    const double two_pi = 6.283185307179586476925;
    for (int k = threadIdx.x; k < ks.K; k += blockDim.x) {
#pragma unroll
      for (int q = 0; q < 7; ++q) v[NP + q] += ks.part[(size_t)k * 8 + q];
      if (ks.tensor) {
        const int sk = ks.src[k];
        const double kx = two_pi * ks.m[3 * sk] / ks.L[0];
        const double ky = two_pi * ks.m[3 * sk + 1] / ks.L[1];
        const double kz = two_pi * ks.m[3 * sk + 2] / ks.L[2];
        const double2 r = ks.rho[sk];
        const double2 tv = ks.kvir[k];
        double tt = tv.x * (r.x * r.x + r.y * r.y);
        if (tv.y != 0.0) {
          const double k2 = kx * kx + ky * ky + kz * kz;
          tt /= k2 * k2;
        }
        v[NP + 1] += tt * kx * kx;
        v[NP + 2] += tt * ky * ky;
        v[NP + 3] += tt * kz * kz;
        v[NP + 4] += tt * kx * ky;
        v[NP + 5] += tt * kx * kz;
        v[NP + 6] += tt * ky * kz;
      }
    }
  }
  block_sum<NV>(v, smem);
  if (threadIdx.x == 0) {
    if (ntot > 0) {
      for (int q = 0; q < 11; ++q) thermo[q] = q < NP ? v[q] : 0.0;
      for (int q = 0; q < 3; ++q) mean[q] = mean_fix ? v[q] / n_global : 0.0;
    }
    if (KS)
      for (int q = 0; q < 7; ++q) thermo[16 + q] = v[NP + q];
  }
}

// One block after the pair kernels: FP64 sums of the per-block partials in
// a fixed order (thermo[0..10]: force sum, E_lj, E_coul, virial; the mean
// pair force when mean_fix; eflag 0 sums the force only and zeroes the rest),
// the k-space energy and virial (thermo[16..22]) and, after a rebuild, the
// list statistics of the per-block partials of k_lists.
__global__ __launch_bounds__(1024) void k_reduce(
    const double* __restrict__ pair_part, int npair,
    const double* __restrict__ part2, int npair2, int eflag,
    double* thermo, double* mean, int mean_fix, double n_global,
    const int4* lista_part, int nlista, const int2* listb_part, int nlistb,
    Ctrl* ctrl, unsigned thr, KSum ksum) {
  __shared__ double smem[16 * 18];
  if (npair + npair2 > 0 || ksum.K > 0) {
    if (ksum.K > 0) {
      if (eflag)
        reduce_sums<11, true>(pair_part, npair, part2, npair2, ksum, thermo, mean, mean_fix,
                              n_global, smem);
      else
        reduce_sums<3, true>(pair_part, npair, part2, npair2, ksum, thermo, mean, mean_fix,
                             n_global, smem);
    } else {
      if (eflag)
        reduce_sums<11, false>(pair_part, npair, part2, npair2, ksum, thermo, mean,
                               mean_fix, n_global, smem);
      else
        reduce_sums<3, false>(pair_part, npair, part2, npair2, ksum, thermo, mean, mean_fix,
                              n_global, smem);
    }
  }
  if (!rebuild_now(ctrl, thr)) return;
  __syncthreads();
  double s[6] = {0, 0, 0, 0, 0, 0};  // max_a, sum_a, stage, unstaged, max_b, sum_b
  for (int k = threadIdx.x; k < nlista; k += blockDim.x) {
    const int4 v = lista_part[k];
    s[0] = fmax(s[0], (double)v.x);
    s[1] += (double)v.y;
    s[2] = fmax(s[2], (double)v.z);
    s[3] += (double)v.w;
  }
  for (int k = threadIdx.x; k < nlistb; k += blockDim.x) {
    const int2 v = listb_part[k];
    s[4] = fmax(s[4], (double)v.x);
    s[5] += (double)v.y;
  }
  // counts are small integers, exact in FP64; maxima and sums by waves, then
  // across waves
  for (int o = 32; o > 0; o >>= 1) {
    s[0] = fmax(s[0], __shfl_xor(s[0], o, 64));
    s[2] = fmax(s[2], __shfl_xor(s[2], o, 64));
    s[4] = fmax(s[4], __shfl_xor(s[4], o, 64));
    s[1] += __shfl_xor(s[1], o, 64);
    s[3] += __shfl_xor(s[3], o, 64);
    s[5] += __shfl_xor(s[5], o, 64);
  }
  const int nw = blockDim.x >> 6, wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 6; ++q) smem[6 * wid + q] = s[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      s[0] = fmax(s[0], smem[6 * w]);
      s[2] = fmax(s[2], smem[6 * w + 2]);
      s[4] = fmax(s[4], smem[6 * w + 4]);
      s[1] += smem[6 * w + 1];
      s[3] += smem[6 * w + 3];
      s[5] += smem[6 * w + 5];
    }
    ctrl->max_count_a = max(ctrl->max_count_a, (int)s[0]);
    ctrl->sum_count_a = (unsigned long long)s[1];
    ctrl->max_stage = max(ctrl->max_stage, (int)s[2]);
    ctrl->unstaged_bricks = (int)s[3];
    if (nlistb > 0) {
      ctrl->max_count_b = max(ctrl->max_count_b, (int)s[4]);
      ctrl->sum_count_b = (unsigned long long)s[5];
    }
  }
}

// S(k) partial sums of one block: charged atoms [chunk kSkChunk, ...) (in
// the fixed order of xqs) x the kSkTile k's of k tile `tile`.  m_src holds
// the K integer k-vectors, in device memory or, for the sampled stream, in
// host-mapped staging memory; in that case m_copy (device) is given and the
// chunk-0 block of every k tile copies its tile there for the reduction and
// later calls (likewise src_src -> src_copy, the RBSOG virial-chain sources,
// when given).
template <typename R>
__device__ __forceinline__ void sk_block(const Dev<R>& d, const Box<R>& b,
                                         const int* __restrict__ m_arg, int from_ring,
                                         int with_src, int K, int chunk, int tile) {
  // the evaluation's k-sample: either m_arg (device) or its slot of the
  // host-mapped ring (m, then the RBSOG virial-chain sources), copied to HBM
  const int* m_src = m_arg;
  int* m_copy = nullptr;
  const int* src_src = nullptr;
  int* src_copy = nullptr;
  if (from_ring) {
    m_src = d.stage_ring + (size_t)(d.ctrl->step_count % kSlots) * d.stage_stride;
    m_copy = d.m_dev;
    if (with_src) {
      src_src = m_src + 3 * (size_t)K;
      src_copy = d.src_dev;
    }
  }
  __shared__ double smem[16 * 2 * kSkTile];
  const int k0 = tile * kSkTile;
  const int kn = min(kSkTile, K - k0);
  float mx[kSkTile], my[kSkTile], mz[kSkTile];
  int mi[3 * kSkTile];
#pragma unroll
  for (int kk = 0; kk < kSkTile; ++kk) {
    const int k = k0 + (kk < kn ? kk : 0);
    mi[3 * kk] = m_src[3 * k];
    mi[3 * kk + 1] = m_src[3 * k + 1];
    mi[3 * kk + 2] = m_src[3 * k + 2];
    mx[kk] = (float)mi[3 * kk];
    my[kk] = (float)mi[3 * kk + 1];
    mz[kk] = (float)mi[3 * kk + 2];
  }
  if (m_copy && chunk == 0 && threadIdx.x == 0) {
#pragma unroll
    for (int kk = 0; kk < kSkTile; ++kk)
      if (kk < kn)
#pragma unroll
        for (int q = 0; q < 3; ++q) m_copy[3 * (k0 + kk) + q] = mi[3 * kk + q];
  }
  if (src_copy && chunk == 0 && threadIdx.x < kn)
    src_copy[k0 + threadIdx.x] = src_src[k0 + threadIdx.x];
  float re[kSkTile], im[kSkTile];
#pragma unroll
  for (int kk = 0; kk < kSkTile; ++kk) re[kk] = im[kk] = 0.f;
  const int a0 = chunk * kSkChunk, a1 = min(d.nc, a0 + kSkChunk);
  for (int a = a0 + threadIdx.x; a < a1; a += blockDim.x) {
    const R4T<R> pp = d.xqs[a];
    float sx, sy, sz;
    frac_coords<R>(b, pp, sx, sy, sz);
    const float q = (float)pp.w;
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

// S(k) partials as a kernel of its own (grid: chunks x k tiles), for the
// kmax grid of the energy diagnostic.
template <typename R>
__global__ __launch_bounds__(256) void k_sk(Dev<R> d, Box<R> b, const int* __restrict__ m_arg,
                                            int from_ring, int with_src, int K) {
  sk_block<R>(d, b, m_arg, from_ring, with_src, K, (int)blockIdx.x, (int)blockIdx.y);
}

// The energy_rbe_flag diagnostic (ComputeKspaceEnergy,
// kspace_calculator.cpp:816-895): FP64 sums of the S(k) chunk partials of the
// kmax grid, then E = qqr2e (2 pi / V) sum_k exp(-k^2 / 4 alpha) |rho_k|^2 / k^2
// into thermo[24] (the self term is added on the host).  One block.
template <typename R>
__global__ __launch_bounds__(1024) void k_sk_exact(Dev<R> d, Box<R> b, int nchunks,
                                                   Params p) {
  const double two_pi = 6.283185307179586476925;
  double e[1] = {0.0};
  for (int k = threadIdx.x; k < d.kex; k += blockDim.x) {
    double sre = 0.0, sim = 0.0;
#pragma unroll 8
    for (int c = 0; c < nchunks; ++c) {
      const double2 s = d.sk_ex[(size_t)c * d.kex + k];
      sre += s.x;
      sim += s.y;
    }
    const double kx = two_pi * d.m_ex[3 * k] / (double)b.L[0];
    const double ky = two_pi * d.m_ex[3 * k + 1] / (double)b.L[1];
    const double kz = two_pi * d.m_ex[3 * k + 2] / (double)b.L[2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    e[0] += d.kw_ex[k] * (sre * sre + sim * sim) / k2;
  }
  __shared__ double smem[16];
  block_sum<1>(e, smem);
  if (threadIdx.x == 0) d.thermo[24] = p.qqr2e * two_pi / p.volume * e[0];
}

// S(k) partial sums of one block, one lane per k-vector: the kSkAtoms
// charged atoms of chunk `chunk` (in the fixed order of xqs) are staged in
// LDS as fractional coordinates and charge, each wave takes 64 of them with
// broadcast reads, and lane l of every wave accumulates k = tile * kSkK + l
// in FP32 over its atoms (two alternating accumulators); the four waves'
// sums are added in FP64 in wave order into sk_part[chunk][k].  The K
// integer k-vectors are m_arg (device memory) or, for the sampled stream,
// the step's ring slot, which k_skin has copied to m_dev.
template <typename R>
__device__ __forceinline__ void sk2_block(const Dev<R>& d, const Box<R>& b,
                                          const int* __restrict__ m_arg, int from_ring,
                                          int K, int chunk, int tile) {
  const int* m_src = from_ring ? d.m_dev : m_arg;
  __shared__ float4 s_a[kSkAtoms];
  __shared__ double s_p[kSkAtoms / 64][2][kSkK];
  const int t = (int)threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int k0 = tile * kSkK;
  const int kn = min(kSkK, K - k0);
  const int a0 = chunk * kSkAtoms, na = min(kSkAtoms, d.nc - a0);
  if (t < na) {
    const R4T<R> pp = d.xqs[a0 + t];
    float sx, sy, sz;
    frac_coords<R>(b, pp, sx, sy, sz);
    s_a[t] = make_float4(sx, sy, sz, (float)pp.w);
  } else {
    s_a[t] = make_float4(0.f, 0.f, 0.f, 0.f);  // charge 0 adds nothing
  }
  const int kl = k0 + min(lane, kn - 1);
  const float mx = (float)m_src[3 * kl], my = (float)m_src[3 * kl + 1],
              mz = (float)m_src[3 * kl + 2];
  __syncthreads();
  float re0 = 0.f, im0 = 0.f, re1 = 0.f, im1 = 0.f;
  const float4* sa = s_a + 64 * wv;
#pragma unroll 4
  for (int a = 0; a < 64; a += 2) {
    const float4 u = sa[a], v = sa[a + 1];
    float sn, cs, sn1, cs1;
    sincos_rev(fma_r(mx, u.x, fma_r(my, u.y, mz * u.z)), sn, cs);
    sincos_rev(fma_r(mx, v.x, fma_r(my, v.y, mz * v.z)), sn1, cs1);
    re0 = fma_r(u.w, cs, re0);
    im0 = fma_r(u.w, sn, im0);
    re1 = fma_r(v.w, cs1, re1);
    im1 = fma_r(v.w, sn1, im1);
  }
  s_p[wv][0][lane] = (double)re0 + (double)re1;
  s_p[wv][1][lane] = (double)im0 + (double)im1;
  __syncthreads();
  if (t < kn) {
    double sre = 0.0, sim = 0.0;
#pragma unroll
    for (int w = 0; w < kSkAtoms / 64; ++w) {
      sre += s_p[w][0][t];
      sim += s_p[w][1][t];
    }
    d.sk_part[(size_t)chunk * d.kcap + k0 + t] = make_double2(sre, sim);
  }
}

// FP64 reduction of the S(k) chunk partials of the kSkK k's of k tile
// `tile` (16 threads per k, each a contiguous run of chunks, then the 16 in
// order; the kernel boundary after the S(k) blocks makes every partial
// visible) into rho_k and the per-k force coefficients
//   A_k = w_k (-4 pi qqr2e) / (V k^2),  kf0 = (m, A Re rho), kf1 = (k, A Im rho),
// and each k's terms of the k-space energy and virial (without the self
// terms) into kfin_part[k][0..6], which k_reduce adds up:
//   E = qqr2e (2 pi / V) sum_k w_k |rho|^2 / k^2,
// vmode 0 (RBE, EWALD): W_ab = sum_k E_k (delta_ab - 2 k_a k_b (1/(4 alpha)
// + 1/k^2)); vmode 1 (RBSOG, kspace_calculator_op.hip.cu:1133-1222): the
// isotropic part W_aa += E_k here, the tensor part, which reads rho of other
// k's, in k_reduce (ksum_tensor).
template <typename R>
__device__ __forceinline__ void skfin2_block(const Dev<R>& d, const Box<R>& b,
                                             const int* __restrict__ m, int K, int nchunks,
                                             const Params& p, int vmode, int tile,
                                             char* buf) {
  constexpr int kSub = 16;  // threads per k (blockDim = kSkK * kSub)
  // buf: LDS of kSub * 2 * kSkK doubles
  auto s_s = reinterpret_cast<double (*)[2][kSkK]>(buf);
  const int t = (int)threadIdx.x;
  const int kk = t % kSkK, r = t / kSkK;
  const int k = tile * kSkK + kk;
  const int per = (nchunks + kSub - 1) / kSub;
  const int c0 = r * per, c1 = min(nchunks, c0 + per);
  double sre = 0.0, sim = 0.0;
  if (k < K) {
#pragma unroll 4
    for (int c = c0; c < c1; ++c) {
      const double2 v = d.sk_part[(size_t)c * d.kcap + k];
      sre += v.x;
      sim += v.y;
    }
  }
  s_s[r][0][kk] = sre;
  s_s[r][1][kk] = sim;
  __syncthreads();
  const double V = p.volume;
  const double two_pi = 6.283185307179586476925;
  double kv[7] = {0, 0, 0, 0, 0, 0, 0};
  if (r == 0 && k < K) {
    sre = 0.0;
    sim = 0.0;
    for (int q = 0; q < kSub; ++q) {
      sre += s_s[q][0][kk];
      sim += s_s[q][1][kk];
    }
    d.rho[k] = make_double2(sre, sim);
    const double kx = two_pi * m[3 * k] / (double)b.L[0];
    const double ky = two_pi * m[3 * k + 1] / (double)b.L[1];
    const double kz = two_pi * m[3 * k + 2] / (double)b.L[2];
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double w = d.kw[k];
    const double A = w * (-4.0 * 3.141592653589793238463 * p.qqr2e) / (V * k2);
    d.kf0[k] = make_float4((float)m[3 * k], (float)m[3 * k + 1], (float)m[3 * k + 2],
                           (float)(A * sre));
    d.kf1[k] = make_float4((float)kx, (float)ky, (float)kz, (float)(A * sim));
    const double base = p.qqr2e * two_pi / V * w * (sre * sre + sim * sim) / k2;
    kv[0] = base;
    if (vmode == 0) {
      const double cc = 2.0 * (1.0 / (4.0 * p.alpha) + 1.0 / k2);
      kv[1] = base * (1.0 - cc * kx * kx);
      kv[2] = base * (1.0 - cc * ky * ky);
      kv[3] = base * (1.0 - cc * kz * kz);
      kv[4] = base * (-cc * kx * ky);
      kv[5] = base * (-cc * kx * kz);
      kv[6] = base * (-cc * ky * kz);
    } else {
      kv[1] = base;
      kv[2] = base;
      kv[3] = base;
    }
  }
  if (r == 0 && k < K)
    for (int q = 0; q < 7; ++q) d.kfin_part[(size_t)k * 8 + q] = kv[q];
}

// RBE / Ewald force on 64 charged atoms (block `bid`, one lane per atom, the
// fixed order of xqs and fk).  The per-k coefficients are staged in LDS
// kRbeStage k's at a time; wave w takes k's w, w + nw, ... of each stage
// with broadcast reads, and the waves' sums are added in wave order.
template <typename R>
__device__ __forceinline__ void rbe2_block(const Dev<R>& d, const Box<R>& b, int K,
                                           int bid) {
  __shared__ float4 s_k0[kRbeStage], s_k1[kRbeStage];
  __shared__ float s_f[kRbeWaves][3][64];
  const int nw = (int)blockDim.x >> 6;
  const int t = (int)threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const int c = bid * 64 + lane;
  float sx = 0.f, sy = 0.f, sz = 0.f, q = 0.f;
  if (c < d.nc) {
    const R4T<R> pp = d.xqs[c];
    frac_coords<R>(b, pp, sx, sy, sz);
    q = (float)pp.w;
  }
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for (int s0 = 0; s0 < K; s0 += kRbeStage) {
    const int ns = min(kRbeStage, K - s0);
    __syncthreads();  // the previous stage is consumed
    for (int k = t; k < ns; k += blockDim.x) {
      s_k0[k] = d.kf0[s0 + k];
      s_k1[k] = d.kf1[s0 + k];
    }
    __syncthreads();
#pragma unroll 4
    for (int k = wv; k < ns; k += nw) {
      const float4 a = s_k0[k];
      const float4 e = s_k1[k];
      const float ph = fma_r(a.x, sx, fma_r(a.y, sy, a.z * sz));
      float sn, cs;
      sincos_rev(ph, sn, cs);
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
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (int w = 0; w < nw; ++w) {
      ax += s_f[w][0][lane];
      ay += s_f[w][1][lane];
      az += s_f[w][2][lane];
    }
    d.fk[d.qcall[c]] = make4<R>((R)(q * ax), (R)(q * ay), (R)(q * az), R(0));
  }
}

// k-space work that rides on the rebuild kernels' launches (their blocks
// exit at once unless the evaluation rebuilds), so that one evaluation is one
// stream of kernels: the S(k) partials after the skin kernel has written the
// fixed-order charged copy, their reduction next, the RBE force after that.
// The k-space kernels index the charged atoms in that fixed order, which a
// rebuild does not change.
struct KsArgs {
  int K = 0;                    // k-vectors of the evaluation (0: none)
  const int* m_arg = nullptr;   // device k-vectors for the S(k) blocks, or
  int from_ring = 0;            // the step's ring slot (k_skin copies it to m_dev)
  const int* m_fin = nullptr;   // k-vectors as the reduction reads them
  const int* src = nullptr;     // RBSOG virial-chain sources (device)
  int vmode = 0;                // virial form (0 RBE / EWALD, 1 RBSOG)
  int nchunks = 0, ntiles = 0;  // S(k) blocks: kSkAtoms-atom chunks x kSkK-k tiles
  int nrbe = 0;                 // RBE force blocks (64 charged atoms each)
  Params p;
};

// Binning (blocks [0, nbin)) and the S(k) partials (the blocks after).
template <typename R>
__global__ __launch_bounds__(256) void k_bin_sk(Dev<R> d, Box<R> b, Grid g, int split,
                                                unsigned thr, int nskin, int nbin,
                                                KsArgs ks) {
  if ((int)blockIdx.x < nbin) {
    bin_block<R>(d, b, g, split, thr, nskin, blockIdx.x, nbin);
    return;
  }
  const int q = (int)blockIdx.x - nbin;
  sk2_block<R>(d, b, ks.m_arg, ks.from_ring, ks.K, q % ks.nchunks, q / ks.nchunks);
}

// The cell scans (block 0) and the FP64 S(k) reduction (blocks 1 .. ntiles,
// one k tile each).
template <typename R>
__global__ __launch_bounds__(1024) void k_scan_fin(Dev<R> d, Box<R> b, int ncell,
                                                   unsigned thr, KsArgs ks) {
  // one LDS buffer for either role of the block
  constexpr size_t kScanBytes = (size_t)kScanPer * 1024 * (sizeof(int2) + sizeof(int));
  constexpr size_t kFinBytes = (size_t)16 * 2 * kSkK * sizeof(double);
  __shared__ __align__(16) char buf[kScanBytes > kFinBytes ? kScanBytes : kFinBytes];
  if (blockIdx.x == 0)
    scan_block(d.key_count, d.key_start, d.cstart, ncell, d.ctrl, thr, buf);
  else
    skfin2_block<R>(d, b, ks.m_fin, ks.K, ks.nchunks, ks.p, ks.vmode, (int)blockIdx.x - 1,
                    buf);
}

// The sorted-order scatter (blocks [0, nsc)) and the RBE force (the blocks
// after).
template <typename R>
__global__ __launch_bounds__(256) void k_scatter_rbe(Dev<R> d, Box<R> b, unsigned thr,
                                                     int nsc, KsArgs ks) {
  if ((int)blockIdx.x < nsc)
    scatter_block<R>(d, thr, blockIdx.x, nsc);
  else
    rbe2_block<R>(d, b, ks.K, (int)blockIdx.x - nsc);
}

// v + 0.5 F / m dt ftm2v evaluated left to right, as RBMD's UpdateVelocity
// (update_velocity_op.hip.cu:19-21), without contraction into FMA.
template <typename R>
__device__ __forceinline__ R md_half_kick(R v, R f, R m, R dt, R ftm2v) {
#pragma clang fp contract(off)
  return v + (R)0.5 * f / m * dt * ftm2v;
}

// Total force in caller order (one thread per caller atom); resets the skin
// accumulator for the next call and mirrors the counters to the host.  With
// velocities bound it also applies the leapfrog update
// v += dt ftm2v F / m, x += dt v (RBMD's UpdateVelocity half-kicks of two
// consecutive steps merged, src/simulate_pipeline/src/controller/
// velocity_controller/velocity_controller_op/rocm/update_velocity_op.hip.cu:5-27).
// With md_kick (the molecular-dynamics step) it applies the closing half-kick
// v += 0.5 F / m dt ftm2v instead and writes the workgroup's FP64 sum of
// m v^2 to d.md_part[blockIdx.x].
template <typename R>
__global__ __launch_bounds__(256) void k_assemble(Dev<R> d, int mean_fix, int has_a,
                                                  int has_k, int has_b, R dt,
                                                  R dt_ftm, int md_kick, R ftm2v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    d.ctrl->maxdisp_bits = 0u;
    d.ctrl->step_count += 1;
    if (d.ctrl_host) {
      d.ctrl_host->max_count_a = d.ctrl->max_count_a;
      d.ctrl_host->max_count_b = d.ctrl->max_count_b;
      d.ctrl_host->step_count = d.ctrl->step_count;
      d.ctrl_host->max_stage = d.ctrl->max_stage;
      d.ctrl_host->unstaged_bricks = d.ctrl->unstaged_bricks;
    }
  }
  double ke = 0.0;
  if (i < d.n) {
    // the force kernels write in caller order (zero rows of fb and fk for
    // the neutral atoms): coalesced reads, no gathers through the maps
    R4T<R> f = has_a ? d.fa[i] : make4<R>(R(0), R(0), R(0), R(0));
    if (mean_fix) {
      f.x -= (R)d.mean[0];
      f.y -= (R)d.mean[1];
      f.z -= (R)d.mean[2];
    }
    if (has_b) {
      const R4T<R> g = d.fb[i];
      f.x += g.x;
      f.y += g.y;
      f.z += g.z;
    }
    if (has_k) {
      const R4T<R> g = d.fk[i];
      f.x += g.x;
      f.y += g.y;
      f.z += g.z;
    }
    d.fx[i] = f.x;
    d.fy[i] = f.y;
    d.fz[i] = f.z;
    if (md_kick) {
      const double md = d.md_mass[d.type[i]];
      const R m = (R)md;
      const R vx = md_half_kick(d.vx[i], f.x, m, dt, ftm2v);
      const R vy = md_half_kick(d.vy[i], f.y, m, dt, ftm2v);
      const R vz = md_half_kick(d.vz[i], f.z, m, dt, ftm2v);
      d.vx[i] = vx;
      d.vy[i] = vy;
      d.vz[i] = vz;
      const double x = vx, y = vy, z = vz;
      ke = md * (x * x + y * y + z * z);
    } else if (d.vx) {
      const R s = dt_ftm * d.inv_mass[d.type[i]];
      const R vx = d.vx[i] + s * f.x, vy = d.vy[i] + s * f.y, vz = d.vz[i] + s * f.z;
      d.vx[i] = vx;
      d.vy[i] = vy;
      d.vz[i] = vz;
      d.xw[i] = d.px[i] + dt * vx;
      d.yw[i] = d.py[i] + dt * vy;
      d.zw[i] = d.pz[i] + dt * vz;
    }
  }
  if (md_kick) {  // uniform: every thread of the workgroup reaches the sum
    __shared__ double smem[16];
    double v[1] = {ke};
    block_sum<1>(v, smem);
    if (threadIdx.x == 0) d.md_part[blockIdx.x] = v[0];
  }
}

// Parity export (block per brick): classified pairs (or raw candidates) as
// (gid_i, gid_j, cls).
template <typename R, bool RBL>
__global__ __launch_bounds__(512) void k_export(Dev<R> d, Box<R> b, Grid g,
                                                PairConst<R> pc,
                                                int candidates, int cap) {
  __shared__ SlotTables<kMaxSlots> t;
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
    if (d2 < pc.t_shell && rbl_selected(pc, gi, gj)) return 2;
    return 0;
  };
  const int e1[3] = {1, 1, 1};
  const Neigh h = brick_neigh(g, blockIdx.x, e1);
  build_slots(g, h, d.key_start, nullptr, false, t);
  const int i0 = d.key_start[2 * h.cell0], i1 = d.key_start[2 * h.cell1];
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const R4T<R> xi = d.xs[i];
    const int gi = d.gid_s[i];
    const int cnt = min(d.cnt_a[i], d.cap_a);
    for (int k = 0; k < cnt; ++k) {
      const int l = d.list_a[list_a_index(k, i, d.n)] >> StageShift<R>::value;
      const int s = slot_of_pos(t, h.nslots, l);
      const int j = t.first[s] + (l - t.off[s]);
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
  const int c0 = d.cstart[h.cell0], c1 = d.cstart[h.cell1];
  for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
    const R4T<R> xi = d.xqc[c];
    const int gi = d.gidc[c];
    const int cnt = min(d.cnt_b[c], d.cap_b);
    for (int k = 0; k < cnt; ++k) {
      const int e = d.list_b[(size_t)c * d.cap_b + k];
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

int list_b_grid(const Params& p, int nc, const Grid& g) {
  return (p.split && nc > 0) ? list_b_count(g, nc) : 0;
}

template <typename R>
hipError_t build_part(const Dev<R>& d, const StepArgs<R>& a, int nskin,
                      hipStream_t s, const KsArgs& ks = KsArgs()) {
  const Grid& g = a.grid;
  const unsigned thr = a.rebuild_thr;
  const int gb = blocks_for(d.n, kBlock, 2048);
  const int cap_a = a.count_only ? 0 : d.cap_a;
  const int cap_b = a.count_only ? 0 : d.cap_b;
  const bool kon = ks.K > 0;
  hipLaunchKernelGGL((k_bin_sk<R>), dim3(gb + (kon ? ks.nchunks * ks.ntiles : 0)),
                     dim3(kBlock), 0, s, d, a.box, g, a.p.split, thr, nskin, gb, ks);
  hipLaunchKernelGGL((k_scan_fin<R>), dim3(kon ? 1 + ks.ntiles : 1), dim3(1024), 0, s,
                     d, a.box,
                     g.ncell, thr, ks);
  hipLaunchKernelGGL((k_scatter_rbe<R>), dim3(gb + (kon ? ks.nrbe : 0)), dim3(kBlock), 0,
                     s, d, a.box, thr, gb, ks);
  hipLaunchKernelGGL((k_cellsort<R>), dim3(blocks_for(g.ncell * 64, kBlock)),
                     dim3(kBlock), 0, s, d, a.box, g.ncell, thr, a.pc.ms);
  const bool dense = a.p.split && d.nc > 0 && g.qblocks;
  hipLaunchKernelGGL((k_lists<R>), dim3(g.nbrick + (dense ? 0 : list_b_grid(a.p, d.nc, g))),
                     dim3(g.block), 0, s, d, a.box, g, a.pc.r2_list_a,
                     a.pc.r2_list_b, a.p.split, cap_a, cap_b, thr);
  if (dense)
    hipLaunchKernelGGL((k_lists_dense<R>), dim3(g.ncell * g.dsplit), dim3(kDenseBlock), 0, s,
                       d, a.box, g,
                       a.pc.r2_list_b, cap_b, thr);
  return hipGetLastError();
}

}  // namespace

template <typename R>
hipError_t launch_build(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s) {
  hipLaunchKernelGGL((k_skin<R>), dim3(1), dim3(kBlock), 0, s, d, a.box, 1, 0, 0, 0);
  hipError_t e = build_part<R>(d, a, 1, s);
  if (e != hipSuccess) return e;
  // list statistics only (no pair partials)
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, d.pair_part, 0,
                     (const double*)nullptr, 0, 0, d.thermo, d.mean, 0, 1.0, d.lista_part,
                     a.grid.nbrick, d.listb_part, list_b_grid(a.p, d.nc, a.grid), d.ctrl,
                     a.rebuild_thr, KSum());
  return hipGetLastError();
}

template <typename R>
hipError_t launch_force(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s) {
  const Params& p = a.p;
  // skin check and gather: one thread per atom up to kSkinBlocksMax blocks
  // (grid-stride beyond; their maxima are folded by every block of k_bin)
  const int sb = blocks_for(d.n, kBlock, kSkinBlocksMax);
  if (a.ev[0]) hipEventRecord(a.ev[0], s);
  const int K = p.kspace ? p.num_k : 0;
  const bool ks_on = K > 0 && d.nc > 0;
  // (a forced rebuild comes through the step ring, like the epoch, and so
  // does the sampled k-sample, which k_skin copies to HBM)
  const int kcopy = (ks_on && a.m_from_ring) ? K : 0;
  const int ctl = d.hdr_ring ? 1 : 0;
  hipLaunchKernelGGL((k_skin<R>), dim3(sb + ctl), dim3(kBlock), 0, s, d, a.box, 0, ctl,
                     kcopy, p.kspace == 3 ? 1 : 0);
  // k-space blocks beside the rebuild kernels' blocks (one stream, no fork)
  KsArgs ks;
  if (ks_on) {
    const bool sog = p.kspace == 3;
    ks.K = K;
    ks.from_ring = a.m_from_ring;
    ks.m_arg = a.m_use;
    ks.m_fin = a.m_from_ring ? (const int*)d.m_dev : a.m_use;
    ks.src = sog ? (const int*)d.src_dev : nullptr;
    ks.vmode = sog ? 1 : 0;
    ks.nchunks = (d.nc + kSkAtoms - 1) / kSkAtoms;
    ks.ntiles = (K + kSkK - 1) / kSkK;
    ks.nrbe = blocks_for(d.nc, 64);
    ks.p = p;
  }
  if (a.ev[3]) hipEventRecord(a.ev[3], s);
  hipError_t e = build_part<R>(d, a, sb + ctl, s, ks);
  if (e != hipSuccess) return e;
  if (ks_on && p.eflag && d.kex > 0) {
    // energy_rbe_flag: the exact Ewald energy over the kmax grid
    const int nchunks = (d.nc + kSkChunk - 1) / kSkChunk;
    Dev<R> dx = d;
    dx.sk_part = d.sk_ex;
    dx.kcap = d.kex;
    hipLaunchKernelGGL((k_sk<R>), dim3(nchunks, (d.kex + kSkTile - 1) / kSkTile),
                       dim3(kBlock), 0, s, dx, a.box, (const int*)d.m_ex, 0, 0, d.kex);
    hipLaunchKernelGGL((k_sk_exact<R>), dim3(1), dim3(1024), 0, s, d, a.box, nchunks, p);
  }
  if (a.ev[4]) hipEventRecord(a.ev[4], s);
  const Grid& g = a.grid;
  const size_t lj_bytes = 4 * sizeof(R) * (size_t)a.pc.ntypes * a.pc.ntypes;
  // ev[1] .. ev[2] bracket the pair kernel alone
  if (a.ev[1]) hipEventRecord(a.ev[1], s);
#define RBMDX_PAIR(RBLv, EFv, SHv, SPv)                                     \
  hipLaunchKernelGGL((k_pair<R, RBLv, EFv, SHv, SPv>), dim3(g.nbrick * g.psub), \
                     dim3(g.block + g.qwave), lj_bytes, s, d, a.box, g, a.pc, \
                     a.lj_tab, p.split)
  // list A holds RBL shell pairs only when the LJ cutoff exceeds r_core
  const bool shell_a = a.pc.rclj2 > a.pc.rs2;
  // every atom charged (water, molten salts): list A is empty and every pair
  // is k_pair_b's, so the brick kernel is not launched at all
  const bool has_a = !(p.split && d.nc == d.n && g.qblocks > 0);
  if (!has_a) {
    // nothing: k_assemble reads no list-A force
  } else if (p.special) {
    // CVFF special bonds: the general (shell-capable) list-A loop
    if (p.rbl) {
      if (p.eflag)
        RBMDX_PAIR(true, true, true, true);
      else
        RBMDX_PAIR(true, false, true, true);
    } else {
      if (p.eflag)
        RBMDX_PAIR(false, true, true, true);
      else
        RBMDX_PAIR(false, false, true, true);
    }
  } else if (p.rbl) {
    if (p.eflag) {
      if (shell_a)
        RBMDX_PAIR(true, true, true, false);
      else
        RBMDX_PAIR(true, true, false, false);
    } else {
      if (shell_a)
        RBMDX_PAIR(true, false, true, false);
      else
        RBMDX_PAIR(true, false, false, false);
    }
  } else {
    if (p.eflag)
      RBMDX_PAIR(false, true, false, false);
    else
      RBMDX_PAIR(false, false, false, false);
  }
#undef RBMDX_PAIR
  if (g.qblocks > 0) {
#define RBMDX_PAIRB(RBLv, EFv, SPv)                                             \
  hipLaunchKernelGGL((k_pair_b<R, RBLv, EFv, SPv>), dim3(g.qblocks),           \
                     dim3(64 * kQWaves), lj_bytes, s, d, a.box, g, a.pc, a.lj_tab, \
                     d.pairb_part)
    if (p.special) {
      if (p.rbl) {
        if (p.eflag)
          RBMDX_PAIRB(true, true, true);
        else
          RBMDX_PAIRB(true, false, true);
      } else {
        if (p.eflag)
          RBMDX_PAIRB(false, true, true);
        else
          RBMDX_PAIRB(false, false, true);
      }
    } else if (p.rbl) {
      if (p.eflag)
        RBMDX_PAIRB(true, true, false);
      else
        RBMDX_PAIRB(true, false, false);
    } else {
      if (p.eflag)
        RBMDX_PAIRB(false, true, false);
      else
        RBMDX_PAIRB(false, false, false);
    }
#undef RBMDX_PAIRB
  }
  if (a.ev[2]) hipEventRecord(a.ev[2], s);
  KSum ksum;
  if (ks_on) {
    ksum.part = d.kfin_part;
    ksum.tensor = ks.vmode == 1;
    ksum.K = K;
    ksum.m = ks.m_fin;
    ksum.src = ks.src;
    ksum.kvir = d.kvir;
    ksum.rho = d.rho;
    for (int q = 0; q < 3; ++q) ksum.L[q] = (double)a.box.L[q];
  }
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, s, d.pair_part,
                     has_a ? g.nbrick * g.psub : 0, (const double*)d.pairb_part, g.qblocks,
                     p.eflag ? 1 : 0, d.thermo, d.mean, p.mean_fix, p.n_global,
                     d.lista_part, g.nbrick, d.listb_part,
                     list_b_grid(p, d.nc, g), d.ctrl, a.rebuild_thr, ksum);
  hipLaunchKernelGGL((k_assemble<R>), dim3(blocks_for(d.n)), dim3(kBlock), 0,
                     s, d, p.mean_fix, has_a ? 1 : 0, ks_on ? 1 : 0, p.split ? 1 : 0, (R)a.dt,
                     (R)a.dt_ftm, a.md_kick, (R)a.ftm2v);
  if (a.ev[5]) hipEventRecord(a.ev[5], s);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_export(const Dev<R>& d, const StepArgs<R>& a, int candidates,
                         int cap, hipStream_t s) {
  const Grid& g = a.grid;
  if (a.p.rbl)
    hipLaunchKernelGGL((k_export<R, true>), dim3(g.nbrick), dim3(g.block), 0, s,
                       d, a.box, g, a.pc, candidates, cap);
  else
    hipLaunchKernelGGL((k_export<R, false>), dim3(g.nbrick), dim3(g.block), 0,
                       s, d, a.box, g, a.pc, candidates, cap);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Force-matching adjoint: the vector-Jacobian product of one evaluation's
// force F(r, Q, eps, sigma) with a cotangent G = dL/dF, for positions and the
// per-type charge (a common shift of the charges of every atom of the type),
// LJ eps and sigma.  The pair set, RBL selections and k-sample are those of
// the evaluation.  Per directed pair (i, j), d = r_j - r_i, u = |d|^2,
// weights w_ij (i's set) and w_ji (j's set), unweighted pair scalar f0(u)
// (LJ below rc_LJ, erfc Coulomb) with u-derivative f0':
//   dL/dr_i += f0 H + 2 f0' (H . d) d,  H = w_ji G'_j - w_ij G'_i,
// G' = G - mean(G) under the mean-force fix; the owner i of (i, j) adds
// w_ij (G'_i . d) times df0/d eps_ij, df0/d sigma_ij, df0/d(q_i q_j) to its
// class sums, which the mixing rules carry to the types.  k-space (weights
// w_k fixed): A_k = w_k (-4 pi qqr2e) / (V k^2), rho = sum q e^{i theta},
// beta = sum q (G . k) e^{i theta}, per type T_a = sum e^{i theta},
// U_a = sum (G . k) e^{i theta}:
//   dL/dQ_a += sum_k A_k [Im(conj(U_a) rho) + Im(conj(beta) T_a)],
//   dL/dr_m += q_m sum_k A_k [Re(conj(beta) e^{i theta_m})
//                            - (G_m . k) Re(e^{-i theta_m} rho)] k.
// S(k) is recomputed here (no per atom-k storage).  Parameter sums use LDS
// float atomics inside a workgroup and FP64 atomics across workgroups (their
// last bits may vary between runs); everything else is deterministic.
// ---------------------------------------------------------------------------

// LJ scalars of one pair from its adjoint table row (-24 eps, sigma^2, eps,
// 1/sigma or 0): f = -24 eps (2 s6 - 1) s6 / u (= eps df/deps), fp = df/du,
// fs = df/d sigma_ij.
template <typename R>
__device__ __forceinline__ void lj_vjp(const R* row, R u, R& f, R& fp, R& fs) {
  const R inv = pair_rcp(u);
  const R s2 = row[1] * inv;
  const R s6 = s2 * s2 * s2;
  f = row[0] * (R(2) * s6 - R(1)) * s6 * inv;
  fp = R(24) * row[2] * (R(14) * s6 * s6 - R(4) * s6) * inv * inv;
  fs = R(-144) * row[2] * (R(4) * s6 - R(1)) * s6 * inv * row[3];
}

// erfc Coulomb without the charges: f_C = -qqr2e q_i q_j h(u),
// h = erfc(g r) / (u r) + (2 g / sqrt(pi)) e^{-alpha u} / u, hp = dh/du.
// With the SOG near field (VERLET-SOG, rbsog_near_field "sog"; coul_sog):
// h = u^{-3/2} + sum_n T_n u^n, hp = -1.5 u^{-5/2} + sum_n n T_n u^{n-1}.
template <typename R>
__device__ __forceinline__ void coul_vjp(const PairConst<R>& pc, R u, R& h,
                                         R& hp) {
  // 1/r and 1/u from the hardware reciprocal square root and reciprocal
  // (1 ulp) in FP32, as in the forward's pair terms
  const R rinv = pair_rsqrt(u);
  const R inv = pair_rcp(u);
  if (pc.sog) {
    const R r3inv = rinv * inv;
    R poly = pc.T[5], dpoly = R(5) * pc.T[5];
#pragma unroll
    for (int n = 4; n >= 0; --n) poly = poly * u + pc.T[n];
#pragma unroll
    for (int n = 4; n >= 1; --n) dpoly = dpoly * u + R(n) * pc.T[n];
    h = r3inv + poly;
    hp = R(-1.5) * r3inv * inv + dpoly;
    return;
  }
  const R r = u * rinv;
  const R er = erfc(pc.g * r), ex = exp(-pc.alpha * u);
  h = er * inv * rinv + pc.two_g_over_sqrtpi * ex * inv;
  hp = R(-1.5) * er * inv * inv * rinv -
       R(1.5) * pc.two_g_over_sqrtpi * ex * inv * inv -
       pc.alpha * pc.two_g_over_sqrtpi * ex * inv;
}

// Directed pair weight (the forward's pair_weight with an explicit hash pair).
template <typename R, bool RBL>
__device__ __forceinline__ R vjp_w(const PairConst<R>& pc, R u,
                                   unsigned long long h_own,
                                   unsigned long long hj_other) {
  if (!RBL) return u < pc.rc2 ? R(1) : R(0);
  if (u < pc.t_core) return (u < pc.rs2 && u > pc.eps_min) ? R(1) : R(0);
  if (u < pc.t_shell && u > pc.rs2 &&
      rbl_keep(pc.ifreq, pc.fmask, pc.c32, pc.epoch, h_own, hj_other))
    return pc.freq;
  return R(0);
}

// Parameter cotangents of one owner atom.  Every pair is listed once from
// each side, so the owner i adds what both directed pairs (i, j) and (j, i)
// contribute to its own type a: with m = w_ij G'_i . d - w_ji G'_j . d
// (= -H . d), dL/dQ_a += q_j (-qqr2e h) m, dL/deps_a += f_LJ m / (2 eps_a)
// (f_LJ = eps_ab df/deps_ab and deps_ab/deps_a = eps_ab / (2 eps_a)),
// dL/dsigma_a += 0.5 (df/dsigma_ab) m.
struct VjpAcc {
  float dq = 0.f, de = 0.f, ds = 0.f;
};

// Adds an owner's sums to the block's per-type table ([3][nt]: Q, eps,
// sigma); eps_a from the LJ table diagonal.
template <typename R>
__device__ __forceinline__ void vjp_flush(float* tab, int nt, int a, const R* s_lj,
                                          const VjpAcc& c) {
  const R ea = s_lj[4 * (a * nt + a) + 2];
  if (c.dq != 0.f) atomicAdd(&tab[a], c.dq);
  if (c.de != 0.f && ea > R(0)) atomicAdd(&tab[nt + a], (float)(c.de / (R(2) * ea)));
  if (c.ds != 0.f) atomicAdd(&tab[2 * nt + a], c.ds);
}

// Cotangent to sorted and charged order, with block sums for its mean.
template <typename R>
__global__ __launch_bounds__(256) void k_vjp_gather(Dev<R> d, VjpDev<R> v) {
  if (blockIdx.x == 0 && threadIdx.x < 3 * kVjpMaxTypes) v.pacc[threadIdx.x] = 0.0;
  double s[3] = {0, 0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n;
       i += gridDim.x * blockDim.x) {
    const R gx = v.gx[i], gy = v.gy[i], gz = v.gz[i];
    s[0] += (double)gx;
    s[1] += (double)gy;
    s[2] += (double)gz;
    v.gs[d.inv_perm[i]] = make4<R>(gx, gy, gz, R(0));
    const int c = d.c_of[i];
    if (c >= 0) v.gc[c] = make4<R>(gx, gy, gz, R(0));
  }
  __shared__ double smem[16 * 3];
  block_sum<3>(s, smem);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) v.gpart[3 * blockIdx.x + q] = s[q];
}

__global__ __launch_bounds__(256) void k_vjp_mean(const double* gpart, int nb,
                                                  double n, int mean_fix,
                                                  double* gmean) {
  double s[3] = {0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int q = 0; q < 3; ++q) s[q] += gpart[3 * b + q];
  __shared__ double smem[16 * 3];
  block_sum<3>(s, smem);
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) gmean[q] = mean_fix ? s[q] / n : 0.0;
}

// One owner pair of the adjoint pair pass.  The caller provides d and u;
// gi/gj are the mean-shifted cotangents, row the LJ row (null beyond rc_LJ),
// coul whether Coulomb applies, qj the neighbour's charge.
//
// CVFF special bonds (ws < 1, the same weight both ways): the LJ terms carry
// ws, and a kept charged pair also removes (1 - ws) of the full Coulomb force
// without the RBL multiplier (interact_b): f_c = (1 - ws) qqr2e q_i q_j u^-3/2
// along d, weighted by cij / cji = 1 where the directed pair is kept.
template <typename R>
__device__ __forceinline__ void vjp_pair(R dx, R dy, R dz, R u, const R* row,
                                         bool coul, const PairConst<R>& pc,
                                         R qi, R qj, R wij, R wji,
                                         const R gi[3], const R gj[3], R r[3],
                                         VjpAcc& acc, R ws = R(1), R cij = R(0),
                                         R cji = R(0)) {
  R flj = 0, fp = 0, fs = 0;
  if (row) lj_vjp<R>(row, u, flj, fp, fs);
  if (ws != R(1)) {
    flj *= ws;
    fp *= ws;
    fs *= ws;
  }
  R f = flj, gqb = 0;
  if (coul) {
    R h, hp;
    coul_vjp<R>(pc, u, h, hp);
    f += -pc.qqr2e * qi * qj * h;
    fp += -pc.qqr2e * qi * qj * hp;
    gqb = -pc.qqr2e * h;
  }
  const R hx = wji * gj[0] - wij * gi[0], hy = wji * gj[1] - wij * gi[1],
          hz = wji * gj[2] - wij * gi[2];
  const R hd = hx * dx + hy * dy + hz * dz;
  r[0] += f * hx + R(2) * fp * hd * dx;
  r[1] += f * hy + R(2) * fp * hd * dy;
  r[2] += f * hz + R(2) * fp * hd * dz;
  if (row) {
    acc.de -= (float)(flj * hd);
    acc.ds -= (float)(R(0.5) * fs * hd);
  }
  if (coul) acc.dq -= (float)(qj * gqb * hd);
  if (coul && ws != R(1)) {
    const R rinv = pair_rsqrt(u);
    const R gc = (R(1) - ws) * pc.qqr2e * rinv * rinv * rinv;  // f_c / (q_i q_j)
    const R fc = qi * qj * gc, fcp = R(-1.5) * fc * pair_rcp(u);
    const R kx = cji * gj[0] - cij * gi[0], ky = cji * gj[1] - cij * gi[1],
            kz = cji * gj[2] - cij * gi[2];
    const R kd = kx * dx + ky * dy + kz * dz;
    r[0] += fc * kx + R(2) * fcp * kd * dx;
    r[1] += fc * ky + R(2) * fcp * kd * dy;
    r[2] += fc * kz + R(2) * fcp * kd * dz;
    acc.dq -= (float)(qj * gc * kd);
  }
}

// Adjoint pair pass, one workgroup per brick (the forward's lists and stage;
// the cotangent and global ids staged beside the positions).
template <typename R, bool RBL, bool SPECIAL>
__global__ __launch_bounds__(512) void k_pair_vjp(Dev<R> d, VjpDev<R> v, Box<R> b,
                                                  Grid g, PairConst<R> pc_arg,
                                                  const R* lj_tab, int split) {
  PairConst<R> pc = pc_arg;
  if (RBL) pc.epoch = d.ctrl->epoch;
  __shared__ R4T<R> stage[StageCap<R>::value];
  __shared__ R4T<R> gstage[StageCap<R>::value];
  __shared__ SlotTables<kMaxSlots> t;
  __shared__ float tab[3 * kVjpMaxTypes];
  extern __shared__ char dyn_lds[];
  const int nt = pc.ntypes;
  R* s_lj = reinterpret_cast<R*>(dyn_lds);
  // the LJ table with 1/sigma_ab in place of sigma_ab
  for (int q = threadIdx.x; q < 4 * nt * nt; q += blockDim.x) {
    const R x = lj_tab[q];
    s_lj[q] = (q & 3) == 3 ? (x > R(0) ? R(1) / x : R(0)) : x;
  }
  for (int q = threadIdx.x; q < 3 * nt; q += blockDim.x) tab[q] = 0.f;
  const R gm[3] = {(R)v.gmean[0], (R)v.gmean[1], (R)v.gmean[2]};
  const int e1[3] = {1, 1, 1};
  const int brick = xcd_brick(blockIdx.x, g.nbrick);
  const Neigh h = brick_neigh(g, brick, e1);
  build_slots(g, h, d.key_start, nullptr, false, t);
  const int total = t.off[h.nslots];
  const bool staged = total < StageCap<R>::value;
  if (staged) {
    stage_atoms<R, true>(g, b, h, t, d.xs, stage);
    const int lane = threadIdx.x & 31;
    for (int s = threadIdx.x >> 5; s < h.nslots; s += blockDim.x >> 5) {
      const int o = t.off[s], len = t.off[s + 1] - o, f = t.first[s];
      for (int j = lane; j < len; j += 32) {
        const R4T<R> q = v.gs[f + j];
        R4T<R> w;
        w.x = q.x - gm[0];
        w.y = q.y - gm[1];
        w.z = q.z - gm[2];
        w.w = (R)0;
        reinterpret_cast<int*>(&w.w)[0] = d.gid_s[f + j];
        gstage[o + j] = w;
      }
    }
    __syncthreads();
  }
  constexpr int sh = StageShift<R>::value;
  const int i0 = d.key_start[2 * h.cell0], i1 = d.key_start[2 * h.cell1];
  // the brick's atoms in the pair kernel's order (longest list first)
  for (int k = (int)threadIdx.x; k < i1 - i0; k += blockDim.x) {
    const int i = d.order_a[i0 + k];
    const R4T<R> xi = d.xs[i];
    const int ti = (int)xi.w;
    const R4T<R> g4 = v.gs[i];
    const R gi[3] = {g4.x - gm[0], g4.y - gm[1], g4.z - gm[2]};
    const unsigned long long hi = RBL ? rbl_hi(d.gid_s[i]) : 0ull;
    const unsigned long long hji = RBL ? d.hj_s[i] : 0ull;
    const R* lrow = s_lj + 4 * ti * nt;
    const SpecRow sr = SPECIAL ? spec_row<R>(d, pc.ms, i) : SpecRow();
    R r[3] = {0, 0, 0};
    VjpAcc acc;
    const int cnt = d.cnt_a[i];
    // the hashes enter only for shell pairs (core pairs have weight 1 both
    // ways), so they are formed there alone
    auto one = [&](R dx, R dy, R dz, R u, int tj, const R gj[3], int gidj) {
      if (!(u < pc.rclj2) || !(u < pc.rc2)) return;
      R wij, wji;
      if (!RBL || u < pc.t_core) {
        wij = wji = vjp_w<R, RBL>(pc, u, 0ull, 0ull);
      } else {
        wij = vjp_w<R, RBL>(pc, u, hi, rbl_hj(gidj));
        wji = vjp_w<R, RBL>(pc, u, rbl_hi(gidj), hji);
      }
      if (wij == R(0) && wji == R(0)) return;
      const R ws = SPECIAL ? special_w<R>(sr, gidj, pc) : R(1);
      vjp_pair<R>(dx, dy, dz, u, lrow + 4 * tj, false, pc, R(0), R(0), wij, wji, gi,
                  gj, r, acc, ws);
    };
    // Without special bonds or list-A shell pairs, an entry's weight both
    // ways is the force window a_lo < u < a_hi: the entries (and the
    // sentinels that pad the last group) go without branches, a masked one
    // with finite stand-ins (u = 1, G'_j = 0, weights 0) that add exact zeros.
    const bool flat = !SPECIAL && (!RBL || !(pc.rclj2 > pc.rs2));
    if (flat && staged && cnt <= d.cap_a) {
      auto run = [&](auto mi) {
        constexpr bool MI = decltype(mi)::value;
        const uint4* lp = reinterpret_cast<const uint4*>(d.list_a) + i;
        for (int k0 = 0; k0 < cnt; k0 += 8) {
          const uint4 q8 = lp[(size_t)(k0 >> 3) * d.n];
          const unsigned wv[4] = {q8.x, q8.y, q8.z, q8.w};
#pragma unroll
          for (int kk = 0; kk < 8; ++kk) {
            const int l = (int)((wv[kk >> 1] >> (16 * (kk & 1))) & 0xffffu) >> sh;
            const R4T<R> xj = stage[l];
            const R4T<R> gq = gstage[l];
            R dx, dy, dz;
            const R u = pair_d2<R, MI>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
            const bool in = (u > pc.a_lo) & (u < pc.a_hi);
            const R w = in ? R(1) : R(0);
            const R gj[3] = {in ? gq.x : R(0), in ? gq.y : R(0), in ? gq.z : R(0)};
            const int tj = pair_woff(xj.w) / (4 * (int)sizeof(R));
            vjp_pair<R>(dx, dy, dz, in ? u : R(1), lrow + 4 * tj, false, pc, R(0), R(0), w,
                        w, gi, gj, r, acc, R(1));
          }
        }
      };
      if (h.mi)
        run(std::true_type());
      else
        run(std::false_type());
    } else if (staged && cnt <= d.cap_a) {
      const uint4* lp = reinterpret_cast<const uint4*>(d.list_a) + i;
      for (int k0 = 0; k0 < cnt; k0 += 8) {
        const uint4 q8 = lp[(size_t)(k0 >> 3) * d.n];
        const unsigned wv[4] = {q8.x, q8.y, q8.z, q8.w};
        const int kn = min(8, cnt - k0);
        for (int kk = 0; kk < kn; ++kk) {
          const int l = (int)((wv[kk >> 1] >> (16 * (kk & 1))) & 0xffffu) >> sh;
          const R4T<R> xj = stage[l];
          const R4T<R> gq = gstage[l];
          const int gidj = reinterpret_cast<const int*>(&gq.w)[0];
          const R gj[3] = {gq.x, gq.y, gq.z};
          R dx, dy, dz;
          const R u = h.mi ? pair_d2<R, true>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz)
                           : pair_d2<R, false>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
          const int tj = pair_woff(xj.w) / (4 * (int)sizeof(R));
          one(dx, dy, dz, u, tj, gj, gidj);
        }
      }
    } else {
      // unstaged brick or list overflow: scan the 27 cells directly
      int c[3];
      unpack_coord(g.coord_of_cell[d.cell_s[i]], c);
      const Stencil st = atom_stencil(g, h, c);
      const int os = own_slot(h, c);
      const bool qi = split && (i - t.first[os]) < t.nq[os];
      for (int kz = 0; kz < st.n[2]; ++kz)
        for (int ky = 0; ky < st.n[1]; ++ky)
          for (int kx = 0; kx < st.n[0]; ++kx) {
            const int sl_ = (st.at(2, kz) * h.cnt[1] + st.at(1, ky)) * h.cnt[0] +
                            st.at(0, kx);
            const int len = t.off[sl_ + 1] - t.off[sl_];
            for (int kk = qi ? t.nq[sl_] : 0; kk < len; ++kk) {
              const int jg = t.first[sl_] + kk;
              if (jg == i) continue;
              const R4T<R> xj = d.xs[jg];
              const R4T<R> gq = v.gs[jg];
              const R gj[3] = {gq.x - gm[0], gq.y - gm[1], gq.z - gm[2]};
              R dx, dy, dz;
              const R u = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
              one(dx, dy, dz, u, (int)xj.w, gj, d.gid_s[jg]);
            }
          }
    }
    v.drs[i] = make4<R>(r[0], r[1], r[2], R(0));
    vjp_flush<R>(tab, nt, ti, s_lj, acc);
  }
  // charged atoms over list B (LJ + Coulomb), kChargedLanes lanes per atom
  if (split) {
    __syncthreads();
    const int c0 = d.cstart[h.cell0], c1 = d.cstart[h.cell1];
    const int grp = threadIdx.x / kChargedLanes, gl = threadIdx.x % kChargedLanes;
    const int ngrp = blockDim.x / kChargedLanes;
    for (int c = c0 + grp; c < c1; c += ngrp) {
      const R4T<R> xi = d.xqc[c];
      const int ti = d.typec[c];
      const R4T<R> g4 = v.gc[c];
      const R gi[3] = {g4.x - gm[0], g4.y - gm[1], g4.z - gm[2]};
      const unsigned long long hi = RBL ? rbl_hi(d.gidc[c]) : 0ull;
      const unsigned long long hji = RBL ? d.hjc[c] : 0ull;
      const R* lrow = s_lj + 4 * ti * nt;
      const SpecRow sr = SPECIAL ? spec_row<R>(d, pc.ms, d.c2s[c]) : SpecRow();
      R r[3] = {0, 0, 0};
      VjpAcc acc;
      auto one = [&](int e) {
        const R4T<R> xj = d.xqc[e];
        R dx, dy, dz;
        const R u = pair_d2<R>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
        if (!(u < pc.rc2)) return;
        R wij, wji;
        if (!RBL || u < pc.t_core) {
          wij = wji = vjp_w<R, RBL>(pc, u, 0ull, 0ull);
        } else {
          wij = vjp_w<R, RBL>(pc, u, hi, d.hjc[e]);
          wji = vjp_w<R, RBL>(pc, u, rbl_hi(d.gidc[e]), hji);
        }
        if (wij == R(0) && wji == R(0)) return;
        const R4T<R> gq = v.gc[e];
        const R gj[3] = {gq.x - gm[0], gq.y - gm[1], gq.z - gm[2]};
        const int tj = d.typec[e];
        const R ws = SPECIAL ? special_w<R>(sr, d.gidc[e], pc) : R(1);
        vjp_pair<R>(dx, dy, dz, u, u < pc.rclj2 ? lrow + 4 * tj : nullptr, true, pc,
                    xi.w, xj.w, wij, wji, gi, gj, r, acc, ws,
                    wij != R(0) ? R(1) : R(0), wji != R(0) ? R(1) : R(0));
      };
      const int cnt = d.cnt_b[c];
      if (cnt <= d.cap_b) {
        const int* lp = d.list_b + (size_t)c * d.cap_b;
        for (int k = gl; k < cnt; k += kChargedLanes) one(lp[k]);
      } else {
        const int i = d.c2s[c];
        int cc[3];
        unpack_coord(g.coord_of_cell[d.cell_s[i]], cc);
        const Span sx = make_span(cc[0], g.ext_b[0], g.nc[0]);
        const Span sy = make_span(cc[1], g.ext_b[1], g.nc[1]);
        const Span sz = make_span(cc[2], g.ext_b[2], g.nc[2]);
        const int nst = sx.count * sy.count * sz.count;
        for (int s = gl; s < nst; s += kChargedLanes) {
          const int kx = s % sx.count, ky = (s / sx.count) % sy.count,
                    kz = s / (sx.count * sy.count);
          const int cell =
              g.cell_of_coord[(sz.at(kz) * g.nc[1] + sy.at(ky)) * g.nc[0] + sx.at(kx)];
          for (int e = d.cstart[cell]; e < d.cstart[cell + 1]; ++e)
            if (e != c) one(e);
        }
      }
      R fx = r[0], fy = r[1], fz = r[2];
      for (int o = kChargedLanes / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o, kChargedLanes);
        fy += __shfl_xor(fy, o, kChargedLanes);
        fz += __shfl_xor(fz, o, kChargedLanes);
        acc.dq += __shfl_xor(acc.dq, o, kChargedLanes);
        acc.de += __shfl_xor(acc.de, o, kChargedLanes);
        acc.ds += __shfl_xor(acc.ds, o, kChargedLanes);
      }
      if (gl == 0) {
        const int i = d.c2s[c];
        R4T<R> q = v.drs[i];
        q.x += fx;
        q.y += fy;
        q.z += fz;
        v.drs[i] = q;
        vjp_flush<R>(tab, nt, ti, s_lj, acc);
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * nt; q += blockDim.x)
    if (tab[q] != 0.f) atomicAdd(&v.pacc[q], (double)tab[q]);
}

// Charge gradient through Coulomb pairs of a charged and a neutral atom (the
// forward evaluates them without Coulomb, their charge product being 0): only
// Q of the neutral atom's type moves.  One workgroup per brick: the charged
// atoms of the brick's cells widened by the charged-pair stencil (ext_b) are
// staged in LDS (shifted across box edges as stage_atoms does), and each
// neutral atom of the brick, kQncLanes lanes per atom, scans the staged rows
// of cells within ext_b of its own cell.  A neighbourhood beyond the stage
// reads the same atoms from global memory.  With CVFF special bonds
// (SPECIAL, ws the neutral atom's weight for the pair, the same both ways) a
// kept pair also carries the charge derivative of the (1 - ws) full-Coulomb
// removal, (1 - ws) qqr2e q_j r^-3 along d for each directed pair the lists
// keep, as vjp_pair adds it for charged pairs.
constexpr int kQncSlots = 512;   // neighbourhood cells per brick
constexpr int kQncCap = 1024;    // staged charged atoms
constexpr int kQncLanes = 2;     // lanes per neutral atom

template <typename R, bool RBL, bool SPECIAL>
__global__ __launch_bounds__(256) void k_vjp_qnc(Dev<R> d, VjpDev<R> v, Box<R> b,
                                                 Grid g, PairConst<R> pc_arg) {
  PairConst<R> pc = pc_arg;
  if (RBL) pc.epoch = d.ctrl->epoch;
  __shared__ SlotTables<kQncSlots> t;
  __shared__ R4T<R> sx[kQncCap];  // (x, y, z, q), edge-shifted
  __shared__ R4T<R> sg[kQncCap];  // (G' x, y, z, global id bits)
  __shared__ float tab[kVjpMaxTypes];
  const int nt = pc.ntypes;
  for (int q = threadIdx.x; q < nt; q += blockDim.x) tab[q] = 0.f;
  const R gm[3] = {(R)v.gmean[0], (R)v.gmean[1], (R)v.gmean[2]};
  const int brick = xcd_brick(blockIdx.x, g.nbrick);
  const Neigh h = brick_neigh(g, brick, g.ext_b);
  const bool fits = h.nslots <= kQncSlots;
  int total = 0;
  if (fits) {
    build_slots(g, h, d.key_start, d.cstart, true, t);
    total = t.off[h.nslots];
  }
  const bool staged = fits && total <= kQncCap;
  if (staged) {
    const int lane = threadIdx.x & 31;
    for (int s = threadIdx.x >> 5; s < h.nslots; s += blockDim.x >> 5) {
      const int o = t.off[s], len = t.off[s + 1] - o, cf = t.cfirst[s];
      const int e = t.edge[s];
      R sh[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int w = (e >> (2 * q)) & 3;
        sh[q] = w == 0 ? -b.L[q] : (w == 2 ? b.L[q] : R(0));
      }
      for (int j = lane; j < len; j += 32) {
        R4T<R> p = d.xqc[cf + j];
        p.x += sh[0];
        p.y += sh[1];
        p.z += sh[2];
        sx[o + j] = p;
        const R4T<R> gq = v.gc[cf + j];
        R4T<R> w;
        w.x = gq.x - gm[0];
        w.y = gq.y - gm[1];
        w.z = gq.z - gm[2];
        w.w = (R)0;
        reinterpret_cast<int*>(&w.w)[0] = d.gidc[cf + j];
        sg[o + j] = w;
      }
    }
  }
  __syncthreads();
  const int i0 = d.key_start[2 * h.cell0], i1 = d.key_start[2 * h.cell1];
  const int sub = threadIdx.x % kQncLanes;
  for (int ib = i0; ib < i1; ib += blockDim.x / kQncLanes) {
    const int i = ib + threadIdx.x / kQncLanes;
    float acc = 0.f;
    int ti = 0;
    if (i < i1 && d.s2c[i] < 0) {
      const R4T<R> xi = d.xs[i];
      ti = (int)xi.w;
      const R4T<R> g4 = v.gs[i];
      const R gi[3] = {g4.x - gm[0], g4.y - gm[1], g4.z - gm[2]};
      const unsigned long long hi = RBL ? rbl_hi(d.gid_s[i]) : 0ull;
      const unsigned long long hji = RBL ? d.hj_s[i] : 0ull;
      const SpecRow sr = SPECIAL ? spec_row<R>(d, pc.ms, i) : SpecRow();
      // (owner i neutral, j charged): m = w_ij G'_i . d - w_ji G'_j . d
      auto one = [&](const R4T<R>& xj, const R gj[3], int gidj, bool mi) {
        R dx, dy, dz;
        const R u = mi ? pair_d2<R, true>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz)
                       : pair_d2<R, false>(b, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z, dx, dy, dz);
        if (!(u < pc.rc2)) return;
        R wij, wji;
        if (!RBL || u < pc.t_core) {
          wij = wji = vjp_w<R, RBL>(pc, u, 0ull, 0ull);
        } else {
          wij = vjp_w<R, RBL>(pc, u, hi, rbl_hj(gidj));
          wji = vjp_w<R, RBL>(pc, u, rbl_hi(gidj), hji);
        }
        if (wij == R(0) && wji == R(0)) return;
        R hh, hp;
        coul_vjp<R>(pc, u, hh, hp);
        const R m = wij * (gi[0] * dx + gi[1] * dy + gi[2] * dz) -
                    wji * (gj[0] * dx + gj[1] * dy + gj[2] * dz);
        acc += (float)(xj.w * (-pc.qqr2e * hh) * m);
        if (SPECIAL) {
          const R ws = special_w<R>(sr, gidj, pc);
          if (ws != R(1)) {
            const R rinv = pair_rsqrt(u);
            const R gc = (R(1) - ws) * pc.qqr2e * rinv * rinv * rinv;
            const R ms = (wij != R(0) ? R(1) : R(0)) * (gi[0] * dx + gi[1] * dy + gi[2] * dz) -
                         (wji != R(0) ? R(1) : R(0)) * (gj[0] * dx + gj[1] * dy + gj[2] * dz);
            acc += (float)(xj.w * gc * ms);
          }
        }
      };
      if (staged && !h.mi) {
        // rows of cells within ext_b of the owner's cell: contiguous along x
        int c[3];
        unpack_coord(g.coord_of_cell[d.cell_s[i]], c);
        int k0[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) k0[q] = wrapi(c[q] - h.start[q], g.nc[q]) - g.ext_b[q];
        const int nx = 2 * g.ext_b[0] + 1;
        for (int kz = 0; kz <= 2 * g.ext_b[2]; ++kz)
          for (int ky = 0; ky <= 2 * g.ext_b[1]; ++ky) {
            const int s0 = ((k0[2] + kz) * h.cnt[1] + k0[1] + ky) * h.cnt[0] + k0[0];
            const int l1 = t.off[s0 + nx];
            for (int l = t.off[s0] + sub; l < l1; l += kQncLanes) {
              const R4T<R> gq = sg[l];
              const R gj[3] = {gq.x, gq.y, gq.z};
              one(sx[l], gj, reinterpret_cast<const int*>(&gq.w)[0], false);
            }
          }
      } else if (staged) {
        for (int l = sub; l < total; l += kQncLanes) {
          const R4T<R> gq = sg[l];
          const R gj[3] = {gq.x, gq.y, gq.z};
          one(sx[l], gj, reinterpret_cast<const int*>(&gq.w)[0], true);
        }
      } else {
        // the stencil of the owner's cell from global memory, minimum image
        int cc[3];
        unpack_coord(g.coord_of_cell[d.cell_s[i]], cc);
        const Span spx = make_span(cc[0], g.ext_b[0], g.nc[0]);
        const Span spy = make_span(cc[1], g.ext_b[1], g.nc[1]);
        const Span spz = make_span(cc[2], g.ext_b[2], g.nc[2]);
        const int nst = spx.count * spy.count * spz.count;
        for (int s = sub; s < nst; s += kQncLanes) {
          const int kx = s % spx.count, ky = (s / spx.count) % spy.count,
                    kz = s / (spx.count * spy.count);
          const int cell = g.cell_of_coord[(spz.at(kz) * g.nc[1] + spy.at(ky)) * g.nc[0] +
                                           spx.at(kx)];
          for (int e = d.cstart[cell]; e < d.cstart[cell + 1]; ++e) {
            const R4T<R> gq = v.gc[e];
            const R gj[3] = {gq.x - gm[0], gq.y - gm[1], gq.z - gm[2]};
            one(d.xqc[e], gj, d.gidc[e], true);
          }
        }
      }
    }
    for (int o = kQncLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kQncLanes);
    if (sub == 0 && acc != 0.f) atomicAdd(&tab[ti], acc);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nt; q += blockDim.x)
    if (tab[q] != 0.f) atomicAdd(&v.pacc[q], (double)tab[q]);
}

// Adjoint structure sums: one k-vector per thread, atoms staged through LDS.
constexpr int kVjpSkBlock = 128;  // k-vectors per block
constexpr int kVjpSkTile = 256;   // atoms per LDS tile

// Per chunk (type-major runs of at most kSkChunk atoms of one type, the
// charged atoms' chunks first) and k: (T, U, rho, beta) with
// T = sum e^{i theta}, U = sum (G . k) e^{i theta}, rho = sum q e^{i theta},
// beta = sum q (G . k) e^{i theta}; a neutral chunk has rho = beta = 0.
// Float sums over a tile, FP64 across tiles.
template <typename R>
__global__ __launch_bounds__(kVjpSkBlock) void k_vjp_sk(Dev<R> d, VjpDev<R> v, Box<R> b,
                                                        const int* __restrict__ m, int K) {
  __shared__ float4 sa[kVjpSkTile];  // (s_x, s_y, s_z, q)
  __shared__ float4 sg[kVjpSkTile];  // (G_x, G_y, G_z, 0)
  const int ch = blockIdx.x;
  const int k = blockIdx.y * kVjpSkBlock + threadIdx.x;
  const int kk = k < K ? k : K - 1;
  const double two_pi = 6.283185307179586476925;
  const float mx = (float)m[3 * kk], my = (float)m[3 * kk + 1], mz = (float)m[3 * kk + 2];
  const float kx = (float)(two_pi * m[3 * kk] / (double)b.L[0]);
  const float ky = (float)(two_pi * m[3 * kk + 1] / (double)b.L[1]);
  const float kz = (float)(two_pi * m[3 * kk + 2] / (double)b.L[2]);
  const bool charged = ch < v.nch_c;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int a0 = v.ch_beg[ch], a1 = v.ch_beg[ch + 1];
  for (int t0 = a0; t0 < a1; t0 += kVjpSkTile) {
    const int nt = min(kVjpSkTile, a1 - t0);
    __syncthreads();
    for (int j = threadIdx.x; j < nt; j += blockDim.x) {
      const int i = v.tord[t0 + j];
      sa[j] = make_float4((float)((d.px[i] - b.lo[0]) * b.invL[0]),
                          (float)((d.py[i] - b.lo[1]) * b.invL[1]),
                          (float)((d.pz[i] - b.lo[2]) * b.invL[2]), (float)d.charge[i]);
      sg[j] = make_float4((float)v.gx[i], (float)v.gy[i], (float)v.gz[i], 0.f);
    }
    __syncthreads();
    float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (charged) {
      for (int j = 0; j < nt; ++j) {
        const float4 a = sa[j], g = sg[j];
        float sn, cs;
        sincos_rev(fma_r(mx, a.x, fma_r(my, a.y, mz * a.z)), sn, cs);
        const float gk = fma_r(g.x, kx, fma_r(g.y, ky, g.z * kz));
        const float qg = a.w * gk;
        f[0] += cs;
        f[1] += sn;
        f[2] = fma_r(gk, cs, f[2]);
        f[3] = fma_r(gk, sn, f[3]);
        f[4] = fma_r(a.w, cs, f[4]);
        f[5] = fma_r(a.w, sn, f[5]);
        f[6] = fma_r(qg, cs, f[6]);
        f[7] = fma_r(qg, sn, f[7]);
      }
    } else {
      for (int j = 0; j < nt; ++j) {
        const float4 a = sa[j], g = sg[j];
        float sn, cs;
        sincos_rev(fma_r(mx, a.x, fma_r(my, a.y, mz * a.z)), sn, cs);
        const float gk = fma_r(g.x, kx, fma_r(g.y, ky, g.z * kz));
        f[0] += cs;
        f[1] += sn;
        f[2] = fma_r(gk, cs, f[2]);
        f[3] = fma_r(gk, sn, f[3]);
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] += (double)f[q];
  }
  if (k < K) {
    double* o = v.vsk + ((size_t)ch * K + k) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = acc[q];
  }
}

// One block per k: FP64 sums of the chunk partials per group (the chunks of
// one type, charged or neutral), then the k-space charge gradient's terms of
// this k per type and the k's coefficients of the position pass.
template <typename R>
__global__ __launch_bounds__(256) void k_vjp_sk_red(Dev<R> d, VjpDev<R> v, Box<R> b,
                                                    const int* m, int K, Params p,
                                                    int ngrp) {
  __shared__ double gsum[2 * kVjpMaxTypes][8];
  __shared__ double smem[16 * 8];
  const int k = blockIdx.x;
  for (int g = 0; g < ngrp; ++g) {
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int ch = v.grp_beg[g] + threadIdx.x; ch < v.grp_beg[g + 1]; ch += blockDim.x) {
      const double* x = v.vsk + ((size_t)ch * K + k) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) s[q] += x[q];
    }
    block_sum<8>(s, smem);
    if (threadIdx.x == 0)
#pragma unroll
      for (int q = 0; q < 8; ++q) gsum[g][q] = s[q];
  }
  if (threadIdx.x != 0) return;
  double rr = 0, ri = 0, br = 0, bi = 0;
  for (int g = 0; g < ngrp; ++g) {
    rr += gsum[g][4];
    ri += gsum[g][5];
    br += gsum[g][6];
    bi += gsum[g][7];
  }
  const double two_pi = 6.283185307179586476925;
  const double kx = two_pi * m[3 * k] / (double)b.L[0];
  const double ky = two_pi * m[3 * k + 1] / (double)b.L[1];
  const double kz = two_pi * m[3 * k + 2] / (double)b.L[2];
  const double k2 = kx * kx + ky * ky + kz * kz;
  const double A = d.kw[k] * (-4.0 * 3.141592653589793238463 * p.qqr2e) / (p.volume * k2);
  double dq[kVjpMaxTypes];
  for (int a = 0; a < kVjpMaxTypes; ++a) dq[a] = 0.0;
  for (int g = 0; g < ngrp; ++g) {
    const double* s = gsum[g];
    const double c = A * ((s[2] * ri - s[3] * rr) + (br * s[1] - bi * s[0]));
    const int a = v.grp_type[g];
    for (int q = 0; q < kVjpMaxTypes; ++q)
      if (q == a) dq[q] += c;
  }
  for (int q = 0; q < kVjpMaxTypes; ++q) v.dqk[(size_t)k * kVjpMaxTypes + q] = dq[q];
  v.kv0[k] = make_float4((float)m[3 * k], (float)m[3 * k + 1], (float)m[3 * k + 2],
                         (float)(A * br));
  v.kv1[k] = make_float4((float)kx, (float)ky, (float)kz, (float)(A * bi));
  v.kv2[k] = make_float2((float)(A * rr), (float)(A * ri));
}

// k-space part of dL/dr for the charged atoms (charged order); the per-k
// coefficients pass through LDS in tiles of kVjpKTile.
constexpr int kVjpKTile = 256;
template <typename R>
__global__ __launch_bounds__(256) void k_vjp_kforce(Dev<R> d, VjpDev<R> v,
                                                    Box<R> b, int K) {
  __shared__ float4 s0[kVjpKTile], s1[kVjpKTile];
  __shared__ float2 s2[kVjpKTile];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = c < d.nc;
  float sx = 0.f, sy = 0.f, sz = 0.f, gx = 0.f, gy = 0.f, gz = 0.f, q = 0.f;
  if (live) {
    const R4T<R> p = d.xqc[c];
    frac_coords<R>(b, p, sx, sy, sz);
    const R4T<R> g4 = v.gc[c];
    gx = (float)g4.x;
    gy = (float)g4.y;
    gz = (float)g4.z;
    q = (float)p.w;
  }
  float rx = 0.f, ry = 0.f, rz = 0.f;
  for (int k0 = 0; k0 < K; k0 += kVjpKTile) {
    const int kn = min(kVjpKTile, K - k0);
    __syncthreads();
    for (int j = threadIdx.x; j < kn; j += blockDim.x) {
      s0[j] = v.kv0[k0 + j];
      s1[j] = v.kv1[k0 + j];
      s2[j] = v.kv2[k0 + j];
    }
    __syncthreads();
    for (int k = 0; k < kn; ++k) {
      const float4 a = s0[k];
      const float4 e = s1[k];
      const float2 rh = s2[k];
      const float t = fma_r(a.x, sx, fma_r(a.y, sy, a.z * sz));
      float sn, cs;
      sincos_rev(t, sn, cs);
      const float gk = fma_r(gx, e.x, fma_r(gy, e.y, gz * e.z));
      const float s = fma_r(a.w, cs, e.w * sn) - gk * fma_r(rh.x, cs, rh.y * sn);
      rx = fma_r(s, e.x, rx);
      ry = fma_r(s, e.y, ry);
      rz = fma_r(s, e.z, rz);
    }
  }
  if (live) v.drk[c] = make4<R>((R)(q * rx), (R)(q * ry), (R)(q * rz), R(0));
}

// dL/dr in caller order; keeps the evaluation counter in step with the
// forward path (the step ring is indexed by it).
template <typename R>
__global__ __launch_bounds__(256) void k_vjp_assemble(Dev<R> d, VjpDev<R> v, int has_k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    d.ctrl->maxdisp_bits = 0u;
    d.ctrl->step_count += 1;
    if (d.ctrl_host) d.ctrl_host->step_count = d.ctrl->step_count;
  }
  if (i >= d.n) return;
  R4T<R> r = v.drs[d.inv_perm[i]];
  const int c = d.c_of[i];
  if (has_k && c >= 0) {
    const R4T<R> k = v.drk[c];
    r.x += k.x;
    r.y += k.y;
    r.z += k.z;
  }
  v.dx[i] = r.x;
  v.dy[i] = r.y;
  v.dz[i] = r.z;
}

// Parameter gradients: the pair passes' FP64 per-type sums (each workgroup
// adds its LDS table with FP64 atomics) plus the k-space charge terms per k,
// summed in a fixed order.  dparam = [dL/dQ | dL/deps | dL/dsigma].
template <typename R>
__global__ __launch_bounds__(256) void k_vjp_params(VjpDev<R> v, int nt, int K) {
  __shared__ double smem[16 * kVjpMaxTypes];
  double s[kVjpMaxTypes];
#pragma unroll
  for (int q = 0; q < kVjpMaxTypes; ++q) s[q] = 0.0;
  for (int k = threadIdx.x; k < K; k += blockDim.x)
#pragma unroll
    for (int q = 0; q < kVjpMaxTypes; ++q) s[q] += v.dqk[(size_t)k * kVjpMaxTypes + q];
  block_sum<kVjpMaxTypes>(s, smem);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 3 * kVjpMaxTypes; ++q)
      if (q < 3 * nt) v.dparam[q] = v.pacc[q] + (q < nt ? s[q % kVjpMaxTypes] : 0.0);
}

// The rebuild part of an evaluation alone: skin check and gather, and the
// rebuild when the decision says so (lists and sorted tiles current).
template <typename R>
hipError_t launch_skin_build(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s) {
  const int sb = blocks_for(d.n, kBlock, kSkinBlocksMax);
  const int ctl = d.hdr_ring ? 1 : 0;
  hipLaunchKernelGGL((k_skin<R>), dim3(sb + ctl), dim3(kBlock), 0, s, d, a.box, 0, ctl, 0, 0);
  return build_part<R>(d, a, sb + ctl, s);
}

template <typename R>
hipError_t launch_vjp(const Dev<R>& d, const VjpDev<R>& v, const StepArgs<R>& a,
                      hipStream_t s) {
  const Params& p = a.p;
  const Grid& g = a.grid;
  const int nt = a.pc.ntypes;
  hipLaunchKernelGGL((k_vjp_gather<R>), dim3(v.gather_blocks), dim3(kBlock), 0, s, d, v);
  hipLaunchKernelGGL(k_vjp_mean, dim3(1), dim3(256), 0, s, v.gpart, v.gather_blocks,
                     (double)d.n, p.mean_fix, v.gmean);
  const size_t lds = 4 * sizeof(R) * (size_t)nt * nt;
#define RBMDX_PAIR_VJP(RBL, SP)                                                 \
  hipLaunchKernelGGL((k_pair_vjp<R, RBL, SP>), dim3(g.nbrick), dim3(g.block), lds, s, \
                     d, v, a.box, g, a.pc, a.lj_tab, p.split)
  if (p.special) {
    if (p.rbl)
      RBMDX_PAIR_VJP(true, true);
    else
      RBMDX_PAIR_VJP(false, true);
  } else if (p.rbl) {
    RBMDX_PAIR_VJP(true, false);
  } else {
    RBMDX_PAIR_VJP(false, false);
  }
#undef RBMDX_PAIR_VJP
  if (v.neutral_q && v.nq_blocks > 0) {
#define RBMDX_QNC(RBLv, SPv)                                                            \
  hipLaunchKernelGGL((k_vjp_qnc<R, RBLv, SPv>), dim3(g.nbrick), dim3(256), 0, s, d, v, \
                     a.box, g, a.pc)
    if (p.special) {
      if (p.rbl)
        RBMDX_QNC(true, true);
      else
        RBMDX_QNC(false, true);
    } else if (p.rbl) {
      RBMDX_QNC(true, false);
    } else {
      RBMDX_QNC(false, false);
    }
#undef RBMDX_QNC
  }
  const int K = p.kspace ? p.num_k : 0;
  const bool has_k = K > 0 && d.nc > 0;
  // with the neutral atoms' charges held at zero, only the charged chunks
  const int nch = v.neutral_q ? v.nch : v.nch_c;
  const int ngrp = v.neutral_q ? v.ngrp : v.ngrp_c;
  if (has_k) {
    hipLaunchKernelGGL((k_vjp_sk<R>), dim3(nch, (K + kVjpSkBlock - 1) / kVjpSkBlock),
                       dim3(kVjpSkBlock), 0, s, d, v, a.box, (const int*)d.m_dev, K);
    hipLaunchKernelGGL((k_vjp_sk_red<R>), dim3(K), dim3(256), 0, s, d, v, a.box,
                       (const int*)d.m_dev, K, p, ngrp);
    hipLaunchKernelGGL((k_vjp_kforce<R>), dim3(blocks_for(d.nc)), dim3(kBlock), 0, s, d,
                       v, a.box, K);
  }
  hipLaunchKernelGGL((k_vjp_assemble<R>), dim3(blocks_for(d.n)), dim3(kBlock), 0, s, d,
                     v, has_k ? 1 : 0);
  hipLaunchKernelGGL((k_vjp_params<R>), dim3(1), dim3(256), 0, s, v, nt, has_k ? K : 0);
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

namespace rbmdx {
template hipError_t launch_skin_build<float>(const Dev<float>&, const StepArgs<float>&,
                                             hipStream_t);
template hipError_t launch_skin_build<double>(const Dev<double>&, const StepArgs<double>&,
                                              hipStream_t);
template hipError_t launch_vjp<float>(const Dev<float>&, const VjpDev<float>&,
                                      const StepArgs<float>&, hipStream_t);
template hipError_t launch_vjp<double>(const Dev<double>&, const VjpDev<double>&,
                                       const StepArgs<double>&, hipStream_t);
}  // namespace rbmdx

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free force evaluation.  The array's name
// is only a host-side stand-in: host code uses the address that
// hipGetSymbolAddress returns on each device (each device holds its own copy
// of the array), kept per device below.
// ---------------------------------------------------------------------------
namespace rbmdx {
namespace {
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kScratchDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kScratchDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  scratch_ptr[device] = ptr;
  return hipSuccess;
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_ptr[device];
}
}  // namespace rbmdx
