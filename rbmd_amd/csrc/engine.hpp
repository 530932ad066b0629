// Internal interface between the host engine (api.cpp) and the HIP kernels
// (kernels.hip).  Plain structs of device pointers and scalars; every launch
// wrapper takes the stream it runs on.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <type_traits>

namespace rbmdx {

constexpr int kBlock = 256;      // threads per block for per-atom kernels
constexpr int kSkinBlocksMax = 1024;  // skin-check grid (grid-stride beyond)
constexpr int kOrderMax = 512;        // bricks of more atoms keep the cell order
constexpr int kMaxTypes = 32;    // LJ type table held in LDS
constexpr int kMaxSpecial = 64;  // CVFF special partners per atom
constexpr int kSkTile = 8;       // k-vectors per S(k) block
constexpr int kSkChunk = 1024;   // charged atoms per S(k) block
constexpr int kRbeWaves = 8;     // waves per RBE-force block (k split)
constexpr int kRbeStage = 256;   // k coefficients per LDS stage of the RBE force
constexpr int kSkAtoms = 256;    // charged atoms per S(k) block (4 waves of 64)
constexpr int kSkK = 64;         // k-vectors per S(k) block and per reduction block
constexpr int kMaxK = 1 << 20;   // k-vector capacity bound
constexpr int kMaxSlots = 64;    // neighbourhood cells staged per brick
constexpr int kChargedLanes = 16;  // lanes per charged atom (Coulomb pairs)
constexpr int kQWaves = 4;  // waves per block of the dense charged-pair kernel
// Per-brick slot table of the last rebuild (written by the list build, read
// by the pair kernel): slot offsets [kMaxSlots + 1], first sorted rows, charged
// counts and box-edge shifts [kMaxSlots each].
constexpr int kBrickTabInts = 4 * kMaxSlots + 1;

// Device control block: rebuild decision, overflow bookkeeping.
struct Ctrl {
  unsigned int maxdisp_bits;  // max |dr|^2 since the last rebuild (float bits)
  int rebuild;                // unused by kernels; host mirror
  int max_count_a;            // largest all-atom candidate count seen
  int max_count_b;            // largest charged-pair candidate count seen
  unsigned long long rebuild_count;
  unsigned long long step_count;
  unsigned long long sum_count_a;  // stored candidates of the last rebuild
  unsigned long long sum_count_b;
  int max_stage;              // largest brick neighbourhood seen
  int unstaged_bricks;        // bricks whose neighbourhood exceeded LDS
  long long epoch;            // RBL sampling epoch of the current evaluation
};

// Per-evaluation parameters, written by the host into a host-mapped ring of
// kSlots entries; evaluation c reads entry c mod kSlots (c = step_count), so
// the same launch sequence (or captured graph) serves every evaluation.
struct StepHdr {
  long long epoch;  // RBL sampling epoch (the caller's step)
  int force;        // force a neighbour rebuild
  int pad;
};
constexpr int kSlots = 16;

// Blocks of the charged-pair list build (kChargedLanes lanes per atom) with
// `block` threads per block.
inline int list_b_blocks(int nc, int block) {
  return (nc + block / kChargedLanes - 1) / (block / kChargedLanes);
}


template <typename R>
struct Box {
  R lo[3];
  R hi[3];  // lo + L as stored by the caller's box (upper bound, exclusive)
  R L[3];
  R half[3];
  R invL[3];
};

// Cell grid.  Cells are numbered brick-major: the cells of one brick (a
// bw x bw x bw block of cells, ragged at the upper edges) are consecutive, so
// the cell-sorted atoms of a brick form one contiguous range.
struct Grid {
  int nc[3];       // cells per dimension
  int ncell;
  int ext_b[3];    // stencil half-width for the charged-pair list
  float scale[3];  // nc / L (for cell index)
  int bw;          // brick width in cells
  int nb[3];       // bricks per dimension
  int nbrick;
  int block;       // threads per brick block
  // a wave of the pair kernel's brick block that takes the brick's charged
  // pairs while the others take list A (0 or 64 threads; mostly neutral
  // systems with one workgroup per brick)
  int qwave = 0;
  int psub = 1;    // pair-kernel workgroups per brick (they share list B)
  // blocks of the dense charged-pair kernel (one wave per charged atom) that
  // takes list B in mostly charged systems; 0: list B inside the pair kernel
  int qblocks = 0;
  // blocks per cell of the dense charged-pair list build (64 charged atoms
  // of the cell per block and pass)
  int dsplit = 1;
  const int* cell_of_coord = nullptr;  // [ncell] (z*ny+y)*nx+x -> cell id
  const int* coord_of_cell = nullptr;  // [ncell] cell id -> packed x|y<<10|z<<20
  const int* brick_cell0 = nullptr;    // [nbrick+1] first cell id of a brick
};

// Blocks of the charged-pair list build: dsplit per cell in mostly charged
// systems, else kChargedLanes lanes per charged atom.
inline int list_b_count(const Grid& g, int nc) {
  return g.qblocks ? g.ncell * g.dsplit : list_b_blocks(nc, g.block);
}

template <typename R>
struct PairConst {
  R t_core, t_shell;  // RBL classification thresholds r_core^2-eps, rc^2-eps
  R rs2, rc2, rclj2;  // force windows
  R eps_min;          // coincident-pair guard (EPSILON)
  // force window of list-A pairs when the list holds no RBL shell pairs:
  // a_lo < d^2 < a_hi (min of the LJ and Coulomb cutoffs and, with RBL, of
  // the core thresholds; a_lo the guard with RBL, else -1)
  R a_lo, a_hi;
  R r2_list_a, r2_list_b;  // (cut + skin)^2 of the two lists
  R alpha, g, qqr2e, two_g_over_sqrtpi;
  R freq;             // RBL selection frequency as a multiplier
  int ifreq;
  // SOG near field (VERLET-SOG) instead of erfc: force polynomial T[0..5] in
  // r^2 and, for energies, far_m Gaussians (omega_l, 1/(2 s_l^2)) in far_tab
  int sog = 0;
  R T[6];
  const R* far_tab = nullptr;
  int far_m = 0;
  unsigned fmask;     // ifreq - 1 when ifreq is a power of two, else 0
  unsigned c32;       // 2^32 mod ifreq
  long long epoch;
  int ntypes;
  // CVFF special bonds: weights of 1-2, 1-3, 1-4 partners and the row width
  R w12 = 1, w13 = 1, w14 = 1;
  int ms = 0;
};

// All device buffers of one context (sorted order unless noted).
template <typename R>
struct Dev {
  using R4 = typename std::conditional<sizeof(R) == 4, float4, double4>::type;
  int n = 0, nc = 0;  // atoms, charged atoms
  // caller order
  const R* px = nullptr;
  const R* py = nullptr;
  const R* pz = nullptr;
  R* fx = nullptr;
  R* fy = nullptr;
  R* fz = nullptr;
  R* vx = nullptr;         // [n] velocities, fused leapfrog update only
  R* vy = nullptr;
  R* vz = nullptr;
  R* xw = nullptr;         // [n] writable positions (same arrays as px..pz)
  R* yw = nullptr;
  R* zw = nullptr;
  R* inv_mass = nullptr;   // [ntypes]
  const double* md_mass = nullptr;  // [ntypes] FP64 masses (MD closing half-kick)
  double* md_part = nullptr;        // [assemble workgroups] sum m v^2 (MD step)
  int* type = nullptr;     // [n]
  int* gid = nullptr;      // [n]
  R* charge = nullptr;     // [n]
  R4* ref = nullptr;       // [n] positions at the last rebuild
  int4* spn = nullptr;     // [n] CVFF special counts (n12, n13, n14, 0)
  int* spid = nullptr;     // [n * ms] CVFF special partners (global ids)
  int* key = nullptr;      // [n] cell key of the last rebuild
  int* slot = nullptr;     // [n] arrival slot inside the key
  int* inv_perm = nullptr; // [n] caller -> sorted
  int* c_of = nullptr;     // [n] caller -> charged compact, -1 if neutral
  // grid
  int* key_count = nullptr;  // [2*ncell padded]
  int* key_start = nullptr;  // [2*ncell+1]
  int* cstart = nullptr;     // [ncell+1] charged atoms before the cell
  // sorted order
  int* perm = nullptr;     // [n] sorted -> caller
  int* cell_s = nullptr;   // [n]
  int* gid_s = nullptr;    // [n]
  unsigned long long* hj_s = nullptr;  // [n] mix64(gid + 0xd1b54a32d192ed03)
  unsigned long long* hjc = nullptr;   // [nc] the same, charged order
  int* s2c = nullptr;      // [n] sorted -> charged compact, -1 if neutral
  R4* xs = nullptr;        // [n] (x, y, z, type)
  int4* spn_s = nullptr;   // [n] spn in sorted order
  int* spid_s = nullptr;   // [n * ms] spid in sorted order
  R4* fa = nullptr;        // [n] list-A pair force, caller order
  // charged compact
  int* c2s = nullptr;      // [nc]
  int* c2i = nullptr;      // [nc] caller index of charged atom c (rebuild order)
  int* qcall = nullptr;    // [nc] caller index of the k-space order's atom
  int* gidc = nullptr;     // [nc]
  int* typec = nullptr;    // [nc]
  R4* xqc = nullptr;       // [nc] (x, y, z, q)
  R4* fb = nullptr;        // [n] charged-pair (list B) force, caller order (0: neutral)
  // charged atoms in a fixed order (ascending caller index), independent of
  // the cell sort: the k-space kernels' input and output
  int* qpos = nullptr;     // [n] caller -> fixed charged index, -1 if neutral
  R4* xqs = nullptr;       // [nc] (x, y, z, q) in the fixed order
  R4* fk = nullptr;        // [n] k-space force, caller order (0: neutral)
  // lists
  int cap_a = 0, cap_b = 0;
  int* brick_tab = nullptr;    // [nbrick * kBrickTabInts] slot tables (list build)
  uint16_t* list_a = nullptr;  // [cap_a * n] groups of 8 (list_a_index), stage byte offsets
  int* cnt_a = nullptr;
  int* order_a = nullptr;  // [n] per brick: its atoms by list-A length, longest first
  int* list_b = nullptr;       // [nc * cap_b] row-major, charged compact index
  int* cnt_b = nullptr;
  // list B of a mostly charged system: the type of each entry's atom
  unsigned char* list_t = nullptr;  // [nc * cap_b]
  // pair partial sums per block: 3 force + 2 energy + 6 virial
  float* skin_part = nullptr;   // [skin blocks] per-block max |dr|^2
  int4* lista_part = nullptr;   // [nbrick] (max count, sum, stage size, unstaged)
  int2* listb_part = nullptr;   // [list-B blocks] (max count, sum)
  double* pair_part = nullptr;
  double* pairb_part = nullptr;  // [11 * qblocks] dense charged-pair partials
  int pair_part_blocks = 0;
  // k-space
  int kcap = 0;
  int* m_dev = nullptr;          // [kcap*3]
  int* src_dev = nullptr;        // [kcap] virial-tensor source k (RBSOG)
  double* kw = nullptr;          // [kcap] per-k weight
  double2* kvir = nullptr;       // [kcap] (tau, divide by |k_src|^4) (RBSOG)
  double2* sk_part = nullptr;    // [chunks * kcap]
  double* kfin_part = nullptr;   // [kcap * 8] k-space energy/virial terms per k
  float4* kf0 = nullptr;         // [kcap] (mx, my, mz, A Re rho)
  float4* kf1 = nullptr;         // [kcap] (kx, ky, kz, A Im rho)
  double2* rho = nullptr;        // [kcap]
  // energy_rbe_flag diagnostic: the exact Ewald k-space energy over a kmax
  // grid on energy evaluations (weights exp(-k^2 / 4 alpha))
  int kex = 0;
  int* m_ex = nullptr;           // [kex*3]
  double* kw_ex = nullptr;       // [kex]
  double2* sk_ex = nullptr;      // [chunks * kex]
  double* thermo = nullptr;      // [32]: [0..10] pair sums (F, E_lj, E_coul,
                                 // virial), [16..22] k-space (E, virial)
  double* mean = nullptr;        // [3]
  Ctrl* ctrl = nullptr;
  Ctrl* ctrl_host = nullptr;     // host-mapped mirror written each call
  const StepHdr* hdr_ring = nullptr;  // [kSlots] host-mapped step parameters
  const int* stage_ring = nullptr;    // [kSlots * stage_stride] host-mapped k-samples
  int stage_stride = 0;
  int* export_count = nullptr;
  int* export_buf = nullptr;     // [3*cap]
};

struct Params {
  int rbl = 1;          // RBL split (else full Verlet list)
  int coulomb = 0;      // real-space erfc + k-space present
  int kspace = 0;       // 0 none, 1 RBE, 2 EWALD, 3 RBSOG
  int eflag = 0;
  int mean_fix = 0;
  int split = 0;        // charged/neutral partition of the pair work
  int special = 0;      // CVFF special-bond weights apply
  int num_k = 0;
  double qqr2e = 1.0;
  double alpha = 0.0;
  double volume = 0.0;
  double sum_q2 = 0.0, sum_q = 0.0;
  double n_global = 0.0;
  float skin_half2 = 0.f;  // (skin/2)^2
};

// Everything one force evaluation needs besides the device buffers.
template <typename R>
struct StepArgs {
  Box<R> box;
  Grid grid;
  PairConst<R> pc;
  Params p;
  const R* lj_tab = nullptr;  // [ntypes^2 * 4] (-24 eps_ij, sigma_ij^2, eps_ij, 0), device
  const int* m_use = nullptr; // [num_k * 3] k-vectors for this call, device
  int m_from_ring = 0;        // k-sample from the step's ring slot (else m_use)
  unsigned int rebuild_thr;   // float bits of (skin/2)^2
  int force_rebuild = 0;
  int count_only = 0;         // list builders only count (capacity planning)
  double e_self = 0.0;        // k-space self + background energy
  double dt = 0.0;            // fused leapfrog: x += dt v
  double dt_ftm = 0.0;        //                 v += dt ftm2v F / m
  int md_kick = 0;            // MD step: closing half-kick + kinetic energy
  double ftm2v = 1.0;         //          (dt above is the time step)
  hipEvent_t ev[6] = {};      // optional section timing events
};

// Launch one complete force evaluation (kernels.hip).
template <typename R>
hipError_t launch_force(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s);

// Static per-device scratch of the allocation-free force evaluation
// (kernels.hip): a file-scope __device__ array of this many bytes holds the
// per-evaluation partial results (reduction partials, S(k) chunk partials,
// per-k coefficients) that a context otherwise allocates.  Every call on one
// device shares the device's copy.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
// Resolves the array's address on `device`, which must be the current
// device; the first call per device does the lookup, later calls keep it.
hipError_t noalloc_scratch_init(int device);
// The address resolved for `device` by noalloc_scratch_init, else nullptr.
void* noalloc_scratch(int device);
// Launch only the (forced) rebuild: skin, binning, sort, gather, lists.
template <typename R>
hipError_t launch_build(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s);
// Parity export of classified pairs (candidates = 0) or raw candidate lists.
template <typename R>
hipError_t launch_export(const Dev<R>& d, const StepArgs<R>& a, int candidates,
                         int cap, hipStream_t s);

// Force-matching adjoint: buffers of one context (allocated on first use).
constexpr int kVjpMaxTypes = 8;  // types of the adjoint's fixed-size per-type tables
template <typename R>
struct VjpDev {
  using R4 = typename Dev<R>::R4;
  const R* gx = nullptr;   // [n] cotangent dL/dF, caller order
  const R* gy = nullptr;
  const R* gz = nullptr;
  R* dx = nullptr;         // [n] dL/dr out, caller order
  R* dy = nullptr;
  R* dz = nullptr;
  double* dparam = nullptr;  // [3 * ntypes] out: dL/dQ, dL/deps, dL/dsigma
  R4* gs = nullptr;        // [n] cotangent, sorted order
  R4* gc = nullptr;        // [nc] cotangent, charged order
  R4* drs = nullptr;       // [n] pair part of dL/dr, sorted order
  R4* drk = nullptr;       // [nc] k-space part of dL/dr, charged order
  double* gpart = nullptr; // [gather blocks * 3] partial sums of G
  double* gmean = nullptr; // [3] mean cotangent (mean-force fix)
  double* pacc = nullptr;  // [3 * kVjpMaxTypes] pair passes' per-type sums
  int nq_blocks = 0;       // workgroups of the neutral-charged charge pass
  int neutral_q = 0;       // this call: also dL/dQ through the neutral atoms
  int* tord = nullptr;     // [n] caller indices: charged atoms type-major, then
                           // neutral atoms type-major
  int* ch_beg = nullptr;   // [nch + 1] chunk starts (one type, one of the two parts)
  int nch = 0, nch_c = 0;  // chunks, of them charged (first)
  int* grp_beg = nullptr;  // [ngrp + 1] chunk ranges of one (part, type)
  int* grp_type = nullptr; // [ngrp]
  int ngrp = 0, ngrp_c = 0;
  double* vsk = nullptr;   // [nch * kcap * 8] per chunk and k: T, U, rho, beta
  float4* kv0 = nullptr;   // [kcap] (m, A Re beta)
  float4* kv1 = nullptr;   // [kcap] (k, A Im beta)
  float2* kv2 = nullptr;   // [kcap] (A Re rho, A Im rho)
  double* dqk = nullptr;   // [kcap * kVjpMaxTypes] k-space dL/dQ terms per k
  int gather_blocks = 0;
};
template <typename R>
hipError_t launch_vjp(const Dev<R>& d, const VjpDev<R>& v, const StepArgs<R>& a,
                      hipStream_t s);
// Skin check, gather and (when decided) rebuild of one evaluation.
template <typename R>
hipError_t launch_skin_build(const Dev<R>& d, const StepArgs<R>& a, hipStream_t s);

}  // namespace rbmdx
