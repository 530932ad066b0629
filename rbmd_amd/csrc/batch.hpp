// Internal interface between the replica-batch host engine (batch.cpp) and
// its HIP kernels (batch.hip).  A batch holds many independent periodic
// boxes ("replicas") of one force field; every stage kernel covers a range of
// replicas in one launch.  FP32 positions and forces, FP64 energies and
// virials per replica.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "engine.hpp"

namespace rbmdx {

constexpr int kBatchStageCap = 1536;  // atoms staged per brick (LDS)
constexpr int kBatchMaxSlots = 64;    // neighbourhood cells per brick (bw <= 2)
constexpr int kBatchQSlots = 512;     // cells of the charged pass's neighbourhood
constexpr int kBatchMaxK = 512;       // k-vectors per replica
constexpr int kBatchMaxTypes = 8;     // atom types (LDS LJ table)
constexpr int kBatchAcc = 11;         // per-replica pair sums: F (3), E_lj, E_c, W (6)
constexpr int kBatchKAcc = 7;         // per-replica k-space sums: E, W (6)

// Constants of one replica (device array, one entry per replica).
struct RepInfo {
  int a0, n;        // atoms [a0, a0 + n), caller and sorted order alike
  int q0, nq;       // charged atoms [q0, q0 + nq) of the compact charged order
  int c0;           // first global cell id; cells are x-fastest per replica
  int nc[3];        // cells per dimension (side >= the LJ pair radius)
  int ext_c[3];     // charged-pair stencil half-width in cells
  int nb[3];        // bricks per dimension (bw^3 cells each, ragged at the top)
  Box<float> box;
  float freq;       // RBL shell-selection frequency of this replica
  int ifreq;
  unsigned fmask, c32;
  double volume;
  double kw;        // weight of every sampled k: S / P (RBE), S / (4 pi P) (RBSOG)
  double ktau;      // RBSOG: virial coefficient of a sampled k, -(S_npt / P) qqr2e / (4 V)
};

// One work item of the pair stage: a brick of one replica.
struct BatchWork {
  int rep, brick;
};

struct BatchDev {
  int R = 0, N = 0, NQ = 0, ncell = 0, nwork = 0, P = 0;
  const RepInfo* rep = nullptr;   // [R]
  const int* rep_of = nullptr;    // [N] caller atom -> replica
  const int* type = nullptr;      // [N] caller order
  const int* gid = nullptr;       // [N] caller order, ids within the replica
  const float* charge = nullptr;  // [N] caller order
  const BatchWork* work = nullptr;  // [nwork], replica-major
  const float* lj = nullptr;      // [4 ntypes^2] (-24 eps, sigma^2, eps, sigma)
  // per evaluation
  const float* px = nullptr;
  const float* py = nullptr;
  const float* pz = nullptr;
  float* fx = nullptr;
  float* fy = nullptr;
  float* fz = nullptr;
  const int* m = nullptr;         // [R * mstride] k-vectors (and RBSOG sources), replica-major
  int mstride = 0;                // ints per replica in m: 3 P, or 4 P for RBSOG (the P
                                  // k-vectors, then their virial-chain sources)
  int kmode = 0;                  // 0 RBE (weight kw), 1 EWALD (the kmax grid weighted
                                  // exp(-k^2 / 4 alpha)), 2 RBSOG (P sampled k's with kw
                                  // and nd direct vectors with per-replica weights)
  int nd = 0;                     // RBSOG direct vectors (|m|^2 <= 1)
  const int* dm = nullptr;        // [3 nd]
  const double* kwd = nullptr;    // [R * nd] direct-vector weights f(k) k^2 / (4 pi)
  const double* kvd = nullptr;    // [R * nd] their virial coefficients -qqr2e f_npt(k) / (4 V)
  // CVFF special bonds (caller order; partner ids are ids within the replica)
  const int4* spn = nullptr;      // [N] (n12, n13, n14, 0)
  const int* spid = nullptr;      // [N * ms]
  // scratch
  int* cnt = nullptr;             // [ncell] atoms per cell (zero between calls)
  int* cntq = nullptr;            // [ncell] charged atoms per cell (likewise)
  int* start = nullptr;           // [ncell + 1] first sorted atom of a cell
  int* qstart = nullptr;          // [ncell + 1] first compact charged atom
  int* cell_of = nullptr;         // [N] caller atom -> global cell
  int* slot = nullptr;            // [N] caller atom -> arrival rank in its cell
  int* tmp = nullptr;             // [N] cell-bucketed caller atoms
  float4* xs = nullptr;           // [N] sorted (x, y, z wrapped into the box, q)
  int2* tg = nullptr;             // [N] sorted (type, id)
  unsigned long long* hj = nullptr;  // [N] sorted RBL hash of the id
  int* perm = nullptr;            // [N] sorted -> caller
  int* qidx = nullptr;            // [NQ] compact charged -> sorted
  float4* fs = nullptr;           // [N] sorted pair + k-space force
  double* racc = nullptr;         // [R * kBatchAcc]
  double* kacc = nullptr;         // [R * kBatchKAcc]
  long long* epoch = nullptr;     // [1] RBL sampling epoch of the evaluation
};

// Launch shapes chosen by the autotuner.
struct BatchTune {
  int bw = 2;            // brick width in cells (1 or 2)
  int pair_block = 128;  // threads per pair workgroup
  int grid_per_cu = 2;   // persistent pair workgroups per CU
  int qlanes = 8;        // lanes per charged atom in the Coulomb pass
};

// Replica range [r0, r1): atoms, cells and work items are the contiguous
// ranges of those replicas (offsets given by the host).
struct BatchRange {
  int r0, r1;
  int a0, a1;  // atoms
  int c0, c1;  // cells
  int w0, w1;  // work items
  int q0;      // first compact charged atom
};

hipError_t batch_set_epoch(long long* epoch, long long step, hipStream_t s);
hipError_t batch_launch(const BatchDev& d, const PairConst<float>& pc, int rbl,
                        int coulomb, int eflag, int special, const BatchRange& rg,
                        const BatchTune& t, int num_cu, hipStream_t s);

}  // namespace rbmdx
