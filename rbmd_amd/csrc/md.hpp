// The molecular-dynamics step either side of the force: velocity Verlet with
// periodic wrapping and image flags, the Nose-Hoover chain thermostat (NVT)
// and the energy-stable velocity rescale of NVE runs with random-batch lists.
// Shared by md.hip (kernels) and api.cpp (host engine).
//
// Every thermostat scalar lives in device memory (MdState, FP64), so a run of
// many steps needs no host round trip: the chain is advanced by the kernels
// themselves.  Two copies of the running state alternate: a kernel that
// advances it reads copy `cur` and writes copy `cur ^ 1` (all workgroups read
// the same input, workgroup 0 writes the output), and the host flips `cur`;
// a third copy holds the next step's opened state.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rbmdx {

constexpr int kMdChain = 3;     // thermostat chain length (RBMD's default)
constexpr int kMdParts = 1024;  // k_md_second's kinetic-energy partials (at most)
constexpr int kMdBlock = 256;

enum MdMode { kMdNve = 0, kMdNveStable = 1, kMdNvtNh = 2 };

struct MdState {
  double temp_sum;   // mvv2e * sum m v^2 of the current velocities
  double T;          // temperature of the current velocities
  double scale;      // velocity factor not yet applied to the caller's arrays
  double t_target;   // target temperature of the step in progress
  double eta[kMdChain], eta_dot[kMdChain + 1], eta_dotdot[kMdChain];
  double eta_mass[kMdChain];
  double ke, pe;     // kinetic and potential energy per atom (last step)
  double ke_init, pe_init;  // energy-stable reference (step 1)
  double xi;         // last energy-stable factor
  int64_t step;      // last completed step
  int64_t pad;
};

struct MdParams {
  int mode = kMdNve;
  int has_k = 0;         // the force includes a k-space energy
  int64_t natoms = 0;
  double dt = 0, ftm2v = 1, mvv2e = 1, kb = 1, tdof = 0;
  double t_start = 0, t_stop = 0, t_freq = 0;
  int64_t num_steps = 0;  // temperature ramp length (execution.num_steps)
  double e_self = 0;      // k-space self / background energy (host constant)
};

template <typename R>
struct MdArrays {
  R *px = nullptr, *py = nullptr, *pz = nullptr;
  R *vx = nullptr, *vy = nullptr, *vz = nullptr;
  const R *fx = nullptr, *fy = nullptr, *fz = nullptr;
  int32_t *ix = nullptr, *iy = nullptr, *iz = nullptr;  // image flags (may be null)
  const int32_t* type = nullptr;  // caller order
  const double* mass = nullptr;   // [ntypes]
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  int n = 0;
  double* part = nullptr;          // [kMdParts]
  const double* thermo = nullptr;  // the force engine's device energy sums
};

// Wrap positions into the box once, updating image flags (start of a run).
template <typename R>
hipError_t md_launch_wrap(const MdArrays<R>& a, hipStream_t s);
// The opening half of step `next` (Nose-Hoover) on a closed state: in -> out
// (the opened state, its scale the product of the pending and opening factors).
hipError_t md_launch_open(const MdParams& p, const MdState* in, MdState* out,
                          int64_t next, hipStream_t s);
// Scale by the opened state's factor, half-kick, drift and wrap; the opened
// state (scale applied) is written to out.
template <typename R>
hipError_t md_launch_first(const MdArrays<R>& a, const MdParams& p,
                           const MdState* open, MdState* out, hipStream_t s);
// Second half-kick (kick = 0: none) and the kinetic-energy partial sums.
template <typename R>
hipError_t md_launch_second(const MdArrays<R>& a, const MdParams& p, int kick,
                            hipStream_t s);
// Temperature and the end-of-step thermostat (setup = 1: the chain's set-up
// at step 0 instead; with_pe = 1: the force engine's energy sums are valid):
// state in -> out.
// nparts partials are summed; the opening half of step + 1 goes to `open`.
hipError_t md_launch_thermostat(const MdParams& p, const double* part, int nparts,
                                const double* thermo, const MdState* in,
                                MdState* out, MdState* open, int64_t step, int setup,
                                int with_pe, hipStream_t s);
// Number of partials md_launch_second writes for n atoms.
int md_init_parts(int64_t n);
// Apply the pending velocity factor: state in -> out with scale 1.
template <typename R>
hipError_t md_launch_apply(const MdArrays<R>& a, const MdState* in, MdState* out,
                           hipStream_t s);

}  // namespace rbmdx
