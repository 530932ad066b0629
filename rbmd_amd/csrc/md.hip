// Kernels of the molecular-dynamics step around the force (velocity Verlet,
// periodic wrap with image flags, Nose-Hoover chain, energy-stable rescale).
//
// RBMD's NVE / NVT step (src/simulate_pipeline/src/nve_ensemble.cpp:58-88,
// src/simulate_pipeline/src/nvt_ensemble.cpp:139-167 and
// src/simulate_pipeline/src/controller/pressure_controller/
// nose_hoover_controller.cpp:633-700) is, per step n:
//   [NVT] chain half-step (scales v),  v += dt/2 F/m,  x += dt v,  wrap,
//   F = force(x),  v += dt/2 F/m,
//   [NVT] temperature, chain half-step (scales v)
//   [NVE + RBL] temperature, energy-stable rescale of v.
// Here that is two launches around the force: k_md_first (one scale -- the
// pending end-of-step factor times the chain's opening factor of step n --
// then kick, drift and wrap) and k_md_thermostat (one workgroup: temperature
// from the kinetic-energy partials that the force's assembly kernel wrote
// with the closing half-kick, the end-of-step chain half or the
// energy-stable factor, left pending, and the opening half of step n + 1
// computed ahead for the next k_md_first).  k_md_open opens the first step
// of a run; k_md_apply applies the pending factor at its end; k_md_second
// (kick = 0) gives the partials of the starting velocities.  All thermostat
// scalars stay in device memory, so a run of many steps needs no host round
// trip.
//
// Arithmetic follows the reference's expression order (no contraction into
// FMA), so the FP64 path agrees with a plain restatement to rounding.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "md.hpp"

#pragma clang fp contract(off)

namespace rbmdx {
namespace {

// ComputeTempTarget (nose_hoover_controller.cpp:473-497): a linear ramp from
// t_start to t_stop over num_steps.
__device__ double target_temperature(const MdParams& p, int64_t step) {
  if (p.t_stop == p.t_start) return p.t_start;
  double delta = (double)step;
  if (delta != 0.0) delta = delta / (double)p.num_steps;
  return p.t_start + delta * (p.t_stop - p.t_start);
}

// One half-step of the thermostat chain (NHCTempIntegrate,
// nose_hoover_controller.cpp:732-840, one loop, no drag, masses refreshed
// from the target).  Returns the velocity factor; s.T becomes the
// temperature of the scaled velocities.
__device__ double nhc_half(MdState& s, const MdParams& p) {
  const double dthalf = 0.5 * p.dt, dt4 = 0.25 * p.dt, dt8 = 0.125 * p.dt;
  const double tt = s.t_target;
  const double tdkb = p.tdof * p.kb;
  const double ke_target = tdkb * tt;
  const double ff = p.t_freq * p.t_freq;
  double ke = tdkb * s.T;
  s.eta_mass[0] = tdkb * tt / ff;
  for (int c = 1; c < kMdChain; ++c) s.eta_mass[c] = p.kb * tt / ff;
  s.eta_dotdot[0] = s.eta_mass[0] > 0.0 ? (ke - ke_target) / s.eta_mass[0] : 0.0;
  for (int c = kMdChain - 1; c > 0; --c) {
    const double e = exp(-dt8 * s.eta_dot[c + 1]);
    s.eta_dot[c] *= e;
    s.eta_dot[c] += s.eta_dotdot[c] * dt4;
    s.eta_dot[c] *= e;
  }
  const double e0 = exp(-dt8 * s.eta_dot[1]);
  s.eta_dot[0] *= e0;
  s.eta_dot[0] += s.eta_dotdot[0] * dt4;
  s.eta_dot[0] *= e0;
  const double factor = exp(-dthalf * s.eta_dot[0]);
  s.T = s.T * factor * factor;
  ke = tdkb * s.T;
  s.eta_dotdot[0] = s.eta_mass[0] > 0.0 ? (ke - ke_target) / s.eta_mass[0] : 0.0;
  for (int c = 0; c < kMdChain; ++c) s.eta[c] += dthalf * s.eta_dot[c];
  s.eta_dot[0] *= e0;
  s.eta_dot[0] += s.eta_dotdot[0] * dt4;
  s.eta_dot[0] *= e0;
  for (int c = 1; c < kMdChain; ++c) {
    const double e = exp(-dt8 * s.eta_dot[c + 1]);
    s.eta_dot[c] *= e;
    s.eta_dotdot[c] = (s.eta_mass[c - 1] * s.eta_dot[c - 1] * s.eta_dot[c - 1] -
                       p.kb * tt) / s.eta_mass[c];
    s.eta_dot[c] += s.eta_dotdot[c] * dt4;
    s.eta_dot[c] *= e;
  }
  return factor;
}

// ApplyPBC (src/data_manager/include/model/box.h:133-173): one image shift
// per call, flags counted.
template <typename R>
__device__ __forceinline__ void wrap_flag(R& x, int32_t* img, int i, R lo, R hi,
                                          R L) {
  if (x >= hi) {
    x -= L;
    if (img) img[i] += 1;
  }
  if (x < lo) {
    x += L;
    if (img) img[i] -= 1;
  }
}

template <typename R>
__global__ __launch_bounds__(kMdBlock) void k_md_wrap(MdArrays<R> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  R x = a.px[i], y = a.py[i], z = a.pz[i];
  wrap_flag(x, a.ix, i, (R)a.lo[0], (R)a.hi[0], (R)a.L[0]);
  wrap_flag(y, a.iy, i, (R)a.lo[1], (R)a.hi[1], (R)a.L[1]);
  wrap_flag(z, a.iz, i, (R)a.lo[2], (R)a.hi[2], (R)a.L[2]);
  a.px[i] = x;
  a.py[i] = y;
  a.pz[i] = z;
}

// Half-kick v += 0.5 F / m dt ftm2v in UpdateVelocity's order
// (update_velocity_op.hip.cu:19-21).
template <typename R>
__device__ __forceinline__ R half_kick(R v, R f, R m, R dt, R ftm2v) {
  return v + (R)0.5 * f / m * dt * ftm2v;
}

// The opening half of step `next` on a closed state whose scale holds the
// pending end-of-step factor: NVT folds the chain's opening factor into it.
__device__ void open_step(MdState& st, const MdParams& p, int64_t next) {
  if (p.mode == kMdNvtNh) {
    st.t_target = target_temperature(p, next);
    st.scale *= nhc_half(st, p);
  }
}

// One thread: the opened state of the first step of a run.
__global__ void k_md_open(MdParams p, const MdState* __restrict__ in,
                          MdState* __restrict__ out, int64_t next) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  MdState st = *in;
  open_step(st, p, next);
  *out = st;
}

// Scale by the opened state's factor, half-kick, drift, wrap.  Workgroup 0
// stores the opened state (scale applied) as the step's running state.
template <typename R>
__global__ __launch_bounds__(kMdBlock) void k_md_first(
    MdArrays<R> a, MdParams p, const MdState* __restrict__ open,
    MdState* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MdState st = *open;
    st.scale = 1.0;
    *out = st;
  }
  const R f = (R)open->scale, dt = (R)p.dt, c = (R)p.ftm2v;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += gridDim.x * blockDim.x) {
    const R m = (R)a.mass[a.type[i]];
    R vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    if (f != (R)1) {
      vx *= f;
      vy *= f;
      vz *= f;
    }
    vx = half_kick(vx, a.fx[i], m, dt, c);
    vy = half_kick(vy, a.fy[i], m, dt, c);
    vz = half_kick(vz, a.fz[i], m, dt, c);
    a.vx[i] = vx;
    a.vy[i] = vy;
    a.vz[i] = vz;
    R x = a.px[i] + vx * dt, y = a.py[i] + vy * dt, z = a.pz[i] + vz * dt;
    wrap_flag(x, a.ix, i, (R)a.lo[0], (R)a.hi[0], (R)a.L[0]);
    wrap_flag(y, a.iy, i, (R)a.lo[1], (R)a.hi[1], (R)a.L[1]);
    wrap_flag(z, a.iz, i, (R)a.lo[2], (R)a.hi[2], (R)a.L[2]);
    a.px[i] = x;
    a.py[i] = y;
    a.pz[i] = z;
  }
}

template <typename R>
__global__ __launch_bounds__(kMdBlock) void k_md_second(MdArrays<R> a, MdParams p,
                                                        int kick) {
  __shared__ double smem[16];
  const R dt = (R)p.dt, c = (R)p.ftm2v;
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += gridDim.x * blockDim.x) {
    const double md = a.mass[a.type[i]];
    R vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    if (kick) {
      const R m = (R)md;
      vx = half_kick(vx, a.fx[i], m, dt, c);
      vy = half_kick(vy, a.fy[i], m, dt, c);
      vz = half_kick(vz, a.fz[i], m, dt, c);
      a.vx[i] = vx;
      a.vy[i] = vy;
      a.vz[i] = vz;
    }
    const double x = vx, y = vy, z = vz;
    acc += md * (x * x + y * y + z * z);
  }
  double v[1] = {acc};
  block_sum<1>(v, smem);
  if (threadIdx.x == 0) a.part[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(1024) void k_md_thermostat(
    MdParams p, const double* __restrict__ part, int nparts,
    const double* __restrict__ thermo, const MdState* __restrict__ in,
    MdState* __restrict__ out, MdState* __restrict__ open, int64_t step, int setup,
    int with_pe) {
  __shared__ double smem[16];
  double acc = 0.0;
  for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += part[q];
  double v[1] = {acc};
  block_sum<1>(v, smem);
  if (threadIdx.x != 0) return;
  MdState st = *in;
  // temperature as SampleCurrentTemperatureState computes it
  // (nose_hoover_controller.cpp:256-287; energy_stable_scheme_controller.cpp:98-101)
  st.temp_sum = v[0] * p.mvv2e;
  st.T = 0.5 * st.temp_sum / (p.tdof * p.kb / 2.0);
  st.ke = st.temp_sum / (2.0 * (double)p.natoms);
  st.pe = 0.0;
  if (with_pe)
    st.pe = (thermo[3] + thermo[4] + (p.has_k ? thermo[16] + p.e_self : 0.0)) /
            (double)p.natoms;
  st.step = step;
  if (setup) {
    // SetUp (nose_hoover_controller.cpp:415-470): chain at rest, masses and
    // the forces on links 1.. from the target at step 0
    st.scale = 1.0;
    st.xi = 1.0;
    st.ke_init = st.pe_init = 0.0;
    for (int c = 0; c < kMdChain; ++c)
      st.eta[c] = st.eta_dot[c] = st.eta_dotdot[c] = st.eta_mass[c] = 0.0;
    st.eta_dot[kMdChain] = 0.0;
    if (p.mode == kMdNvtNh) {
      const double tt = target_temperature(p, 0);
      const double ff = p.t_freq * p.t_freq;
      st.t_target = tt;
      st.eta_mass[0] = p.tdof * p.kb * tt / ff;
      for (int c = 1; c < kMdChain; ++c) st.eta_mass[c] = p.kb * tt / ff;
      for (int c = 1; c < kMdChain; ++c)
        st.eta_dotdot[c] = (st.eta_mass[c - 1] * st.eta_dot[c - 1] *
                                st.eta_dot[c - 1] - p.kb * tt) / st.eta_mass[c];
    }
  } else if (p.mode == kMdNveStable) {
    // energy_stable_scheme_controller.cpp:111-137 with the step-1 energies
    // kept as the reference Hamiltonian
    if (step == 1) {
      st.ke_init = st.ke;
      st.pe_init = st.pe;
    }
    const double h_init = st.pe_init + st.ke_init;
    const double h = st.pe + st.ke;
    const double gamma = 10.0 * p.dt;
    // a system at rest has no kinetic energy to rescale
    st.xi = st.ke > 0.0 ? sqrt(1.0 + p.dt / (gamma * st.ke) * (h_init - h)) : 1.0;
    st.scale = st.xi;
  } else if (p.mode == kMdNvtNh) {
    st.scale = nhc_half(st, p);
    st.ke = st.ke * st.scale * st.scale;
  }
  *out = st;
  // the next step's opening half, ready for its k_md_first
  open_step(st, p, step + 1);
  *open = st;
}

template <typename R>
__global__ __launch_bounds__(kMdBlock) void k_md_apply(MdArrays<R> a,
                                                       const MdState* __restrict__ in,
                                                       MdState* __restrict__ out) {
  const double fd = in->scale;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MdState st = *in;
    st.scale = 1.0;
    *out = st;
  }
  if (fd == 1.0) return;
  const R f = (R)fd;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += gridDim.x * blockDim.x) {
    a.vx[i] *= f;
    a.vy[i] *= f;
    a.vz[i] *= f;
  }
}

int md_blocks(int64_t n) {
  const int64_t b = (n + kMdBlock - 1) / kMdBlock;
  return (int)(b < kMdParts ? (b > 0 ? b : 1) : kMdParts);
}

}  // namespace

template <typename R>
hipError_t md_launch_wrap(const MdArrays<R>& a, hipStream_t s) {
  hipLaunchKernelGGL((k_md_wrap<R>), dim3((a.n + kMdBlock - 1) / kMdBlock),
                     dim3(kMdBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t md_launch_open(const MdParams& p, const MdState* in, MdState* out,
                          int64_t next, hipStream_t s) {
  hipLaunchKernelGGL(k_md_open, dim3(1), dim3(64), 0, s, p, in, out, next);
  return hipGetLastError();
}

template <typename R>
hipError_t md_launch_first(const MdArrays<R>& a, const MdParams& p,
                           const MdState* open, MdState* out, hipStream_t s) {
  hipLaunchKernelGGL((k_md_first<R>), dim3(md_blocks(a.n)), dim3(kMdBlock), 0, s,
                     a, p, open, out);
  return hipGetLastError();
}

template <typename R>
hipError_t md_launch_second(const MdArrays<R>& a, const MdParams& p, int kick,
                            hipStream_t s) {
  hipLaunchKernelGGL((k_md_second<R>), dim3(md_blocks(a.n)), dim3(kMdBlock), 0, s,
                     a, p, kick);
  return hipGetLastError();
}

int md_init_parts(int64_t n) { return md_blocks(n); }

hipError_t md_launch_thermostat(const MdParams& p, const double* part, int nparts,
                                const double* thermo, const MdState* in,
                                MdState* out, MdState* open, int64_t step, int setup,
                                int with_pe, hipStream_t s) {
  hipLaunchKernelGGL(k_md_thermostat, dim3(1), dim3(1024), 0, s, p, part, nparts,
                     thermo, in, out, open, step, setup, with_pe);
  return hipGetLastError();
}

template <typename R>
hipError_t md_launch_apply(const MdArrays<R>& a, const MdState* in, MdState* out,
                           hipStream_t s) {
  hipLaunchKernelGGL((k_md_apply<R>), dim3(md_blocks(a.n)), dim3(kMdBlock), 0, s,
                     a, in, out);
  return hipGetLastError();
}

#define RBMDX_MD_INST(R)                                                        \
  template hipError_t md_launch_wrap<R>(const MdArrays<R>&, hipStream_t);      \
  template hipError_t md_launch_first<R>(const MdArrays<R>&, const MdParams&,  \
                                         const MdState*, MdState*, hipStream_t); \
  template hipError_t md_launch_second<R>(const MdArrays<R>&, const MdParams&, \
                                          int, hipStream_t);                   \
  template hipError_t md_launch_apply<R>(const MdArrays<R>&, const MdState*,   \
                                         MdState*, hipStream_t);
RBMDX_MD_INST(float)
RBMDX_MD_INST(double)

}  // namespace rbmdx
