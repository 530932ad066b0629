// Host engine behind the C ABI of include/rbmdx.h.
//
// A context parses RBMD's JSON configuration (the keys RBMD's force modules
// read: src/force/src/lj_cut_coul_kspace.cpp:424-462,
// src/force/src/kspace_calculator.cpp:159-355,
// src/neighbor_list/src/neighbor_list_builder/rbl_full_neighbor_list_builder.cpp:43-77),
// owns every device buffer of one system, and launches one stream-ordered
// force evaluation per rbmdx_force call.  Nothing is allocated inside a force
// call except on the first call (list capacity planning) or when a list
// outgrows its capacity (growth, which the kernels survive by falling back to
// a direct stencil scan for the affected atoms).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <numeric>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"
#include "json.hpp"
#include "rbmdx.h"
#include "sampler.hpp"
#include "md.hpp"
#include "sog.hpp"

using namespace rbmdx;

namespace {

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
// Caller arrays of a molecular-dynamics call.
struct MdCall {
  void *px, *py, *pz, *vx, *vy, *vz, *fx, *fy, *fz;
  int32_t *ix, *iy, *iz;
};

struct EngineBase {
  virtual ~EngineBase() = default;
  virtual void force(int64_t step, const void* px, const void* py,
                     const void* pz, void* fx, void* fy, void* fz,
                     const int32_t* m, int32_t nk, uint32_t flags,
                     hipStream_t s) = 0;
  virtual void force_noalloc(int64_t step, const void* px, const void* py,
                             const void* pz, void* fx, void* fy, void* fz,
                             const int32_t* m, int32_t nk, uint32_t flags,
                             hipStream_t s) = 0;
  virtual void force_vjp(int64_t step, const void* px, const void* py,
                         const void* pz, const void* gx, const void* gy,
                         const void* gz, void* dx, void* dy, void* dz,
                         double* dparam, uint32_t flags, hipStream_t s) = 0;
  virtual void set_lj(const double* eps, const double* sig) = 0;
  virtual void force_leapfrog(int64_t step, void* px, void* py, void* pz,
                              void* vx, void* vy, void* vz, void* fx, void* fy,
                              void* fz, double dt, uint32_t flags,
                              hipStream_t s) = 0;
  virtual void set_masses(const double* m) = 0;
  virtual void md_init(const MdCall& c, uint32_t flags, hipStream_t s) = 0;
  virtual void md_run(int32_t nsteps, const MdCall& c, uint32_t flags,
                      hipStream_t s) = 0;
  virtual void md_state(rbmdx_md_state_t* out) = 0;
  virtual void thermo(rbmdx_thermo_t* out) = 0;
  virtual int64_t export_pairs(int64_t step, const void* px, const void* py,
                               const void* pz, int candidates, int32_t* gi,
                               int32_t* gj, int32_t* cls, int64_t cap) = 0;
  virtual void set_charges(const double* q) = 0;
  virtual void set_special(int32_t ms, const int32_t* nspecial,
                           const int32_t* special) = 0;
  virtual int kernel_times(const char** names, double* usec, int n) = 0;
  virtual void rbe_sample(int64_t call, int32_t* m) = 0;
  virtual int num_samples() const = 0;
};

template <typename T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  return (T*)p;
}

template <typename R>
class Engine : public EngineBase {
 public:
  using R4 = typename Dev<R>::R4;

  Engine(const Config& c, int device) : cfg_(c), device_(device) {}

  ~Engine() override {
    if (!ready_ && allocs_.empty() && !stage_ && !ctrl_host_ && !hdr_) return;
    hipSetDevice(device_);
    if (last_stream_set_) hipStreamSynchronize(last_stream_);
    pipeline_.reset();
    for (void* p : allocs_) hipFree(p);
    for (auto& ev : timing_events_)
      for (auto e : ev)
        if (e) hipEventDestroy(e);
    for (auto e : slot_events_)
      if (e) hipEventDestroy(e);
    for (auto& ge : graphs_)
      for (auto x : ge.exec)
        if (x) hipGraphExecDestroy(x);
    if (cap_stream_) hipStreamDestroy(cap_stream_);
    if (hdr_) hipHostFree(hdr_);
    if (stage_) hipHostFree(stage_);
    if (ctrl_host_) hipHostFree(ctrl_host_);
    // leave no sticky error behind for the caller's next HIP call
    (void)hipGetLastError();
  }

  void set_system(int64_t n, const double* lo, const double* L,
                  const int32_t* type, const int32_t* id, const double* q,
                  int32_t ntypes, const double* eps, const double* sig) {
    if (n <= 0 || n > (1LL << 30)) throw Error(RBMDX_E_ARG, "bad atom count");
    if (ntypes <= 0 || ntypes > kMaxTypes)
      throw Error(RBMDX_E_ARG, "ntypes must be in [1, 32]");
    for (int d = 0; d < 3; ++d)
      if (!(L[d] > 0) || !std::isfinite(L[d]) || !std::isfinite(lo[d]))
        throw Error(RBMDX_E_ARG, "box lengths must be finite and > 0");
    n_ = (int)n;
    ntypes_ = ntypes;
    for (int d = 0; d < 3; ++d) {
      lo_[d] = lo[d];
      L_[d] = L[d];
    }
    type_.assign(type, type + n);
    id_.assign(id, id + n);
    q_.assign(q, q + n);
    for (int64_t i = 0; i < n; ++i)
      if (type_[i] < 0 || type_[i] >= ntypes)
        throw Error(RBMDX_E_ARG, "atom type out of range");
    eps_.assign(eps, eps + ntypes);
    sig_.assign(sig, sig + ntypes);
    setup();
  }

  void set_charges(const double* q) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "set_system first");
    const bool was_split = split_;
    std::vector<double> nq(q, q + n_);
    bool any = false;
    for (double v : nq) any |= v != 0.0;
    // The charged/neutral partition is fixed at set-up; allow value changes
    // that keep it (zero stays zero).
    for (int i = 0; i < n_; ++i)
      if (((R)nq[i] != R(0)) != ((R)q_[i] != R(0)) && was_split)
        throw Error(RBMDX_E_ARG,
                    "set_charges may not change which atoms are charged");
    (void)any;
    q_ = nq;
    upload_charges();
    charge_sums();
    hipDeviceSynchronize();
  }

  // CVFF special-bond topology (DeviceData d_nspecial / d_special /
  // maxspecial, device_data.h:270-277), caller order.  Counts are clamped as
  // LoadSpecialCounts clamps them (cvff_op.hip.cu:51-77); the rows are
  // copied to sorted order at the next rebuild, which this call forces.
  void set_special(int32_t ms, const int32_t* nspecial,
                   const int32_t* special) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "set_system first");
    if (cfg_.force_type != "CVFF")
      throw Error(RBMDX_E_STATE, "special bonds apply to force_field.type CVFF only");
    if (ms < 0 || ms > kMaxSpecial)
      throw Error(RBMDX_E_ARG, "maxspecial must be in [0, 64]");
    if (ms > 0 && (!nspecial || !special))
      throw Error(RBMDX_E_ARG, "null special-bond arrays");
    HIPCHK(hipSetDevice(device_));
    // evaluations in flight may still read the current rows
    HIPCHK(hipDeviceSynchronize());
    std::vector<int4> cnt(n_);
    bool any = false;
    for (int i = 0; i < n_ && ms > 0; ++i) {
      int n12 = std::max(0, nspecial[3 * i]), n13 = std::max(0, nspecial[3 * i + 1]),
          n14 = std::max(0, nspecial[3 * i + 2]);
      if (n12 > n13) n13 = n12;
      if (n13 > n14) n14 = n13;
      if (n14 > ms) n14 = ms;
      if (n13 > n14) n13 = n14;
      if (n12 > n13) n12 = n13;
      cnt[i] = make_int4(n12, n13, n14, 0);
      any |= n14 > 0;
    }
    for (int* p : {d_.spid, d_.spid_s})
      if (p) release(p);
    for (int4* p : {d_.spn, d_.spn_s})
      if (p) release(p);
    d_.spn = d_.spn_s = nullptr;
    d_.spid = d_.spid_s = nullptr;
    params_.special = 0;
    pc_.ms = 0;
    if (any) {
      d_.spn = alloc<int4>(n_);
      d_.spn_s = alloc<int4>(n_);
      d_.spid = alloc<int>((size_t)n_ * ms);
      d_.spid_s = alloc<int>((size_t)n_ * ms);
      HIPCHK(hipMemcpy(d_.spn, cnt.data(), sizeof(int4) * n_, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_.spid, special, sizeof(int) * (size_t)n_ * ms,
                       hipMemcpyHostToDevice));
      params_.special = 1;
      pc_.ms = ms;
    }
    lists_stale_ = true;
    ++gen_;
  }

  int num_samples() const override {
    return (cfg_.coulomb_type == "RBE" || cfg_.coulomb_type == "RBSOG")
               ? cfg_.coulomb_sample_num
               : 0;
  }

  void rbe_sample(int64_t call, int32_t* m) override {
    if (!sampler_) throw Error(RBMDX_E_STATE, "no sampled k-space");
    sampler_->draw(rbe_seed_of_call(base_seed_, call), m);
  }

  void set_masses(const double* m) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_set_system first");
    std::vector<R> inv(ntypes_);
    for (int t = 0; t < ntypes_; ++t) {
      if (!(m[t] > 0)) throw Error(RBMDX_E_ARG, "masses must be > 0");
      inv[t] = (R)(1.0 / m[t]);
    }
    HIPCHK(hipMemcpy(d_.inv_mass, inv.data(), sizeof(R) * ntypes_,
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(md_mass_, m, sizeof(double) * ntypes_, hipMemcpyHostToDevice));
  }

  // --- molecular-dynamics step (md.hip) ------------------------------------
  MdArrays<R> md_arrays(const MdCall& c) const {
    if (!c.px || !c.py || !c.pz || !c.vx || !c.vy || !c.vz || !c.fx || !c.fy ||
        !c.fz)
      throw Error(RBMDX_E_ARG, "null position, velocity or force pointer");
    MdArrays<R> a;
    a.px = (R*)c.px;
    a.py = (R*)c.py;
    a.pz = (R*)c.pz;
    a.vx = (R*)c.vx;
    a.vy = (R*)c.vy;
    a.vz = (R*)c.vz;
    a.fx = (const R*)c.fx;
    a.fy = (const R*)c.fy;
    a.fz = (const R*)c.fz;
    a.ix = c.ix;
    a.iy = c.iy;
    a.iz = c.iz;
    a.type = d_.type;
    a.mass = md_mass_;
    for (int d = 0; d < 3; ++d) {
      a.lo[d] = lo_[d];
      a.L[d] = L_[d];
      a.hi[d] = lo_[d] + L_[d];
    }
    a.n = n_;
    a.part = md_part_;
    a.thermo = d_.thermo;
    return a;
  }

  uint32_t md_force_flags(uint32_t flags) const {
    return flags | (md_.mode == kMdNveStable ? (uint32_t)RBMDX_FLAG_ENERGY : 0u);
  }

  void md_init(const MdCall& c, uint32_t flags, hipStream_t s) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_set_system first");
    if (n_ < 2) throw Error(RBMDX_E_ARG, "molecular dynamics needs at least 2 atoms");
    HIPCHK(hipSetDevice(device_));
    const MdConfig m = parse_md(cfg_);
    MdParams p;
    p.mode = m.nose_hoover ? kMdNvtNh : (m.energy_stable ? kMdNveStable : kMdNve);
    p.has_k = (kspace_ != 0 && d_.nc > 0) ? 1 : 0;
    p.natoms = n_;
    p.dt = m.dt;
    p.ftm2v = cfg_.ftm2v;
    p.mvv2e = m.mvv2e;
    p.kb = m.kb;
    p.tdof = 3.0 * (double)n_ - 3.0;
    p.t_start = m.t_start;
    p.t_stop = m.t_stop;
    p.t_freq = m.nose_hoover ? 1.0 / m.t_damp : 0.0;
    p.num_steps = m.num_steps;
    p.e_self = e_self_;
    if (!md_state_) {
      // [0], [1]: the alternating running states; [2]: the next step's opened state
      md_state_ = alloc<MdState>(3);
      md_part_ = alloc<double>(std::max<int64_t>(kMdParts, md_assemble_parts()));
    }
    const MdArrays<R> a = md_arrays(c);
    md_ = p;
    md_ready_ = false;
    HIPCHK(hipMemsetAsync(md_state_, 0, 3 * sizeof(MdState), s));
    md_cur_ = 0;
    HIPCHK(md_launch_wrap(a, s));
    const uint32_t ff = md_force_flags(flags);
    force(0, c.px, c.py, c.pz, c.fx, c.fy, c.fz, nullptr, 0, ff, s);
    HIPCHK(md_launch_second(a, md_, 0, s));
    HIPCHK(md_launch_thermostat(md_, md_part_, md_init_parts(n_), d_.thermo,
                                md_state_ + md_cur_, md_state_ + (md_cur_ ^ 1),
                                md_state_ + 2, 0, 1, (ff & RBMDX_FLAG_ENERGY) ? 1 : 0, s));
    md_cur_ ^= 1;
    md_step_ = 0;
    md_ready_ = true;
  }

  // k_assemble's workgroups: one partial of the kinetic energy each
  int64_t md_assemble_parts() const { return ((int64_t)n_ + kBlock - 1) / kBlock; }

  void md_run(int32_t nsteps, const MdCall& c, uint32_t flags,
              hipStream_t s) override {
    if (!md_ready_) throw Error(RBMDX_E_STATE, "rbmdx_md_init first");
    if (nsteps < 0) throw Error(RBMDX_E_ARG, "nsteps must be >= 0");
    HIPCHK(hipSetDevice(device_));
    const MdArrays<R> a = md_arrays(c);
    const uint32_t ff = md_force_flags(flags);
    const int with_pe = (ff & RBMDX_FLAG_ENERGY) ? 1 : 0;
    MdState* open = md_state_ + 2;
    if (nsteps > 0) HIPCHK(md_launch_open(md_, md_state_ + md_cur_, open, md_step_ + 1, s));
    const int parts = (int)md_assemble_parts();
    for (int32_t k = 0; k < nsteps; ++k) {
      const int64_t step = md_step_ + 1;
      HIPCHK(md_launch_first(a, md_, open, md_state_ + (md_cur_ ^ 1), s));
      md_cur_ ^= 1;
      // the force with the closing half-kick and kinetic-energy partials
      // folded into its assembly kernel
      d_.vx = (R*)c.vx;
      d_.vy = (R*)c.vy;
      d_.vz = (R*)c.vz;
      d_.md_mass = md_mass_;
      d_.md_part = md_part_;
      md_kick_ = true;
      try {
        force(step, c.px, c.py, c.pz, c.fx, c.fy, c.fz, nullptr, 0, ff, s);
      } catch (...) {
        d_.vx = d_.vy = d_.vz = nullptr;
        md_kick_ = false;
        throw;
      }
      d_.vx = d_.vy = d_.vz = nullptr;
      md_kick_ = false;
      HIPCHK(md_launch_thermostat(md_, md_part_, parts, d_.thermo, md_state_ + md_cur_,
                                  md_state_ + (md_cur_ ^ 1), open, step, 0, with_pe, s));
      md_cur_ ^= 1;
      md_step_ = step;
    }
    // leave the caller's velocities with the last thermostat factor applied
    HIPCHK(md_launch_apply(a, md_state_ + md_cur_, md_state_ + (md_cur_ ^ 1), s));
    md_cur_ ^= 1;
  }

  void md_state(rbmdx_md_state_t* out) override {
    if (!md_ready_) throw Error(RBMDX_E_STATE, "rbmdx_md_init first");
    HIPCHK(hipSetDevice(device_));
    if (last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
    MdState st;
    HIPCHK(hipMemcpy(&st, md_state_ + md_cur_, sizeof(MdState), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->step = md_step_;
    out->mode = md_.mode;
    out->temperature = st.T;
    out->ke = st.ke;
    out->pe = st.pe;
    out->t_target = st.t_target;
    out->xi = md_.mode == kMdNveStable ? st.xi : 1.0;
    // FixNH-style chain energy: ke_target eta_0 + kT sum_{i>0} eta_i +
    // sum Q_i eta_dot_i^2 / 2
    const double kt = md_.kb * st.t_target;
    double e = 0.0;
    for (int q = 0; q < 3; ++q) {
      out->eta[q] = st.eta[q];
      out->eta_dot[q] = st.eta_dot[q];
      out->eta_mass[q] = st.eta_mass[q];
      e += 0.5 * st.eta_mass[q] * st.eta_dot[q] * st.eta_dot[q];
      e += (q == 0 ? md_.tdof * kt : kt) * st.eta[q];
    }
    out->thermostat_energy = md_.mode == kMdNvtNh ? e : 0.0;
  }

  void force_leapfrog(int64_t step, void* px, void* py, void* pz, void* vx,
                      void* vy, void* vz, void* fx, void* fy, void* fz,
                      double dt, uint32_t flags, hipStream_t s) override {
    if (!vx || !vy || !vz) throw Error(RBMDX_E_ARG, "null velocity pointer");
    d_.vx = (R*)vx;
    d_.vy = (R*)vy;
    d_.vz = (R*)vz;
    d_.xw = (R*)px;
    d_.yw = (R*)py;
    d_.zw = (R*)pz;
    leap_dt_ = dt;
    try {
      force(step, px, py, pz, fx, fy, fz, nullptr, 0, flags, s);
    } catch (...) {
      d_.vx = d_.vy = d_.vz = nullptr;
      leap_dt_ = 0;
      throw;
    }
    d_.vx = d_.vy = d_.vz = nullptr;
    leap_dt_ = 0;
  }

  // Replace the per-type LJ parameters (keeps lists and the charges).
  void set_lj(const double* eps, const double* sig) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "set_system first");
    HIPCHK(hipSetDevice(device_));
    HIPCHK(hipDeviceSynchronize());
    eps_.assign(eps, eps + ntypes_);
    sig_.assign(sig, sig + ntypes_);
    upload_lj();
  }

  // Vector-Jacobian product of one evaluation's force with the cotangent
  // (gx, gy, gz) = dL/dF: dL/dr into (dx, dy, dz) (caller order) and
  // [dL/dQ | dL/deps | dL/dsigma] per type into dparam (device, 3 ntypes
  // doubles).  Same pair set (step = RBL epoch) and the k-sample of the last
  // forward evaluation.  Eager launches on stream s.
  void force_vjp(int64_t step, const void* px, const void* py, const void* pz,
                 const void* gx, const void* gy, const void* gz, void* dx,
                 void* dy, void* dz, double* dparam, uint32_t flags,
                 hipStream_t s) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_set_system first");
    if (!px || !py || !pz || !gx || !gy || !gz || !dx || !dy || !dz || !dparam)
      throw Error(RBMDX_E_ARG, "null pointer");
    if (d_.cap_a == 0 || calls_ == 0)
      throw Error(RBMDX_E_STATE, "rbmdx_force_vjp needs a forward evaluation first");
    if (ntypes_ > kVjpMaxTypes)
      throw Error(RBMDX_E_CONFIG, "the adjoint supports at most 8 atom types");
    HIPCHK(hipSetDevice(device_));
    bind_positions(px, py, pz);
    grow_if_needed(s);
    int force_rebuild = (flags & RBMDX_FLAG_REBUILD) ? 1 : 0;
    if (lists_stale_) force_rebuild = 1;
    lists_stale_ = false;
    alloc_vjp();
    StepArgs<R> a = args(step, 0);
    a.force_rebuild = force_rebuild;
    const int slot = (int)(calls_ % kSlots);
    if (slot_used_[slot]) HIPCHK(hipEventSynchronize(slot_events_[slot]));
    hdr_[slot].epoch = (long long)step;
    hdr_[slot].force = force_rebuild;
    vjp_.gx = (const R*)gx;
    vjp_.gy = (const R*)gy;
    vjp_.gz = (const R*)gz;
    vjp_.dx = (R*)dx;
    vjp_.dy = (R*)dy;
    vjp_.dz = (R*)dz;
    vjp_.dparam = dparam;
    vjp_.neutral_q = (flags & RBMDX_FLAG_VJP_NEUTRAL_Q) ? 1 : 0;
    a.p.num_k = last_num_k_;
    // the evaluation's lists and sorted tiles (as the forward path builds them)
    HIPCHK(launch_skin_build<R>(d_, a, s));
    HIPCHK(launch_vjp<R>(d_, vjp_, a, s));
    HIPCHK(hipEventRecord(slot_events_[slot], s));
    slot_used_[slot] = true;
    last_stream_ = s;
    last_stream_set_ = true;
    ++calls_;
    if (flags & RBMDX_FLAG_SYNC) HIPCHK(hipStreamSynchronize(s));
  }

  // Adjoint buffers, sized for the system on first use.
  void alloc_vjp() {
    if (vjp_.gs) return;
    VjpDev<R>& v = vjp_;
    v.gs = alloc<R4>(n_);
    v.gc = alloc<R4>(std::max(1, d_.nc));
    v.drs = alloc<R4>(n_);
    v.drk = alloc<R4>(std::max(1, d_.nc));
    v.gather_blocks = std::min(1024, std::max(1, (n_ + kBlock - 1) / kBlock));
    v.gpart = alloc<double>(3 * (size_t)v.gather_blocks);
    v.gmean = alloc<double>(3);
    // the neutral-charged charge pass: one row per brick
    v.nq_blocks = (params_.split && d_.nc > 0 && d_.nc < n_) ? grid_.nbrick : 0;
    v.pacc = alloc<double>(3 * kVjpMaxTypes);
    // caller indices, charged atoms type-major then neutral atoms type-major,
    // cut into chunks of at most kSkChunk atoms of one type and part
    auto neutral = [&](int i) { return (R)q_[i] == R(0) ? 1 : 0; };
    std::vector<int> ord(n_), beg, gbeg, gtyp;
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
      const int nx = neutral(x), ny = neutral(y);
      return nx != ny ? nx < ny : type_[x] < type_[y];
    });
    int nch_c = -1, ngrp_c = -1;
    for (int s = 0; s < n_;) {
      const int i0 = ord[s];
      if (neutral(i0) && nch_c < 0) {
        nch_c = (int)beg.size();
        ngrp_c = (int)gtyp.size();
      }
      if (s == 0 || neutral(i0) != neutral(ord[s - 1]) || type_[i0] != type_[ord[s - 1]]) {
        gbeg.push_back((int)beg.size());
        gtyp.push_back(type_[i0]);
      }
      int e = s;
      while (e < n_ && neutral(ord[e]) == neutral(i0) && type_[ord[e]] == type_[i0] &&
             e - s < kSkChunk)
        ++e;
      beg.push_back(s);
      s = e;
    }
    v.nch = (int)beg.size();
    v.ngrp = (int)gtyp.size();
    v.nch_c = nch_c < 0 ? v.nch : nch_c;
    v.ngrp_c = ngrp_c < 0 ? v.ngrp : ngrp_c;
    beg.push_back(n_);
    gbeg.push_back(v.nch);
    v.tord = alloc<int>(n_);
    v.ch_beg = alloc<int>(beg.size());
    v.grp_beg = alloc<int>(gbeg.size());
    v.grp_type = alloc<int>(std::max<size_t>(1, gtyp.size()));
    HIPCHK(hipMemcpy(v.tord, ord.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(v.ch_beg, beg.data(), sizeof(int) * beg.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(v.grp_beg, gbeg.data(), sizeof(int) * gbeg.size(),
                     hipMemcpyHostToDevice));
    if (!gtyp.empty())
      HIPCHK(hipMemcpy(v.grp_type, gtyp.data(), sizeof(int) * gtyp.size(),
                       hipMemcpyHostToDevice));
    const int K = std::max(1, d_.kcap);
    v.vsk = alloc<double>((size_t)v.nch * K * 8);
    v.kv0 = alloc<float4>(K);
    v.kv1 = alloc<float4>(K);
    v.kv2 = alloc<float2>(K);
    v.dqk = alloc<double>((size_t)K * kVjpMaxTypes);
  }

  void force(int64_t step, const void* px, const void* py, const void* pz,
             void* fx, void* fy, void* fz, const int32_t* m, int32_t nk,
             uint32_t flags, hipStream_t s) override {
    evaluate(step, px, py, pz, fx, fy, fz, m, nk, flags, s, false);
  }

  // The same evaluation without any device allocation: eager launches on
  // the caller's stream (no graph instantiation), no list growth (an
  // overflowed list is covered exactly by the pair kernel's stencil
  // fallback), and the per-evaluation scratch in the static per-device array
  // instead of the context's buffers.
  void force_noalloc(int64_t step, const void* px, const void* py, const void* pz,
                     void* fx, void* fy, void* fz, const int32_t* m, int32_t nk,
                     uint32_t flags, hipStream_t s) override {
    evaluate(step, px, py, pz, fx, fy, fz, m, nk, flags, s, true);
  }

  // Lays out the per-evaluation scratch of one force evaluation from `base`
  // (nullptr: sizes only), every buffer 256-byte aligned, and points dd's
  // buffers there: the pair kernel's and k_lists' per-block partials, the
  // skin maxima, the mean pair force, and with k-space the S(k) chunk
  // partials, rho and the per-k force coefficients.  Each is written and
  // read within one evaluation (the list statistics only on a rebuild, which
  // writes them first), so none needs clearing between calls.  Returns the
  // bytes used; the sizes are those of the context's own buffers.
  size_t scratch_layout(Dev<R>& dd, char* base) const {
    size_t off = 0;
    auto take = [&](auto*& p, size_t count) {
      using T = std::remove_reference_t<decltype(*p)>;
      off = (off + 255) / 256 * 256;
      p = base ? reinterpret_cast<T*>(base + off) : nullptr;
      off += sizeof(T) * std::max<size_t>(1, count);
    };
    take(dd.pair_part, (size_t)11 * grid_.nbrick * grid_.psub);
    if (grid_.qblocks) take(dd.pairb_part, (size_t)11 * grid_.qblocks);
    take(dd.skin_part, kSkinBlocksMax + 1);
    take(dd.lista_part, grid_.nbrick);
    take(dd.listb_part, std::max(1, list_b_count(grid_, dd.nc)));
    take(dd.mean, 3);
    const size_t nchunks = std::max(1, (d_.nc + kSkChunk - 1) / kSkChunk);
    if (d_.sk_part) {
      take(dd.sk_part, (size_t)std::max(1, (d_.nc + kSkAtoms - 1) / kSkAtoms) * d_.kcap);
      take(dd.kfin_part, (size_t)8 * d_.kcap);
      take(dd.kf0, d_.kcap);
      take(dd.kf1, d_.kcap);
      take(dd.rho, d_.kcap);
    }
    if (d_.sk_ex) take(dd.sk_ex, nchunks * d_.kex);
    return off;
  }

  void evaluate(int64_t step, const void* px, const void* py, const void* pz,
                void* fx, void* fy, void* fz, const int32_t* m, int32_t nk,
                uint32_t flags, hipStream_t s, bool noalloc) {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_set_system first");
    if (!px || !py || !pz || !fx || !fy || !fz)
      throw Error(RBMDX_E_ARG, "null position or force pointer");
    HIPCHK(hipSetDevice(device_));
    Dev<R> dd;
    if (noalloc) {
      Dev<R> probe = d_;
      const size_t need = scratch_layout(probe, nullptr);
      if (need > kNoAllocScratchBytes)
        throw Error(RBMDX_E_CAPACITY,
                    "rbmdx_force_noalloc: this system needs " + std::to_string(need) +
                        " bytes of scratch; the static scratch holds " +
                        std::to_string(kNoAllocScratchBytes));
      if (d_.cap_a == 0)
        throw Error(RBMDX_E_STATE,
                    "rbmdx_force_noalloc: the context has no neighbour lists yet "
                    "(its first rbmdx_force call sizes them)");
      if (flags & RBMDX_FLAG_TIMING)
        throw Error(RBMDX_E_ARG, "rbmdx_force_noalloc: RBMDX_FLAG_TIMING is not supported");
      if (!m && (kspace_ == 1 || kspace_ == 3) && !pipeline_)
        throw Error(RBMDX_E_STATE,
                    "rbmdx_force_noalloc: the seeded k-sample stream starts on an "
                    "rbmdx_force call; pass m_samples");
      char* base = static_cast<char*>(noalloc_scratch(device_));
      if (!base)
        throw Error(RBMDX_E_STATE, "static scratch not resolved on this device");
      dd = d_;
      scratch_layout(dd, base);
    }
    bind_positions(px, py, pz);
    d_.fx = (R*)fx;
    d_.fy = (R*)fy;
    d_.fz = (R*)fz;
    int force_rebuild = (flags & RBMDX_FLAG_REBUILD) ? 1 : 0;
    if (noalloc) {
      // capacity stays as it is
    } else if (d_.cap_a == 0) {
      plan_capacity(s);
    } else {
      grow_if_needed(s);
    }
    // freshly (re)allocated lists hold no entries yet: rebuild before use
    if (lists_stale_) force_rebuild = 1;
    lists_stale_ = false;
    StepArgs<R> a = args(step, flags);
    a.force_rebuild = force_rebuild;
    // this evaluation's slot of the host-mapped rings: step parameters and,
    // for sampled k-space, the k-sample (read on the device through
    // step_count, so one captured graph serves every evaluation)
    const int slot = (int)(calls_ % kSlots);
    if (slot_used_[slot]) HIPCHK(hipEventSynchronize(slot_events_[slot]));
    hdr_[slot].epoch = (long long)step;
    hdr_[slot].force = force_rebuild;
    bool explicit_m = false;
    if (kspace_ == 1 || kspace_ == 3) {
      const bool sog = kspace_ == 3;
      const int P = num_samples();
      if (m) {
        explicit_m = true;
        if (!sog) {
          if (nk <= 0 || nk > d_.kcap)
            throw Error(RBMDX_E_ARG, "num_k out of range for explicit samples");
          // kept in device memory: the adjoint reuses the evaluation's sample
          HIPCHK(hipMemcpyAsync(d_.m_dev, m, sizeof(int32_t) * 3 * (size_t)nk,
                                hipMemcpyDeviceToDevice, s));
          a.m_use = d_.m_dev;
          a.p.num_k = nk;
        } else {
          // RBSOG: P m-vectors followed by P virial-chain indices; the direct
          // vectors stay in place behind them
          if (nk != P)
            throw Error(RBMDX_E_ARG, "RBSOG explicit samples must hold exactly P vectors");
          HIPCHK(hipMemcpyAsync(d_.m_dev, m, sizeof(int32_t) * 3 * P,
                                hipMemcpyDeviceToDevice, s));
          HIPCHK(hipMemcpyAsync(d_.src_dev, m + 3 * P, sizeof(int32_t) * P,
                                hipMemcpyDeviceToDevice, s));
          a.m_use = d_.m_dev;
          a.p.num_k = d_.kcap;
        }
      } else {
        int32_t* dst = stage_ + (size_t)slot * stage_stride_;
        if (!pipeline_) {
          const char* th = std::getenv("RBMDX_SAMPLER_THREADS");
          const int nth = th ? std::max(1, std::atoi(th)) : 4;
          pipeline_.reset(new SamplePipeline(sampler_.get(), base_seed_, nth, 64));
          pipeline_call_ = 0;
        }
        if (!sog) {
          pipeline_->take(pipeline_call_++, dst);
          a.p.num_k = P;
        } else {
          // slot layout: m of the P samples and the direct vectors, then the
          // virial-chain source of every k (direct ones are their own)
          const int K = d_.kcap;
          sog_buf_.resize((size_t)4 * P);
          pipeline_->take(pipeline_call_++, sog_buf_.data());
          std::copy(sog_buf_.begin(), sog_buf_.begin() + 3 * P, dst);
          std::copy(sog_direct_m_.begin(), sog_direct_m_.end(), dst + 3 * P);
          int32_t* src = dst + 3 * K;
          std::copy(sog_buf_.begin() + 3 * P, sog_buf_.end(), src);
          for (int e = P; e < K; ++e) src[e] = e;
          a.p.num_k = K;
        }
        a.m_from_ring = 1;
        a.m_use = d_.m_dev;
      }
    } else if (kspace_ == 2) {
      a.m_use = d_.m_dev;
    }
    const bool timed = !noalloc && (timing_ || (flags & RBMDX_FLAG_TIMING));
    if (noalloc) {
      // the context's buffers with the scratch redirected, bound to this
      // call's caller arrays
      Dev<R> dn = d_;
      dn.pair_part = dd.pair_part;
      dn.pairb_part = dd.pairb_part;
      dn.skin_part = dd.skin_part;
      dn.lista_part = dd.lista_part;
      dn.listb_part = dd.listb_part;
      dn.mean = dd.mean;
      dn.sk_part = dd.sk_part;
      dn.kfin_part = dd.kfin_part;
      dn.kf0 = dd.kf0;
      dn.kf1 = dd.kf1;
      dn.rho = dd.rho;
      dn.sk_ex = dd.sk_ex;
      HIPCHK(launch_force<R>(dn, a, s));
    } else if (graph_on_ && !timed && !explicit_m && !(flags & RBMDX_FLAG_EAGER)) {
      const GraphKey key = graph_key(a);
      GraphEntry* ge = nullptr;
      for (auto& e : graphs_)
        if (std::memcmp(&key, &e.key, sizeof key) == 0) ge = &e;
      if (!ge) ge = &capture(a, key);
      ge->last_use = calls_;
      // consecutive calls use different instantiated copies: relaunching an
      // exec whose previous launch is still running costs the host about
      // twice as much
      HIPCHK(hipGraphLaunch(ge->exec[ge->next], s));
      ge->next = (ge->next + 1) % kGraphCopies;
    } else {
      if (timed) {
        if (timing_used_ == timing_events_.size()) {
          timing_events_.emplace_back();
          for (auto& e : timing_events_.back()) HIPCHK(hipEventCreate(&e));
        }
        auto& ev = timing_events_[timing_used_++];
        for (int q = 0; q < 6; ++q) a.ev[q] = ev[q];
      }
      HIPCHK(launch_force<R>(d_, a, s));
    }
    HIPCHK(hipEventRecord(slot_events_[slot], s));
    slot_used_[slot] = true;
    last_stream_ = s;
    last_stream_set_ = true;
    ++calls_;
    if (force_rebuild) ++forced_;
    last_num_k_ = (kspace_ ? a.p.num_k : 0);
    last_eflag_ = a.p.eflag != 0;
    if (flags & RBMDX_FLAG_SYNC) {
      HIPCHK(hipStreamSynchronize(s));
      if (!noalloc && grow_if_needed(s)) {
        // lists were outgrown: the result is still exact (stencil fallback);
        // capacity is now larger for the next call.
      }
    }
  }

  // Everything a captured evaluation bakes into its kernel arguments besides
  // the device buffers (whose reallocation bumps gen_).
  struct GraphKey {
    const void* ptr[12];
    int eflag, num_k, m_from_ring, md_kick;
    double dt;
    uint64_t gen;
  };
  static constexpr int kGraphCopies = 2;  // executable copies launched in turn
  static constexpr int kGraphCache = 4;   // captured evaluation shapes kept
  struct GraphEntry {
    GraphKey key{};
    hipGraphExec_t exec[kGraphCopies] = {};
    int next = 0;
    int64_t last_use = 0;
  };
  GraphKey graph_key(const StepArgs<R>& a) const {
    GraphKey g;
    std::memset(&g, 0, sizeof g);
    const void* p[12] = {d_.px, d_.py, d_.pz, d_.fx, d_.fy, d_.fz,
                         d_.vx, d_.vy, d_.vz, d_.xw, d_.yw, d_.zw};
    for (int q = 0; q < 12; ++q) g.ptr[q] = p[q];
    g.eflag = a.p.eflag;
    g.num_k = a.p.num_k;
    g.m_from_ring = a.m_from_ring;
    g.md_kick = a.md_kick;
    g.dt = a.dt;
    g.gen = gen_;
    return g;
  }

  // Captures one evaluation (both streams) into a graph and instantiates
  // kGraphCopies executable copies of it, in a cache of kGraphCache keys (a
  // periodic energy evaluation, say, keeps its own entry); the least
  // recently used entry makes room, after the device has finished with it.
  GraphEntry& capture(const StepArgs<R>& a, const GraphKey& key) {
    if ((int)graphs_.size() >= kGraphCache) {
      size_t lru = 0;
      for (size_t q = 1; q < graphs_.size(); ++q)
        if (graphs_[q].last_use < graphs_[lru].last_use) lru = q;
      if (last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
      for (auto x : graphs_[lru].exec)
        if (x) HIPCHK(hipGraphExecDestroy(x));
      graphs_.erase(graphs_.begin() + lru);
    }
    if (!cap_stream_) HIPCHK(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeRelaxed));
    const hipError_t e1 = launch_force<R>(d_, a, cap_stream_);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cap_stream_, &g);
    if (e1 != hipSuccess || e2 != hipSuccess || !g) {
      if (g) hipGraphDestroy(g);
      (void)hipGetLastError();
      throw Error(RBMDX_E_HIP, std::string("graph capture failed: ") +
                                   hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    }
    GraphEntry ge;
    ge.key = key;
    hipError_t e3 = hipSuccess;
    for (int c = 0; c < kGraphCopies && e3 == hipSuccess; ++c)
      e3 = hipGraphInstantiate(&ge.exec[c], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e3 != hipSuccess) {
      for (auto x : ge.exec)
        if (x) hipGraphExecDestroy(x);
      throw Error(RBMDX_E_HIP, std::string("graph instantiation failed: ") +
                                   hipGetErrorString(e3));
    }
    ++captures_;
    graphs_.push_back(ge);
    return graphs_.back();
  }

  void thermo(rbmdx_thermo_t* out) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_set_system first");
    HIPCHK(hipSetDevice(device_));
    if (last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
    // device layout: [0..2] sum of pair forces, [3] E_lj, [4] E_coul,
    // [5..10] pair virial, [16] k-space energy without the self terms,
    // [17..22] k-space virial
    double t[32];
    HIPCHK(hipMemcpy(t, d_.thermo, sizeof(double) * 32, hipMemcpyDeviceToHost));
    Ctrl c;
    HIPCHK(hipMemcpy(&c, d_.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    const bool has_k = kspace_ != 0 && d_.nc > 0 && last_num_k_ > 0;
    out->e_vdwl = t[3];
    out->e_coul = t[4];
    out->e_kspace = has_k ? t[16] + e_self_ : 0.0;
    for (int k = 0; k < 6; ++k) {
      out->virial_kspace[k] = has_k ? t[17 + k] : 0.0;
      out->virial[k] = t[5 + k] + out->virial_kspace[k];
    }
    if (params_.mean_fix)
      for (int k = 0; k < 3; ++k) out->mean_force[k] = t[k] / params_.n_global;
    out->steps = (int64_t)c.step_count;
    out->rebuilds = (int64_t)c.rebuild_count;
    out->rbl_frequency = freq_;
    out->num_k = last_num_k_;
    out->alpha = alpha_;
    out->rbe_S = rbe_S_;
    out->pairs_a = (int64_t)c.sum_count_a;
    out->pairs_b = (int64_t)c.sum_count_b;
    out->atoms = n_;
    out->charged = d_.nc;
    // exact Ewald energy of the last energy evaluation, with the self term
    // of ComputeSelfEnergy (kspace_calculator.cpp:897-907)
    if (d_.kex > 0 && last_eflag_)
      out->e_kspace_exact =
          t[24] - cfg_.qqr2e * std::sqrt(alpha_ / M_PI) * sum_q2_;
  }

  int64_t export_pairs(int64_t step, const void* px, const void* py,
                       const void* pz, int candidates, int32_t* gi,
                       int32_t* gj, int32_t* cls, int64_t cap) override {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_set_system first");
    HIPCHK(hipSetDevice(device_));
    hipStream_t s = nullptr;
    if (!candidates) {
      bind_positions(px, py, pz);
      if (d_.cap_a == 0) plan_capacity(s);
      StepArgs<R> a = args(step, 0);
      HIPCHK(launch_build<R>(d_, a, s));
      HIPCHK(hipStreamSynchronize(s));
      Ctrl c;
      HIPCHK(hipMemcpy(&c, d_.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
      if (c.max_count_a > d_.cap_a || c.max_count_b > d_.cap_b) {
        alloc_lists(std::max(c.max_count_a, d_.cap_a),
                    std::max(c.max_count_b, d_.cap_b));
        a = args(step, 0);
        HIPCHK(launch_build<R>(d_, a, s));
      }
    } else if (d_.cap_a == 0) {
      throw Error(RBMDX_E_STATE, "no lists built yet");
    }
    if (!candidates) lists_stale_ = false;
    else if (lists_stale_) throw Error(RBMDX_E_STATE, "no lists built yet");
    StepArgs<R> a = args(step, 0);
    const int ecap = (int)std::min<int64_t>(cap, 1LL << 30);
    int* buf = dalloc<int>((size_t)3 * std::max(ecap, 1));
    d_.export_buf = buf;
    HIPCHK(hipMemsetAsync(d_.export_count, 0, sizeof(int), s));
    HIPCHK(launch_export<R>(d_, a, candidates, ecap, s));
    HIPCHK(hipStreamSynchronize(s));
    int count = 0;
    HIPCHK(hipMemcpy(&count, d_.export_count, sizeof(int), hipMemcpyDeviceToHost));
    const int ncopy = std::min(count, ecap);
    std::vector<int> h((size_t)3 * std::max(ncopy, 1));
    if (ncopy > 0)
      HIPCHK(hipMemcpy(h.data(), buf, sizeof(int) * 3 * ncopy, hipMemcpyDeviceToHost));
    hipFree(buf);
    d_.export_buf = nullptr;
    for (int k = 0; k < ncopy; ++k) {
      gi[k] = h[3 * k];
      gj[k] = h[3 * k + 1];
      cls[k] = h[3 * k + 2];
    }
    return count;
  }

  int kernel_times(const char** names, double* usec, int n) override {
    // phase -> (start event, end event); kspace runs on the second stream
    static const char* kNames[4] = {"rebuild", "pair", "kspace", "assemble"};
    static const int kFrom[4] = {0, 1, 3, 2}, kTo[4] = {1, 2, 4, 5};
    if (last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
    double acc[4] = {0, 0, 0, 0};
    const size_t calls = timing_used_;
    for (size_t c = 0; c < calls; ++c) {
      auto& ev = timing_events_[c];
      for (int k = 0; k < 4; ++k) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev[kFrom[k]], ev[kTo[k]]));
        acc[k] += ms * 1000.0;
      }
    }
    timing_used_ = 0;
    const int m = std::min(n, 4);
    for (int k = 0; k < m; ++k) {
      if (names) names[k] = kNames[k];
      usec[k] = calls ? acc[k] / (double)calls : 0.0;
    }
    return m;
  }

 private:
  Config cfg_;
  int device_;
  int num_cu_ = 256;  // compute units of the device (dense charged-pair grid)
  bool ready_ = false;
  int n_ = 0, ntypes_ = 0;
  double lo_[3], L_[3];
  std::vector<int32_t> type_, id_;
  std::vector<double> q_, eps_, sig_;
  bool split_ = false;
  int kspace_ = 0;  // 0 none, 1 RBE, 2 EWALD
  double alpha_ = 0, rbe_S_ = 0, sum_q_ = 0, sum_q2_ = 0, e_self_ = 0;
  bool last_eflag_ = false;
  int freq_ = 1;
  Grid grid_{};
  Box<R> box_{};
  PairConst<R> pc_{};
  Params params_{};
  Dev<R> d_{};
  R* lj_tab_ = nullptr;
  VjpDev<R> vjp_;
  std::vector<void*> allocs_;
  std::unique_ptr<KSampler> sampler_;
  std::unique_ptr<SamplePipeline> pipeline_;
  int64_t pipeline_call_ = 0;
  uint32_t base_seed_ = 0;
  int32_t* stage_ = nullptr;       // host view of the k-sample staging ring
  int32_t* stage_dev_ = nullptr;   // device view of the same memory
  size_t stage_stride_ = 0;        // ints per slot
  SogSet sog_;                     // RBSOG / VERLET-SOG set-up
  bool sog_ready_ = false;
  int sog_ndirect_ = 0;
  std::vector<int> sog_direct_m_;
  std::vector<int32_t> sog_buf_;
  Ctrl* ctrl_host_ = nullptr;      // host view of the control mirror
  hipEvent_t slot_events_[kSlots] = {};
  bool slot_used_[kSlots] = {};
  StepHdr* hdr_ = nullptr;          // host view of the step-parameter ring
  bool graph_on_ = true;            // replay evaluations through a hipGraph
  std::vector<GraphEntry> graphs_;
  hipStream_t cap_stream_ = nullptr;
  uint64_t gen_ = 0;                // bumped whenever device buffers change
  int64_t captures_ = 0;
  hipStream_t last_stream_ = nullptr;
  bool last_stream_set_ = false;
  int64_t calls_ = 0, forced_ = 0;
  int last_num_k_ = 0;
  bool timing_ = false;
  std::vector<std::array<hipEvent_t, 6>> timing_events_;  // reusable pool
  size_t timing_used_ = 0;
  unsigned rebuild_thr_ = 0;
  bool lists_stale_ = true;  // lists (re)allocated since the last rebuild
  double leap_dt_ = 0;        // fused leapfrog time step of the current call
  // molecular-dynamics step: per-type masses (FP64), the two alternating
  // thermostat states, kinetic-energy partials, settings and progress
  double* md_mass_ = nullptr;
  MdState* md_state_ = nullptr;
  double* md_part_ = nullptr;
  MdParams md_;
  int md_cur_ = 0;
  int64_t md_step_ = 0;
  bool md_ready_ = false;
  bool md_kick_ = false;  // the current evaluation is an MD step's force

  template <typename T>
  T* alloc(size_t n) {
    T* p = dalloc<T>(n);
    allocs_.push_back(p);
    return p;
  }


  // LJ table: (-24 eps_ij, sigma_ij^2, eps_ij, sigma_ij), Lorentz-Berthelot
  // (lj_op.hip.cu:130-131); -24 eps is formed in R, as the force term forms it.
  void upload_lj() {
    std::vector<R> t(4 * (size_t)ntypes_ * ntypes_);
    for (int a = 0; a < ntypes_; ++a)
      for (int b = 0; b < ntypes_; ++b) {
        const double e = std::sqrt(eps_[a] * eps_[b]);
        const double s = (sig_[a] + sig_[b]) / 2.0;
        R* row = &t[4 * (a * ntypes_ + b)];
        row[0] = (R)-24 * (R)e;
        row[1] = (R)(s * s);
        row[2] = (R)e;
        row[3] = (R)s;
      }
    if (!lj_tab_) lj_tab_ = alloc<R>(t.size());
    HIPCHK(hipMemcpy(lj_tab_, t.data(), sizeof(R) * t.size(), hipMemcpyHostToDevice));
  }

  void bind_positions(const void* px, const void* py, const void* pz) {
    d_.px = (const R*)px;
    d_.py = (const R*)py;
    d_.pz = (const R*)pz;
  }

  void charge_sums() {
    sum_q_ = 0;
    sum_q2_ = 0;
    for (double v : q_) {
      sum_q_ += v;
      sum_q2_ += v * v;
    }
    params_.sum_q = sum_q_;
    params_.sum_q2 = sum_q2_;
    const double V = L_[0] * L_[1] * L_[2];
    if (kspace_ == 3)  // RBSOG (kspace_calculator.cpp:1524-1528)
      e_self_ = sog_ready_ ? sog_self_energy(sog_, sum_q2_, cfg_.qqr2e) : 0.0;
    else if (kspace_)
      e_self_ = -cfg_.qqr2e * (std::sqrt(alpha_ / M_PI) * sum_q2_ +
                               0.5 * M_PI * sum_q_ * sum_q_ / (alpha_ * V));
  }

  void upload_charges() {
    std::vector<R> qr(n_);
    for (int i = 0; i < n_; ++i) qr[i] = (R)q_[i];
    HIPCHK(hipMemcpy(d_.charge, qr.data(), sizeof(R) * n_, hipMemcpyHostToDevice));
  }

  void setup() {
    HIPCHK(hipSetDevice(device_));
    const Config& c = cfg_;
    const bool coulomb = c.coulomb();
    kspace_ = !coulomb ? 0
              : c.coulomb_type == "RBE" ? 1
              : c.coulomb_type == "EWALD" ? 2
                                          : 3;
    // charged = nonzero in the device precision (the kernels test (R)q != 0)
    int nc = 0;
    if (coulomb)
      for (double v : q_) nc += (R)v != R(0);
    split_ = coulomb && nc > 0;
    if (!split_) nc = 0;
    const double V = L_[0] * L_[1] * L_[2];
    charge_sums();
    // alpha: manual override wins, else DeriveAlphaFromAccuracy
    // (kspace_calculator.cpp:105-157)
    if (coulomb) {
      if (c.has_alpha) {
        alpha_ = c.alpha;
      } else {
        // q2 = qqr2e * sum q^2 as the float product of float factors
        const double q2 = (double)((float)c.qqr2e * (float)sum_q2_);
        if (!(q2 > 0))
          throw Error(RBMDX_E_CONFIG,
                      "automatic alpha derivation requires a charged system");
        // the reference's types: double accuracy, g and q2, float cutoff,
        // sqrtf / logf of double arguments rounded to float, float alpha
        const float rcf = (float)c.cut_off;
        const double acc = (double)(float)c.accuracy * (double)(float)c.qqr2e;
        double g = acc * (double)sqrtf((float)((double)n_ * (double)rcf * V)) / (2.0 * q2);
        g = g >= 1.0 ? (1.35 - 0.15 * (double)logf((float)acc)) / (double)rcf
                     : (double)(sqrtf(-logf((float)g)) / rcf);
        alpha_ = (double)(float)(g * g);
      }
    }
    charge_sums();  // self energy needs alpha
    {
      hipDeviceProp_t prop;
      HIPCHK(hipGetDeviceProperties(&prop, device_));
      num_cu_ = std::max(1, prop.multiProcessorCount);
    }
    // interaction radii and grid
    const double rc = c.cut_off;
    const double rclj = c.lj_cut_off;
    const double ra = std::min(rclj, rc) + c.skin;  // all-atom list radius
    const double rb = rc + c.skin;                    // charged-pair radius
    const double rmax = split_ ? std::max(ra, rb) : ra;
    for (int d = 0; d < 3; ++d)
      if (!(rmax < 0.5 * L_[d]))
        throw Error(RBMDX_E_CONFIG,
                    "neighbor cutoff + skin must be below half the shortest box edge");
    // cells of side >= ra (the 27-cell stencil covers the all-atom list)
    for (int d = 0; d < 3; ++d) {
      int ncell = (int)std::floor(L_[d] / (ra * (1.0 + 1e-5)));
      if (ncell < 1) ncell = 1;
      if (ncell > 1023) ncell = 1023;
      grid_.nc[d] = ncell;
      const double side = L_[d] / ncell;
      grid_.ext_b[d] = std::max(1, (int)std::ceil(rb * (1.0 + 1e-5) / side));
      grid_.scale[d] = (float)(ncell / L_[d]);
    }
    grid_.ncell = grid_.nc[0] * grid_.nc[1] * grid_.nc[2];
    // bricks: 2x2x2 cells when the 4x4x4-cell neighbourhood fits the LDS
    // stage with margin, else single cells
    {
      const double apc = (double)n_ / grid_.ncell;
      const int stage_cap = (24 * 1024) / (int)sizeof(R4);  // the kernels' LDS stage
      grid_.bw = (64.0 * apc * 1.15 <= stage_cap) ? 2 : 1;
      int nb = 1;
      for (int d = 0; d < 3; ++d) {
        grid_.nb[d] = (grid_.nc[d] + grid_.bw - 1) / grid_.bw;
        nb *= grid_.nb[d];
      }
      grid_.nbrick = nb;
      const int per = (int)std::ceil(1.25 * apc * grid_.bw * grid_.bw * grid_.bw);
      grid_.block = std::min(512, std::max(64, (per + 63) / 64 * 64));
      // mostly charged systems put nearly all pairs in list B: kernels of
      // their own take them, one wave per charged atom, with one resident
      // round of 5 waves per SIMD (contiguous runs of atoms per wave)
      grid_.psub = 1;
      grid_.qblocks = 0;
      grid_.dsplit = 1;
      if (split_ && 2 * (long)nc > (long)n_) {
        grid_.qblocks = std::max(1, std::min((nc + kQWaves - 1) / kQWaves,
                                             num_cu_ * 4 * 5 / kQWaves));
        // the list build's groups of 64 charged atoms per cell, with margin
        // for cells above the mean
        const double qpc = (double)nc / grid_.ncell;
        grid_.dsplit = std::max(1, std::min(16, (int)std::ceil(1.25 * qpc / 64.0)));
      }
      // mostly neutral: one more wave per brick block takes the charged
      // pairs beside the list-A waves
      grid_.qwave = (split_ && grid_.qblocks == 0 && grid_.block <= 448) ? 64 : 0;
      // brick-major cell numbering
      std::vector<int> c_of_xyz(grid_.ncell), xyz_of_c(grid_.ncell),
          b0(nb + 1);
      int next = 0;
      const int* ncd = grid_.nc;
      for (int bz = 0; bz < grid_.nb[2]; ++bz)
        for (int by = 0; by < grid_.nb[1]; ++by)
          for (int bx = 0; bx < grid_.nb[0]; ++bx) {
            const int b = (bz * grid_.nb[1] + by) * grid_.nb[0] + bx;
            b0[b] = next;
            const int o[3] = {bx * grid_.bw, by * grid_.bw, bz * grid_.bw};
            const int w[3] = {std::min(grid_.bw, ncd[0] - o[0]),
                              std::min(grid_.bw, ncd[1] - o[1]),
                              std::min(grid_.bw, ncd[2] - o[2])};
            for (int lz = 0; lz < w[2]; ++lz)
              for (int ly = 0; ly < w[1]; ++ly)
                for (int lx = 0; lx < w[0]; ++lx) {
                  const int cx = o[0] + lx, cy = o[1] + ly, cz = o[2] + lz;
                  c_of_xyz[(cz * ncd[1] + cy) * ncd[0] + cx] = next;
                  xyz_of_c[next] = cx | (cy << 10) | (cz << 20);
                  ++next;
                }
          }
      b0[nb] = next;
      int* t0 = alloc<int>(grid_.ncell);
      int* t1 = alloc<int>(grid_.ncell);
      int* t2 = alloc<int>(nb + 1);
      HIPCHK(hipMemcpy(t0, c_of_xyz.data(), sizeof(int) * grid_.ncell,
                       hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(t1, xyz_of_c.data(), sizeof(int) * grid_.ncell,
                       hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(t2, b0.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
      grid_.cell_of_coord = t0;
      grid_.coord_of_cell = t1;
      grid_.brick_cell0 = t2;
    }
    for (int d = 0; d < 3; ++d) {
      box_.lo[d] = (R)lo_[d];
      box_.L[d] = (R)L_[d];
      box_.hi[d] = (R)(lo_[d] + L_[d]);
      box_.half[d] = (R)0.5 * box_.L[d];
      box_.invL[d] = (R)(1.0 / L_[d]);
    }
    // RBL frequency (pice_num)
    freq_ = c.neighbor_type == "RBL"
                ? rbl_frequency(n_, L_, c.r_core, rc, c.sample_num)
                : 1;
    // pair constants, thresholds formed in R as the reference forms them
    {
      const R rs = (R)c.r_core, rcr = (R)rc, rl = (R)rclj, eps = (R)1e-4;
      pc_.rs2 = rs * rs;
      pc_.rc2 = rcr * rcr;
      pc_.rclj2 = rl * rl;
      pc_.t_core = pc_.rs2 - eps;
      pc_.t_shell = pc_.rc2 - eps;
      pc_.eps_min = eps;
      {
        // one interval for list-A pairs without shell pairs: the LJ and
        // Coulomb cutoffs and, with RBL, core (d^2 < r_core^2 - EPSILON,
        // d^2 < r_core^2) and the guard d^2 > EPSILON
        const bool rbl = c.neighbor_type == "RBL";
        R hi = std::min(pc_.rclj2, pc_.rc2);
        if (rbl) hi = std::min(hi, std::min(pc_.t_core, pc_.rs2));
        pc_.a_hi = hi;
        pc_.a_lo = rbl ? eps : (R)-1;
      }
      pc_.r2_list_a = (R)(ra * ra);
      pc_.r2_list_b = (R)(rb * rb);
      pc_.alpha = (R)alpha_;
      pc_.g = (R)std::sqrt(alpha_);
      pc_.qqr2e = (R)c.qqr2e;
      pc_.two_g_over_sqrtpi = (R)(2.0 * std::sqrt(alpha_) / std::sqrt(M_PI));
      pc_.freq = (R)freq_;
      pc_.ifreq = freq_;
      pc_.fmask = (freq_ >= 2 && (freq_ & (freq_ - 1)) == 0) ? (unsigned)(freq_ - 1) : 0u;
      pc_.c32 = freq_ >= 2 ? (unsigned)((1ULL << 32) % (unsigned long long)freq_) : 0u;
      pc_.ntypes = ntypes_;
      pc_.sog = 0;
      for (int q = 0; q < 6; ++q) pc_.T[q] = (R)0;
    }
    // SOG far field (RBSOG k-space) and, for VERLET-SOG, its near field
    if (c.rbsog_level >= 0) {
      sog_ = sog_set(c.rbsog_level, rc, L_, 1, c.rbsog_exact);
      sog_ready_ = true;
      if (c.neighbor_type == "VERLET-SOG" || c.sog_near) {
        const SogNear nf = sog_near(sog_);
        pc_.sog = 1;
        for (int q = 0; q < 6; ++q) pc_.T[q] = (R)nf.T[q];
        std::vector<R> tab(2 * nf.omega.size());
        for (size_t l = 0; l < nf.omega.size(); ++l) {
          tab[2 * l] = (R)nf.omega[l];
          tab[2 * l + 1] = (R)nf.inv2s2[l];
        }
        R* dt = alloc<R>(tab.size());
        HIPCHK(hipMemcpy(dt, tab.data(), sizeof(R) * tab.size(), hipMemcpyHostToDevice));
        pc_.far_tab = dt;
        pc_.far_m = (int)nf.omega.size();
      }
      charge_sums();
    }
    params_.rbl = c.neighbor_type == "RBL" ? 1 : 0;
    params_.coulomb = coulomb ? 1 : 0;
    params_.kspace = kspace_;
    params_.mean_fix = params_.rbl;
    params_.split = split_ ? 1 : 0;
    params_.special = 0;
    d_.spn = d_.spn_s = nullptr;
    d_.spid = d_.spid_s = nullptr;
    pc_.ms = 0;
    pc_.w12 = (R)c.w_special[0];
    pc_.w13 = (R)c.w_special[1];
    pc_.w14 = (R)c.w_special[2];
    params_.qqr2e = c.qqr2e;
    params_.alpha = alpha_;
    params_.volume = V;
    params_.n_global = (double)n_;
    {
      const float h = (float)(0.5 * c.skin);
      const float thr = h * h;
      std::memcpy(&rebuild_thr_, &thr, sizeof(thr));
    }
    // ---- device buffers
    d_.n = n_;
    d_.nc = nc;
    d_.type = alloc<int>(n_);
    d_.gid = alloc<int>(n_);
    d_.charge = alloc<R>(n_);
    d_.ref = alloc<R4>(n_);
    d_.key = alloc<int>(n_);
    d_.slot = alloc<int>(n_);
    d_.inv_perm = alloc<int>(n_);
    d_.c_of = alloc<int>(n_);
    d_.key_count = alloc<int>(2 * (size_t)grid_.ncell);
    d_.key_start = alloc<int>(2 * (size_t)grid_.ncell + 1);
    d_.cstart = alloc<int>((size_t)grid_.ncell + 1);
    d_.perm = alloc<int>(n_);
    d_.cell_s = alloc<int>(n_);
    d_.gid_s = alloc<int>(n_);
    d_.hj_s = alloc<unsigned long long>(n_);
    d_.s2c = alloc<int>(n_);
    d_.xs = alloc<R4>(n_);
    d_.fa = alloc<R4>(n_);
    d_.c2s = alloc<int>(nc);
    d_.gidc = alloc<int>(nc);
    d_.hjc = alloc<unsigned long long>(nc);
    d_.typec = alloc<int>(nc);
    d_.xqc = alloc<R4>(nc);
    d_.fk = alloc<R4>(n_);
    d_.fb = alloc<R4>(n_);
    d_.c2i = alloc<int>(nc);
    d_.qcall = alloc<int>(nc);
    // the k-space kernels' fixed order of the charged atoms: ascending caller
    // index (zero charges stay zero, so the set is fixed for the system)
    d_.xqs = alloc<R4>(nc);
    d_.qpos = alloc<int>(n_);
    {
      std::vector<int> qp(n_, -1);
      int k = 0;
      if (nc > 0)
        for (int i = 0; i < n_; ++i)
          if ((R)q_[i] != R(0)) qp[i] = k++;
      if (k != nc) throw Error(RBMDX_E_STATE, "charged-atom count mismatch");
      HIPCHK(hipMemcpy(d_.qpos, qp.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
      std::vector<int> qc(std::max(nc, 1), 0);
      for (int i = 0; i < n_; ++i)
        if (qp[i] >= 0) qc[qp[i]] = i;
      if (nc > 0)
        HIPCHK(hipMemcpy(d_.qcall, qc.data(), sizeof(int) * nc, hipMemcpyHostToDevice));
    }
    d_.cnt_a = alloc<int>(n_);
    d_.order_a = alloc<int>(n_);
    d_.cnt_b = alloc<int>(nc);
    d_.inv_mass = alloc<R>(ntypes_);
    md_mass_ = alloc<double>(ntypes_);
    {
      std::vector<R> one(ntypes_, (R)1);
      HIPCHK(hipMemcpy(d_.inv_mass, one.data(), sizeof(R) * ntypes_,
                       hipMemcpyHostToDevice));
      std::vector<double> oned(ntypes_, 1.0);
      HIPCHK(hipMemcpy(md_mass_, oned.data(), sizeof(double) * ntypes_,
                       hipMemcpyHostToDevice));
    }
    md_ready_ = false;
    md_state_ = nullptr;
    d_.brick_tab = alloc<int>((size_t)grid_.nbrick * kBrickTabInts);
    d_.pair_part_blocks = grid_.nbrick * grid_.psub;
    d_.pair_part = alloc<double>((size_t)11 * grid_.nbrick * grid_.psub);
    d_.pairb_part = grid_.qblocks ? alloc<double>((size_t)11 * grid_.qblocks) : nullptr;
    d_.skin_part = alloc<float>(kSkinBlocksMax + 1);  // + the control block
    d_.lista_part = alloc<int4>(grid_.nbrick);
    d_.listb_part = alloc<int2>(std::max(1, list_b_count(grid_, nc)));
    d_.thermo = alloc<double>(32);
    d_.mean = alloc<double>(3);
    d_.ctrl = alloc<Ctrl>(1);
    d_.export_count = alloc<int>(1);
    {
      // a valid order before the first rebuild (k_skin gathers through it on
      // every evaluation): the identity, with no charged copies
      std::vector<int> iota(n_);
      for (int i = 0; i < n_; ++i) iota[i] = i;
      HIPCHK(hipMemcpy(d_.inv_perm, iota.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_.order_a, iota.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
      HIPCHK(hipMemset(d_.c_of, 0xff, sizeof(int) * n_));
    }
    HIPCHK(hipMemset(d_.key_count, 0, sizeof(int) * 2 * (size_t)grid_.ncell));
    HIPCHK(hipMemset(d_.ctrl, 0, sizeof(Ctrl)));
    HIPCHK(hipMemset(d_.thermo, 0, sizeof(double) * 32));
    HIPCHK(hipMemset(d_.mean, 0, sizeof(double) * 3));
    HIPCHK(hipMemset(d_.ref, 0, sizeof(R4) * n_));
    // caller-order force rows: the neutral atoms' rows of fb and fk stay 0
    HIPCHK(hipMemset(d_.fk, 0, sizeof(R4) * n_));
    HIPCHK(hipMemset(d_.fb, 0, sizeof(R4) * n_));
    {
      void* hp = nullptr;
      HIPCHK(hipHostMalloc(&hp, sizeof(Ctrl), hipHostMallocMapped));
      std::memset(hp, 0, sizeof(Ctrl));
      ctrl_host_ = (Ctrl*)hp;
      void* dp = nullptr;
      HIPCHK(hipHostGetDevicePointer(&dp, hp, 0));
      d_.ctrl_host = (Ctrl*)dp;
    }
    HIPCHK(hipMemcpy(d_.type, type_.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_.gid, id_.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
    upload_charges();
    lj_tab_ = nullptr;  // sized for this system's ntypes
    upload_lj();
    // ---- k-space
    if (kspace_ == 3) {
      // RBSOG: P sampled vectors (weights S/(4 pi P)) followed by the direct
      // vectors |m|^2 <= Kcut (weights f(k) k^2 / (4 pi)), so that every k
      // enters the force and energy in the RBE form
      // E = qqr2e (2 pi / V) w |rho|^2 / k^2 (kspace_calculator.cpp:1331-1543)
      const int P = c.coulomb_sample_num;
      const double qqr2e = c.qqr2e;
      std::vector<int> dm;
      for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
          for (int kk = -1; kk <= 1; ++kk)
            if ((i || j || kk) && i * i + j * j + kk * kk <= sog_.kcut) {
              dm.push_back(i);
              dm.push_back(j);
              dm.push_back(kk);
            }
      sog_ndirect_ = (int)dm.size() / 3;
      d_.kcap = P + sog_ndirect_;
      std::vector<double> w(d_.kcap);
      std::vector<double2> vir(d_.kcap);
      for (int k = 0; k < P; ++k) {
        w[k] = sog_.S / (4 * M_PI * P);
        vir[k] = make_double2(-(sog_.S_npt / P) * qqr2e / (4.0 * V), 1.0);
      }
      sog_direct_m_.assign(dm.begin(), dm.end());
      for (int e = 0; e < sog_ndirect_; ++e) {
        double k2 = 0;
        for (int d = 0; d < 3; ++d) {
          const double kd = 2 * M_PI * dm[3 * e + d] / L_[d];
          k2 += kd * kd;
        }
        w[P + e] = sog_f(sog_, sog_.coef, k2) * k2 / (4 * M_PI);
        vir[P + e] = make_double2(-qqr2e / (4.0 * V) * sog_f(sog_, sog_.coef_npt, k2), 0.0);
      }
      rbe_S_ = sog_.S;
      sampler_.reset(new SogSampler(sog_, L_, P));
      if (c.has_seed) {
        base_seed_ = c.seed;
      } else {
        std::random_device rd;
        base_seed_ = rd();
      }
      alloc_kspace(w, nullptr);
      d_.kvir = alloc<double2>(d_.kcap);
      d_.src_dev = alloc<int>(d_.kcap);
      HIPCHK(hipMemcpy(d_.kvir, vir.data(), sizeof(double2) * d_.kcap, hipMemcpyHostToDevice));
      // the direct vectors and their (own) virial sources stay resident
      HIPCHK(hipMemcpy(d_.m_dev + 3 * P, dm.data(), sizeof(int) * dm.size(),
                       hipMemcpyHostToDevice));
      std::vector<int> src(d_.kcap);
      for (int e = 0; e < d_.kcap; ++e) src[e] = e;
      HIPCHK(hipMemcpy(d_.src_dev, src.data(), sizeof(int) * d_.kcap, hipMemcpyHostToDevice));
      alloc_stage(4);
    } else if (kspace_ == 1) {
      const int P = c.coulomb_sample_num;
      d_.kcap = std::max(P, 1);
      // S = prod_d H_d - 1 (kspace_calculator.cpp:68-78), in FP64
      double h[3];
      for (int d = 0; d < 3; ++d) {
        const double f = -(alpha_ * L_[d] * L_[d]);
        h[d] = 0;
        for (int mm = -10; mm <= 10; ++mm) h[d] += std::exp(mm * mm * f);
        h[d] *= std::sqrt(-f / M_PI);
      }
      rbe_S_ = h[0] * h[1] * h[2] - 1.0;
      const float Lf[3] = {(float)L_[0], (float)L_[1], (float)L_[2]};
      sampler_.reset(new RbeSampler((float)alpha_, Lf, P));
      if (c.has_seed) {
        base_seed_ = c.seed;
      } else {
        std::random_device rd;
        base_seed_ = rd();
      }
      std::vector<double> w(d_.kcap, rbe_S_ / P);
      alloc_kspace(w, nullptr);
      alloc_stage(3);
    } else if (kspace_ == 2) {
      std::vector<int> mv;
      std::vector<double> w;
      ewald_grid(mv, w);
      if ((int)w.size() > kMaxK)
        throw Error(RBMDX_E_CONFIG, "EWALD kmax grid too large");
      d_.kcap = (int)w.size();
      alloc_kspace(w, &mv);
      params_.num_k = d_.kcap;
    }
    // energy_rbe_flag = "yes" (RBE): the exact Ewald energy on energy
    // evaluations, reported beside the RBE estimate (kspace_calculator.cpp:
    // 628-645, which then overwrites it)
    d_.kex = 0;
    if (kspace_ == 1 && c.energy_rbe && d_.nc > 0) {
      std::vector<int> mv;
      std::vector<double> w;
      ewald_grid(mv, w);
      if ((int)w.size() > kMaxK)
        throw Error(RBMDX_E_CONFIG, "energy_rbe_flag kmax grid too large");
      const int nchunks = std::max(1, (d_.nc + kSkChunk - 1) / kSkChunk);
      d_.kex = (int)w.size();
      d_.m_ex = alloc<int>(mv.size());
      d_.kw_ex = alloc<double>(w.size());
      d_.sk_ex = alloc<double2>((size_t)nchunks * d_.kex);
      HIPCHK(hipMemcpy(d_.m_ex, mv.data(), sizeof(int) * mv.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_.kw_ex, w.data(), sizeof(double) * w.size(),
                       hipMemcpyHostToDevice));
    }
    params_.num_k = kspace_ == 1 ? c.coulomb_sample_num
                    : kspace_ == 3 ? d_.kcap
                                   : params_.num_k;
    // step-parameter ring and the per-slot reuse events
    if (!hdr_) {
      void* hp = nullptr;
      HIPCHK(hipHostMalloc(&hp, sizeof(StepHdr) * kSlots, hipHostMallocMapped));
      hdr_ = (StepHdr*)hp;
      std::memset(hdr_, 0, sizeof(StepHdr) * kSlots);
      void* dp = nullptr;
      HIPCHK(hipHostGetDevicePointer(&dp, hp, 0));
      d_.hdr_ring = (const StepHdr*)dp;
      for (int q = 0; q < kSlots; ++q)
        HIPCHK(hipEventCreateWithFlags(&slot_events_[q], hipEventDisableTiming));
    }
    const char* gr = std::getenv("RBMDX_GRAPH");
    graph_on_ = !(gr && std::atoi(gr) == 0);
    ++gen_;
    timing_ = std::getenv("RBMDX_TIMING") && std::atoi(std::getenv("RBMDX_TIMING"));
    HIPCHK(hipDeviceSynchronize());
    ready_ = true;
  }

  // Host-mapped ring of per-evaluation k-samples: ints_per_k values per k
  // (m, or m then the virial-chain index for RBSOG), kSlots evaluations.
  void alloc_stage(int ints_per_k) {
    stage_stride_ = ints_per_k * (size_t)d_.kcap;
    void* hp = nullptr;
    HIPCHK(hipHostMalloc(&hp, sizeof(int32_t) * stage_stride_ * kSlots,
                         hipHostMallocMapped));
    stage_ = (int32_t*)hp;
    void* dp = nullptr;
    HIPCHK(hipHostGetDevicePointer(&dp, hp, 0));
    stage_dev_ = (int32_t*)dp;
    d_.stage_ring = stage_dev_;
    d_.stage_stride = (int)stage_stride_;
  }

  // The kmax grid of exact Ewald (manual coulomb.kmax, else SetKspacePara's
  // rms rule, kspace_calculator.cpp:909-986) with weights exp(-k^2 / 4 alpha).
  void ewald_grid(std::vector<int>& mv, std::vector<double>& w) const {
    const Config& c = cfg_;
    int km[3];
    if (c.has_kmax) {
      for (int d = 0; d < 3; ++d) km[d] = c.kmax[d];
    } else {
      const double g = std::sqrt(alpha_);
      const double acc = c.has_alpha ? 1e-5 * c.qqr2e : c.accuracy * c.qqr2e;
      const double q2 = c.qqr2e * sum_q2_;
      for (int d = 0; d < 3; ++d) {
        int k = 1;
        auto rms = [&](int kk) {
          return 2.0 * q2 * g / L_[d] * std::sqrt(1.0 / (M_PI * kk * n_)) *
                 std::exp(-M_PI * M_PI * kk * kk / (g * g * L_[d] * L_[d]));
        };
        while (rms(k) > acc && k < 512) ++k;
        km[d] = k;
      }
    }
    mv.clear();
    w.clear();
    for (int a = -km[0]; a <= km[0]; ++a)
      for (int b = -km[1]; b <= km[1]; ++b)
        for (int e = -km[2]; e <= km[2]; ++e) {
          if (!a && !b && !e) continue;
          mv.push_back(a);
          mv.push_back(b);
          mv.push_back(e);
          const double kx = 2 * M_PI * a / L_[0], ky = 2 * M_PI * b / L_[1],
                       kz = 2 * M_PI * e / L_[2];
          w.push_back(std::exp(-(kx * kx + ky * ky + kz * kz) / (4 * alpha_)));
        }
  }

  void alloc_kspace(const std::vector<double>& w, const std::vector<int>* mv) {
    // S(k) partials: one row per kSkAtoms charged atoms
    const int nchunks = std::max(1, (d_.nc + kSkAtoms - 1) / kSkAtoms);
    d_.m_dev = alloc<int>(3 * (size_t)d_.kcap);
    d_.kw = alloc<double>(d_.kcap);
    d_.sk_part = alloc<double2>((size_t)nchunks * d_.kcap);
    d_.kfin_part = alloc<double>((size_t)8 * d_.kcap);
    d_.kf0 = alloc<float4>(d_.kcap);
    d_.kf1 = alloc<float4>(d_.kcap);
    d_.rho = alloc<double2>(d_.kcap);
    HIPCHK(hipMemcpy(d_.kw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    if (mv)
      HIPCHK(hipMemcpy(d_.m_dev, mv->data(), sizeof(int) * mv->size(),
                       hipMemcpyHostToDevice));
    else
      HIPCHK(hipMemset(d_.m_dev, 0, sizeof(int) * 3 * (size_t)d_.kcap));
  }

  StepArgs<R> args(int64_t step, uint32_t flags) {
    StepArgs<R> a;
    a.box = box_;
    a.grid = grid_;
    a.pc = pc_;
    a.pc.epoch = (long long)step;
    a.p = params_;
    a.p.eflag = (flags & RBMDX_FLAG_ENERGY) ? 1 : 0;
    a.lj_tab = lj_tab_;
    a.rebuild_thr = rebuild_thr_;
    a.dt = leap_dt_;
    a.dt_ftm = leap_dt_ * cfg_.ftm2v;
    if (md_kick_) {  // the MD step's closing half-kick in the assembly kernel
      a.md_kick = 1;
      a.dt = md_.dt;
      a.ftm2v = cfg_.ftm2v;
    }
    a.e_self = e_self_;
    a.m_use = d_.m_dev;
    return a;
  }

  // First call: count-only build to size the candidate lists.
  void plan_capacity(hipStream_t s) {
    StepArgs<R> a = args(0, 0);
    a.count_only = 1;
    HIPCHK(hipMemset(&d_.ctrl->max_count_a, 0, 2 * sizeof(int)));
    HIPCHK(launch_build<R>(d_, a, s));
    HIPCHK(hipStreamSynchronize(s));
    Ctrl c;
    HIPCHK(hipMemcpy(&c, d_.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    alloc_lists(c.max_count_a, c.max_count_b);
  }

  void alloc_lists(int max_a, int max_b) {
    lists_stale_ = true;
    ++gen_;
    // list A is read 8 entries at a time: capacity is a multiple of 8
    const int cap_a = ((int)std::ceil(max_a * 1.2) + 16 + 7) / 8 * 8;
    // list B rows start 16-byte aligned (the dense build stores 4 entries
    // at a time): capacity is a multiple of 4
    const int cap_b = ((int)std::ceil(max_b * 1.2) + 16 + 3) / 4 * 4;
    if (d_.list_a) release(d_.list_a);
    if (d_.list_b) release(d_.list_b);
    if (d_.list_t) release(d_.list_t);
    d_.list_t = nullptr;
    d_.list_a = alloc<uint16_t>((size_t)cap_a * n_);
    d_.cap_a = cap_a;
    if (d_.nc > 0) {
      d_.list_b = alloc<int>((size_t)cap_b * d_.nc);
      d_.cap_b = cap_b;
      if (grid_.qblocks) d_.list_t = alloc<unsigned char>((size_t)cap_b * d_.nc);
    } else {
      d_.cap_b = 0;
    }
  }

  void release(void* p) {
    auto it = std::find(allocs_.begin(), allocs_.end(), p);
    if (it != allocs_.end()) allocs_.erase(it);
    hipFree(p);
  }

  // Reads the host-mapped counters written by the last completed call.
  bool grow_if_needed(hipStream_t s) {
    const volatile Ctrl* h = ctrl_host_;
    const int ma = h->max_count_a, mb = h->max_count_b;
    if (ma <= d_.cap_a && mb <= d_.cap_b) return false;
    HIPCHK(hipStreamSynchronize(s));
    alloc_lists(std::max(ma, d_.cap_a), std::max(mb, d_.cap_b));
    return true;
  }
};

}  // namespace

struct rbmdx_ctx {
  std::string err;
  int device = 0;
  Config cfg;
  std::unique_ptr<EngineBase> eng;
  Engine<float>* ef = nullptr;
  Engine<double>* ed = nullptr;
};

namespace {
template <typename F>
int guarded(rbmdx_ctx* c, F&& f) {
  try {
    f();
    // Every HIP status this call produced was checked; do not leak one into
    // the caller's hipGetLastError().
    (void)hipGetLastError();
    return RBMDX_OK;
  } catch (const Error& e) {
    if (c) c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return RBMDX_E_HIP;
  }
}
}  // namespace

extern "C" {

int rbmdx_create(const char* json_text, int device, rbmdx_ctx** out) {
  if (!out) return RBMDX_E_ARG;
  *out = nullptr;
  auto* c = new rbmdx_ctx();
  const int rc = guarded(c, [&] {
    if (!json_text) throw Error(RBMDX_E_ARG, "null JSON text");
    c->cfg = parse_config(json_text);
    c->device = device;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
      throw Error(RBMDX_E_ARG, "device ordinal out of range");
    HIPCHK(hipSetDevice(device));
    // the allocation-free evaluation's static scratch on this device
    HIPCHK(noalloc_scratch_init(device));
    if (c->cfg.fp64) {
      c->ed = new Engine<double>(c->cfg, device);
      c->eng.reset(c->ed);
    } else {
      c->ef = new Engine<float>(c->cfg, device);
      c->eng.reset(c->ef);
    }
  });
  *out = c;  // returned even on error so rbmdx_last_error can explain
  return rc;
}

void rbmdx_destroy(rbmdx_ctx* ctx) { delete ctx; }

const char* rbmdx_last_error(const rbmdx_ctx* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int rbmdx_set_system(rbmdx_ctx* ctx, int64_t n, const double* box_lo,
                     const double* box_len, const int32_t* type,
                     const int32_t* id, const double* charge, int32_t ntypes,
                     const double* eps, const double* sigma) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded(ctx, [&] {
    if (!ctx->eng) throw Error(RBMDX_E_STATE, "context was not created");
    if (!box_lo || !box_len || !type || !id || !charge || !eps || !sigma)
      throw Error(RBMDX_E_ARG, "null argument");
    // a new system replaces the old engine wholesale
    if (ctx->cfg.fp64) {
      ctx->ed = new Engine<double>(ctx->cfg, ctx->device);
      ctx->eng.reset(ctx->ed);
      ctx->ed->set_system(n, box_lo, box_len, type, id, charge, ntypes, eps, sigma);
    } else {
      ctx->ef = new Engine<float>(ctx->cfg, ctx->device);
      ctx->eng.reset(ctx->ef);
      ctx->ef->set_system(n, box_lo, box_len, type, id, charge, ntypes, eps, sigma);
    }
  });
}

int rbmdx_set_special(rbmdx_ctx* ctx, int32_t maxspecial, const int32_t* nspecial,
                      const int32_t* special) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded(ctx, [&] { ctx->eng->set_special(maxspecial, nspecial, special); });
}

int rbmdx_set_charges(rbmdx_ctx* ctx, const double* charge) {
  if (!ctx || !charge) return RBMDX_E_ARG;
  return guarded(ctx, [&] { ctx->eng->set_charges(charge); });
}

int rbmdx_force(rbmdx_ctx* ctx, int64_t step, const void* px, const void* py,
                const void* pz, void* fx, void* fy, void* fz,
                const int32_t* m_samples, int32_t num_k, uint32_t flags,
                void* stream) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded(ctx, [&] {
    ctx->eng->force(step, px, py, pz, fx, fy, fz, m_samples, num_k, flags,
                    (hipStream_t)stream);
  });
}

int rbmdx_force_noalloc(rbmdx_ctx* ctx, int64_t step, const void* px, const void* py,
                        const void* pz, void* fx, void* fy, void* fz,
                        const int32_t* m_samples, int32_t num_k, uint32_t flags,
                        void* stream) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded(ctx, [&] {
    if (!ctx->eng) throw Error(RBMDX_E_STATE, "context was not created");
    ctx->eng->force_noalloc(step, px, py, pz, fx, fy, fz, m_samples, num_k, flags,
                            (hipStream_t)stream);
  });
}

int rbmdx_force_vjp(rbmdx_ctx* ctx, int64_t step, const void* px, const void* py,
                    const void* pz, const void* gx, const void* gy, const void* gz,
                    void* dx, void* dy, void* dz, double* dparam, uint32_t flags,
                    void* stream) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded(ctx, [&] {
    ctx->eng->force_vjp(step, px, py, pz, gx, gy, gz, dx, dy, dz, dparam, flags,
                        (hipStream_t)stream);
  });
}

int rbmdx_set_lj(rbmdx_ctx* ctx, const double* eps, const double* sigma) {
  if (!ctx || !eps || !sigma) return RBMDX_E_ARG;
  return guarded(ctx, [&] { ctx->eng->set_lj(eps, sigma); });
}

int rbmdx_set_masses(rbmdx_ctx* ctx, const double* mass) {
  if (!ctx || !mass) return RBMDX_E_ARG;
  return guarded(ctx, [&] { ctx->eng->set_masses(mass); });
}

int rbmdx_force_leapfrog(rbmdx_ctx* ctx, int64_t step, void* px, void* py,
                         void* pz, void* vx, void* vy, void* vz, void* fx,
                         void* fy, void* fz, double dt, uint32_t flags,
                         void* stream) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded(ctx, [&] {
    ctx->eng->force_leapfrog(step, px, py, pz, vx, vy, vz, fx, fy, fz, dt, flags,
                             (hipStream_t)stream);
  });
}

int rbmdx_md_init(rbmdx_ctx* ctx, void* px, void* py, void* pz, void* vx,
                  void* vy, void* vz, int32_t* ix, int32_t* iy, int32_t* iz,
                  void* fx, void* fy, void* fz, uint32_t flags, void* stream) {
  if (!ctx) return RBMDX_E_ARG;
  const MdCall c{px, py, pz, vx, vy, vz, fx, fy, fz, ix, iy, iz};
  return guarded(ctx, [&] { ctx->eng->md_init(c, flags, (hipStream_t)stream); });
}

int rbmdx_md_run(rbmdx_ctx* ctx, int32_t nsteps, void* px, void* py, void* pz,
                 void* vx, void* vy, void* vz, int32_t* ix, int32_t* iy,
                 int32_t* iz, void* fx, void* fy, void* fz, uint32_t flags,
                 void* stream) {
  if (!ctx) return RBMDX_E_ARG;
  const MdCall c{px, py, pz, vx, vy, vz, fx, fy, fz, ix, iy, iz};
  return guarded(ctx,
                 [&] { ctx->eng->md_run(nsteps, c, flags, (hipStream_t)stream); });
}

int rbmdx_md_state(rbmdx_ctx* ctx, rbmdx_md_state_t* out) {
  if (!ctx || !out) return RBMDX_E_ARG;
  return guarded(ctx, [&] { ctx->eng->md_state(out); });
}

int rbmdx_thermo(rbmdx_ctx* ctx, rbmdx_thermo_t* out) {
  if (!ctx || !out) return RBMDX_E_ARG;
  return guarded(ctx, [&] { ctx->eng->thermo(out); });
}

int rbmdx_rbe_sample(rbmdx_ctx* ctx, int64_t call_index, int32_t* m_out) {
  if (!ctx || !m_out || call_index < 0) return RBMDX_E_ARG;
  return guarded(ctx, [&] { ctx->eng->rbe_sample(call_index, m_out); });
}

int32_t rbmdx_num_samples(const rbmdx_ctx* ctx) {
  return (ctx && ctx->eng) ? ctx->eng->num_samples() : 0;
}

int64_t rbmdx_export_pairs(rbmdx_ctx* ctx, int64_t step, const void* px,
                           const void* py, const void* pz, int32_t* gi,
                           int32_t* gj, int32_t* cls, int64_t cap) {
  if (!ctx || !gi || !gj || !cls || cap < 0) return -RBMDX_E_ARG;
  int64_t r = 0;
  const int rc = guarded(ctx, [&] {
    r = ctx->eng->export_pairs(step, px, py, pz, 0, gi, gj, cls, cap);
  });
  return rc ? -rc : r;
}

int64_t rbmdx_export_candidates(rbmdx_ctx* ctx, int32_t* gi, int32_t* gj,
                                int32_t* cls, int64_t cap) {
  if (!ctx || !gi || !gj || !cls || cap < 0) return -RBMDX_E_ARG;
  int64_t r = 0;
  const int rc = guarded(ctx, [&] {
    r = ctx->eng->export_pairs(0, nullptr, nullptr, nullptr, 1, gi, gj, cls, cap);
  });
  return rc ? -rc : r;
}

int rbmdx_rbe_sample_host(double alpha, const double* L, int32_t P,
                          uint32_t base_seed, int64_t call_index,
                          int32_t* m_out) {
  if (!L || !m_out || P <= 0 || call_index < 0 || !(alpha > 0)) return RBMDX_E_ARG;
  return guarded(nullptr, [&] {
    const float Lf[3] = {(float)L[0], (float)L[1], (float)L[2]};
    RbeSampler s((float)alpha, Lf, P);
    s.draw(rbe_seed_of_call(base_seed, call_index), m_out);
  });
}

int rbmdx_sog_setup_host(int32_t level, double cut_off, const double* L,
                         int32_t exact, double* scal, double* coef,
                         double* coef_npt) {
  SogPresetValues pv;
  if (!L || !scal || !coef || !coef_npt || !sog_preset(level, &pv) ||
      !(cut_off > 0))
    return RBMDX_E_ARG;
  return guarded(nullptr, [&] {
    const SogSet s = sog_set(level, cut_off, L, 1, exact != 0);
    const double v[6] = {s.b, s.sigma, s.w0, s.S, s.S_npt, (double)s.mmax};
    for (int q = 0; q < 6; ++q) scal[q] = v[q];
    for (int l = 0; l < s.mmax; ++l) {
      coef[l] = s.coef[l];
      coef_npt[l] = s.coef_npt[l];
    }
  });
}

int rbmdx_sog_sample_host(int32_t level, double cut_off, const double* L,
                          int32_t exact, int32_t P, uint32_t base_seed,
                          int64_t call_index, int32_t* m_out, int32_t* idx_out) {
  SogPresetValues pv;
  if (!L || !m_out || !idx_out || P <= 0 || call_index < 0 ||
      !sog_preset(level, &pv) || !(cut_off > 0))
    return RBMDX_E_ARG;
  return guarded(nullptr, [&] {
    const SogSet s = sog_set(level, cut_off, L, 1, exact != 0);
    SogSampler smp(s, L, P);
    std::vector<int32_t> out((size_t)smp.out_size());
    smp.draw(rbe_seed_of_call(base_seed, call_index), out.data());
    std::copy(out.begin(), out.begin() + 3 * P, m_out);
    std::copy(out.begin() + 3 * P, out.end(), idx_out);
  });
}

int rbmdx_check_config(const char* json_text, char* err, int32_t errlen) {
  if (!json_text) return RBMDX_E_ARG;
  rbmdx_ctx tmp;
  const int rc = guarded(&tmp, [&] { parse_config(json_text); });
  if (err && errlen > 0) {
    std::strncpy(err, tmp.err.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = 0;
  }
  return rc;
}

int rbmdx_kernel_times(rbmdx_ctx* ctx, const char** names, double* usec,
                       int n) {
  if (!ctx || !usec) return -RBMDX_E_ARG;
  int r = 0;
  const int rc = guarded(ctx, [&] { r = ctx->eng->kernel_times(names, usec, n); });
  return rc ? -rc : r;
}

}  // extern "C"
