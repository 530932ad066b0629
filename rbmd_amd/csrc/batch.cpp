// Replica-batch host engine behind the rbmdx_batch_* entry points of
// include/rbmdx.h: many independent periodic boxes of one force field
// (RBMD's JSON surface, LJ/CUT or LJ/CUT/COUL/LONG with RBL or a full pair
// set and RBE k-space) evaluated together.  One evaluation of every replica
// is one captured hipGraph (seven stage kernels over the whole batch, plus a
// one-thread epoch update launched before it); the same kernels launched one
// replica at a time are the per-replica path.  Launch shapes come from a
// measuring autotuner whose choice is cached as JSON per device and batch
// shape.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "batch.hpp"
#include "host_common.hpp"
#include "rbmdx.h"
#include "sog.hpp"

using namespace rbmdx;

namespace {

// RBSOG direct vectors: every m != 0 with |m|^2 <= 1 (the single-system
// engine's kcut = 1 set, in its order)
constexpr int kSogDirect = 6;
constexpr int kSogDirectM[3 * kSogDirect] = {-1, 0, 0, 0, -1, 0, 0, 0, -1,
                                              0,  0, 1, 0, 1,  0, 1, 0, 0};

class Batch {
 public:
  Batch(const Config& c, int device) : cfg_(c), device_(device) {
    if (c.fp64)
      throw Error(RBMDX_E_CONFIG, "the replica batch computes in FP32 (precision fp32)");
    if (c.force_type != "LJ/CUT" && c.force_type != "LJ/CUT/COUL/LONG" &&
        c.force_type != "CVFF")
      throw Error(RBMDX_E_CONFIG,
                  "the replica batch supports LJ/CUT, LJ/CUT/COUL/LONG and CVFF");
    if (c.neighbor_type != "RBL" && c.neighbor_type != "Verlet-List")
      throw Error(RBMDX_E_CONFIG, "the replica batch supports RBL and Verlet-List");
    if (c.coulomb() && c.coulomb_type != "RBE" && c.coulomb_type != "EWALD" &&
        c.coulomb_type != "RBSOG")
      throw Error(RBMDX_E_CONFIG, "the replica batch supports RBE, EWALD and RBSOG k-space");
    if (c.coulomb() && c.coulomb_type == "RBSOG" && c.sog_near)
      throw Error(RBMDX_E_CONFIG,
                  "the replica batch's RBSOG uses the erfc near field (no rbsog_near_field)");
    if (c.coulomb() && c.coulomb_type == "EWALD") {
      if (!c.has_kmax) throw Error(RBMDX_E_CONFIG, "EWALD needs coulomb.kmax");
      const long K = (2L * c.kmax[0] + 1) * (2L * c.kmax[1] + 1) * (2L * c.kmax[2] + 1) - 1;
      if (K < 1 || K > kBatchMaxK)
        throw Error(RBMDX_E_CONFIG,
                    "the replica batch's EWALD grid must hold 1 to 512 k-vectors");
    }
    if (c.coulomb() && !c.has_alpha)
      throw Error(RBMDX_E_CONFIG,
                  "the replica batch needs coulomb.alpha (one alpha for every replica)");
    if (c.coulomb() && c.coulomb_type == "RBE" &&
        (c.coulomb_sample_num < 1 || c.coulomb_sample_num > kBatchMaxK))
      throw Error(RBMDX_E_CONFIG, "coulomb.sample_num must be in [1, 512] for the batch");
    // RBSOG: the P sampled k's and the 6 direct vectors |m|^2 <= 1
    if (c.coulomb() && c.coulomb_type == "RBSOG" &&
        (c.coulomb_sample_num < 1 || c.coulomb_sample_num + 6 > kBatchMaxK))
      throw Error(RBMDX_E_CONFIG, "coulomb.sample_num must be in [1, 506] for the batch's RBSOG");
    HIPCHK(hipSetDevice(device_));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device_));
    num_cu_ = prop.multiProcessorCount;
    arch_ = prop.gcnArchName;
  }

  ~Batch() {
    hipSetDevice(device_);
    if (last_stream_set_) hipStreamSynchronize(last_stream_);
    for (auto& kv : graphs_) hipGraphExecDestroy(kv.second);
    if (cap_stream_) hipStreamDestroy(cap_stream_);
    if (ev_[0]) hipEventDestroy(ev_[0]);
    if (ev_[1]) hipEventDestroy(ev_[1]);
    for (void* p : allocs_) hipFree(p);
    (void)hipGetLastError();
  }

  void set_systems(int R, const int32_t* natoms, const double* lo, const double* len,
                   const int32_t* type, const int32_t* id, const double* q, int ntypes,
                   const double* eps, const double* sig) {
    if (ready_) throw Error(RBMDX_E_STATE, "rbmdx_batch_set_systems was already called");
    if (R < 1) throw Error(RBMDX_E_ARG, "at least one replica");
    if (ntypes < 1 || ntypes > kBatchMaxTypes)
      throw Error(RBMDX_E_ARG, "ntypes must be in [1, 8] for the batch");
    const bool coul = cfg_.coulomb();
    const bool rbl = cfg_.neighbor_type == "RBL";
    const double rc = cfg_.cut_off;
    const double rl = std::min(cfg_.lj_cut_off > 0 ? cfg_.lj_cut_off : rc, rc);
    R_ = R;
    ntypes_ = ntypes;
    rep_.assign(R, RepInfo{});
    a0_.assign(R + 1, 0);
    q0_.assign(R + 1, 0);
    c0_.assign(R + 1, 0);
    self_.assign(R, 0.0);
    for (int r = 0; r < R; ++r) {
      if (natoms[r] < 1) throw Error(RBMDX_E_ARG, "every replica needs atoms");
      a0_[r + 1] = a0_[r] + natoms[r];
      if (a0_[r + 1] > (1 << 28)) throw Error(RBMDX_E_ARG, "batch too large");
    }
    N_ = a0_[R];
    for (int i = 0; i < N_; ++i)
      if (type[i] < 0 || type[i] >= ntypes)
        throw Error(RBMDX_E_ARG, "atom type out of range");
    std::vector<int> rep_of(N_);
    std::vector<double> kwd, kvd;  // RBSOG direct vectors, replica-major
    for (int r = 0; r < R; ++r) {
      RepInfo& p = rep_[r];
      const int n = natoms[r];
      p.a0 = a0_[r];
      p.n = n;
      double Ld[3];
      for (int d = 0; d < 3; ++d) {
        Ld[d] = len[3 * r + d];
        if (!(Ld[d] > 0) || !std::isfinite(Ld[d]) || !std::isfinite(lo[3 * r + d]))
          throw Error(RBMDX_E_ARG, "box lengths must be finite and > 0");
        if (!(rc < 0.5 * Ld[d]))
          throw Error(RBMDX_E_CONFIG, "cut_off must be below half the shortest box edge "
                                      "of every replica");
        p.box.lo[d] = (float)lo[3 * r + d];
        p.box.L[d] = (float)Ld[d];
        p.box.hi[d] = (float)(lo[3 * r + d] + Ld[d]);
        p.box.half[d] = 0.5f * p.box.L[d];
        p.box.invL[d] = (float)(1.0 / Ld[d]);
        int nc = (int)std::floor(Ld[d] / (rl * (1.0 + 1e-5)));
        nc = std::max(1, std::min(nc, 1023));
        p.nc[d] = nc;
        const double side = Ld[d] / nc;
        p.ext_c[d] = std::max(1, (int)std::ceil(rc * (1.0 + 1e-5) / side));
      }
      p.c0 = c0_[r];
      c0_[r + 1] = c0_[r] + p.nc[0] * p.nc[1] * p.nc[2];
      p.volume = Ld[0] * Ld[1] * Ld[2];
      int nq = 0;
      double sq = 0, sq2 = 0;
      for (int i = p.a0; i < p.a0 + n; ++i) {
        rep_of[i] = r;
        const bool charged = coul && (float)q[i] != 0.f;
        nq += charged;
        if (coul) {
          sq += q[i];
          sq2 += q[i] * q[i];
        }
      }
      p.q0 = q0_[r];
      p.nq = nq;
      q0_[r + 1] = q0_[r] + nq;
      const int f = rbl ? rbl_frequency(n, Ld, cfg_.r_core, rc, cfg_.sample_num) : 1;
      p.freq = (float)f;
      p.ifreq = f;
      p.fmask = (f >= 2 && (f & (f - 1)) == 0) ? (unsigned)(f - 1) : 0u;
      p.c32 = f >= 2 ? (unsigned)((1ULL << 32) % (unsigned long long)f) : 0u;
      if (coul && cfg_.coulomb_type == "RBSOG") {
        // this box's SOG set (coefficients and normalisation), as the
        // single-system engine forms it: sampled k's weigh S / (4 pi P),
        // the direct vectors f(k) k^2 / (4 pi)
        const SogSet ss = sog_set(cfg_.rbsog_level, rc, Ld, 1, cfg_.rbsog_exact);
        const int P = cfg_.coulomb_sample_num;
        p.kw = ss.S / (4 * M_PI * P);
        p.ktau = -(ss.S_npt / P) * cfg_.qqr2e / (4.0 * p.volume);
        for (int e = 0; e < kSogDirect; ++e) {
          double k2 = 0;
          for (int d = 0; d < 3; ++d) {
            const double kd = 2 * M_PI * kSogDirectM[3 * e + d] / Ld[d];
            k2 += kd * kd;
          }
          kwd.push_back(sog_f(ss, ss.coef, k2) * k2 / (4 * M_PI));
          kvd.push_back(-cfg_.qqr2e / (4.0 * p.volume) * sog_f(ss, ss.coef_npt, k2));
        }
        self_[r] = sog_self_energy(ss, sq2, cfg_.qqr2e);
      } else if (coul) {
        const double a = cfg_.alpha;
        // RBE normalisation S = prod_d H_d - 1 (kspace_calculator.cpp:68-78)
        double h[3];
        for (int d = 0; d < 3; ++d) {
          const double ff = -(a * Ld[d] * Ld[d]);
          h[d] = 0;
          for (int mm = -10; mm <= 10; ++mm) h[d] += std::exp(mm * mm * ff);
          h[d] *= std::sqrt(-ff / M_PI);
        }
        const double S = h[0] * h[1] * h[2] - 1.0;
        p.kw = S / cfg_.coulomb_sample_num;
        self_[r] = -cfg_.qqr2e * (std::sqrt(a / M_PI) * sq2 +
                                  0.5 * M_PI * sq * sq / (a * p.volume));
      }
    }
    NQ_ = q0_[R];
    ncell_ = c0_[R];
    P_ = coul ? cfg_.coulomb_sample_num : 0;
    d_.kmode = 0;
    d_.mstride = 3 * P_;
    d_.nd = 0;
    if (coul && cfg_.coulomb_type == "RBSOG") {
      d_.kmode = 2;
      d_.mstride = 4 * P_;
      d_.nd = kSogDirect;
      d_.dm = upload(std::vector<int>(kSogDirectM, kSogDirectM + 3 * kSogDirect));
      d_.kwd = upload(kwd);
      d_.kvd = upload(kvd);
    }
    if (coul && cfg_.coulomb_type == "EWALD") {
      // every m != 0 with |m_d| <= kmax_d, the same integer grid in every
      // replica (KSpaceCalculator's EWALD k-set, weights on the device)
      std::vector<int> grid;
      for (int z = -cfg_.kmax[2]; z <= cfg_.kmax[2]; ++z)
        for (int y = -cfg_.kmax[1]; y <= cfg_.kmax[1]; ++y)
          for (int x = -cfg_.kmax[0]; x <= cfg_.kmax[0]; ++x)
            if (x || y || z) {
              grid.push_back(x);
              grid.push_back(y);
              grid.push_back(z);
            }
      P_ = (int)(grid.size() / 3);
      std::vector<int> all;
      all.reserve(grid.size() * R);
      for (int r = 0; r < R; ++r) all.insert(all.end(), grid.begin(), grid.end());
      ewald_m_ = upload(all);
      d_.kmode = 1;
      d_.mstride = 3 * P_;
    }
    // pair constants shared by every replica (thresholds formed in FP32 as
    // the single-system engine forms them)
    {
      const float rs = (float)cfg_.r_core, rcr = (float)rc, rlj = (float)rl, e = 1e-4f;
      pc_.rs2 = rs * rs;
      pc_.rc2 = rcr * rcr;
      pc_.rclj2 = rlj * rlj;
      pc_.t_core = pc_.rs2 - e;
      pc_.t_shell = pc_.rc2 - e;
      pc_.eps_min = e;
      pc_.a_lo = -1.f;
      pc_.a_hi = pc_.rclj2;
      pc_.alpha = (float)cfg_.alpha;
      pc_.g = (float)std::sqrt(cfg_.alpha);
      pc_.qqr2e = (float)cfg_.qqr2e;
      pc_.two_g_over_sqrtpi = (float)(2.0 * std::sqrt(cfg_.alpha) / std::sqrt(M_PI));
      pc_.freq = 1.f;
      pc_.ifreq = 1;
      pc_.ntypes = ntypes;
      pc_.sog = 0;
      pc_.ms = 0;
      pc_.w12 = (float)cfg_.w_special[0];
      pc_.w13 = (float)cfg_.w_special[1];
      pc_.w14 = (float)cfg_.w_special[2];
    }
    HIPCHK(hipSetDevice(device_));
    // device arrays
    std::vector<float> qf(N_);
    for (int i = 0; i < N_; ++i) qf[i] = coul ? (float)q[i] : 0.f;
    std::vector<float> lj(4 * ntypes * ntypes);
    for (int a = 0; a < ntypes; ++a)
      for (int b = 0; b < ntypes; ++b) {
        const double e = std::sqrt(eps[a] * eps[b]), s = 0.5 * (sig[a] + sig[b]);
        float* row = &lj[4 * (a * ntypes + b)];
        row[0] = -24.f * (float)e;
        row[1] = (float)(s * s);
        row[2] = (float)e;
        row[3] = (float)s;
      }
    d_.R = R;
    d_.N = N_;
    d_.NQ = NQ_;
    d_.ncell = ncell_;
    d_.P = P_;
    d_.rep_of = upload(rep_of);
    d_.type = upload(std::vector<int>(type, type + N_));
    d_.gid = upload(std::vector<int>(id, id + N_));
    d_.charge = upload(qf);
    d_.lj = upload(lj);
    d_.cnt = alloc<int>(ncell_);
    d_.cntq = alloc<int>(ncell_);
    HIPCHK(hipMemset(d_.cnt, 0, sizeof(int) * ncell_));
    HIPCHK(hipMemset(d_.cntq, 0, sizeof(int) * ncell_));
    d_.start = alloc<int>(ncell_ + 1);
    d_.qstart = alloc<int>(ncell_ + 1);
    d_.cell_of = alloc<int>(N_);
    d_.slot = alloc<int>(N_);
    d_.tmp = alloc<int>(N_);
    d_.xs = alloc<float4>(N_);
    d_.tg = alloc<int2>(N_);
    d_.hj = alloc<unsigned long long>(N_);
    d_.perm = alloc<int>(N_);
    d_.qidx = alloc<int>(std::max(1, NQ_));
    d_.fs = alloc<float4>(N_);
    d_.racc = alloc<double>((size_t)R * kBatchAcc);
    d_.kacc = alloc<double>((size_t)R * kBatchKAcc);
    HIPCHK(hipMemset(d_.racc, 0, sizeof(double) * R * kBatchAcc));
    HIPCHK(hipMemset(d_.kacc, 0, sizeof(double) * R * kBatchKAcc));
    d_.epoch = alloc<long long>(1);
    // work list sized for the finest bricks (bw = 1)
    size_t wmax = 0;
    for (const RepInfo& p : rep_) wmax += (size_t)p.nc[0] * p.nc[1] * p.nc[2];
    work_dev_ = alloc<BatchWork>(wmax);
    rep_dev_ = alloc<RepInfo>(R);
    d_.work = work_dev_;
    d_.rep = rep_dev_;
    layout(tune_.bw);
    ready_ = true;
  }

  // CVFF special-bond topology of every replica, caller order (the atoms of
  // all replicas concatenated): counts (n12, n13, n14) per atom, clamped as
  // the single-system engine clamps them, and ms partner ids per atom, ids
  // within the atom's replica.  Caller-order rows need no re-sort.
  void set_special(int32_t ms, const int32_t* nspecial, const int32_t* special) {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_batch_set_systems first");
    if (cfg_.force_type != "CVFF")
      throw Error(RBMDX_E_STATE, "special bonds apply to force_field.type CVFF only");
    if (ms < 0 || ms > kMaxSpecial)
      throw Error(RBMDX_E_ARG, "maxspecial must be in [0, 64]");
    if (ms > 0 && (!nspecial || !special))
      throw Error(RBMDX_E_ARG, "null special-bond arrays");
    HIPCHK(hipSetDevice(device_));
    if (last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
    std::vector<int4> cnt(N_);
    bool any = false;
    for (int i = 0; i < N_ && ms > 0; ++i) {
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
    special_ = 0;
    pc_.ms = 0;
    if (any) {
      d_.spn = upload(cnt);
      d_.spid = upload(std::vector<int>(special, special + (size_t)N_ * ms));
      special_ = 1;
      pc_.ms = ms;
    }
    drop_graphs();
  }

  // Bricks of width bw: per-replica brick counts and the replica-major work
  // list (one item per brick).
  void layout(int bw) {
    w0_.assign(R_ + 1, 0);
    std::vector<BatchWork> work;
    for (int r = 0; r < R_; ++r) {
      RepInfo& p = rep_[r];
      for (int d = 0; d < 3; ++d) p.nb[d] = (p.nc[d] + bw - 1) / bw;
      const int nb = p.nb[0] * p.nb[1] * p.nb[2];
      for (int b = 0; b < nb; ++b) work.push_back(BatchWork{r, b});
      w0_[r + 1] = w0_[r] + nb;
    }
    HIPCHK(hipMemcpy(rep_dev_, rep_.data(), sizeof(RepInfo) * R_, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(work_dev_, work.data(), sizeof(BatchWork) * work.size(),
                     hipMemcpyHostToDevice));
    d_.nwork = (int)work.size();
    layout_bw_ = bw;
    drop_graphs();
  }

  BatchRange range(int r0, int r1) const {
    BatchRange g;
    g.r0 = r0;
    g.r1 = r1;
    g.a0 = a0_[r0];
    g.a1 = a0_[r1];
    g.c0 = c0_[r0];
    g.c1 = c0_[r1];
    g.w0 = w0_[r0];
    g.w1 = w0_[r1];
    g.q0 = q0_[r0];
    return g;
  }

  void bind(const float* px, const float* py, const float* pz, float* fx, float* fy,
            float* fz, const int32_t* m) {
    if (!px || !py || !pz || !fx || !fy || !fz)
      throw Error(RBMDX_E_ARG, "null position or force pointer");
    if (d_.kmode == 1) {
      if (m) throw Error(RBMDX_E_ARG, "EWALD uses the kmax grid: pass no k-samples");
      m = ewald_m_;
    }
    if (P_ > 0 && !m) throw Error(RBMDX_E_ARG, "RBE needs the k-samples of every replica");
    d_.px = px;
    d_.py = py;
    d_.pz = pz;
    d_.fx = fx;
    d_.fy = fy;
    d_.fz = fz;
    d_.m = m;
  }

  void launch_all(int eflag, hipStream_t s) {
    HIPCHK(batch_launch(d_, pc_, rbl(), cfg_.coulomb() ? 1 : 0, eflag, special_, range(0, R_), tune_,
                        num_cu_, s));
  }

  void force(int64_t step, const float* px, const float* py, const float* pz, float* fx,
             float* fy, float* fz, const int32_t* m, uint32_t flags, hipStream_t s) {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_batch_set_systems first");
    HIPCHK(hipSetDevice(device_));
    bind(px, py, pz, fx, fy, fz, m);
    const int eflag = (flags & RBMDX_FLAG_ENERGY) ? 1 : 0;
    HIPCHK(batch_set_epoch(d_.epoch, (long long)step, s));
    if (flags & RBMDX_FLAG_PER_REPLICA) {
      for (int r = 0; r < R_; ++r)
        HIPCHK(batch_launch(d_, pc_, rbl(), cfg_.coulomb() ? 1 : 0, eflag, special_,
                            range(r, r + 1), tune_, num_cu_, s));
    } else if (flags & RBMDX_FLAG_EAGER) {
      launch_all(eflag, s);
    } else {
      HIPCHK(hipGraphLaunch(graph_for(eflag), s));
    }
    last_stream_ = s;
    last_stream_set_ = true;
    last_eflag_ = eflag;
    ++calls_;
    if (flags & RBMDX_FLAG_SYNC) HIPCHK(hipStreamSynchronize(s));
  }

  // Per replica: E_vdwl, E_coul, E_kspace (with the self term), virial xx yy
  // zz xy xz yz (pair + k-space) of the last evaluation made with
  // RBMDX_FLAG_ENERGY.
  void energies(double* out) {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_batch_set_systems first");
    HIPCHK(hipSetDevice(device_));
    if (last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
    if (!last_eflag_)
      throw Error(RBMDX_E_STATE, "the last evaluation was made without RBMDX_FLAG_ENERGY");
    std::vector<double> ra((size_t)R_ * kBatchAcc), ka((size_t)R_ * kBatchKAcc);
    HIPCHK(hipMemcpy(ra.data(), d_.racc, sizeof(double) * ra.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ka.data(), d_.kacc, sizeof(double) * ka.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < R_; ++r) {
      const double* a = &ra[(size_t)r * kBatchAcc];
      const double* k = &ka[(size_t)r * kBatchKAcc];
      double* o = out + 9 * (size_t)r;
      o[0] = a[3];
      o[1] = a[4];
      o[2] = P_ > 0 ? k[0] + self_[r] : 0.0;
      for (int q = 0; q < 6; ++q) o[3 + q] = a[5 + q] + (P_ > 0 ? k[1 + q] : 0.0);
    }
  }

  void set_tune(const BatchTune& t) {
    if (t.bw < 1 || t.bw > 2 || (t.pair_block != 64 && t.pair_block != 128 &&
                                 t.pair_block != 256) ||
        t.grid_per_cu < 1 || t.grid_per_cu > 16 ||
        (t.qlanes != 1 && t.qlanes != 2 && t.qlanes != 4 && t.qlanes != 8 &&
         t.qlanes != 16 && t.qlanes != 32 && t.qlanes != 64) ||
        t.qlanes > t.pair_block)
      throw Error(RBMDX_E_ARG, "launch shape out of range");
    tune_ = t;
    if (ready_ && t.bw != layout_bw_) {
      if (last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
      layout(t.bw);
    }
    drop_graphs();
  }
  BatchTune tune() const { return tune_; }

  // Measure candidate launch shapes on the caller's inputs (coordinate
  // descent: brick width x block, then the persistent grid, then the
  // Coulomb lanes; each candidate one warm-up and 5 timed eager evaluations
  // without energies) and keep the fastest; with a cache path, a stored choice for
  // this device and batch shape is used without measuring, and a new one is
  // stored.  Returns the chosen evaluation time in microseconds (0 from the
  // cache).
  double autotune(int64_t step, const float* px, const float* py, const float* pz,
                  float* fx, float* fy, float* fz, const int32_t* m, const char* cache,
                  hipStream_t s) {
    if (!ready_) throw Error(RBMDX_E_STATE, "rbmdx_batch_set_systems first");
    HIPCHK(hipSetDevice(device_));
    bind(px, py, pz, fx, fy, fz, m);
    const std::string key = cache_key();
    std::map<std::string, std::string> entries;
    if (cache && *cache) {
      entries = read_cache(cache);
      auto it = entries.find(key);
      if (it != entries.end()) {
        BatchTune t;
        if (parse_tune(it->second, t)) {
          set_tune(t);
          return 0.0;
        }
      }
    }
    if (!ev_[0]) {
      HIPCHK(hipEventCreate(&ev_[0]));
      HIPCHK(hipEventCreate(&ev_[1]));
    }
    HIPCHK(batch_set_epoch(d_.epoch, (long long)step, s));
    auto measure = [&](const BatchTune& t) {
      HIPCHK(hipStreamSynchronize(s));  // the layout upload below is synchronous
      set_tune(t);
      launch_all(0, s);
      HIPCHK(hipEventRecord(ev_[0], s));
      for (int k = 0; k < 5; ++k) launch_all(0, s);
      HIPCHK(hipEventRecord(ev_[1], s));
      HIPCHK(hipEventSynchronize(ev_[1]));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
      return 1000.0 * ms / 5.0;
    };
    BatchTune best = tune_;
    double best_us = measure(best);
    auto consider = [&](BatchTune t) {
      if (t.qlanes > t.pair_block) return;
      const double us = measure(t);
      if (us < best_us) {
        best_us = us;
        best = t;
      }
    };
    for (int bw : {1, 2})
      for (int blk : {64, 128, 256}) {
        BatchTune t = best;
        t.bw = bw;
        t.pair_block = blk;
        consider(t);
      }
    for (int g : {1, 2, 4, 8}) {
      BatchTune t = best;
      t.grid_per_cu = g;
      consider(t);
    }
    for (int ql : {4, 8, 16}) {
      BatchTune t = best;
      t.qlanes = ql;
      consider(t);
    }
    set_tune(best);
    last_stream_ = s;
    last_stream_set_ = true;
    if (cache && *cache) {
      char buf[160];
      std::snprintf(buf, sizeof(buf),
                    "{\"bw\": %d, \"pair_block\": %d, \"grid_per_cu\": %d, \"qlanes\": %d, "
                    "\"us\": %.2f}",
                    best.bw, best.pair_block, best.grid_per_cu, best.qlanes, best_us);
      entries[key] = buf;
      write_cache(cache, entries);
    }
    return best_us;
  }

  int num_samples() const { return P_; }
  int replicas() const { return R_; }

 private:
  int rbl() const { return cfg_.neighbor_type == "RBL" ? 1 : 0; }
  int special_ = 0;  // CVFF special-bond rows are set

  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)));
    allocs_.push_back(p);
    return (T*)p;
  }
  const int* ewald_m_ = nullptr;  // EWALD: the kmax grid for every replica

  template <typename T>
  const T* upload(const std::vector<T>& v) {
    T* p = alloc<T>(v.size());
    if (!v.empty())
      HIPCHK(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return p;
  }

  void drop_graphs() {
    // an exec may still run from the last call: wait for it before destroying
    if (!graphs_.empty() && last_stream_set_) HIPCHK(hipStreamSynchronize(last_stream_));
    for (auto& kv : graphs_) hipGraphExecDestroy(kv.second);
    graphs_.clear();
  }

  // The batch evaluation captured once per (pointers, energy flag, shape).
  hipGraphExec_t graph_for(int eflag) {
    std::ostringstream k;
    k << (const void*)d_.px << ' ' << (const void*)d_.py << ' ' << (const void*)d_.pz << ' '
      << (const void*)d_.fx << ' ' << (const void*)d_.fy << ' ' << (const void*)d_.fz << ' '
      << (const void*)d_.m << ' ' << eflag;
    auto it = graphs_.find(k.str());
    if (it != graphs_.end()) return it->second;
    if (graphs_.size() >= 8) drop_graphs();
    if (!cap_stream_) HIPCHK(hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking));
    HIPCHK(hipStreamBeginCapture(cap_stream_, hipStreamCaptureModeRelaxed));
    const hipError_t e1 = batch_launch(d_, pc_, rbl(), cfg_.coulomb() ? 1 : 0, eflag,
                                       special_, range(0, R_), tune_, num_cu_, cap_stream_);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cap_stream_, &g);
    if (e1 != hipSuccess || e2 != hipSuccess || !g) {
      if (g) hipGraphDestroy(g);
      (void)hipGetLastError();
      throw Error(RBMDX_E_HIP, std::string("graph capture failed: ") +
                                   hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t e3 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e3 != hipSuccess)
      throw Error(RBMDX_E_HIP, std::string("graph instantiation failed: ") +
                                   hipGetErrorString(e3));
    graphs_[k.str()] = x;
    return x;
  }

  std::string cache_key() const {
    std::ostringstream k;
    k << arch_ << "|cu" << num_cu_ << "|R" << R_ << "|N" << N_ << "|NQ" << NQ_ << "|P"
      << P_ << "|rc" << cfg_.cut_off << "|rl" << cfg_.lj_cut_off << "|" << cfg_.neighbor_type
      << "|t" << ntypes_ << "|k" << d_.kmode << "|s" << special_;
    return k.str();
  }

  static std::map<std::string, std::string> read_cache(const char* path) {
    std::map<std::string, std::string> out;
    std::ifstream f(path);
    if (!f) return out;
    std::stringstream ss;
    ss << f.rdbuf();
    try {
      const Json root = parse_json(ss.str());
      if (root.kind != Json::Object) return out;
      for (const auto& kv : root.obj) {
        const Json& v = kv.second;
        if (v.kind != Json::Object) continue;
        auto num = [&](const char* k) {
          const Json* j = v.get(k);
          return j && j->kind == Json::Number ? j->num : -1.0;
        };
        char buf[160];
        std::snprintf(buf, sizeof(buf),
                      "{\"bw\": %d, \"pair_block\": %d, \"grid_per_cu\": %d, "
                      "\"qlanes\": %d, \"us\": %.2f}",
                      (int)num("bw"), (int)num("pair_block"), (int)num("grid_per_cu"),
                      (int)num("qlanes"), num("us"));
        out[kv.first] = buf;
      }
    } catch (...) {
      out.clear();  // an unreadable cache is ignored (and rewritten)
    }
    return out;
  }

  static bool parse_tune(const std::string& text, BatchTune& t) {
    try {
      const Json v = parse_json(text);
      auto num = [&](const char* k) {
        const Json* j = v.get(k);
        return j && j->kind == Json::Number ? (int)j->num : -1;
      };
      t.bw = num("bw");
      t.pair_block = num("pair_block");
      t.grid_per_cu = num("grid_per_cu");
      t.qlanes = num("qlanes");
      return t.bw >= 1 && t.pair_block >= 64 && t.grid_per_cu >= 1 && t.qlanes >= 1;
    } catch (...) {
      return false;
    }
  }

  static void write_cache(const char* path, const std::map<std::string, std::string>& e) {
    const std::string tmp = std::string(path) + ".tmp";
    {
      std::ofstream f(tmp);
      if (!f) throw Error(RBMDX_E_ARG, std::string("cannot write the tuning cache ") + path);
      f << "{\n";
      size_t i = 0;
      for (const auto& kv : e) {
        f << "  \"" << kv.first << "\": " << kv.second << (++i < e.size() ? ",\n" : "\n");
      }
      f << "}\n";
    }
    if (std::rename(tmp.c_str(), path) != 0)
      throw Error(RBMDX_E_ARG, std::string("cannot write the tuning cache ") + path);
  }

  Config cfg_;
  int device_ = 0;
  int num_cu_ = 256;
  std::string arch_;
  bool ready_ = false;
  int R_ = 0, N_ = 0, NQ_ = 0, ncell_ = 0, P_ = 0, ntypes_ = 0;
  std::vector<RepInfo> rep_;
  std::vector<int> a0_, q0_, c0_, w0_;
  std::vector<double> self_;
  RepInfo* rep_dev_ = nullptr;
  BatchWork* work_dev_ = nullptr;
  int layout_bw_ = 0;
  PairConst<float> pc_{};
  BatchDev d_{};
  BatchTune tune_{};
  std::vector<void*> allocs_;
  std::map<std::string, hipGraphExec_t> graphs_;
  hipStream_t cap_stream_ = nullptr;
  hipStream_t last_stream_ = nullptr;
  bool last_stream_set_ = false;
  int last_eflag_ = 0;
  int64_t calls_ = 0;
  hipEvent_t ev_[2] = {nullptr, nullptr};
};

}  // namespace

struct rbmdx_batch {
  std::string err;
  std::unique_ptr<Batch> b;
};

namespace {
template <typename F>
int guarded_b(rbmdx_batch* c, F&& f) {
  try {
    f();
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

int rbmdx_batch_create(const char* json_text, int device, rbmdx_batch** out) {
  if (!out) return RBMDX_E_ARG;
  *out = nullptr;
  if (!json_text) return RBMDX_E_ARG;
  rbmdx_batch* c = new rbmdx_batch();
  const int rc = guarded_b(c, [&] { c->b.reset(new Batch(parse_config(json_text), device)); });
  *out = c;  // returned even on error so rbmdx_batch_last_error can explain
  return rc;
}

void rbmdx_batch_destroy(rbmdx_batch* ctx) { delete ctx; }

const char* rbmdx_batch_last_error(const rbmdx_batch* ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int rbmdx_batch_set_systems(rbmdx_batch* ctx, int32_t nrep, const int32_t* natoms,
                            const double* box_lo, const double* box_len,
                            const int32_t* type, const int32_t* id, const double* charge,
                            int32_t ntypes, const double* eps, const double* sigma) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded_b(ctx, [&] {
    if (!ctx->b) throw Error(RBMDX_E_STATE, "context was not created");
    if (!natoms || !box_lo || !box_len || !type || !id || !charge || !eps || !sigma)
      throw Error(RBMDX_E_ARG, "null argument");
    ctx->b->set_systems(nrep, natoms, box_lo, box_len, type, id, charge, ntypes, eps, sigma);
  });
}

int rbmdx_batch_set_special(rbmdx_batch* ctx, int32_t maxspecial, const int32_t* nspecial,
                            const int32_t* special) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded_b(ctx, [&] {
    if (!ctx->b) throw Error(RBMDX_E_STATE, "context was not created");
    ctx->b->set_special(maxspecial, nspecial, special);
  });
}

int rbmdx_batch_force(rbmdx_batch* ctx, int64_t step, const float* px, const float* py,
                      const float* pz, float* fx, float* fy, float* fz,
                      const int32_t* m_samples, uint32_t flags, void* stream) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded_b(ctx, [&] {
    if (!ctx->b) throw Error(RBMDX_E_STATE, "context was not created");
    ctx->b->force(step, px, py, pz, fx, fy, fz, m_samples, flags, (hipStream_t)stream);
  });
}

int rbmdx_batch_energies(rbmdx_batch* ctx, double* out) {
  if (!ctx || !out) return RBMDX_E_ARG;
  return guarded_b(ctx, [&] {
    if (!ctx->b) throw Error(RBMDX_E_STATE, "context was not created");
    ctx->b->energies(out);
  });
}

int rbmdx_batch_autotune(rbmdx_batch* ctx, int64_t step, const float* px, const float* py,
                         const float* pz, float* fx, float* fy, float* fz,
                         const int32_t* m_samples, const char* cache_path, void* stream,
                         double* us_out) {
  if (!ctx) return RBMDX_E_ARG;
  return guarded_b(ctx, [&] {
    if (!ctx->b) throw Error(RBMDX_E_STATE, "context was not created");
    const double us = ctx->b->autotune(step, px, py, pz, fx, fy, fz, m_samples, cache_path,
                                       (hipStream_t)stream);
    if (us_out) *us_out = us;
  });
}

int rbmdx_batch_set_shape(rbmdx_batch* ctx, const int32_t* shape) {
  if (!ctx || !shape) return RBMDX_E_ARG;
  return guarded_b(ctx, [&] {
    if (!ctx->b) throw Error(RBMDX_E_STATE, "context was not created");
    BatchTune t;
    t.bw = shape[0];
    t.pair_block = shape[1];
    t.grid_per_cu = shape[2];
    t.qlanes = shape[3];
    ctx->b->set_tune(t);
  });
}

int rbmdx_batch_get_shape(rbmdx_batch* ctx, int32_t* shape) {
  if (!ctx || !shape) return RBMDX_E_ARG;
  return guarded_b(ctx, [&] {
    if (!ctx->b) throw Error(RBMDX_E_STATE, "context was not created");
    const BatchTune t = ctx->b->tune();
    shape[0] = t.bw;
    shape[1] = t.pair_block;
    shape[2] = t.grid_per_cu;
    shape[3] = t.qlanes;
  });
}

int32_t rbmdx_batch_num_samples(const rbmdx_batch* ctx) {
  return ctx && ctx->b ? ctx->b->num_samples() : 0;
}

}  // extern "C"
