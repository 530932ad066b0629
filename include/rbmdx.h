/*
 * rbmdx.h - C ABI of the MI355X-native RBMD force path (librbmdx.so).
 *
 * One context owns the device state of one periodic system on one GPU:
 * cell-sorted copies of positions/charges/types, neighbour candidate lists,
 * k-space scratch and the RBE k-sample stream.  Caller arrays are never
 * reordered: inputs and outputs are in the caller's atom order, device
 * pointers, SoA, exactly like RBMD's DeviceData (_d_px/_d_py/_d_pz,
 * _d_fx/_d_fy/_d_fz).
 *
 * Every call that takes a `stream` is stream-ordered and performs no host
 * synchronisation unless RBMDX_FLAG_SYNC is given.  A context is not
 * thread-safe; use one per host thread.  No function calls exit(): errors are
 * returned as status codes and described by rbmdx_last_error().
 *
 * Reference interfaces replaced (SOG-AITech/rbmd source tree):
 *   rbmdx_create        Force plugin construction from JSON
 *                       (src/force/include/force.h:10-23, factory map at
 *                        src/simulate_pipeline/src/nve_ensemble.cpp:19-26,
 *                        config keys of docs/json-configuration.en.md:75-167)
 *   rbmdx_set_system    DeviceData upload (src/data_manager/src/memory_scheduler.cpp:16-220,
 *                        src/data_manager/src/lj_memory_scheduler.cpp:3-36) and
 *                       Force::Init (src/force/src/lj_cut_coul_kspace.cpp:424-462,
 *                        src/force/src/kspace_calculator.cpp:159-355)
 *   rbmdx_force         Force::Execute (src/force/src/lj_cut_coul_kspace.cpp:464-471,
 *                        src/force/src/lj.cpp:54-69) including the neighbour
 *                        builders it calls and the k-space calculator
 *   rbmdx_thermo        Force::EvaluatePotentialEnergy + virial read-back
 *                       (src/force/src/lj_cut_coul_kspace.cpp:867-897,
 *                        src/force/src/kspace_calculator.cpp:648-669)
 */
#ifndef RBMDX_H
#define RBMDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rbmdx_ctx rbmdx_ctx;

/* status codes */
enum {
  RBMDX_OK = 0,
  RBMDX_E_CONFIG = 1,    /* bad or unsupported JSON configuration          */
  RBMDX_E_CAPACITY = 2,  /* a list outgrew its capacity (auto-grown; the
                            forces of the call are still correct)          */
  RBMDX_E_HIP = 3,       /* HIP runtime / launch failure                   */
  RBMDX_E_NONFINITE = 4, /* non-finite force or energy detected            */
  RBMDX_E_ARG = 5,       /* invalid argument                               */
  RBMDX_E_STATE = 6      /* call out of order (e.g. force before system)   */
};

/* rbmdx_force flags */
enum {
  RBMDX_FLAG_ENERGY = 1,   /* also accumulate energies and the virial      */
  RBMDX_FLAG_SYNC = 2,     /* synchronise and validate before returning    */
  RBMDX_FLAG_REBUILD = 4,  /* force a neighbour rebuild this call          */
  RBMDX_FLAG_TIMING = 8,   /* record per-phase GPU timing events this call
                              (read with rbmdx_kernel_times; the events
                              themselves cost a few microseconds per call;
                              implies RBMDX_FLAG_EAGER) */
  RBMDX_FLAG_EAGER = 16,   /* launch the kernels one by one instead of
                              replaying the context's captured hipGraph
                              (also RBMDX_GRAPH=0 in the environment) */
  RBMDX_FLAG_VJP_NEUTRAL_Q = 32, /* rbmdx_force_vjp: let dL/dQ_t include the
                              atoms of zero charge (see there) */
  RBMDX_FLAG_PER_REPLICA = 64 /* rbmdx_batch_force: launch the stage kernels
                              one replica at a time (eager) instead of the
                              batch's captured graph */
};

/*
 * Force-matching adjoint (no counterpart in the reference, which has no
 * differentiation): the vector-Jacobian product of one evaluation's force F
 * with a cotangent G = dL/dF (device arrays gx/gy/gz, caller order, the
 * force's precision):
 *   dx/dy/dz  dL/dr, device arrays in caller order;
 *   dparam    device array of 3 ntypes doubles: dL/dQ_t (a common shift of
 *             the charges of the charged atoms of type t), dL/deps_t,
 *             dL/dsigma_t through the mixing rules.
 * The charged/neutral partition is fixed for a system (rbmdx_set_charges
 * keeps zeros zero), so by default the atoms of zero charge stay out of
 * dL/dQ_t.  With RBMDX_FLAG_VJP_NEUTRAL_Q the shift covers every atom of
 * type t: the Coulomb pairs of a charged and a neutral atom and the neutral
 * atoms' structure sums enter as well (several times the cost when most atoms
 * are neutral).
 * The pair set (step is the RBL sampling epoch, as for rbmdx_force) and the
 * k-sample are those of the last rbmdx_force call; the RBL mean-force fix is
 * differentiated too.  LJ + erfc Coulomb pairs with RBE, RBSOG or EWALD
 * k-space, CVFF special bonds (with RBMDX_FLAG_VJP_NEUTRAL_Q the special
 * weight of a charged-neutral pair also scales its full-Coulomb removal) and
 * the SOG near field included; at most 8 types.  Eager launches on stream; parameter sums use LDS float atomics
 * inside a workgroup and FP64 atomics across workgroups, so their last bits
 * may vary between runs.  Flags:
 * RBMDX_FLAG_SYNC, RBMDX_FLAG_REBUILD, RBMDX_FLAG_VJP_NEUTRAL_Q.
 */
int rbmdx_force_vjp(rbmdx_ctx* ctx, int64_t step, const void* px,
                    const void* py, const void* pz, const void* gx,
                    const void* gy, const void* gz, void* dx, void* dy,
                    void* dz, double* dparam, uint32_t flags, void* stream);

/* Thermo record filled by rbmdx_thermo (all FP64, extensive totals). */
typedef struct rbmdx_thermo_t {
  double e_vdwl;      /* LJ pair energy (full list within cut_off)          */
  double e_coul;      /* real-space erfc Coulomb energy                     */
  double e_kspace;    /* k-space energy incl. self/background terms         */
  double virial[6];   /* pair + k-space virial: xx yy zz xy xz yz           */
  double virial_kspace[6];
  double mean_force[3]; /* RBL mean pair force that was removed             */
  int64_t steps;      /* force evaluations performed                        */
  int64_t rebuilds;   /* neighbour rebuilds performed                       */
  int32_t rbl_frequency; /* RBL shell selection frequency (pice_num)        */
  int32_t num_k;      /* k-vectors used by the last evaluation              */
  double alpha;       /* Ewald splitting parameter (RBMD "alpha" = g^2)    */
  double rbe_S;       /* RBE normalisation S = sum_{k!=0} exp(-k^2/4alpha)  */
  int64_t pairs_a;    /* stored all-atom candidate entries (last rebuild)  */
  int64_t pairs_b;    /* stored charged-pair candidate entries              */
  int64_t atoms;      /* atoms in the system                                */
  int64_t charged;    /* charged atoms handled by the k-space kernels       */
  double e_kspace_exact; /* with coulomb.energy_rbe_flag "yes" (RBE): the
                            exact Ewald k-space energy over the kmax grid,
                            incl. the self term, of the last evaluation made
                            with RBMDX_FLAG_ENERGY; else 0.  RBMD computes it
                            and then overwrites it with the RBE estimate
                            (kspace_calculator.cpp:628-669)               */
} rbmdx_thermo_t;

/* State of the molecular-dynamics step (rbmdx_md_*). */
typedef struct rbmdx_md_state_t {
  int64_t step;          /* steps completed since rbmdx_md_init              */
  int32_t mode;          /* 0 NVE, 1 NVE + energy-stable rescale, 2 NVT
                            Nose-Hoover chain                                */
  int32_t pad;
  double temperature;    /* as the thermostat last sampled it: NVT after the
                            chain's scaling (the current velocities), NVE
                            energy-stable before the rescale by xi          */
  double ke;             /* kinetic energy per atom, sampled with temperature */
  double pe;             /* potential energy per atom of the last evaluation
                            (0 unless energies were computed)                */
  double t_target;       /* thermostat target of the last step               */
  double xi;             /* last energy-stable factor (1 when not used)      */
  double eta[3], eta_dot[3], eta_mass[3];  /* thermostat chain              */
  double thermostat_energy;  /* chain energy (total): sum Q eta_dot^2 / 2 +
                                tdof kT eta_0 + kT sum_{i>0} eta_i           */
} rbmdx_md_state_t;

/*
 * Create a context from an RBMD JSON configuration (the whole document, or
 * just its "hyper_parameters" object).  Recognised keys:
 *   init_configuration.read_data.unit   "LJ" | "REAL" | "METAL"
 *   hyper_parameters.force_field.type   "LJ/CUT" | "LJ/CUT/COUL/LONG" | "CVFF"
 *       (CVFF: the non-bonded part with special-bond weights; topology from
 *        rbmdx_set_special)
 *   hyper_parameters.extend.special_bonds  [w12, w13, w14] (CVFF; default 1)
 *   hyper_parameters.neighbor.{type,cut_off,skin,r_core,neighbor_sample_num,
 *                              energy_rbl_flag}
 *       type: "RBL" | "Verlet-List" | "VERLET-SOG"
 *   hyper_parameters.neighbor.lj_cut_off   (extension: separate LJ cutoff,
 *                                           default cut_off)
 *   hyper_parameters.coulomb.{type,alpha,accuracy,coulomb_sample_num,kmax,
 *                             energy_rbe_flag,random_seed,rbsog_level}
 *       type: "RBE" | "EWALD" | "RBSOG"; rbsog_level 0-4 selects the
 *       reference's SOG presets (RBSOG, VERLET-SOG)
 *   hyper_parameters.coulomb.rbsog_exact   (extension) SOG coefficients and
 *                                           normalisation without the
 *                                           reference's float defects
 *   hyper_parameters.precision             (extension) "fp32" | "fp64"
 * `device` is the HIP device ordinal.
 */
int rbmdx_create(const char* json_text, int device, rbmdx_ctx** out);
void rbmdx_destroy(rbmdx_ctx* ctx);
const char* rbmdx_last_error(const rbmdx_ctx* ctx);

/*
 * Describe the system (host arrays, copied).  Box is orthogonal and periodic:
 * [lo_d, lo_d + L_d).  type is 0-based (< ntypes); id is the 0-based global
 * atom id used by the RBL shell hash; eps/sigma are per type (mixing:
 * eps_ij = sqrt(eps_i eps_j), sigma_ij = (sigma_i + sigma_j)/2).
 */
int rbmdx_set_system(rbmdx_ctx* ctx, int64_t n, const double* box_lo,
                     const double* box_len, const int32_t* type,
                     const int32_t* id, const double* charge, int32_t ntypes,
                     const double* eps, const double* sigma);

/* Replace the charges (host array, n entries); keeps lists. */
int rbmdx_set_charges(rbmdx_ctx* ctx, const double* charge);

/* Replace the per-type LJ parameters (host arrays, ntypes entries); keeps
 * lists.  Synchronises the device. */
int rbmdx_set_lj(rbmdx_ctx* ctx, const double* eps, const double* sigma);

/*
 * Special-bond topology for force_field.type "CVFF" (host arrays, caller
 * order, copied): nspecial[3 i + 0..2] are the cumulative counts
 * n12 <= n13 <= n14 of atom i's 1-2, 1-3 and 1-4 partners, and
 * special[i * maxspecial + k] their global ids (the same ids as
 * rbmdx_set_system's id).  A pair (i, j) with j among i's first n12 / n13 /
 * n14 partners takes the weight w12 / w13 / w14 of
 * hyper_parameters.extend.special_bonds: its LJ force is scaled by it, and
 * (1 - w) of the full Coulomb force -qqr2e q_i q_j / r^3 is removed from every
 * pair the lists keep.  maxspecial 0 (or all counts 0) clears the topology.
 * Replaces the d_nspecial / d_special / maxspecial fields of DeviceData
 * (src/data_manager/include/model/device_data.h:270-277) read by
 * LookupSpecialWeight (src/force/src/cvff_op/rocm/cvff_op.hip.cu:51-110).
 * maxspecial <= 64; call after rbmdx_set_system; synchronises the device.
 */
int rbmdx_set_special(rbmdx_ctx* ctx, int32_t maxspecial,
                      const int32_t* nspecial, const int32_t* special);

/*
 * One force evaluation.  px/py/pz are device arrays of float (precision
 * fp32) or double (precision fp64); fx/fy/fz are device arrays of the same
 * type and receive the total force (pair + k-space).  `step` is the RBL
 * sampling epoch (RBMD's test_current_step).  `m_samples` is NULL (draw the
 * next RBE k-sample from the seeded stream) or a device array of num_k*3
 * int32 integer k-vectors m (k = 2 pi m / L) used for this call only.
 * `stream` is a hipStream_t (NULL = default stream).
 */
int rbmdx_force(rbmdx_ctx* ctx, int64_t step, const void* px, const void* py,
                const void* pz, void* fx, void* fy, void* fz,
                const int32_t* m_samples, int32_t num_k, uint32_t flags,
                void* stream);

/*
 * rbmdx_force without any device allocation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory during the call (tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as rbmdx_force; the kernels are
 * the same and are launched one by one on `stream` (no graph instantiation).
 * The per-evaluation scratch (per-block reduction partials, S(k) chunk
 * partials, rho and the per-k coefficients) lives in a static per-device
 * array of 4 MiB (held by every device the library runs on until the
 * process ends) instead of the context's buffers; the neighbour lists and
 * sorted arrays are the context's, allocated by rbmdx_set_system and the
 * first rbmdx_force call, and are never grown here (an overflowed list is
 * still evaluated exactly; the next rbmdx_force grows it).
 * Errors: RBMDX_E_CAPACITY when the system's scratch exceeds 4 MiB (the
 * 200,000-atom electrolyte at P = 200 needs about 0.23 MiB, 2,000,000 atoms
 * at P = 400 about 2.9 MiB), RBMDX_E_STATE before the context's first
 * rbmdx_force call, or with m_samples NULL before the seeded k-sample stream
 * has started on an rbmdx_force call; RBMDX_FLAG_TIMING is rejected.
 * Every call on one device shares that device's array and leaves its last
 * contents there: calls that overlap in time on one device (other streams,
 * other contexts, other host threads) race and must be serialised by the
 * caller.
 */
int rbmdx_force_noalloc(rbmdx_ctx* ctx, int64_t step, const void* px,
                        const void* py, const void* pz, void* fx, void* fy,
                        void* fz, const int32_t* m_samples, int32_t num_k,
                        uint32_t flags, void* stream);

/*
 * Per-type masses (ntypes entries, default 1) for rbmdx_force_leapfrog.
 * Replaces the mass table RBMD's integrator reads (DeviceData _d_mass).
 */
int rbmdx_set_masses(rbmdx_ctx* ctx, const double* mass);

/*
 * One force evaluation fused with the leapfrog update of the caller's arrays,
 * in the same kernel that assembles the force:
 *   F = force(x);  v += dt * ftm2v * F / m;  x += dt * v
 * (positions and velocities updated in place; forces written to fx..fz).
 * This is velocity Verlet with the second half-kick of one step and the first
 * half-kick of the next merged (RBMD's UpdateVelocity / position update,
 * src/simulate_pipeline/src/controller/velocity_controller/
 * velocity_controller_op/rocm/update_velocity_op.hip.cu:5-27); ftm2v follows
 * the unit style (src/common/unit_factor.cpp:30-50).  The k-sample comes from
 * the seeded stream.
 */
int rbmdx_force_leapfrog(rbmdx_ctx* ctx, int64_t step, void* px, void* py,
                         void* pz, void* vx, void* vy, void* vz, void* fx,
                         void* fy, void* fz, double dt, uint32_t flags,
                         void* stream);

/* Synchronise the context's last stream and read energies/virial. */
int rbmdx_thermo(rbmdx_ctx* ctx, rbmdx_thermo_t* out);

/*
 * Molecular dynamics on the context's system: RBMD's step either side of the
 * force, velocity Verlet with periodic wrapping and image flags
 * (src/simulate_pipeline/src/controller/velocity_controller/
 * velocity_controller_op/rocm/update_velocity_op.hip.cu:5-27,
 * src/simulate_pipeline/src/controller/position_controller/
 * position_controller_op/rocm/update_position_op.hip.cu:17-37,
 * src/data_manager/include/model/box.h:133-173), with
 *   - execution.ensemble "NVT", temp_ctrl_type "NOSE_HOOVER": the Nose-Hoover
 *     chain thermostat (3 links) with temperature [start, stop, damp]
 *     (src/simulate_pipeline/src/controller/pressure_controller/
 *     nose_hoover_controller.cpp:89-250,415-470,633-840);
 *   - execution.ensemble "NVE" with RBL lists: the energy-stable velocity
 *     rescale (src/simulate_pipeline/src/controller/energy_stable_controller/
 *     energy_stable_scheme_controller.cpp:56-144), against the Hamiltonian of
 *     step 1; execution.energy_stable "no" switches it off (extension).
 * Settings come from the context's JSON (execution.timestep, ensemble,
 * temp_ctrl_type, temperature, num_steps; init_configuration.read_data.unit)
 * and the masses from rbmdx_set_masses.  Arrays are the caller's, in caller
 * order and the context's precision; image flags (int32, may be null) are
 * updated with every wrap.  The thermostat state stays on the device: a run
 * of many steps does not synchronise.
 *
 * rbmdx_md_init: wrap the positions into the box, evaluate the force of
 * step 0 into fx..fz and set the thermostat up (NoseHooverController::SetUp).
 * rbmdx_md_run: nsteps steps; positions, velocities, image flags and forces
 * are updated in place, the force of the last step is left in fx..fz.  flags:
 * RBMDX_FLAG_ENERGY computes energies every step (always on with the
 * energy-stable rescale); other force flags pass through.
 * rbmdx_md_state: synchronise and read the thermostat state.
 */
int rbmdx_md_init(rbmdx_ctx* ctx, void* px, void* py, void* pz, void* vx,
                  void* vy, void* vz, int32_t* ix, int32_t* iy, int32_t* iz,
                  void* fx, void* fy, void* fz, uint32_t flags, void* stream);
int rbmdx_md_run(rbmdx_ctx* ctx, int32_t nsteps, void* px, void* py, void* pz,
                 void* vx, void* vy, void* vz, int32_t* ix, int32_t* iy,
                 int32_t* iz, void* fx, void* fy, void* fz, uint32_t flags,
                 void* stream);
int rbmdx_md_state(rbmdx_ctx* ctx, rbmdx_md_state_t* out);

/*
 * Host copy of the k-sample the seeded stream yields for force evaluation
 * `call_index` (0 = first rbmdx_force call).  RBE: m_out holds P*3 int32.
 * RBSOG: m_out holds P*4 int32, the P m-vectors followed by the P indices of
 * the virial chain (the same layout rbmdx_force accepts as m_samples with
 * num_k = P).
 */
int rbmdx_rbe_sample(rbmdx_ctx* ctx, int64_t call_index, int32_t* m_out);

/* Number of sampled k-vectors P per evaluation (RBE, RBSOG; 0 otherwise).
 * For RBSOG this is coulomb_sample_num rounded up to a multiple of 16. */
int32_t rbmdx_num_samples(const rbmdx_ctx* ctx);

/*
 * Parity export: rebuild the lists for the given positions (device arrays)
 * and write, for every directed pair the RBL classification keeps at `step`,
 * (gid_i, gid_j, cls) with cls = 1 core, 2 selected shell (Verlet mode: 1
 * for every pair inside cut_off).  Host arrays of capacity `cap`; the return
 * value is the pair count (negative status on error; if the count exceeds
 * cap only cap pairs are written).
 */
int64_t rbmdx_export_pairs(rbmdx_ctx* ctx, int64_t step, const void* px,
                           const void* py, const void* pz, int32_t* gi,
                           int32_t* gj, int32_t* cls, int64_t cap);

/*
 * Parity export of the neighbour candidate lists (every directed pair with
 * d^2 < (cut + skin)^2 as stored on the device) for the positions of the
 * last rebuild.  Same conventions as rbmdx_export_pairs; cls = 1 for pairs
 * of the all-atom list, 2 for pairs of the charged-pair list.
 */
int64_t rbmdx_export_candidates(rbmdx_ctx* ctx, int32_t* gi, int32_t* gj,
                                int32_t* cls, int64_t cap);

/*
 * Host-only helpers (no device needed).
 *
 * rbmdx_rbe_sample_host: the k-sample the seeded stream yields for force
 * evaluation `call_index` of a system with Ewald parameter `alpha` (RBMD
 * "alpha" = g_ewald^2) and box lengths L[3]; replaces
 * RbmdParallelUntil::GetRbeRadomM (src/parallel/src/rbmd_parallel_until.cpp:557-724)
 * together with its seed chain (:305-325, :484-501).  m_out holds P*3 int32.
 *
 * rbmdx_check_config: parse and validate a JSON configuration the way
 * rbmdx_create does (the checks of LJCutCoulKspace::Init and
 * KSpaceCalculator::Init, src/force/src/lj_cut_coul_kspace.cpp:424-462 and
 * src/force/src/kspace_calculator.cpp:159-355); writes the message of a
 * failure into err (errlen bytes).  Returns a status code.
 */
int rbmdx_rbe_sample_host(double alpha, const double* L, int32_t P,
                          uint32_t base_seed, int64_t call_index,
                          int32_t* m_out);
int rbmdx_check_config(const char* json_text, char* err, int32_t errlen);

/*
 * rbmdx_sog_setup_host: the RBSOG far-field set-up for coulomb.rbsog_level
 * `level`, neighbor.cut_off and box lengths L[3] (replaces
 * KSpaceCalculator::RBSOGInit / RBSOGSetup, src/force/src/kspace_calculator.cpp:
 * 988-1140): scal = {b, sigma, w0, S, S_npt, mmax}, coef / coef_npt receive
 * mmax values (at most 128).  exact = 0 reproduces the reference's float
 * set-up; 1 the corrected one (extension key coulomb.rbsog_exact).
 *
 * rbmdx_sog_sample_host: the RBSOG k-sample of force evaluation call_index
 * (RBSOGSampleKSpace, kspace_calculator.cpp:1142-1329, seeded from the same
 * per-evaluation seed stream as RBE): m_out holds P*3 int32, idx_out the P
 * indices of the virial chain.
 */
int rbmdx_sog_setup_host(int32_t level, double cut_off, const double* L,
                         int32_t exact, double* scal, double* coef,
                         double* coef_npt);
int rbmdx_sog_sample_host(int32_t level, double cut_off, const double* L,
                          int32_t exact, int32_t P, uint32_t base_seed,
                          int64_t call_index, int32_t* m_out, int32_t* idx_out);

/* Per-phase GPU times, averaged
 * over the calls timed since the previous rbmdx_kernel_times: calls made with
 * RBMDX_FLAG_TIMING, or every call when the context was created with
 * RBMDX_TIMING=1 in the environment.  Fills up to n (name, microseconds)
 * entries and returns their number (synchronises). */
int rbmdx_kernel_times(rbmdx_ctx* ctx, const char** names, double* usec,
                       int n);
/* The phases: "rebuild" (skin check, the conditional rebuild kernels and the
 * gather), "pair" (the pair kernel alone), "kspace" (S(k), its reduction and
 * the k-space force, on the second stream when that runs beside the pair
 * kernel), "assemble" (the pair-sum reduction, the wait for k-space and the
 * assembly). */

/*
 * Replica batch (no counterpart in the reference, which runs one system per
 * process): many independent periodic boxes of one force field evaluated
 * together, FP32 positions and forces, FP64 energies and virials per
 * replica.  The JSON surface is rbmdx_create's; the batch takes LJ/CUT or
 * LJ/CUT/COUL/LONG, RBL or Verlet-List, RBE with a given coulomb.alpha,
 * precision fp32, at most 8 types and 512 k-samples.  Pairs are found from
 * cells rebuilt at every evaluation (neighbor.skin is not used); boxes with
 * fewer than three cells per dimension are handled exactly.  RBL selection
 * and its mean-force removal are per replica (frequency from the replica's
 * size), keyed by the replica's own atom ids and the step.
 */
typedef struct rbmdx_batch rbmdx_batch;

int rbmdx_batch_create(const char* json_text, int device, rbmdx_batch** out);
void rbmdx_batch_destroy(rbmdx_batch* ctx);
const char* rbmdx_batch_last_error(const rbmdx_batch* ctx);

/* nrep replicas of natoms[r] atoms each; per-atom arrays are the replicas'
 * atoms concatenated in replica order (that order is also the order of
 * positions and forces); box_lo/box_len hold 3 doubles per replica; eps and
 * sigma are per type and shared.  Once per context. */
int rbmdx_batch_set_systems(rbmdx_batch* ctx, int32_t nrep, const int32_t* natoms,
                            const double* box_lo, const double* box_len,
                            const int32_t* type, const int32_t* id,
                            const double* charge, int32_t ntypes,
                            const double* eps, const double* sigma);

/* CVFF special bonds of the batch (the counterpart of rbmdx_set_special,
 * which replaces SpecialBondsManager's topology for one system,
 * cvff.cpp:1091-1191): per atom of the concatenated replicas (caller order)
 * the cumulative counts n12 <= n13 <= n14 (3 ints, clamped to maxspecial)
 * and maxspecial partner ids, which are ids within the atom's own replica.
 * force_field.type "CVFF" only; maxspecial in [0, 64]. */
int rbmdx_batch_set_special(rbmdx_batch* ctx, int32_t maxspecial,
                            const int32_t* nspecial, const int32_t* special);

/* One evaluation of every replica.  Device arrays; m_samples holds
 * rbmdx_batch_num_samples() integer k-vectors (3 ints) per replica, replica
 * after replica (RBE), or for RBSOG per replica the P k-vectors followed by
 * their P virial-chain source indices (as rbmdx_sog_sample_host returns them;
 * the batch adds the direct vectors |m|^2 <= 1 itself); null without Coulomb
 * and with EWALD, whose kmax grid weighted exp(-k^2 / 4 alpha) the batch holds
 * itself.  step is the RBL
 * sampling epoch.  Flags: RBMDX_FLAG_ENERGY, RBMDX_FLAG_SYNC,
 * RBMDX_FLAG_EAGER (the batch's kernels launched directly),
 * RBMDX_FLAG_PER_REPLICA (the same kernels, one replica per launch).  By
 * default the evaluation is one replay of a graph captured on first use for
 * these pointers. */
int rbmdx_batch_force(rbmdx_batch* ctx, int64_t step, const float* px,
                      const float* py, const float* pz, float* fx, float* fy,
                      float* fz, const int32_t* m_samples, uint32_t flags,
                      void* stream);

/* 9 doubles per replica: E_vdwl, E_coul, E_kspace (with the self term),
 * virial xx yy zz xy xz yz (pair + k-space), of the last evaluation, which
 * must have been made with RBMDX_FLAG_ENERGY.  Synchronises. */
int rbmdx_batch_energies(rbmdx_batch* ctx, double* out);

/* Launch shape (brick width in cells, threads per pair workgroup, persistent
 * pair workgroups per CU, lanes per charged atom): measured on the caller's
 * inputs by coordinate descent over {1, 2} x {64, 128, 256}, {1, 2, 4, 8},
 * {4, 8, 16} and kept; with cache_path, a choice stored for this device and
 * batch shape is reused without measuring and a new one is written there
 * (JSON).  us_out (may be null) receives the chosen evaluation time in
 * microseconds, 0 when it came from the cache.  The forces written are those
 * of step. */
int rbmdx_batch_autotune(rbmdx_batch* ctx, int64_t step, const float* px,
                         const float* py, const float* pz, float* fx, float* fy,
                         float* fz, const int32_t* m_samples,
                         const char* cache_path, void* stream, double* us_out);
int rbmdx_batch_set_shape(rbmdx_batch* ctx, const int32_t* shape /* [4] */);
int rbmdx_batch_get_shape(rbmdx_batch* ctx, int32_t* shape /* [4] */);
int32_t rbmdx_batch_num_samples(const rbmdx_batch* ctx);

/*
 * LAMMPS data files (replaces the reader path of RBMD's input,
 * src/reader/src/format/lammps/lammps_full_reader.cpp: ReadData :454-633,
 * ParseAtom :1496-1563, ApplyReplicate :899-1000, and the special-bond
 * lists of src/reader/src/mpi/special_bond_builder.cpp).  atom_style is
 * "atomic", "charge" or "full" (init_configuration.read_data.atom_style);
 * replicate (may be null) holds read_data.replicate [nx, ny, nz] (systems
 * without bonds).  Ids, types and molecules become 0-based; atoms keep the
 * file's order.  The special-bond topology follows LAMMPS' rules as RBMD
 * builds it: 1-2 partners from the bonds, 1-3 through a common partner and
 * held by an angle or dihedral, 1-4 held by a dihedral, each partner once.
 * Errors (RBMDX_E_CONFIG) name the file and line.  Host-only; no device.
 */
typedef struct rbmdx_data rbmdx_data;
int rbmdx_data_read(const char* path, const char* atom_style,
                    const int32_t* replicate, rbmdx_data** out);
void rbmdx_data_destroy(rbmdx_data* data);
/*
 * Replicate the data read so far nx x ny x nz times (RBMD's ApplyReplicate,
 * src/reader/src/format/lammps/lammps_full_reader.cpp:899-1050) and rebuild
 * the special-bond lists.  bond_periodic = 1 (RBMD's default, read_data
 * "bond_periodic"): topology partners are taken from the tile of their nearest
 * seed image; 0: positions are unwrapped by their image flags (required on
 * every atom of molecular data) and partners stay in their tile.
 * rbmdx_data_read's replicate argument is this call with bond_periodic = 1.
 */
int rbmdx_data_replicate(rbmdx_data* data, const int32_t* factors /* [3] */,
                         int32_t bond_periodic);
/*
 * GAUSS velocities (init_configuration.read_data.velocity_type "GAUSS",
 * RBMD's LammpsFullReader::ApplyGaussianVelocitiesIfRequested,
 * src/reader/src/format/lammps/lammps_full_reader.cpp:710-866): per atom
 * std::mt19937_64 seeded by MixVelocitySeed(seed, id + 1), three
 * std::normal_distribution draws scaled by 1/sqrt(m), centre-of-mass velocity
 * removed, rescaled to `temperature` with 3N - 3 degrees of freedom in the
 * unit style `unit` ("LJ", "REAL", "METAL").  Masses from the Masses section;
 * nothing changes when the file had a Velocities section.  Read the result
 * with rbmdx_data_get.
 */
int rbmdx_data_gauss_velocities(rbmdx_data* data, double temperature, int32_t seed,
                                 const char* unit);
const char* rbmdx_data_last_error(const rbmdx_data* data);
/* flags: 1 Pair Coeffs present, 2 Velocities present, 4 bonds present */
int rbmdx_data_info(const rbmdx_data* data, int64_t* natoms, int32_t* ntypes,
                    int32_t* maxspecial, int32_t* flags);
/* Any output may be null.  pos, vel, image: 3 per atom (x y z); mass, eps,
 * sigma: per type (0 where the file has none). */
int rbmdx_data_get(const rbmdx_data* data, double* box_lo, double* box_len,
                   int32_t* type, int32_t* id, double* charge, double* pos,
                   double* vel, int32_t* image, double* mass, double* eps,
                   double* sigma);
/* nspecial: 3 per atom; special: maxspecial per atom (rbmdx_set_special's
 * layout, unused entries -1). */
int rbmdx_data_special(const rbmdx_data* data, int32_t* nspecial, int32_t* special);

/*
 * Output formats (host arrays, caller order; ids and types 0-based, written
 * 1-based; pos / vel / force are N x 3 row-major; image (N x 3) may be null).
 * rbmdx_write_dump: one LAMMPS-style trajectory frame as RBMD's trajectory
 * output writes it (src/output/src/TrajectoryOutput.cpp:433-500): header
 * with TIMESTEP, NUMBER OF ATOMS, BOX BOUNDS pp pp pp, then
 * "id type q x y z xu yu zu vx vy vz" in id order (xu = x + image L).
 * rbmdx_write_force_samples: rows "sample,id,x,y,z,fx,fy,fz" in id order, the
 * header written when append is 0 -- the CSV RBMD's force-accuracy analysis
 * reads (benchmark/common/analyze_force_samples.py:30-45).
 * Numbers are the shortest decimals that read back exactly.  append = 0
 * truncates the file.  Returns RBMDX_E_ARG on bad arguments or I/O failure.
 */
int rbmdx_write_dump(const char* path, int32_t append, int64_t timestep, int64_t n,
                     const double* box_lo, const double* box_len, const int32_t* id,
                     const int32_t* type, const double* charge, const double* pos,
                     const int32_t* image, const double* vel);
int rbmdx_write_force_samples(const char* path, int32_t append, int64_t sample,
                              int64_t n, const int32_t* id, const double* pos,
                              const double* force);

#ifdef __cplusplus
}
#endif

#endif /* RBMDX_H */
