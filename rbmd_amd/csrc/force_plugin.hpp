// C++ force-module layer over the C ABI (include/rbmdx.h), shaped like RBMD's
// Force plugin interface so that a simulate pipeline can swap it in:
//
//   RBMD                                   here
//   class Force { Init(); Execute();       class Force { Init(); Execute();
//     EvaluatePotentialEnergy(); }           EvaluatePotentialEnergy(); }
//   (src/force/include/force.h:10-23)
//   name -> factory map keyed by           MakeForce(json, data) keyed by the
//   hyper_parameters.force_field.type      same string ("LJ/CUT",
//   (src/simulate_pipeline/src/            "LJ/CUT/COUL/LONG")
//    nve_ensemble.cpp:19-26)
//   DeviceData SoA (_d_px.._d_fz)          DeviceData: device pointers, caller
//   (src/data_manager/include/model/        order, never reordered
//    device_data.h:10-238)
//
// Differences, deliberate: state is passed explicitly instead of through the
// DataManager singleton; errors throw std::runtime_error carrying the ABI's
// message instead of exit(EXIT_FAILURE); the RBL sampling epoch is the
// caller's step counter (RBMD reads the global test_current_step).
#pragma once
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rbmdx.h"

namespace rbmdx {
namespace plugin {

// Per-atom device state the force path reads and writes (SoA, caller order).
// Positions and forces are float for precision "fp32", double for "fp64".
struct DeviceData {
  int64_t num_atoms = 0;
  void* d_px = nullptr;
  void* d_py = nullptr;
  void* d_pz = nullptr;
  void* d_fx = nullptr;
  void* d_fy = nullptr;
  void* d_fz = nullptr;
  // set-up data (host): box, per-atom type / id / charge, per-type LJ
  double box_lo[3] = {0, 0, 0};
  double box_len[3] = {0, 0, 0};
  std::vector<int32_t> type, id;
  std::vector<double> charge, eps, sigma;
  // CVFF special-bond topology (RBMD's d_nspecial / d_special / maxspecial):
  // cumulative 1-2/1-3/1-4 counts, 3 per atom, and maxspecial partner ids per
  // atom; maxspecial 0 when there is none
  int32_t maxspecial = 0;
  std::vector<int32_t> nspecial, special;
  void* stream = nullptr;  // hipStream_t
};

struct PotentialEnergy {
  double vdwl = 0, coul = 0, kspace = 0;  // extensive totals
  double virial[6] = {0, 0, 0, 0, 0, 0};
};

class Force {
 public:
  Force(std::string json, DeviceData* data, int device = 0)
      : json_(std::move(json)), data_(data), device_(device) {}
  virtual ~Force() {
    if (ctx_) rbmdx_destroy(ctx_);
  }
  Force(const Force&) = delete;
  Force& operator=(const Force&) = delete;

  // Parse the configuration and upload the system (Force::Init).
  virtual void Init() {
    if (ctx_) rbmdx_destroy(ctx_);
    ctx_ = nullptr;
    check(rbmdx_create(json_.c_str(), device_, &ctx_));
    const DeviceData& d = *data_;
    check(rbmdx_set_system(ctx_, d.num_atoms, d.box_lo, d.box_len, d.type.data(),
                           d.id.data(), d.charge.data(),
                           (int32_t)d.eps.size(), d.eps.data(), d.sigma.data()));
    if (d.maxspecial > 0)
      check(rbmdx_set_special(ctx_, d.maxspecial, d.nspecial.data(), d.special.data()));
  }

  // One force evaluation into d_fx/d_fy/d_fz (Force::Execute).
  virtual void Execute() {
    const DeviceData& d = *data_;
    const uint32_t flags = want_energy_ ? RBMDX_FLAG_ENERGY : 0u;
    check(rbmdx_force(ctx_, step_, d.d_px, d.d_py, d.d_pz, d.d_fx, d.d_fy,
                      d.d_fz, nullptr, 0, flags, d.stream));
    ++step_;
  }

  // Read energies and virial of the last evaluation (synchronises).
  virtual void EvaluatePotentialEnergy() {
    rbmdx_thermo_t t;
    check(rbmdx_thermo(ctx_, &t));
    pe_.vdwl = t.e_vdwl;
    pe_.coul = t.e_coul;
    pe_.kspace = t.e_kspace;
    for (int k = 0; k < 6; ++k) pe_.virial[k] = t.virial[k];
  }

  void set_step(int64_t s) { step_ = s; }
  int64_t step() const { return step_; }
  void set_energy_every_call(bool on) { want_energy_ = on; }
  const PotentialEnergy& potential_energy() const { return pe_; }
  rbmdx_ctx* context() { return ctx_; }

 protected:
  void check(int rc) const {
    if (rc != RBMDX_OK)
      throw std::runtime_error(std::string("rbmdx: ") +
                               (ctx_ ? rbmdx_last_error(ctx_) : "create failed"));
  }
  std::string json_;
  DeviceData* data_;
  int device_;
  rbmdx_ctx* ctx_ = nullptr;
  int64_t step_ = 0;
  bool want_energy_ = false;
  PotentialEnergy pe_;
};

// "LJ/CUT/COUL/LONG": RBL or Verlet short range + RBE / EWALD k-space.
class LJCutCoulKspace : public Force {
 public:
  using Force::Force;
};

// "LJ/CUT": Lennard-Jones with a full neighbour list.
class LJ : public Force {
 public:
  using Force::Force;
};

// "CVFF": the non-bonded part (LJ + Coulomb with special-bond weights from
// extend.special_bonds and the topology in DeviceData).
class CVFF : public Force {
 public:
  using Force::Force;
};

using ForceFactory =
    std::function<std::unique_ptr<Force>(std::string, DeviceData*, int)>;

inline const std::map<std::string, ForceFactory>& ForceRegistry() {
  static const std::map<std::string, ForceFactory> reg = {
      {"LJ/CUT/COUL/LONG",
       [](std::string j, DeviceData* d, int dev) -> std::unique_ptr<Force> {
         return std::unique_ptr<Force>(new LJCutCoulKspace(std::move(j), d, dev));
       }},
      {"LJ/CUT",
       [](std::string j, DeviceData* d, int dev) -> std::unique_ptr<Force> {
         return std::unique_ptr<Force>(new LJ(std::move(j), d, dev));
       }},
      {"CVFF",
       [](std::string j, DeviceData* d, int dev) -> std::unique_ptr<Force> {
         return std::unique_ptr<Force>(new CVFF(std::move(j), d, dev));
       }},
  };
  return reg;
}

// Build the force module named by force_field_type.
inline std::unique_ptr<Force> MakeForce(const std::string& force_field_type,
                                        std::string json, DeviceData* data,
                                        int device = 0) {
  auto it = ForceRegistry().find(force_field_type);
  if (it == ForceRegistry().end())
    throw std::runtime_error("unsupported force_field.type: " + force_field_type);
  return it->second(std::move(json), data, device);
}

}  // namespace plugin
}  // namespace rbmdx
