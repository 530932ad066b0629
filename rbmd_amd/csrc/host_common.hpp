// Host pieces shared by the single-system engine (api.cpp) and the replica
// batch (batch.cpp): the error type and HIP check, the RBMD JSON
// configuration and its parser, and RBL's shell-selection frequency.
// Internal linkage: each translation unit gets its own copy.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.hpp"
#include "rbmdx.h"
#include "sog.hpp"

namespace {

using namespace rbmdx;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                         \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      throw Error(RBMDX_E_HIP, std::string(#expr) + ": " +                   \
                                   hipGetErrorString(e_));                   \
  } while (0)

// ---------------------------------------------------------------------------
// Configuration (RBMD JSON surface; docs/json-configuration.en.md:75-167)
// ---------------------------------------------------------------------------
struct Config {
  std::string force_type;     // "LJ/CUT" | "LJ/CUT/COUL/LONG" | "CVFF"
  std::string neighbor_type;  // "RBL" | "Verlet-List" | "VERLET-SOG"
  double cut_off = 0, skin = 0, r_core = 0, lj_cut_off = 0;
  int sample_num = 0;         // neighbor_sample_num
  bool energy_rbl = false;
  std::string coulomb_type;   // "" | "RBE" | "EWALD" | "RBSOG"
  int rbsog_level = -1;       // coulomb.rbsog_level (RBSOG, VERLET-SOG)
  bool rbsog_exact = false;   // extension: corrected SOG set-up
  bool sog_near = false;      // extension: SOG-consistent near field with RBSOG
                              // (coulomb.rbsog_near_field "sog"; RBMD: erfc)
  bool has_alpha = false;
  double alpha = 0, accuracy = 0;
  int coulomb_sample_num = 0;
  bool has_kmax = false;
  int kmax[3] = {0, 0, 0};
  bool energy_rbe = false;
  bool has_seed = false;
  uint32_t seed = 0;
  double qqr2e = 1.0;
  double ftm2v = 1.0;  // force/mass -> velocity/time (unit_factor.cpp:30-50)
  bool fp64 = false;
  // CVFF: 1-2 / 1-3 / 1-4 special-bond weights, hyper_parameters.extend
  // .special_bonds; 1 when absent (read_special_bond_weights, cvff.cpp:303-330)
  double w_special[3] = {1.0, 1.0, 1.0};
  std::string text;  // the JSON text (its execution section is read by the
                     // molecular-dynamics step, parse_md)
  bool coulomb() const { return force_type != "LJ/CUT"; }
};

double num_of(const Json* j, const char* what) {
  if (!j || j->kind != Json::Number)
    throw Error(RBMDX_E_CONFIG, std::string("missing or non-numeric ") + what);
  return j->num;
}
std::string str_of(const Json* j, const char* what) {
  if (!j || j->kind != Json::String)
    throw Error(RBMDX_E_CONFIG, std::string("missing or non-string ") + what);
  return j->str;
}

Config parse_config(const std::string& text) {
  Json root;
  try {
    root = parse_json(text);
  } catch (const std::exception& e) {
    throw Error(RBMDX_E_CONFIG, e.what());
  }
  const Json* hp = root.get("hyper_parameters");
  if (!hp) hp = &root;  // a bare hyper_parameters object
  Config c;
  if (const Json* u = root.at({"init_configuration", "read_data", "unit"})) {
    const std::string unit = str_of(u, "unit");
    if (unit == "LJ") {
      c.qqr2e = 1.0;
      c.ftm2v = 1.0;
    } else if (unit == "REAL") {
      c.qqr2e = 332.06371;
      c.ftm2v = 4.186e-4;
    } else if (unit == "METAL") {
      c.qqr2e = 14.399645;
      c.ftm2v = 1.0 / 1.0364269e-4;
    }
    else
      throw Error(RBMDX_E_CONFIG, "unsupported unit style: " + unit);
  }
  c.force_type = str_of(hp->at({"force_field", "type"}), "force_field.type");
  if (c.force_type != "LJ/CUT" && c.force_type != "LJ/CUT/COUL/LONG" &&
      c.force_type != "CVFF")
    throw Error(RBMDX_E_CONFIG,
                "force_field.type not supported by this build: " + c.force_type);
  if (c.force_type == "CVFF") {
    if (const Json* sb = hp->at({"extend", "special_bonds"})) {
      if (sb->kind != Json::Array || sb->arr.size() > 3)
        throw Error(RBMDX_E_CONFIG,
                    "extend.special_bonds must be an array of at most 3 numbers");
      for (size_t q = 0; q < sb->arr.size(); ++q)
        c.w_special[q] = num_of(&sb->arr[q], "extend.special_bonds");
    }
  }
  const Json* nb = hp->get("neighbor");
  if (!nb) throw Error(RBMDX_E_CONFIG, "hyper_parameters.neighbor missing");
  c.neighbor_type = str_of(nb->get("type"), "neighbor.type");
  if (c.neighbor_type != "RBL" && c.neighbor_type != "Verlet-List" &&
      c.neighbor_type != "VERLET-SOG")
    throw Error(RBMDX_E_CONFIG,
                "neighbor.type not supported by this build: " + c.neighbor_type);
  c.cut_off = num_of(nb->get("cut_off"), "neighbor.cut_off");
  if (!(c.cut_off > 0)) throw Error(RBMDX_E_CONFIG, "neighbor.cut_off must be > 0");
  if (const Json* s = nb->get("skin")) c.skin = num_of(s, "neighbor.skin");
  if (!(c.skin >= 0) || !std::isfinite(c.skin))
    throw Error(RBMDX_E_CONFIG, "neighbor.skin must be finite and >= 0");
  c.lj_cut_off = c.cut_off;
  if (const Json* l = nb->get("lj_cut_off"))
    c.lj_cut_off = num_of(l, "neighbor.lj_cut_off");
  if (!(c.lj_cut_off > 0))
    throw Error(RBMDX_E_CONFIG, "neighbor.lj_cut_off must be > 0");
  if (c.neighbor_type == "RBL") {
    c.r_core = num_of(nb->get("r_core"), "neighbor.r_core");
    const double p = num_of(nb->get("neighbor_sample_num"),
                            "neighbor.neighbor_sample_num");
    if (!(p > 0))
      throw Error(RBMDX_E_CONFIG, "neighbor_sample_num must be larger than 0");
    c.sample_num = (int)p;
    if (!(c.r_core > 0 && c.r_core < c.cut_off))
      throw Error(RBMDX_E_CONFIG, "r_core must be > 0 and less than the cutoff");
    const Json* ef = nb->get("energy_rbl_flag");
    if (!ef)
      throw Error(RBMDX_E_CONFIG,
                  "when using RBL for the neighbor type, the key "
                  "'energy_rbl_flag' must be defined");
    c.energy_rbl = str_of(ef, "energy_rbl_flag") == "yes";
  }
  if (const Json* pr = hp->get("precision")) {
    const std::string p = str_of(pr, "precision");
    if (p == "fp64")
      c.fp64 = true;
    else if (p != "fp32")
      throw Error(RBMDX_E_CONFIG, "precision must be fp32 or fp64");
  }
  const Json* cl = hp->get("coulomb");
  if (c.coulomb()) {
    if (!cl)
      throw Error(RBMDX_E_CONFIG, c.force_type + " needs hyper_parameters.coulomb");
    c.coulomb_type = str_of(cl->get("type"), "coulomb.type");
    if (c.coulomb_type != "RBE" && c.coulomb_type != "EWALD" &&
        c.coulomb_type != "RBSOG")
      throw Error(RBMDX_E_CONFIG,
                  "coulomb.type not supported by this build: " + c.coulomb_type);
    // RBSOG: a level preset replaces accuracy / rbsog_* keys
    // (kspace_calculator.cpp:254-300); VERLET-SOG reads the same level
    // (lj_cut_coul_kspace.cpp:651-665)
    const bool sog = c.coulomb_type == "RBSOG" || c.neighbor_type == "VERLET-SOG";
    if (sog) {
      const Json* lv = cl->get("rbsog_level");
      if (!lv) throw Error(RBMDX_E_CONFIG, "RBSOG requires hyper_parameters.coulomb.rbsog_level");
      for (const char* legacy : {"accuracy", "rbsog_b", "rbsog_sigma", "rbsog_Mmax",
                                 "rbsog_Kcut"})
        if (cl->get(legacy))
          throw Error(RBMDX_E_CONFIG, std::string("Legacy RBSOG config key '") + legacy +
                                          "' is not supported. Use "
                                          "hyper_parameters.coulomb.rbsog_level instead.");
      const double lvv = num_of(lv, "coulomb.rbsog_level");
      if (!(lvv >= 0 && lvv <= 4) || lvv != std::floor(lvv))
        throw Error(RBMDX_E_CONFIG, "RBSOG rbsog_level must be in [0, 4]");
      c.rbsog_level = (int)lvv;
      if (const Json* ex = cl->get("rbsog_exact")) {
        if (ex->kind != Json::Bool)
          throw Error(RBMDX_E_CONFIG, "coulomb.rbsog_exact must be true or false");
        c.rbsog_exact = ex->b;
      }
    }
    // RBMD screens the RBL / Verlet near field with erfc even under RBSOG
    // (lj_cut_coul_kspace.cpp:473-492); "sog" pairs the SOG far field with its
    // own near field, as VERLET-SOG does
    if (const Json* nf = cl->get("rbsog_near_field")) {
      const std::string v = str_of(nf, "coulomb.rbsog_near_field");
      if (v != "erfc" && v != "sog")
        throw Error(RBMDX_E_CONFIG, "coulomb.rbsog_near_field must be \"erfc\" or \"sog\"");
      if (c.coulomb_type != "RBSOG")
        throw Error(RBMDX_E_CONFIG, "coulomb.rbsog_near_field applies to RBSOG only");
      if (c.neighbor_type == "VERLET-SOG" && v == "erfc")
        throw Error(RBMDX_E_CONFIG, "VERLET-SOG always uses the SOG near field");
      c.sog_near = v == "sog";
    }
    if (const Json* a = cl->get("alpha")) {
      c.has_alpha = true;
      c.alpha = num_of(a, "coulomb.alpha");
      if (!(c.alpha > 0))
        throw Error(RBMDX_E_CONFIG, "coulomb alpha override must be > 0");
    } else if (sog) {
      SogPresetValues pv;
      sog_preset(c.rbsog_level, &pv);
      c.accuracy = pv.accuracy;  // DeriveAlphaFromAccuracy(target_accuracy)
    } else {
      const Json* acc = cl->get("accuracy");
      if (!acc)
        throw Error(RBMDX_E_CONFIG,
                    "coulomb accuracy must be defined when alpha is derived");
      c.accuracy = num_of(acc, "coulomb.accuracy");
      if (!(c.accuracy > 0))
        throw Error(RBMDX_E_CONFIG, "coulomb accuracy must be > 0");
    }
    if (const Json* k = cl->get("kmax")) {
      if (k->kind != Json::Array || k->arr.size() != 3)
        throw Error(RBMDX_E_CONFIG, "'kmax' must contain exactly 3 integers");
      c.has_kmax = true;
      for (int d = 0; d < 3; ++d) c.kmax[d] = (int)num_of(&k->arr[d], "kmax");
    }
    if (c.coulomb_type == "RBSOG") {
      const int P =
          (int)num_of(cl->get("coulomb_sample_num"), "coulomb.coulomb_sample_num");
      if (P <= 0)
        throw Error(RBMDX_E_CONFIG, "the RBSOG sample number must be greater than 0");
      c.coulomb_sample_num = ((P + 15) / 16) * 16;  // kspace_calculator.cpp:300
    }
    if (c.coulomb_type == "RBE") {
      c.coulomb_sample_num =
          (int)num_of(cl->get("coulomb_sample_num"), "coulomb.coulomb_sample_num");
      if (c.coulomb_sample_num <= 0)
        throw Error(RBMDX_E_CONFIG, "the RBE sample number must be greater than 0");
      const Json* ef = cl->get("energy_rbe_flag");
      if (!ef)
        throw Error(RBMDX_E_CONFIG,
                    "when using RBE for the coulomb type, the key "
                    "'energy_rbe_flag' must be defined");
      c.energy_rbe = str_of(ef, "energy_rbe_flag") == "yes";
    }
    if (const Json* s = cl->get("random_seed")) {
      const double v = num_of(s, "coulomb.random_seed");
      if (v < 0 || v > 4294967295.0)
        throw Error(RBMDX_E_CONFIG, "coulomb.random_seed exceeds UINT_MAX");
      c.has_seed = true;
      c.seed = (uint32_t)v;
    }
  }
  c.text = text;
  return c;
}

// ---------------------------------------------------------------------------
// Molecular-dynamics step settings: the execution section RBMD's ensembles
// read (src/simulate_pipeline/src/nvt_ensemble.cpp:66-97,
// src/simulate_pipeline/src/controller/pressure_controller/
// nose_hoover_controller.cpp:89-120) and the unit constants
// (src/common/unit_factor.cpp:23-37).
// ---------------------------------------------------------------------------
struct MdConfig {
  std::string ensemble;        // "NVE" | "NVT"
  bool nose_hoover = false;    // NVT with temp_ctrl_type NOSE_HOOVER
  double dt = 0;               // execution.timestep
  double t_start = 0, t_stop = 0, t_damp = 0;  // execution.temperature
  int64_t num_steps = 0;       // execution.num_steps (temperature ramp)
  bool energy_stable = false;  // NVE with RBL lists (nve_ensemble.cpp:81-84)
  double mvv2e = 1.0, kb = 1.0;
};

MdConfig parse_md(const Config& c) {
  Json root;
  try {
    root = parse_json(c.text);
  } catch (const std::exception& e) {
    throw Error(RBMDX_E_CONFIG, e.what());
  }
  MdConfig m;
  const Json* ex = root.get("execution");
  if (!ex) throw Error(RBMDX_E_CONFIG, "execution section missing");
  if (const Json* u = root.at({"init_configuration", "read_data", "unit"})) {
    const std::string unit = str_of(u, "unit");
    if (unit == "REAL") {
      m.kb = 1.9872067 * std::pow(10.0, -3.0);
      m.mvv2e = 1.0 / (4.186 * std::pow(10.0, -4.0));
    } else if (unit == "METAL") {
      m.kb = 8.617343e-5;
      m.mvv2e = 1.0364269e-4;
    }
  }
  m.ensemble = str_of(ex->get("ensemble"), "execution.ensemble");
  m.dt = num_of(ex->get("timestep"), "execution.timestep");
  if (!(m.dt > 0) || !std::isfinite(m.dt))
    throw Error(RBMDX_E_CONFIG, "execution.timestep must be > 0");
  if (const Json* it = ex->get("integration_type"))
    if (str_of(it, "execution.integration_type") != "vv")
      throw Error(RBMDX_E_CONFIG,
                  "execution.integration_type not supported by this build "
                  "(only \"vv\")");
  if (const Json* cb = ex->get("com_bias"))
    if (str_of(cb, "execution.com_bias") == "yes")
      throw Error(RBMDX_E_CONFIG, "execution.com_bias \"yes\" not supported by this build");
  if (const Json* sh = root.at({"hyper_parameters", "extend", "fix_shake"}))
    if (sh->kind == Json::Bool && sh->b)
      throw Error(RBMDX_E_CONFIG, "extend.fix_shake not supported by this build");
  if (const Json* ns = ex->get("num_steps"))
    m.num_steps = (int64_t)num_of(ns, "execution.num_steps");
  if (m.ensemble == "NVE") {
    m.energy_stable = c.neighbor_type == "RBL";
    // extension: the energy-stable rescale can be switched off
    if (const Json* es = ex->get("energy_stable"))
      m.energy_stable = m.energy_stable && str_of(es, "execution.energy_stable") != "no";
  } else if (m.ensemble == "NVT") {
    const std::string tc = str_of(ex->get("temp_ctrl_type"), "execution.temp_ctrl_type");
    if (tc != "NOSE_HOOVER")
      throw Error(RBMDX_E_CONFIG,
                  "execution.temp_ctrl_type not supported by this build: " + tc);
    m.nose_hoover = true;
    const Json* t = ex->get("temperature");
    if (!t || t->kind != Json::Array || t->arr.size() < 3)
      throw Error(RBMDX_E_CONFIG,
                  "execution.temperature must be [start, stop, damp]");
    m.t_start = num_of(&t->arr[0], "execution.temperature");
    m.t_stop = num_of(&t->arr[1], "execution.temperature");
    m.t_damp = num_of(&t->arr[2], "execution.temperature");
    if (!(m.t_start > 0) || !(m.t_stop > 0) || !(m.t_damp > 0))
      throw Error(RBMDX_E_CONFIG, "execution.temperature entries must be > 0");
    if (m.t_start != m.t_stop && m.num_steps <= 0)
      throw Error(RBMDX_E_CONFIG, "a temperature ramp needs execution.num_steps > 0");
  } else {
    throw Error(RBMDX_E_CONFIG,
                "execution.ensemble not supported by this build: " + m.ensemble);
  }
  return m;
}

// RblFullNeighborListBuilder::GetRblParams
// (rbl_full_neighbor_list_builder.cpp:193-228), same float/double mix.
int rbl_frequency(long n_global, const double L[3], double r_core, double rc,
                  int sample_num) {
  const float volume = (float)L[0] * (float)L[1] * (float)L[2];
  const float rho = (float)n_global / volume;
  const float coeff = (float)(1.0 + (0.05 / rho - 0.05));
  const int icoeff = (int)std::round((double)coeff);
  const float rsf = (float)r_core, rcf = (float)rc;
  const int rs_num =
      (int)(icoeff * rho * std::ceil(4.0 / 3.0 * M_PI * std::pow((double)rsf, 3)) + 1);
  const int rc_num =
      (int)(icoeff * rho * std::ceil(4.0 / 3.0 * M_PI * std::pow((double)rcf, 3)) + 1);
  const int shell = rc_num - rs_num;
  if (shell <= sample_num) return 1;
  return (int)std::ceil((float)shell / (float)sample_num);
}

}  // namespace
