// Minimal C++ driver of the force-module layer: runs `steps` evaluations
// through rbmdx::plugin::MakeForce(force_field.type, ...) and prints the
// potential energy, the virial and the forces of the last evaluation.
//
//   rbmdx_plugin_driver config.json system.txt force_field_type [steps]
//     system file: n ntypes / lo[3] / L[3] / ntypes lines "eps sigma" /
//     n lines "id type q x y z";
//   rbmdx_plugin_driver --rbmd config.json [steps]
//     RBMD's own input: init_configuration.read_data.{file, atom_style,
//     replicate} (the file relative to the JSON's directory) read by
//     rbmdx_data_read, eps / sigma from its Pair Coeffs, the special-bond
//     topology from its bonds, angles and dihedrals, and force_field.type.
//   rbmdx_plugin_driver --md config.json [num_steps]
//     RBMD's simulate loop on the same input through the C ABI: GAUSS
//     velocities when read_data.velocity_type asks for them (temperature
//     execution.temperature[0], seed read_data.velocity_seed or seed), the
//     ensemble of the execution section (rbmdx_md_*), one line
//     "thermo step T pe ke" every outputs.thermo_out.interval steps and a
//     trajectory frame appended to rbmd.trj (working directory) every
//     outputs.trajectory_out.interval steps, step 0 included.
//
// Positions are uploaded once; precision follows the JSON ("fp32" default,
// "fp64").
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "force_plugin.hpp"
#include "json.hpp"

using rbmdx::plugin::DeviceData;

static std::string slurp(const char* path) {
  std::ifstream f(path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

template <typename R>
static int run(const std::string& json, const std::string& fftype, DeviceData& d,
               const std::vector<double>& pos, int steps) {
  const size_t n = (size_t)d.num_atoms;
  std::vector<R> h(n);
  R* dev[6];
  for (int k = 0; k < 6; ++k)
    if (hipMalloc((void**)&dev[k], n * sizeof(R)) != hipSuccess) return 3;
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; i < n; ++i) h[i] = (R)pos[3 * i + a];
    hipMemcpy(dev[a], h.data(), n * sizeof(R), hipMemcpyHostToDevice);
  }
  d.d_px = dev[0];
  d.d_py = dev[1];
  d.d_pz = dev[2];
  d.d_fx = dev[3];
  d.d_fy = dev[4];
  d.d_fz = dev[5];
  auto force = rbmdx::plugin::MakeForce(fftype, json, &d, 0);
  force->Init();
  force->set_energy_every_call(true);
  for (int s = 0; s < steps; ++s) force->Execute();
  force->EvaluatePotentialEnergy();
  const auto& pe = force->potential_energy();
  std::printf("energy %.17g %.17g %.17g\n", pe.vdwl, pe.coul, pe.kspace);
  std::printf("virial %.17g %.17g %.17g %.17g %.17g %.17g\n", pe.virial[0],
              pe.virial[1], pe.virial[2], pe.virial[3], pe.virial[4],
              pe.virial[5]);
  std::vector<R> f(3 * n);
  for (int a = 0; a < 3; ++a) {
    hipMemcpy(h.data(), dev[3 + a], n * sizeof(R), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < n; ++i) f[3 * i + a] = h[i];
  }
  for (size_t i = 0; i < n; ++i)
    std::printf("f %.9g %.9g %.9g\n", (double)f[3 * i], (double)f[3 * i + 1],
                (double)f[3 * i + 2]);
  force.reset();
  for (int k = 0; k < 6; ++k) hipFree(dev[k]);
  return 0;
}

// read_data "replicate_bond_periodic", "bond_periodic" or "bond/periodic",
// the first present (default on, as RBMD's
// LammpsFullReader, lammps_full_reader.cpp:680-707).
static int32_t bond_periodic(const rbmdx::Json* rd) {
  for (const char* key : {"replicate_bond_periodic", "bond_periodic", "bond/periodic"})
    if (const rbmdx::Json* b = rd->get(key))
      if (b->kind == rbmdx::Json::Bool) return b->b ? 1 : 0;
  return 1;
}

// RBMD's input: the JSON and the data file it names.
static int run_rbmd(const char* json_path, int steps) {
  const std::string json = slurp(json_path);
  const rbmdx::Json root = rbmdx::parse_json(json);
  const rbmdx::Json* rd = nullptr;
  if (const rbmdx::Json* ic = root.get("init_configuration")) rd = ic->get("read_data");
  const rbmdx::Json* file = rd ? rd->get("file") : nullptr;
  if (!file || file->kind != rbmdx::Json::String) {
    std::fprintf(stderr, "error: init_configuration.read_data.file is missing\n");
    return 2;
  }
  std::string path = file->str;
  const std::string jp = json_path;
  if (!path.empty() && path[0] != '/' && jp.find('/') != std::string::npos)
    path = jp.substr(0, jp.rfind('/') + 1) + path;
  std::string style = "full";
  if (const rbmdx::Json* st = rd->get("atom_style"))
    if (st->kind == rbmdx::Json::String) style = st->str;
  int32_t rep[3] = {1, 1, 1};
  bool has_rep = false;
  if (const rbmdx::Json* r = rd->get("replicate")) {
    if (r->kind != rbmdx::Json::Array || r->arr.size() != 3) {
      std::fprintf(stderr, "error: read_data.replicate must hold 3 integers\n");
      return 2;
    }
    for (int k = 0; k < 3; ++k) rep[k] = (int32_t)r->arr[k].num;
    has_rep = true;
  }
  std::string fftype;
  const rbmdx::Json* hp = root.get("hyper_parameters");
  if (hp)
    if (const rbmdx::Json* ff = hp->get("force_field"))
      if (const rbmdx::Json* t = ff->get("type")) fftype = t->str;
  rbmdx_data* data = nullptr;
  if (rbmdx_data_read(path.c_str(), style.c_str(), nullptr, &data) != RBMDX_OK ||
      (has_rep && rbmdx_data_replicate(data, rep, bond_periodic(rd)) != RBMDX_OK)) {
    std::fprintf(stderr, "error: %s\n", rbmdx_data_last_error(data));
    rbmdx_data_destroy(data);
    return 1;
  }
  int64_t n = 0;
  int32_t nt = 0, maxs = 0, flags = 0;
  rbmdx_data_info(data, &n, &nt, &maxs, &flags);
  DeviceData d;
  d.num_atoms = n;
  d.type.resize(n);
  d.id.resize(n);
  d.charge.resize(n);
  d.eps.resize(nt);
  d.sigma.resize(nt);
  std::vector<double> pos(3 * n);
  rbmdx_data_get(data, d.box_lo, d.box_len, d.type.data(), d.id.data(), d.charge.data(),
                 pos.data(), nullptr, nullptr, nullptr, d.eps.data(), d.sigma.data());
  if (flags & 4) {
    d.maxspecial = maxs;
    d.nspecial.resize(3 * n);
    d.special.resize((size_t)n * maxs);
    rbmdx_data_special(data, d.nspecial.data(), d.special.data());
  }
  rbmdx_data_destroy(data);
  std::printf("system %lld atoms %d types maxspecial %d\n", (long long)n, nt, d.maxspecial);
  const bool fp64 = json.find("\"fp64\"") != std::string::npos;
  try {
    return fp64 ? run<double>(json, fftype, d, pos, steps)
                : run<float>(json, fftype, d, pos, steps);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}

// The simulate loop through the C ABI (rbmdx_create ... rbmdx_md_run).
template <typename R>
static int md_loop(const std::string& json, rbmdx_ctx* ctx, int64_t n,
                   const std::vector<double>& pos, const std::vector<double>& vel,
                   const std::vector<int32_t>& img, const std::vector<int32_t>& type,
                   const std::vector<int32_t>& id, const std::vector<double>& q,
                   const double* lo, const double* L, int64_t nsteps, int64_t thermo_every,
                   int64_t traj_every) {
  std::vector<R*> dev(9, nullptr);
  int32_t* dimg[3] = {nullptr, nullptr, nullptr};
  int rc = 0;
  auto fail = [&](const char* what) {
    std::fprintf(stderr, "error: %s: %s\n", what, rbmdx_last_error(ctx));
    return 1;
  };
  for (int k = 0; k < 9; ++k)
    if (hipMalloc((void**)&dev[k], n * sizeof(R)) != hipSuccess) return 3;
  for (int k = 0; k < 3; ++k)
    if (hipMalloc((void**)&dimg[k], n * sizeof(int32_t)) != hipSuccess) return 3;
  std::vector<R> h(n);
  std::vector<int32_t> hi(n);
  for (int a = 0; a < 3; ++a) {
    for (int64_t i = 0; i < n; ++i) h[i] = (R)pos[3 * i + a];
    hipMemcpy(dev[a], h.data(), n * sizeof(R), hipMemcpyHostToDevice);
    for (int64_t i = 0; i < n; ++i) h[i] = (R)vel[3 * i + a];
    hipMemcpy(dev[3 + a], h.data(), n * sizeof(R), hipMemcpyHostToDevice);
    for (int64_t i = 0; i < n; ++i) hi[i] = img[3 * i + a];
    hipMemcpy(dimg[a], hi.data(), n * sizeof(int32_t), hipMemcpyHostToDevice);
  }
  auto frame = [&](int64_t step) {
    std::vector<double> p(3 * n), v(3 * n);
    std::vector<int32_t> im(3 * n);
    for (int a = 0; a < 3; ++a) {
      hipMemcpy(h.data(), dev[a], n * sizeof(R), hipMemcpyDeviceToHost);
      for (int64_t i = 0; i < n; ++i) p[3 * i + a] = (double)h[i];
      hipMemcpy(h.data(), dev[3 + a], n * sizeof(R), hipMemcpyDeviceToHost);
      for (int64_t i = 0; i < n; ++i) v[3 * i + a] = (double)h[i];
      hipMemcpy(hi.data(), dimg[a], n * sizeof(int32_t), hipMemcpyDeviceToHost);
      for (int64_t i = 0; i < n; ++i) im[3 * i + a] = hi[i];
    }
    return rbmdx_write_dump("rbmd.trj", step > 0, step, n, lo, L, id.data(), type.data(),
                            q.data(), p.data(), im.data(), v.data());
  };
  auto thermo = [&]() {
    rbmdx_md_state_t st;
    if (rbmdx_md_state(ctx, &st) != RBMDX_OK) return fail("rbmdx_md_state");
    std::printf("thermo %lld %.10g %.10g %.10g\n", (long long)st.step, st.temperature,
                st.pe, st.ke);
    return 0;
  };
  const uint32_t eflag = RBMDX_FLAG_ENERGY;
  (void)json;
  if (rbmdx_md_init(ctx, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dimg[0], dimg[1],
                    dimg[2], dev[6], dev[7], dev[8], eflag, nullptr) != RBMDX_OK)
    rc = fail("rbmdx_md_init");
  if (!rc && thermo_every > 0) rc = thermo();
  if (!rc && traj_every > 0 && frame(0) != RBMDX_OK) rc = fail("rbmd.trj");
  int64_t step = 0;
  while (!rc && step < nsteps) {
    int64_t next = nsteps;
    if (thermo_every > 0) next = std::min(next, (step / thermo_every + 1) * thermo_every);
    if (traj_every > 0) next = std::min(next, (step / traj_every + 1) * traj_every);
    const bool th = thermo_every > 0 && next % thermo_every == 0;
    if (rbmdx_md_run(ctx, (int32_t)(next - step), dev[0], dev[1], dev[2], dev[3], dev[4],
                     dev[5], dimg[0], dimg[1], dimg[2], dev[6], dev[7], dev[8],
                     th ? eflag : 0u, nullptr) != RBMDX_OK) {
      rc = fail("rbmdx_md_run");
      break;
    }
    step = next;
    if (th) rc = thermo();
    if (!rc && traj_every > 0 && step % traj_every == 0 && frame(step) != RBMDX_OK)
      rc = fail("rbmd.trj");
  }
  hipDeviceSynchronize();
  for (R* p : dev) hipFree(p);
  for (int32_t* p : dimg) hipFree(p);
  return rc;
}

static int run_md(const char* json_path, int64_t steps_override) {
  const std::string json = slurp(json_path);
  rbmdx::Json root;
  try {
    root = rbmdx::parse_json(json);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  const rbmdx::Json* rd = root.at({"init_configuration", "read_data"});
  const rbmdx::Json* file = rd ? rd->get("file") : nullptr;
  if (!file || file->kind != rbmdx::Json::String) {
    std::fprintf(stderr, "error: init_configuration.read_data.file is missing\n");
    return 2;
  }
  std::string path = file->str;
  const std::string jp = json_path;
  if (!path.empty() && path[0] != '/' && jp.find('/') != std::string::npos)
    path = jp.substr(0, jp.rfind('/') + 1) + path;
  auto str = [&](const rbmdx::Json* j, const char* dflt) {
    return (j && j->kind == rbmdx::Json::String) ? j->str : std::string(dflt);
  };
  auto num = [&](const rbmdx::Json* j, double dflt) {
    return (j && j->kind == rbmdx::Json::Number) ? j->num : dflt;
  };
  const std::string style = str(rd->get("atom_style"), "full");
  const std::string unit = str(rd->get("unit"), "LJ");
  int32_t rep[3] = {1, 1, 1};
  const rbmdx::Json* r = rd->get("replicate");
  if (r && r->kind == rbmdx::Json::Array && r->arr.size() == 3)
    for (int k = 0; k < 3; ++k) rep[k] = (int32_t)r->arr[k].num;
  rbmdx_data* data = nullptr;
  if (rbmdx_data_read(path.c_str(), style.c_str(), nullptr, &data) != RBMDX_OK ||
      (r && rbmdx_data_replicate(data, rep, bond_periodic(rd)) != RBMDX_OK)) {
    std::fprintf(stderr, "error: %s\n", rbmdx_data_last_error(data));
    rbmdx_data_destroy(data);
    return 1;
  }
  const rbmdx::Json* ex = root.get("execution");
  std::string vt = str(rd->get("velocity_type"), "NONE");
  for (char& c : vt) c = (char)std::toupper((unsigned char)c);
  if (vt == "GAUSS" || vt == "GAUSSIAN") {
    double t0 = 0;
    if (const rbmdx::Json* t = ex ? ex->get("temperature") : nullptr)
      if (t->kind == rbmdx::Json::Array && !t->arr.empty()) t0 = t->arr[0].num;
    const int seed = (int)num(rd->get("velocity_seed"), num(rd->get("seed"), 12345));
    if (rbmdx_data_gauss_velocities(data, t0, seed, unit.c_str()) != RBMDX_OK) {
      std::fprintf(stderr, "error: %s\n", rbmdx_data_last_error(data));
      rbmdx_data_destroy(data);
      return 1;
    }
  } else if (vt != "NONE") {
    std::fprintf(stderr, "error: Unsupported velocity_type: %s\n", vt.c_str());
    rbmdx_data_destroy(data);
    return 1;
  }
  int64_t n = 0;
  int32_t nt = 0, maxs = 0, flags = 0;
  rbmdx_data_info(data, &n, &nt, &maxs, &flags);
  double lo[3], L[3];
  std::vector<int32_t> type(n), id(n), img(3 * n), nsp, sp;
  std::vector<double> q(n), pos(3 * n), vel(3 * n), mass(nt), eps(nt), sig(nt);
  rbmdx_data_get(data, lo, L, type.data(), id.data(), q.data(), pos.data(), vel.data(),
                 img.data(), mass.data(), eps.data(), sig.data());
  if (flags & 4) {
    nsp.resize(3 * n);
    sp.resize((size_t)n * maxs);
    rbmdx_data_special(data, nsp.data(), sp.data());
  }
  rbmdx_data_destroy(data);
  const int64_t nsteps =
      steps_override >= 0 ? steps_override : (int64_t)num(ex ? ex->get("num_steps") : nullptr, 0);
  const int64_t thermo_every =
      (int64_t)num(root.at({"outputs", "thermo_out", "interval"}), 0);
  const int64_t traj_every =
      (int64_t)num(root.at({"outputs", "trajectory_out", "interval"}), 0);
  rbmdx_ctx* ctx = nullptr;
  int rc = rbmdx_create(json.c_str(), 0, &ctx);
  if (rc == RBMDX_OK)
    rc = rbmdx_set_system(ctx, n, lo, L, type.data(), id.data(), q.data(), nt, eps.data(),
                          sig.data());
  if (rc == RBMDX_OK && (flags & 4))
    rc = rbmdx_set_special(ctx, maxs, nsp.data(), sp.data());
  if (rc == RBMDX_OK) rc = rbmdx_set_masses(ctx, mass.data());
  if (rc != RBMDX_OK) {
    std::fprintf(stderr, "error: %s\n", ctx ? rbmdx_last_error(ctx) : "rbmdx_create");
    rbmdx_destroy(ctx);
    return 1;
  }
  std::printf("system %lld atoms %d types steps %lld\n", (long long)n, nt,
              (long long)nsteps);
  const bool fp64 = json.find("\"fp64\"") != std::string::npos;
  rc = fp64 ? md_loop<double>(json, ctx, n, pos, vel, img, type, id, q, lo, L, nsteps,
                              thermo_every, traj_every)
            : md_loop<float>(json, ctx, n, pos, vel, img, type, id, q, lo, L, nsteps,
                             thermo_every, traj_every);
  rbmdx_destroy(ctx);
  return rc;
}

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "--md")
    return run_md(argv[2], argc > 3 ? std::atoll(argv[3]) : -1);
  if (argc >= 3 && std::string(argv[1]) == "--rbmd")
    return run_rbmd(argv[2], argc > 3 ? std::atoi(argv[3]) : 1);
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s config.json system.txt force_field_type [steps]\n",
                 argv[0]);
    return 2;
  }
  const std::string json = slurp(argv[1]);
  std::ifstream sf(argv[2]);
  int64_t n;
  int nt;
  sf >> n >> nt;
  DeviceData d;
  d.num_atoms = n;
  for (double& v : d.box_lo) sf >> v;
  for (double& v : d.box_len) sf >> v;
  d.eps.resize(nt);
  d.sigma.resize(nt);
  for (int t = 0; t < nt; ++t) sf >> d.eps[t] >> d.sigma[t];
  d.id.resize(n);
  d.type.resize(n);
  d.charge.resize(n);
  std::vector<double> pos(3 * n);
  for (int64_t i = 0; i < n; ++i)
    sf >> d.id[i] >> d.type[i] >> d.charge[i] >> pos[3 * i] >> pos[3 * i + 1] >>
        pos[3 * i + 2];
  const int steps = argc > 4 ? std::atoi(argv[4]) : 1;
  const bool fp64 = json.find("\"fp64\"") != std::string::npos;
  try {
    return fp64 ? run<double>(json, argv[3], d, pos, steps)
                : run<float>(json, argv[3], d, pos, steps);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
