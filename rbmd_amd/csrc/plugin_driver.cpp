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
//
// Positions are uploaded once; precision follows the JSON ("fp32" default,
// "fp64").
#include <hip/hip_runtime_api.h>

#include <cstdio>
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
  if (rbmdx_data_read(path.c_str(), style.c_str(), has_rep ? rep : nullptr, &data) !=
      RBMDX_OK) {
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

int main(int argc, char** argv) {
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
