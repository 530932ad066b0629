// C ABI of the LAMMPS data reader (rbmdx_data_*, include/rbmdx.h).
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lammps_data.hpp"
#include "rbmdx.h"

struct rbmdx_data {
  std::string err;
  std::unique_ptr<rbmdx::LammpsData> d;
  std::vector<int32_t> nspecial, special;
  int maxs = 1;
};

namespace {
template <typename F>
int guarded_d(rbmdx_data* c, F&& f) {
  try {
    f();
    return RBMDX_OK;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return RBMDX_E_CONFIG;
  }
}
}  // namespace

extern "C" {

int rbmdx_data_read(const char* path, const char* atom_style, const int32_t* replicate,
                    rbmdx_data** out) {
  if (!out) return RBMDX_E_ARG;
  *out = nullptr;
  if (!path || !atom_style) return RBMDX_E_ARG;
  rbmdx_data* c = new rbmdx_data();
  const int rc = guarded_d(c, [&] {
    auto d = std::make_unique<rbmdx::LammpsData>(rbmdx::read_lammps_data(path, atom_style));
    if (replicate) rbmdx::replicate_lammps_data(*d, replicate[0], replicate[1], replicate[2]);
    for (int64_t v : d->id)
      if (v > 0x7fffffffLL) throw std::runtime_error("atom ids beyond 2^31 - 1");
    rbmdx::build_special(*d, c->nspecial, c->special, c->maxs);
    c->d = std::move(d);
  });
  *out = c;  // returned even on error so rbmdx_data_last_error can explain
  return rc;
}

int rbmdx_data_replicate(rbmdx_data* data, const int32_t* factors, int32_t bond_periodic) {
  if (!data || !data->d) return RBMDX_E_STATE;
  if (!factors) return RBMDX_E_ARG;
  return guarded_d(data, [&] {
    rbmdx::LammpsData d = *data->d;
    rbmdx::replicate_lammps_data(d, factors[0], factors[1], factors[2], bond_periodic != 0);
    for (int64_t v : d.id)
      if (v > 0x7fffffffLL) throw std::runtime_error("atom ids beyond 2^31 - 1");
    rbmdx::build_special(d, data->nspecial, data->special, data->maxs);
    *data->d = std::move(d);
  });
}

int rbmdx_data_gauss_velocities(rbmdx_data* data, double temperature, int32_t seed,
                                 const char* unit) {
  if (!data || !data->d) return RBMDX_E_STATE;
  const std::string u = unit ? unit : "LJ";
  double kb = 1.0, mvv2e = 1.0;  // src/common/unit_factor.cpp:23-37
  if (u == "REAL") {
    kb = 1.9872067 * std::pow(10.0, -3.0);
    mvv2e = 1.0 / (4.186 * std::pow(10.0, -4.0));
  } else if (u == "METAL") {
    kb = 8.617343e-5;
    mvv2e = 1.0364269e-4;
  } else if (u != "LJ") {
    data->err = "unsupported unit style: " + u;
    return RBMDX_E_CONFIG;
  }
  return guarded_d(data, [&] {
    rbmdx::gauss_velocities(*data->d, temperature, seed, kb, mvv2e);
  });
}

void rbmdx_data_destroy(rbmdx_data* data) { delete data; }

const char* rbmdx_data_last_error(const rbmdx_data* data) {
  return data ? data->err.c_str() : "null handle";
}

int rbmdx_data_info(const rbmdx_data* data, int64_t* natoms, int32_t* ntypes,
                    int32_t* maxspecial, int32_t* flags) {
  if (!data || !data->d) return RBMDX_E_STATE;
  const rbmdx::LammpsData& d = *data->d;
  if (natoms) *natoms = d.size();
  if (ntypes) *ntypes = d.ntypes;
  if (maxspecial) *maxspecial = data->maxs;
  if (flags)
    *flags = (d.has_pair_coeffs ? 1 : 0) | (d.has_velocities ? 2 : 0) |
             (d.bonds.empty() ? 0 : 4);
  return RBMDX_OK;
}

int rbmdx_data_get(const rbmdx_data* data, double* box_lo, double* box_len, int32_t* type,
                   int32_t* id, double* charge, double* pos, double* vel, int32_t* image,
                   double* mass, double* eps, double* sigma) {
  if (!data || !data->d) return RBMDX_E_STATE;
  const rbmdx::LammpsData& d = *data->d;
  const int64_t n = d.size();
  for (int k = 0; k < 3; ++k) {
    if (box_lo) box_lo[k] = d.lo[k];
    if (box_len) box_len[k] = d.L[k];
  }
  for (int64_t i = 0; i < n; ++i) {
    if (type) type[i] = d.type[i];
    if (id) id[i] = (int32_t)d.id[i];
    if (charge) charge[i] = d.q[i];
    if (pos) {
      pos[3 * i] = d.x[i];
      pos[3 * i + 1] = d.y[i];
      pos[3 * i + 2] = d.z[i];
    }
    if (vel) {
      vel[3 * i] = d.vx[i];
      vel[3 * i + 1] = d.vy[i];
      vel[3 * i + 2] = d.vz[i];
    }
    if (image) {
      image[3 * i] = d.ix[i];
      image[3 * i + 1] = d.iy[i];
      image[3 * i + 2] = d.iz[i];
    }
  }
  for (int t = 0; t < d.ntypes; ++t) {
    if (mass) mass[t] = d.mass[t];
    if (eps) eps[t] = d.eps[t];
    if (sigma) sigma[t] = d.sigma[t];
  }
  return RBMDX_OK;
}

int rbmdx_data_special(const rbmdx_data* data, int32_t* nspecial, int32_t* special) {
  if (!data || !data->d) return RBMDX_E_STATE;
  if (nspecial) std::memcpy(nspecial, data->nspecial.data(), sizeof(int32_t) * data->nspecial.size());
  if (special) std::memcpy(special, data->special.data(), sizeof(int32_t) * data->special.size());
  return RBMDX_OK;
}

}  // extern "C"
