// Output formats (rbmdx_write_*, include/rbmdx.h): RBMD's LAMMPS-style
// trajectory frame (src/output/src/TrajectoryOutput.cpp:433-500: columns
// "id type q x y z xu yu zu vx vy vz", atoms in id order, 1-based ids and
// types) and the force-sample CSV its accuracy tooling reads
// (benchmark/common/analyze_force_samples.py:30-45: "sample,id,x,y,z,fx,fy,fz").
// Host-only C++.  Numbers are written as the shortest decimal that reads back
// to the same double.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "rbmdx.h"

namespace {

void put_num(std::string& s, double v) {
  char buf[40];
  for (int p = 1; p <= 17; ++p) {
    std::snprintf(buf, sizeof buf, "%.*g", p, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  s += buf;
}

void put_int(std::string& s, long long v) {
  char buf[24];
  std::snprintf(buf, sizeof buf, "%lld", v);
  s += buf;
}

std::vector<int64_t> id_order(int64_t n, const int32_t* id) {
  std::vector<int64_t> o(n);
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](int64_t a, int64_t b) { return id[a] < id[b]; });
  return o;
}

bool write_all(const char* path, int32_t append, const std::string& s) {
  FILE* f = std::fopen(path, append ? "ab" : "wb");
  if (!f) return false;
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  return (std::fclose(f) == 0) && ok;
}

}  // namespace

extern "C" {

int rbmdx_write_dump(const char* path, int32_t append, int64_t timestep, int64_t n,
                     const double* box_lo, const double* box_len, const int32_t* id,
                     const int32_t* type, const double* charge, const double* pos,
                     const int32_t* image, const double* vel) {
  if (!path || n < 0 || !box_lo || !box_len || (n > 0 && (!id || !type || !pos)))
    return RBMDX_E_ARG;
  std::string s;
  s.reserve(64 + (size_t)n * 160);
  s += "ITEM: TIMESTEP\n";
  put_int(s, timestep);
  s += "\nITEM: NUMBER OF ATOMS\n";
  put_int(s, n);
  s += "\nITEM: BOX BOUNDS pp pp pp\n";
  for (int d = 0; d < 3; ++d) {
    put_num(s, box_lo[d]);
    s += ' ';
    put_num(s, box_lo[d] + box_len[d]);
    s += '\n';
  }
  s += "ITEM: ATOMS id type q x y z xu yu zu vx vy vz\n";
  for (int64_t i : id_order(n, id)) {
    put_int(s, (long long)id[i] + 1);
    s += ' ';
    put_int(s, (long long)type[i] + 1);
    s += ' ';
    put_num(s, charge ? charge[i] : 0.0);
    for (int d = 0; d < 3; ++d) {
      s += ' ';
      put_num(s, pos[3 * i + d]);
    }
    for (int d = 0; d < 3; ++d) {
      s += ' ';
      put_num(s, pos[3 * i + d] + (image ? (double)image[3 * i + d] * box_len[d] : 0.0));
    }
    for (int d = 0; d < 3; ++d) {
      s += ' ';
      put_num(s, vel ? vel[3 * i + d] : 0.0);
    }
    s += '\n';
  }
  return write_all(path, append, s) ? RBMDX_OK : RBMDX_E_ARG;
}

int rbmdx_write_force_samples(const char* path, int32_t append, int64_t sample,
                              int64_t n, const int32_t* id, const double* pos,
                              const double* force) {
  if (!path || n < 0 || (n > 0 && (!id || !pos || !force))) return RBMDX_E_ARG;
  std::string s;
  s.reserve(64 + (size_t)n * 140);
  if (!append) s += "sample,id,x,y,z,fx,fy,fz\n";
  for (int64_t i : id_order(n, id)) {
    put_int(s, sample);
    s += ',';
    put_int(s, (long long)id[i] + 1);
    for (int d = 0; d < 3; ++d) {
      s += ',';
      put_num(s, pos[3 * i + d]);
    }
    for (int d = 0; d < 3; ++d) {
      s += ',';
      put_num(s, force[3 * i + d]);
    }
    s += '\n';
  }
  return write_all(path, append, s) ? RBMDX_OK : RBMDX_E_ARG;
}

}  // extern "C"
