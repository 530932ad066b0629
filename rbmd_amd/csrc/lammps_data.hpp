// LAMMPS data files as RBMD reads them (atom styles atomic, charge, full):
// the system arrays the force path takes, and the special-bond topology
// built from the bonds, angles and dihedrals.  Host-only C++.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace rbmdx {

struct LammpsData {
  std::string atom_style;           // "atomic" | "charge" | "full"
  double lo[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  int ntypes = 0;
  std::vector<double> mass;         // [ntypes], 0 when absent
  std::vector<double> eps, sigma;   // [ntypes] from Pair Coeffs, 0 when absent
  bool has_pair_coeffs = false;
  // atoms in file order; ids, types and molecules 0-based (molecule -1: none)
  std::vector<int64_t> id;
  std::vector<int32_t> type, mol;
  std::vector<double> q, x, y, z;
  std::vector<int32_t> ix, iy, iz;  // image flags (0 when absent)
  std::vector<double> vx, vy, vz;   // Velocities section (0 when absent)
  bool has_velocities = false;
  bool image_flags_all = true;      // every Atoms line carried ix iy iz
  // topology, atom ids 0-based
  std::vector<std::array<int64_t, 2>> bonds;
  std::vector<std::array<int64_t, 3>> angles;
  std::vector<std::array<int64_t, 4>> dihedrals;
  int64_t size() const { return (int64_t)type.size(); }
};

// Parse a data file (throws std::runtime_error naming the file and line).
LammpsData read_lammps_data(const std::string& path, const std::string& atom_style);

// Tile an orthogonal box nx x ny x nz times (RBMD's ApplyReplicate,
// src/reader/src/format/lammps/lammps_full_reader.cpp:899-1050): copy
// r = (iz ny + iy) nx + ix of atom id a gets id a + r * span (span = largest
// id + 1), molecules likewise; positions shifted by the tile and folded into
// the new box (image flags updated).  Bonds, angles and dihedrals are copied
// per tile; with bond_periodic (RBMD's default) each partner is taken from
// the tile its nearest seed image lies in, so bonds that crossed the seed
// box's edge stay short; without it positions are first unwrapped by their
// image flags (required on every atom) and partners stay in their tile.
void replicate_lammps_data(LammpsData& d, int nx, int ny, int nz,
                           bool bond_periodic = true);

// GAUSS velocities as RBMD's reader creates them when the configuration asks
// for velocity_type GAUSS and the file has no Velocities section: per atom a
// 64-bit Mersenne Twister seeded from (seed, id + 1), three standard normal
// draws scaled by 1 / sqrt(m), the centre-of-mass velocity removed, then a
// rescale to `temperature` with 3 N - 3 degrees of freedom.  kb and mvv2e are
// the unit style's constants.  Throws on a non-positive temperature or seed,
// or a type without a positive mass.  No-op when the file had velocities.
void gauss_velocities(LammpsData& d, double temperature, int seed, double kb,
                      double mvv2e);

// Special-bond topology (1-2, 1-3, 1-4 partners) as RBMD's reader builds
// it: cumulative counts nspecial[3 i + k] and partner ids special[i * maxs
// + k] (global ids as in `id`), in caller (file) order; maxs >= 1.
void build_special(const LammpsData& d, std::vector<int32_t>& nspecial,
                   std::vector<int32_t>& special, int& maxs);

}  // namespace rbmdx
