// LAMMPS data reader and special-bond builder (see lammps_data.hpp).
//
// Follows what RBMD's reader does with a data file
// (src/reader/src/format/lammps/lammps_full_reader.cpp: the header counts and
// box, Masses / Pair Coeffs / Atoms / Velocities / Bonds / Angles / Dihedrals
// sections, ParseAtom :1496-1563 with optional image flags and 1-based ids,
// types and molecules made 0-based; ApplyReplicate :899-1000 for systems
// without bonds) and how it builds the special-bond lists
// (src/reader/src/mpi/special_bond_builder.cpp: 1-2 partners from the bonds
// :325-366, 1-3 candidates through a common partner :370-446 kept only when
// an angle or dihedral supports them :598-700, 1-4 candidates :450-545 kept
// only when a dihedral supports them :737-825, combined per atom with
// duplicates dropped in 1-2, 1-3, 1-4 order :870-930).
#include "lammps_data.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>

namespace rbmdx {
namespace {

std::string strip(const std::string& s) {
  std::string t = s.substr(0, s.find('#'));
  const auto a = t.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  const auto b = t.find_last_not_of(" \t\r\n");
  return t.substr(a, b - a + 1);
}

bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

[[noreturn]] void fail(const std::string& path, long line, const std::string& what) {
  throw std::runtime_error(path + ":" + std::to_string(line) + ": " + what);
}

const char* kSections[] = {"Masses",          "Pair Coeffs",     "PairIJ Coeffs",
                           "Bond Coeffs",     "Angle Coeffs",    "Dihedral Coeffs",
                           "Improper Coeffs", "BondBond Coeffs", "BondAngle Coeffs",
                           "MiddleBondTorsion Coeffs",           "EndBondTorsion Coeffs",
                           "AngleTorsion Coeffs",                "AngleAngleTorsion Coeffs",
                           "BondBond13 Coeffs",                  "AngleAngle Coeffs",
                           "Atoms",           "Velocities",      "Bonds",
                           "Angles",          "Dihedrals",       "Impropers"};

bool section_name(const std::string& content, std::string& name) {
  for (const char* s : kSections)
    if (content == s) {
      name = s;
      return true;
    }
  return false;
}

}  // namespace

LammpsData read_lammps_data(const std::string& path, const std::string& atom_style) {
  if (atom_style != "atomic" && atom_style != "charge" && atom_style != "full")
    throw std::runtime_error("atom_style must be atomic, charge or full (got " +
                             atom_style + ")");
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open data file " + path);
  LammpsData d;
  d.atom_style = atom_style;
  std::vector<std::string> lines;
  for (std::string s; std::getline(f, s);) lines.push_back(s);
  int64_t natoms = -1, nbonds = 0, nangles = 0, ndihedrals = 0, nimpropers = 0;
  int nbondtypes = 0, nangletypes = 0, ndihtypes = 0, nimptypes = 0;
  bool have_box[3] = {false, false, false};
  size_t ln = 1;  // the first line is the title
  std::string sec;
  for (; ln < lines.size(); ++ln) {
    const std::string c = strip(lines[ln]);
    if (c.empty()) continue;
    if (section_name(c, sec)) break;
    std::istringstream is(c);
    if (ends_with(c, "xlo xhi") || ends_with(c, "ylo yhi") || ends_with(c, "zlo zhi")) {
      double a, b;
      if (!(is >> a >> b)) fail(path, (long)ln + 1, "bad box line");
      const int k = ends_with(c, "xlo xhi") ? 0 : ends_with(c, "ylo yhi") ? 1 : 2;
      d.lo[k] = a;
      d.L[k] = b - a;
      have_box[k] = true;
      continue;
    }
    if (ends_with(c, "xy xz yz")) {
      double a, b, e;
      if (!(is >> a >> b >> e)) fail(path, (long)ln + 1, "bad tilt line");
      if (a != 0.0 || b != 0.0 || e != 0.0)
        fail(path, (long)ln + 1, "triclinic boxes are not supported");
      continue;
    }
    long long v;
    if (!(is >> v)) fail(path, (long)ln + 1, "unrecognised header line: " + c);
    std::string rest;
    std::getline(is, rest);
    rest = strip(rest);
    if (rest == "atoms") natoms = v;
    else if (rest == "atom types") d.ntypes = (int)v;
    else if (rest == "bonds") nbonds = v;
    else if (rest == "angles") nangles = v;
    else if (rest == "dihedrals") ndihedrals = v;
    else if (rest == "impropers") nimpropers = v;
    else if (rest == "bond types") nbondtypes = (int)v;
    else if (rest == "angle types") nangletypes = (int)v;
    else if (rest == "dihedral types") ndihtypes = (int)v;
    else if (rest == "improper types") nimptypes = (int)v;
    // other header counts (extra bond per atom, ellipsoids, ...) are ignored
  }
  if (natoms <= 0) throw std::runtime_error(path + ": the header gives no atom count");
  if (d.ntypes <= 0) throw std::runtime_error(path + ": the header gives no atom types");
  for (int k = 0; k < 3; ++k)
    if (!have_box[k] || !(d.L[k] > 0))
      throw std::runtime_error(path + ": missing or empty box dimension");
  d.mass.assign(d.ntypes, 0.0);
  d.eps.assign(d.ntypes, 0.0);
  d.sigma.assign(d.ntypes, 0.0);
  std::unordered_map<int64_t, int64_t> index_of;
  bool seen_atoms = false;
  while (ln < lines.size()) {
    // lines[ln] holds the section header `sec`
    int64_t count = 0;
    if (sec == "Masses" || sec == "Pair Coeffs") count = d.ntypes;
    else if (sec == "PairIJ Coeffs") count = (int64_t)d.ntypes * (d.ntypes + 1) / 2;
    else if (sec == "Atoms" || sec == "Velocities") count = natoms;
    else if (sec == "Bonds") count = nbonds;
    else if (sec == "Angles") count = nangles;
    else if (sec == "Dihedrals") count = ndihedrals;
    else if (sec == "Impropers") count = nimpropers;
    else if (sec == "Bond Coeffs" || sec == "BondBond Coeffs" || sec == "BondAngle Coeffs")
      count = sec == "Bond Coeffs" ? nbondtypes : nangletypes;
    else if (sec == "Angle Coeffs") count = nangletypes;
    else if (sec == "Improper Coeffs" || sec == "AngleAngle Coeffs") count = nimptypes;
    else count = ndihtypes;  // the dihedral class-2 coefficient sections
    if (sec == "Atoms") {
      // an explicit style comment must agree with the configured style
      const std::string raw = lines[ln];
      const auto h = raw.find('#');
      if (h != std::string::npos) {
        const std::string style = strip(raw.substr(h + 1));
        if (!style.empty() && style != atom_style)
          fail(path, (long)ln + 1, "Atoms section is '" + style + "' but atom_style is '" +
                                       atom_style + "'");
      }
    }
    ++ln;
    int64_t got = 0;
    for (; ln < lines.size() && got < count; ++ln) {
      const std::string c = strip(lines[ln]);
      if (c.empty()) continue;
      std::string other;
      if (section_name(c, other)) fail(path, (long)ln + 1, sec + " section is short");
      std::istringstream is(c);
      const long lno = (long)ln + 1;
      if (sec == "Masses") {
        int t;
        double m;
        if (!(is >> t >> m) || t < 1 || t > d.ntypes) fail(path, lno, "bad Masses line");
        d.mass[t - 1] = m;
      } else if (sec == "Pair Coeffs") {
        int t;
        double e, s;
        if (!(is >> t >> e >> s) || t < 1 || t > d.ntypes)
          fail(path, lno, "bad Pair Coeffs line");
        d.eps[t - 1] = e;
        d.sigma[t - 1] = s;
        d.has_pair_coeffs = true;
      } else if (sec == "Atoms") {
        long long id = 0, mol = 0;
        int type = 0;
        double q = 0, x, y, z;
        if (atom_style == "atomic") is >> id >> type >> x >> y >> z;
        else if (atom_style == "charge") is >> id >> type >> q >> x >> y >> z;
        else is >> id >> mol >> type >> q >> x >> y >> z;
        if (is.fail()) fail(path, lno, "bad Atoms line for style " + atom_style);
        int im[3] = {0, 0, 0};
        if (is >> im[0]) {
          if (!(is >> im[1] >> im[2])) fail(path, lno, "incomplete image flags");
        } else {
          d.image_flags_all = false;
        }
        if (id <= 0) fail(path, lno, "atom id must be positive");
        if (type <= 0 || type > d.ntypes) fail(path, lno, "atom type out of range");
        if (!index_of.emplace(id - 1, d.size()).second) fail(path, lno, "duplicate atom id");
        d.id.push_back(id - 1);
        d.type.push_back(type - 1);
        d.mol.push_back(atom_style == "full" && mol > 0 ? (int32_t)(mol - 1) : -1);
        d.q.push_back(q);
        d.x.push_back(x);
        d.y.push_back(y);
        d.z.push_back(z);
        d.ix.push_back(im[0]);
        d.iy.push_back(im[1]);
        d.iz.push_back(im[2]);
      } else if (sec == "Velocities") {
        long long id;
        double a, b, e;
        if (!(is >> id >> a >> b >> e)) fail(path, lno, "bad Velocities line");
        if (!seen_atoms) fail(path, lno, "Velocities before Atoms");
        auto it = index_of.find(id - 1);
        if (it == index_of.end()) fail(path, lno, "velocity of an unknown atom");
        d.vx[it->second] = a;
        d.vy[it->second] = b;
        d.vz[it->second] = e;
        d.has_velocities = true;
      } else if (sec == "Bonds") {
        long long bid, bt, a, b;
        if (!(is >> bid >> bt >> a >> b)) fail(path, lno, "bad Bonds line");
        d.bonds.push_back({a - 1, b - 1});
      } else if (sec == "Angles") {
        long long aid, at, a, b, e;
        if (!(is >> aid >> at >> a >> b >> e)) fail(path, lno, "bad Angles line");
        d.angles.push_back({a - 1, b - 1, e - 1});
      } else if (sec == "Dihedrals") {
        long long did, dt, a, b, e, g;
        if (!(is >> did >> dt >> a >> b >> e >> g)) fail(path, lno, "bad Dihedrals line");
        d.dihedrals.push_back({a - 1, b - 1, e - 1, g - 1});
      }
      // coefficient sections of bonded terms and impropers are not used here
      ++got;
    }
    if (got < count) throw std::runtime_error(path + ": " + sec + " section is short");
    if (sec == "Atoms") {
      seen_atoms = true;
      d.vx.assign(d.size(), 0.0);
      d.vy.assign(d.size(), 0.0);
      d.vz.assign(d.size(), 0.0);
    }
    // next section header
    for (; ln < lines.size(); ++ln) {
      const std::string c = strip(lines[ln]);
      if (c.empty()) continue;
      if (!section_name(c, sec)) fail(path, (long)ln + 1, "unexpected line: " + c);
      break;
    }
  }
  if (!seen_atoms) throw std::runtime_error(path + ": no Atoms section");
  if (d.size() != natoms) throw std::runtime_error(path + ": atom count differs from header");
  for (const auto& b : d.bonds)
    for (int64_t a : b)
      if (!index_of.count(a)) throw std::runtime_error(path + ": bond of an unknown atom");
  return d;
}

namespace {
// Replica tile of a topology partner (BondTargetReplica,
// src/reader/include/core/periodic_replica.h:10-38): the partner's tile
// is the anchor's tile shifted by the partner's nearest image in the seed
// box (no shift at exactly half a box).
int64_t partner_tile(const LammpsData& s, int64_t ia, int64_t it, int64_t anchor_tile,
                     const int n[3]) {
  const double from[3] = {s.x[ia], s.y[ia], s.z[ia]};
  const double to[3] = {s.x[it], s.y[it], s.z[it]};
  int64_t t[3];
  for (int d = 0; d < 3; ++d) {
    const int64_t cur = anchor_tile % n[d];
    anchor_tile /= n[d];
    const double ratio = (to[d] - from[d]) / s.L[d];
    if (!std::isfinite(ratio) || std::abs(ratio) > 1e15)
      throw std::runtime_error("invalid coordinate in periodic replication");
    const long long shift = ratio > 0.5    ? -(long long)std::ceil(ratio - 0.5)
                            : ratio < -0.5 ? -(long long)std::floor(ratio + 0.5)
                                           : 0;
    int64_t w = (cur + shift) % n[d];
    if (w < 0) w += n[d];
    t[d] = w;
  }
  return (t[2] * n[1] + t[1]) * n[0] + t[0];
}
}  // namespace

void replicate_lammps_data(LammpsData& d, int nx, int ny, int nz, bool bond_periodic) {
  if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("replicate factors must be >= 1");
  if (nx * ny * nz == 1) return;
  const bool topo = !d.bonds.empty() || !d.angles.empty() || !d.dihedrals.empty();
  if (!bond_periodic && (topo || d.atom_style == "full") && !d.image_flags_all)
    throw std::runtime_error("replicate of molecular data without bond_periodic needs "
                             "image flags (ix iy iz) on every Atoms line");
  const LammpsData s = d;
  int64_t span = 0;
  int32_t mspan = 0;
  for (int64_t v : s.id) span = std::max(span, v + 1);
  for (int32_t m : s.mol) mspan = std::max(mspan, m + 1);
  std::unordered_map<int64_t, int64_t> index_of;
  if (topo && bond_periodic)
    for (int64_t i = 0; i < s.size(); ++i) index_of.emplace(s.id[i], i);
  const int n[3] = {nx, ny, nz};
  for (int k = 0; k < 3; ++k) d.L[k] = s.L[k] * n[k];
  const int64_t ns = s.size();
  const size_t tot = (size_t)ns * nx * ny * nz;
  for (auto* v : {&d.q, &d.x, &d.y, &d.z, &d.vx, &d.vy, &d.vz}) {
    v->clear();
    v->reserve(tot);
  }
  d.id.clear();
  d.type.clear();
  d.mol.clear();
  d.ix.clear();
  d.iy.clear();
  d.iz.clear();
  d.bonds.clear();
  d.angles.clear();
  d.dihedrals.clear();
  auto remap = [](double v, double lo, double L, int32_t& img) {
    long long i = (long long)std::floor((v - lo) / L);
    double w = v - (double)i * L;
    while (w < lo) {
      w += L;
      --i;
    }
    while (w >= lo + L) {
      w -= L;
      ++i;
    }
    img = (int32_t)i;
    return w;
  };
  // partner ids of one replica (SameReplicaAtomId / BondPeriodicAtomId /
  // BondPeriodicChainEnd, lammps_full_reader.cpp:352-383,988-1040)
  auto same = [&](int64_t a, int64_t r) { return a + r * span; };
  auto idx = [&](int64_t a) {
    auto it = index_of.find(a);
    if (it == index_of.end()) throw std::runtime_error("topology names an unknown atom");
    return it->second;
  };
  auto partner = [&](int64_t anchor, int64_t target, int64_t r) {
    if (!bond_periodic) return same(target, r);
    return same(target, partner_tile(s, idx(anchor), idx(target), r, n));
  };
  for (int kz = 0; kz < nz; ++kz)
    for (int ky = 0; ky < ny; ++ky)
      for (int kx = 0; kx < nx; ++kx) {
        const int64_t r = ((int64_t)kz * ny + ky) * nx + kx;
        for (int64_t i = 0; i < ns; ++i) {
          d.id.push_back(s.id[i] + r * span);
          d.type.push_back(s.type[i]);
          d.mol.push_back(s.mol[i] >= 0 ? (int32_t)(s.mol[i] + r * mspan) : -1);
          d.q.push_back(s.q[i]);
          // BuildReplicatedAtom (lammps_full_reader.cpp:386-421): tile shift,
          // plus the image-flag unwrap when bonds do not pick partner tiles
          const double u[3] = {bond_periodic ? 0.0 : s.ix[i] * s.L[0],
                               bond_periodic ? 0.0 : s.iy[i] * s.L[1],
                               bond_periodic ? 0.0 : s.iz[i] * s.L[2]};
          int32_t im[3];
          d.x.push_back(remap(s.x[i] + u[0] + kx * s.L[0], d.lo[0], d.L[0], im[0]));
          d.y.push_back(remap(s.y[i] + u[1] + ky * s.L[1], d.lo[1], d.L[1], im[1]));
          d.z.push_back(remap(s.z[i] + u[2] + kz * s.L[2], d.lo[2], d.L[2], im[2]));
          d.ix.push_back(im[0]);
          d.iy.push_back(im[1]);
          d.iz.push_back(im[2]);
          d.vx.push_back(s.vx[i]);
          d.vy.push_back(s.vy[i]);
          d.vz.push_back(s.vz[i]);
        }
        for (const auto& b : s.bonds)
          d.bonds.push_back({same(b[0], r), partner(b[0], b[1], r)});
        for (const auto& a : s.angles)
          d.angles.push_back({partner(a[1], a[0], r), same(a[1], r), partner(a[1], a[2], r)});
        for (const auto& h : s.dihedrals) {
          const int64_t a3 = partner(h[1], h[2], r);
          const int64_t a4 = bond_periodic ? partner(h[2], h[3], (a3 - h[2]) / span)
                                           : same(h[3], r);
          d.dihedrals.push_back({partner(h[1], h[0], r), same(h[1], r), a3, a4});
        }
      }
}

// ApplyGaussianVelocitiesIfRequested and MixVelocitySeed
// (src/reader/src/format/lammps/lammps_full_reader.cpp:134-143,710-866): the
// seed mix is a SplitMix64 finaliser over seed + golden-ratio increment + a
// shifted id; the generator and the normal draws are the C++ standard
// library's (std::mt19937_64, std::normal_distribution<double>), as there.
namespace {
uint64_t velocity_seed(uint64_t seed, uint64_t id1) {
  uint64_t z = seed + 0x9e3779b97f4a7c15ULL + (id1 << 6U) + (id1 >> 2U);
  z = (z ^ (z >> 30U)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27U)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31U);
}
}  // namespace

void gauss_velocities(LammpsData& d, double temperature, int seed, double kb,
                      double mvv2e) {
  if (d.has_velocities) return;
  if (!(temperature > 0) || !std::isfinite(temperature))
    throw std::runtime_error("velocity_type GAUSS requires a temperature > 0");
  if (seed <= 0) throw std::runtime_error("velocity_type GAUSS requires a positive seed");
  const int64_t n = d.size();
  if (n < 2) throw std::runtime_error("velocity_type GAUSS needs at least 2 atoms");
  auto mass_of = [&](int32_t t) {
    const double m = (t >= 0 && t < (int32_t)d.mass.size()) ? d.mass[t] : 0.0;
    if (!(m > 0) || !std::isfinite(m))
      throw std::runtime_error("velocity_type GAUSS requires a positive mass for every atom type");
    return m;
  };
  d.vx.assign(n, 0.0);
  d.vy.assign(n, 0.0);
  d.vz.assign(n, 0.0);
  double mt = 0, p[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    const double m = mass_of(d.type[i]);
    std::mt19937_64 gen(velocity_seed((uint64_t)seed, (uint64_t)(d.id[i] + 1)));
    std::normal_distribution<double> normal(0.0, 1.0);
    const double s = 1.0 / std::sqrt(m);
    d.vx[i] = normal(gen) * s;
    d.vy[i] = normal(gen) * s;
    d.vz[i] = normal(gen) * s;
    mt += m;
    p[0] += m * d.vx[i];
    p[1] += m * d.vy[i];
    p[2] += m * d.vz[i];
  }
  const double b[3] = {p[0] / mt, p[1] / mt, p[2] / mt};
  double ts = 0;
  for (int64_t i = 0; i < n; ++i) {
    d.vx[i] -= b[0];
    d.vy[i] -= b[1];
    d.vz[i] -= b[2];
    ts += mass_of(d.type[i]) *
          (d.vx[i] * d.vx[i] + d.vy[i] * d.vy[i] + d.vz[i] * d.vz[i]);
  }
  ts *= mvv2e;
  if (!(ts > 0) || !std::isfinite(ts))
    throw std::runtime_error("velocity_type GAUSS failed to create finite kinetic energy");
  const double dof = (double)(3 * n - 3);
  const double scale = std::sqrt(temperature * dof * kb / ts);
  for (int64_t i = 0; i < n; ++i) {
    d.vx[i] *= scale;
    d.vy[i] *= scale;
    d.vz[i] *= scale;
  }
  d.has_velocities = true;
}

void build_special(const LammpsData& d, std::vector<int32_t>& nspecial,
                   std::vector<int32_t>& special, int& maxs) {
  using Adj = std::unordered_map<int64_t, std::vector<int64_t>>;
  auto dedup = [](std::vector<int64_t>& v, const std::unordered_set<int64_t>& drop) {
    std::unordered_set<int64_t> seen(drop);
    std::vector<int64_t> out;
    for (int64_t x : v)
      if (seen.insert(x).second) out.push_back(x);
    v.swap(out);
  };
  // 1-2: both ends of every bond, file order
  Adj a12;
  for (const auto& b : d.bonds) {
    a12[b[0]].push_back(b[1]);
    a12[b[1]].push_back(b[0]);
  }
  for (auto& kv : a12) dedup(kv.second, {});
  // 1-3: two partners of a common centre
  Adj a13;
  for (int64_t c : d.id) {
    auto it = a12.find(c);
    if (it == a12.end() || it->second.size() < 2) continue;
    const auto& nb = it->second;
    for (size_t j = 0; j < nb.size(); ++j)
      for (size_t k = 0; k < nb.size(); ++k)
        if (k != j) a13[nb[j]].push_back(nb[k]);
  }
  for (auto& kv : a13) dedup(kv.second, {kv.first});
  // kept only when an angle or a dihedral holds the pair
  if (d.angles.empty() && d.dihedrals.empty()) {
    a13.clear();
  } else {
    std::unordered_map<int64_t, std::unordered_set<int64_t>> keep;
    for (const auto& a : d.angles) {
      keep[a[0]].insert(a[2]);
      keep[a[2]].insert(a[0]);
    }
    for (const auto& q : d.dihedrals) {
      keep[q[0]].insert(q[2]);
      keep[q[2]].insert(q[0]);
      keep[q[1]].insert(q[3]);
      keep[q[3]].insert(q[1]);
    }
    for (auto& kv : a13) {
      const auto k = keep.find(kv.first);
      std::vector<int64_t> out;
      if (k != keep.end())
        for (int64_t w : kv.second)
          if (k->second.count(w)) out.push_back(w);
      kv.second.swap(out);
    }
  }
  // 1-4: a 1-3 partner of the centre joined to another partner of it
  Adj a14;
  for (int64_t c : d.id) {
    auto i13 = a13.find(c);
    auto i12 = a12.find(c);
    if (i13 == a13.end() || i12 == a12.end()) continue;
    for (int64_t w : i13->second)
      for (int64_t v : i12->second) a14[w].push_back(v);
  }
  for (auto& kv : a14) {
    std::unordered_set<int64_t> drop{kv.first};
    if (auto it = a12.find(kv.first); it != a12.end()) drop.insert(it->second.begin(), it->second.end());
    if (auto it = a13.find(kv.first); it != a13.end()) drop.insert(it->second.begin(), it->second.end());
    dedup(kv.second, drop);
  }
  if (d.dihedrals.empty()) {
    a14.clear();
  } else {
    std::unordered_map<int64_t, std::unordered_set<int64_t>> keep;
    for (const auto& q : d.dihedrals) {
      keep[q[0]].insert(q[3]);
      keep[q[3]].insert(q[0]);
    }
    for (auto& kv : a14) {
      const auto k = keep.find(kv.first);
      std::vector<int64_t> out;
      if (k != keep.end())
        for (int64_t w : kv.second)
          if (k->second.count(w)) out.push_back(w);
      kv.second.swap(out);
    }
  }
  // combined per atom: 1-2, then new 1-3, then new 1-4 partners
  const int64_t n = d.size();
  std::vector<std::vector<int64_t>> lists(n);
  nspecial.assign(3 * (size_t)n, 0);
  maxs = 1;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t g = d.id[i];
    std::unordered_set<int64_t> seen{g};
    auto& out = lists[i];
    auto add = [&](const Adj& a) {
      auto it = a.find(g);
      if (it != a.end())
        for (int64_t p : it->second)
          if (seen.insert(p).second) out.push_back(p);
    };
    add(a12);
    nspecial[3 * i] = (int32_t)out.size();
    add(a13);
    nspecial[3 * i + 1] = (int32_t)out.size();
    add(a14);
    nspecial[3 * i + 2] = (int32_t)out.size();
    maxs = std::max(maxs, (int)out.size());
  }
  special.assign((size_t)n * maxs, -1);
  for (int64_t i = 0; i < n; ++i)
    for (size_t k = 0; k < lists[i].size(); ++k)
      special[(size_t)i * maxs + k] = (int32_t)lists[i][k];
}

}  // namespace rbmdx
