"""Seeded synthetic systems for the benchmark configurations (no data files).

* ``lj_fcc``        plain LJ fluid: fcc lattice, uniform jitter, Maxwell velocities
                    (4,000 atoms = 10x10x10 fcc cells at rho* = 0.8442).
* ``electrolyte``   LJ electrolyte: 5 % +1, 5 % -1, 90 % neutral LJ solvent,
                    uniform positions with overlap rejection (random sequential
                    addition with a minimum distance d_min) at rho* = 0.8.
* ``ljsalt_config`` the RBMD JSON configuration used with ``electrolyte``.
* ``replica_set``  independent electrolytes of log-uniformly drawn sizes (the
                    replica batch's workload).
* ``spce_water``    rigid SPC/E water (REAL units: Angstrom, kcal/mol) on a cubic
                    lattice with random orientations, and its CVFF special-bond
                    topology (1-2: O-H, 1-3: H-H).
* ``cvff_water_config`` the RBMD JSON for ``spce_water`` (CVFF, RBL, RBE/RBSOG).

Atom ids are 0-based, as RBMD makes them when it reads a LAMMPS data file
(src/reader/src/format/lammps/lammps_full_reader.cpp:1543).
"""
import json
import math

import numpy as np


class System:
    def __init__(self, pos, box_lo, box_len, types, ids, charges, eps, sigma, vel=None,
                 mass=None, name=""):
        self.pos = np.asarray(pos, dtype=np.float64)
        self.box_lo = np.asarray(box_lo, dtype=np.float64)
        self.box_len = np.asarray(box_len, dtype=np.float64)
        self.types = np.asarray(types, dtype=np.int32)
        self.ids = np.asarray(ids, dtype=np.int32)
        self.charges = np.asarray(charges, dtype=np.float64)
        self.eps = np.asarray(eps, dtype=np.float64)
        self.sigma = np.asarray(sigma, dtype=np.float64)
        self.vel = None if vel is None else np.asarray(vel, dtype=np.float64)
        self.mass = np.ones(len(eps)) if mass is None else np.asarray(mass, dtype=np.float64)
        self.name = name

    @property
    def n(self):
        return len(self.types)


def lj_fcc(cells=10, rho=0.8442, jitter=0.05, temperature=1.0, seed=1):
    """fcc LJ fluid: 4 * cells^3 atoms, uniform +-jitter, Maxwell velocities at T*."""
    rng = np.random.default_rng(seed)
    a = (4.0 / rho) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(cells), np.arange(cells), np.arange(cells),
                             indexing="ij"), axis=-1).reshape(-1, 3)
    pos = ((g[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    L = cells * a
    pos = pos + rng.uniform(-jitter, jitter, size=pos.shape)
    pos = np.mod(pos, L)
    n = len(pos)
    vel = rng.normal(0.0, math.sqrt(temperature), size=(n, 3))
    vel -= vel.mean(axis=0)
    ke_t = (vel ** 2).sum() / (3 * n - 3)
    vel *= math.sqrt(temperature / ke_t)
    return System(pos, [0, 0, 0], [L, L, L], np.zeros(n), np.arange(n), np.zeros(n),
                  [1.0], [1.0], vel=vel, name=f"lj_fcc_{n}")


def _rsa(n, L, d_min, rng, batch=None):
    """Random sequential addition in a periodic cube: n points, pairwise >= d_min."""
    from scipy.spatial import cKDTree

    pts = np.zeros((0, 3))
    batch = batch or max(1000, n // 4)
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 10000:
            raise RuntimeError("RSA did not converge; lower d_min")
        need = n - len(pts)
        cand = rng.uniform(0.0, L, size=(max(batch, 2 * need), 3))
        cand = np.minimum(cand, np.nextafter(L, 0))
        if len(pts):
            tree = cKDTree(pts, boxsize=L)
            dist, _ = tree.query(cand, k=1, distance_upper_bound=d_min)
            cand = cand[~(dist < d_min)]
        if len(cand) == 0:
            continue
        # inside the batch keep the first of every conflicting pair (in order)
        ct = cKDTree(cand, boxsize=L)
        pairs = ct.query_pairs(d_min, output_type="ndarray")
        keep = np.ones(len(cand), dtype=bool)
        if len(pairs):
            keep[np.maximum(pairs[:, 0], pairs[:, 1])] = False
        cand = cand[keep][:need]
        pts = np.concatenate([pts, cand], axis=0)
    return pts


def electrolyte(n=200000, rho=0.8, frac_ion=0.05, d_min=0.85, seed=2, temperature=1.0):
    """LJ electrolyte: n*frac_ion cations (+1), as many anions (-1), rest neutral.

    Types: 0 = cation, 1 = anion, 2 = solvent; eps = sigma = 1 for all.
    """
    rng = np.random.default_rng(seed)
    L = (n / rho) ** (1.0 / 3.0)
    pos = _rsa(n, L, d_min, rng)
    nion = int(round(n * frac_ion))
    types = np.full(n, 2, dtype=np.int32)
    perm = rng.permutation(n)
    types[perm[:nion]] = 0
    types[perm[nion:2 * nion]] = 1
    q = np.where(types == 0, 1.0, np.where(types == 1, -1.0, 0.0))
    vel = rng.normal(0.0, math.sqrt(temperature), size=(n, 3))
    vel -= vel.mean(axis=0)
    return System(pos, [0, 0, 0], [L, L, L], types, np.arange(n), q, [1.0, 1.0, 1.0],
                  [1.0, 1.0, 1.0], vel=vel, name=f"electrolyte_{n}")


def replica_set(nrep=256, n_min=500, n_max=8000, seed=4, frac_ion=0.05):
    """Independent LJ electrolytes (composition and density of ``electrolyte``)
    with atom counts N_r = round(exp(U[ln n_min, ln n_max])) drawn from the
    seed, each box placed with its own seed."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), nrep))).astype(int)
    return [electrolyte(int(n), frac_ion=frac_ion, seed=int(1000 * seed + r))
            for r, n in enumerate(sizes)]


def replicate(s, reps):
    """The system tiled reps = (nx, ny, nz) times along the box edges, ids
    offset per copy (the read_data "replicate" option of RBMD's input,
    docs/json-configuration.en.md): periodic images of a valid box stay valid."""
    reps = np.asarray(reps, dtype=int)
    shifts = np.stack(np.meshgrid(*[np.arange(r) for r in reps], indexing="ij"),
                      -1).reshape(-1, 3)
    pos = np.concatenate([s.pos + sh * s.box_len for sh in shifts])
    n = s.n
    k = len(shifts)
    vel = None if s.vel is None else np.tile(s.vel, (k, 1))
    return System(pos, s.box_lo, s.box_len * reps, np.tile(s.types, k),
                  np.concatenate([s.ids + c * n for c in range(k)]), np.tile(s.charges, k),
                  s.eps, s.sigma, vel=vel, mass=s.mass, name=f"{s.name}_x{k}")


def ljsalt_config(neighbor="RBL", coulomb="RBE", cut_off=3.5, lj_cut_off=2.5, r_core=2.5,
                  sample_num=30, skin=0.3, alpha=0.09, P=200, seed=20261202,
                  precision="fp32", kmax=None):
    """RBMD JSON for the LJ electrolyte: LJ cutoff 2.5, Coulomb (real-space) cutoff
    3.5, Ewald alpha = g^2 with g = 0.3 / sigma, P = 200 RBE samples."""
    nb = {"type": neighbor, "cut_off": cut_off, "skin": skin}
    if lj_cut_off is not None:
        nb["lj_cut_off"] = lj_cut_off
    if neighbor == "RBL":
        nb.update({"r_core": r_core, "neighbor_sample_num": sample_num,
                   "energy_rbl_flag": "no"})
    cl = {"type": coulomb, "alpha": alpha, "random_seed": seed}
    if coulomb == "RBE":
        cl.update({"coulomb_sample_num": P, "energy_rbe_flag": "no"})
    if kmax is not None:
        cl["kmax"] = list(kmax)
    return {
        "init_configuration": {"read_data": {"unit": "LJ", "atom_style": "charge"}},
        "hyper_parameters": {
            "force_field": {"type": "LJ/CUT/COUL/LONG"},
            "neighbor": nb,
            "coulomb": cl,
            "precision": precision,
        },
    }


def lj_config(neighbor="Verlet-List", cut_off=2.5, skin=0.3, precision="fp64"):
    """RBMD JSON for the plain LJ fluid (LJ/CUT, full list)."""
    return {
        "init_configuration": {"read_data": {"unit": "LJ", "atom_style": "atomic"}},
        "hyper_parameters": {
            "force_field": {"type": "LJ/CUT"},
            "neighbor": {"type": neighbor, "cut_off": cut_off, "skin": skin},
            "precision": precision,
        },
    }


def _random_rotations(n, rng):
    """n uniformly distributed rotation matrices (unit quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)


class Water(System):
    """A System with CVFF special-bond topology: nspecial (N, 3) cumulative
    1-2/1-3/1-4 counts and special (N, maxspecial) partner ids."""

    def __init__(self, *a, nspecial=None, special=None, **kw):
        super().__init__(*a, **kw)
        self.nspecial = np.asarray(nspecial, dtype=np.int32)
        self.special = np.asarray(special, dtype=np.int32)


def spce_water(n_mol=24327, density=33.4, seed=3):
    """Rigid SPC/E water: O-H 1.0 A, H-O-H 109.47 deg, q_O = -0.8476,
    q_H = +0.4238; O-O eps 0.15535 kcal/mol, sigma 3.166 A; H carries no LJ.
    n_mol molecules on the first sites of a cubic m^3 lattice (m the smallest
    integer with m^3 >= n_mol) in a box of density molecules per nm^3, each
    with a random orientation.  Atoms O, H1, H2 per molecule, ids 0-based.
    Special bonds: O -> (H1, H2) as 1-2; H1 -> O as 1-2 and H2 as 1-3."""
    rng = np.random.default_rng(seed)
    L = (n_mol / (density * 1e-3)) ** (1.0 / 3.0)
    m = int(math.ceil(n_mol ** (1.0 / 3.0) - 1e-9))
    a = L / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"),
                 axis=-1).reshape(-1, 3)[:n_mol]
    o = (g + 0.5) * a
    half = math.radians(109.47) / 2.0
    ref = np.array([[0.0, 0.0, 0.0],
                    [math.sin(half), 0.0, math.cos(half)],
                    [-math.sin(half), 0.0, math.cos(half)]])
    rot = _random_rotations(n_mol, rng)
    pos = o[:, None, :] + np.einsum("nij,kj->nki", rot, ref)
    pos = np.mod(pos.reshape(-1, 3), L)
    n = 3 * n_mol
    types = np.tile(np.array([0, 1, 1], dtype=np.int32), n_mol)
    q = np.tile(np.array([-0.8476, 0.4238, 0.4238]), n_mol)
    ids = np.arange(n, dtype=np.int32)
    base = 3 * np.arange(n_mol, dtype=np.int32)
    nspecial = np.tile(np.array([[2, 2, 2], [1, 2, 2], [1, 2, 2]], dtype=np.int32),
                       (n_mol, 1))
    special = np.empty((n, 2), dtype=np.int32)
    special[0::3] = np.stack([base + 1, base + 2], axis=1)
    special[1::3] = np.stack([base, base + 2], axis=1)
    special[2::3] = np.stack([base, base + 1], axis=1)
    return Water(pos, [0, 0, 0], [L, L, L], types, ids, q, [0.15535, 0.0], [3.166, 0.0],
                 mass=[15.9994, 1.008], name=f"spce_water_{n_mol}", nspecial=nspecial,
                 special=special)


def cvff_water_config(neighbor="RBL", coulomb="RBE", cut_off=10.0, r_core=6.5,
                      sample_num=100, skin=2.0, alpha=None, accuracy=1e-4, P=300,
                      rbsog_level=1, special_bonds=(0.0, 0.0, 0.0), seed=2026102301,
                      precision="fp32"):
    """RBMD JSON for SPC/E water: CVFF in REAL units, rc 10 A, RBL with
    r_core 6.5 A and 100 samples, skin 2 A, special_bonds [0, 0, 0] (the
    reference's benchmark/systems/water.json keys)."""
    nb = {"type": neighbor, "cut_off": cut_off, "skin": skin}
    if neighbor == "RBL":
        nb.update({"r_core": r_core, "neighbor_sample_num": sample_num,
                   "energy_rbl_flag": "no"})
    cl = {"type": coulomb, "random_seed": seed}
    if alpha is not None:
        cl["alpha"] = alpha
    elif coulomb != "RBSOG" and neighbor != "VERLET-SOG":
        cl["accuracy"] = accuracy
    if coulomb == "RBE":
        cl.update({"coulomb_sample_num": P, "energy_rbe_flag": "no"})
    if coulomb == "RBSOG" or neighbor == "VERLET-SOG":
        cl.update({"coulomb_sample_num": P, "rbsog_level": rbsog_level})
    return {
        "init_configuration": {"read_data": {"unit": "REAL", "atom_style": "full"}},
        "hyper_parameters": {
            "force_field": {"type": "CVFF", "mixing_rule": "arithmetic"},
            "neighbor": nb,
            "coulomb": cl,
            "extend": {"special_bonds": list(special_bonds)},
            "precision": precision,
        },
    }


def write_lammps_data(s, path, atom_style="full", bonds=(), angles=(), dihedrals=(),
                      image=None, title="written by rbmd_amd.systems"):
    """Write a System as a LAMMPS data file (ids and types 1-based in the
    file; bonds / angles / dihedrals as tuples of 0-based atom ids, one type
    each).  Positions with %.17g, so a reader recovers them exactly."""
    n, nt = len(s.types), len(s.eps)
    out = [title, "", f"{n} atoms", f"{nt} atom types"]
    for name, lst in (("bonds", bonds), ("angles", angles), ("dihedrals", dihedrals)):
        if len(lst):
            out += [f"{len(lst)} {name}", f"1 {name[:-1]} types"]
    for d, ax in enumerate("xyz"):
        out.append(f"{float(s.box_lo[d])!r} {float(s.box_lo[d] + s.box_len[d])!r} {ax}lo {ax}hi")
    out += ["", "Masses", ""] + [f"{t + 1} {float(s.mass[t])!r}" for t in range(nt)]
    out += ["", "Pair Coeffs # lj/cut/coul/long", ""]
    out += [f"{t + 1} {float(s.eps[t])!r} {float(s.sigma[t])!r}" for t in range(nt)]
    out += ["", f"Atoms # {atom_style}", ""]
    for i in range(n):
        x, y, z = ("%.17g" % v for v in s.pos[i])
        im = "" if image is None else " %d %d %d" % tuple(image[i])
        idx, t = int(s.ids[i]) + 1, int(s.types[i]) + 1
        if atom_style == "atomic":
            out.append(f"{idx} {t} {x} {y} {z}{im}")
        elif atom_style == "charge":
            out.append(f"{idx} {t} {float(s.charges[i])!r} {x} {y} {z}{im}")
        else:
            out.append(f"{idx} {i // 3 + 1} {t} {float(s.charges[i])!r} {x} {y} {z}{im}")
    if s.vel is not None:
        out += ["", "Velocities", ""]
        out += [f"{int(s.ids[i]) + 1} " + " ".join(repr(float(v)) for v in s.vel[i])
                for i in range(n)]
    for name, lst in (("Bonds", bonds), ("Angles", angles), ("Dihedrals", dihedrals)):
        if len(lst):
            out += ["", name, ""]
            out += [f"{k + 1} 1 " + " ".join(str(int(a) + 1) for a in tup)
                    for k, tup in enumerate(lst)]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def water_topology(n_mol):
    """Bonds (O-H1, O-H2) and angles (H1-O-H2) of n_mol molecules laid out
    O, H1, H2 (0-based atom ids)."""
    base = 3 * np.arange(n_mol)
    bonds = [(b, b + 1) for b in base] + [(b, b + 2) for b in base]
    angles = [(b + 1, b, b + 2) for b in base]
    return bonds, angles


def dumps(cfg):
    return json.dumps(cfg)
