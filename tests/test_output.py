"""Output formats on the CPU: the LAMMPS-style trajectory frame
(rbmdx_write_dump) and the force-sample CSV (rbmdx_write_force_samples).

Checks: the frame's header items and column names as RBMD's trajectory output
writes them; atoms in id order with 1-based ids and types; unwrapped
coordinates x + image L; every number reads back to the same double; frames
and samples append; the CSV header and ordered sample indices the accuracy
analysis requires.
"""
import csv

import numpy as np
import pytest

import rbmd_amd
from rbmd_amd import RbmdxError


def _frame(rng, n=40):
    L = np.array([10.0, 12.5, 9.0])
    lo = np.array([-1.0, 0.0, 2.0])
    ids = rng.permutation(n).astype(np.int32)
    types = rng.integers(0, 3, n).astype(np.int32)
    pos = lo + rng.uniform(0, 1, (n, 3)) * L
    img = rng.integers(-2, 3, (n, 3)).astype(np.int32)
    vel = rng.normal(0, 1, (n, 3))
    q = rng.normal(0, 1, n)
    return lo, L, ids, types, pos, img, vel, q


def _read_frames(path):
    lines = path.read_text().splitlines()
    frames, k = [], 0
    while k < len(lines):
        assert lines[k] == "ITEM: TIMESTEP"
        step = int(lines[k + 1])
        assert lines[k + 2] == "ITEM: NUMBER OF ATOMS"
        n = int(lines[k + 3])
        assert lines[k + 4] == "ITEM: BOX BOUNDS pp pp pp"
        bounds = np.array([[float(v) for v in lines[k + 5 + d].split()] for d in range(3)])
        assert lines[k + 8] == "ITEM: ATOMS id type q x y z xu yu zu vx vy vz"
        rows = np.array([[float(v) for v in ln.split()] for ln in lines[k + 9:k + 9 + n]])
        frames.append((step, bounds, rows))
        k += 9 + n
    return frames


def test_dump_frame(tmp_path):
    rng = np.random.default_rng(7)
    lo, L, ids, types, pos, img, vel, q = _frame(rng)
    p = tmp_path / "rbmd.trj"
    rbmd_amd.write_dump(p, 0, lo, L, ids, types, pos, charges=q, image=img, vel=vel)
    rbmd_amd.write_dump(p, 10, lo, L, ids, types, pos * 0.5, charges=q, append=True)
    (s0, b0, r0), (s1, b1, r1) = _read_frames(p)
    assert (s0, s1) == (0, 10)
    np.testing.assert_array_equal(b0, np.stack([lo, lo + L], 1))
    o = np.argsort(ids)
    np.testing.assert_array_equal(r0[:, 0], ids[o] + 1)
    np.testing.assert_array_equal(r0[:, 1], types[o] + 1)
    np.testing.assert_array_equal(r0[:, 2], q[o])
    np.testing.assert_array_equal(r0[:, 3:6], pos[o])
    np.testing.assert_array_equal(r0[:, 6:9], pos[o] + img[o] * L)
    np.testing.assert_array_equal(r0[:, 9:12], vel[o])
    # no image flags: xu = x; no velocities: zeros
    np.testing.assert_array_equal(r1[:, 6:9], r1[:, 3:6])
    np.testing.assert_array_equal(r1[:, 9:12], 0.0)


def test_force_samples_csv(tmp_path):
    rng = np.random.default_rng(8)
    _, _, ids, _, pos, _, vel, _ = _frame(rng, 25)
    p = tmp_path / "forces.csv"
    rbmd_amd.write_force_samples(p, 0, ids, pos, vel)
    rbmd_amd.write_force_samples(p, 1, ids, pos + 1.0, -vel, append=True)
    with open(p, newline="") as f:
        rows = list(csv.DictReader(f))
    assert list(rows[0].keys()) == ["sample", "id", "x", "y", "z", "fx", "fy", "fz"]
    assert [int(r["sample"]) for r in rows] == [0] * 25 + [1] * 25
    o = np.argsort(ids)
    got = np.array([[float(r[k]) for k in ("x", "y", "z", "fx", "fy", "fz")]
                    for r in rows[:25]])
    np.testing.assert_array_equal([int(r["id"]) for r in rows[:25]], ids[o] + 1)
    np.testing.assert_array_equal(got[:, :3], pos[o])
    np.testing.assert_array_equal(got[:, 3:], vel[o])


def test_write_errors(tmp_path):
    with pytest.raises(RbmdxError):
        rbmd_amd.write_force_samples(tmp_path / "missing" / "f.csv", 0, [0], [[0, 0, 0]],
                                     [[1, 2, 3]])
