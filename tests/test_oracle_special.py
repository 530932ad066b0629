"""Pins of the oracle's CVFF special-bond weights (oracle/oracle.cpp,
special_weight and pair_force_f32pos; reference cvff_op.hip.cu:51-110,
:572-599).  The reference's own tests hold no vectors for this path; the pins
are identities that a wrong sign, weight category or missing term breaks:

* all weights 1 reproduce the plain pair force bit for bit;
* two atoms excluded with w = 0 feel exactly the erf part of Coulomb,
  qqr2e q_i q_j [erf(g r) - (2 g r / sqrt(pi)) e^{-g^2 r^2}] / r^3 (a closed
  form written with erf, where the oracle subtracts the full 1/r^3 force from
  the erfc-screened one), energy -qqr2e q_i q_j erf(g r) / r, and no LJ;
* the force is affine in the weights, and each category (1-2, 1-3, 1-4)
  picks its own weight: F(w12, w13, w14) - F(1, 1, 1) is the sum of the
  single-category differences scaled by (1 - w);
* the counts are clamped as LoadSpecialCounts clamps them;
* force = -gradient of the energy with special weights (finite differences).
"""
import math

import numpy as np
import pytest

import oracle

QQR2E = 332.06371


def _params(rc, alpha, rbl=0, coulomb=1, r_core=0.0, freq=1, epoch=0):
    return oracle.pair_params(r_core, rc, rc, alpha, QQR2E, freq, epoch, rbl, coulomb, 0)


def _run(pos, q, types, gid, eps, sig, L, params, special=None):
    return oracle.pair_force(pos.astype(np.float32), q.astype(np.float32), types, gid,
                             np.asarray(eps), np.asarray(sig), np.asarray(L), params,
                             prec=0, special=special)


def _water_cluster(seed=0, n_mol=6, L=40.0):
    rng = np.random.default_rng(seed)
    half = math.radians(109.47) / 2
    ref = np.array([[0, 0, 0], [math.sin(half), 0, math.cos(half)],
                    [-math.sin(half), 0, math.cos(half)]])
    pos = []
    for k in range(n_mol):
        c = rng.uniform(8, L - 8, size=3)
        a = rng.normal(size=(3, 3))
        qm, _ = np.linalg.qr(a)
        pos.append(c + ref @ qm.T)
    pos = np.concatenate(pos)
    n = len(pos)
    types = np.tile([0, 1, 1], n_mol).astype(np.int32)
    q = np.tile([-0.8476, 0.4238, 0.4238], n_mol)
    gid = np.arange(n, dtype=np.int32)
    base = 3 * np.arange(n_mol)
    ns = np.tile([[2, 2, 2], [1, 2, 2], [1, 2, 2]], (n_mol, 1)).astype(np.int32)
    sp = np.empty((n, 2), dtype=np.int32)
    sp[0::3] = np.stack([base + 1, base + 2], 1)
    sp[1::3] = np.stack([base, base + 2], 1)
    sp[2::3] = np.stack([base, base + 1], 1)
    return pos, q, types, gid, ns, sp, [L, L, L]


def test_unit_weights_equal_plain_pair_force():
    pos, q, types, gid, ns, sp, L = _water_cluster()
    p = _params(9.0, 0.08)
    eps, sig = [0.15535, 0.0], [3.166, 0.0]
    plain = _run(pos, q, types, gid, eps, sig, L, p)
    unit = _run(pos, q, types, gid, eps, sig, L, p, special=(ns, sp, (1.0, 1.0, 1.0)))
    for a, b in zip(plain, unit):
        np.testing.assert_array_equal(a, b)


def test_excluded_pair_feels_only_the_erf_part():
    L = [50.0, 50.0, 50.0]
    r, alpha, qi, qj = 1.7, 0.09, 0.8, -0.6
    pos = np.array([[20.0, 20.0, 20.0], [20.0 + r, 20.0, 20.0]])
    q = np.array([qi, qj])
    types = np.array([0, 0], dtype=np.int32)
    gid = np.array([0, 1], dtype=np.int32)
    ns = np.array([[1, 1, 1], [1, 1, 1]], dtype=np.int32)
    sp = np.array([[1], [0]], dtype=np.int32)
    eps, sig = [0.3], [2.0]  # LJ present: w = 0 must remove it
    p = _params(9.0, alpha)
    f, en, vir, _ = _run(pos, q, types, gid, eps, sig, L, p, special=(ns, sp, (0.0, 1.0, 1.0)))
    r32 = float(np.float32(20.0 + r) - np.float32(20.0))
    g = math.sqrt(alpha)
    scal = QQR2E * float(np.float32(qi)) * float(np.float32(qj)) * (
        math.erf(g * r32) - 2 * g * r32 / math.sqrt(math.pi) * math.exp(-alpha * r32 * r32)
    ) / r32 ** 3
    assert f[0, 0] == pytest.approx(scal * r32, rel=1e-12)
    assert f[1, 0] == pytest.approx(-scal * r32, rel=1e-12)
    assert en[0] == 0.0
    e_closed = -QQR2E * float(np.float32(qi)) * float(np.float32(qj)) * math.erf(g * r32) / r32
    assert en[1] == pytest.approx(e_closed, rel=1e-12)


def test_weights_affine_and_categories_distinct():
    rng = np.random.default_rng(5)
    n = 40
    L = [18.0, 18.0, 18.0]
    pos = rng.uniform(0, 18.0, size=(n, 3))
    q = rng.choice([-1.0, 0.0, 1.0], size=n)
    types = rng.integers(0, 2, size=n).astype(np.int32)
    gid = np.arange(n, dtype=np.int32)
    # every atom: three partners, one per category
    sp = np.stack([(gid + 1) % n, (gid + 2) % n, (gid + 3) % n], 1).astype(np.int32)
    ns = np.tile([[1, 2, 3]], (n, 1)).astype(np.int32)
    eps, sig = [0.5, 0.2], [1.6, 1.1]
    p = _params(8.0, 0.12)

    def F(w):
        return _run(pos, q, types, gid, eps, sig, L, p, special=(ns, sp, w))[0]

    base = F((1.0, 1.0, 1.0))
    d12 = F((0.0, 1.0, 1.0)) - base
    d13 = F((1.0, 0.0, 1.0)) - base
    d14 = F((1.0, 1.0, 0.0)) - base
    for arr in (d12, d13, d14):
        assert np.abs(arr).max() > 1e-3  # each category acts on its own pairs
    assert np.abs(d12 - d13).max() > 1e-3 and np.abs(d13 - d14).max() > 1e-3
    w = (0.1, 0.25, 0.7)
    pred = base + (1 - w[0]) * d12 + (1 - w[1]) * d13 + (1 - w[2]) * d14
    got = F(w)
    assert np.abs(got - pred).max() <= 1e-9 * np.abs(base).max()


def test_counts_clamped_like_the_reference():
    pos, q, types, gid, ns, sp, L = _water_cluster(seed=3)
    p = _params(9.0, 0.08)
    eps, sig = [0.15535, 0.0], [3.166, 0.0]
    w = (0.0, 0.5, 1.0)

    def run(counts):
        return _run(pos, q, types, gid, eps, sig, L, p, special=(counts, sp, w))

    def same(a, b):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)

    want = run(ns)
    # n14 beyond maxspecial (2) is cut to it
    c = ns.copy()
    c[:, 2] = 9
    same(run(c), want)
    # n12 > n13 raises n13 and n14, then the cut at maxspecial lowers all three:
    # (5, 1, 1) on the O rows is (2, 2, 2)
    c = ns.copy()
    c[0::3] = [5, 1, 1]
    same(run(c), want)
    # negative counts are 0: (-1, 2, 2) on the H rows equals (0, 2, 2), which
    # moves O from the 1-2 to the 1-3 weight and so changes the forces
    c = ns.copy()
    c[1::3, 0] = -1
    d = ns.copy()
    d[1::3, 0] = 0
    same(run(c), run(d))
    assert np.abs(run(d)[0] - want[0]).max() > 1e-3


def test_special_force_is_minus_energy_gradient():
    pos, q, types, gid, ns, sp, L = _water_cluster(seed=9, n_mol=4)
    eps, sig = [0.15535, 0.0], [3.166, 0.0]
    # Verlet pairs, weights inside (0, 1): both terms are live
    p = _params(12.0, 0.1)
    w = (0.3, 0.6, 1.0)
    f = oracle.pair_force(pos.astype(np.float32), q.astype(np.float32), types, gid,
                          np.asarray(eps), np.asarray(sig), np.asarray(L), p, prec=0,
                          special=(ns, sp, w))[0]
    # finite differences need positions finer than float32: move by whole
    # float32-representable steps of 2^-10 A
    h = 2.0 ** -10
    i = 4
    for d in range(3):
        pp = pos.astype(np.float32).astype(np.float64).copy()
        pm = pp.copy()
        pp[i, d] += h
        pm[i, d] -= h
        ep = oracle.pair_force(pp.astype(np.float32), q.astype(np.float32), types, gid,
                               np.asarray(eps), np.asarray(sig), np.asarray(L), p, prec=0,
                               special=(ns, sp, w))[1].sum()
        em = oracle.pair_force(pm.astype(np.float32), q.astype(np.float32), types, gid,
                               np.asarray(eps), np.asarray(sig), np.asarray(L), p, prec=0,
                               special=(ns, sp, w))[1].sum()
        # each pair is counted from both ends with half the energy: dE/dx_i
        fd = -(ep - em) / (2 * h)
        assert f[i, d] == pytest.approx(fd, rel=2e-5, abs=2e-4)
