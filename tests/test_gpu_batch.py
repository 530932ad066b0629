"""GPU parity of the replica batch (rbmdx_batch_*): every replica's forces
against the FP64 oracle of that box alone (its own RBL frequency, k-sample and
self term), energies and virials per replica, the batched graph against the
eager and per-replica launches, positions outside the box, boxes with fewer
than three cells per dimension, and the autotuner's JSON cache.

Tolerances: forces relative vector RMS <= 2e-5 per replica (FP32 pair and
phase terms against FP64; the single-system engine's bar); pair energies and
virials relative 1e-5, k-space energy relative 1e-5 (FP64 S(k) sums of FP32
phases).
"""
import json

import numpy as np
import pytest

from rbmd_amd import BatchContext, rbe_sample_host, systems
from tests.refs import reference_force, rel_rms

pytestmark = pytest.mark.gpu

SIZES = [500, 650, 1200, 2600, 4100, 800]


def _batch(cfg, sizes, seed=7):
    reps = [systems.electrolyte(n, seed=seed + r) for r, n in enumerate(sizes)]
    b = BatchContext(cfg, device=0)
    b.set_systems(reps)
    return b, reps


def _tensors(device, reps, P, step, seed=11):
    import torch

    pos = np.concatenate([s.pos for s in reps]).astype(np.float32)
    p = torch.tensor(pos, device=device)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    f = [torch.zeros(len(pos), dtype=torch.float32, device=device) for _ in range(3)]
    ms = [rbe_sample_host(0.09, s.box_len, P, seed + r, step) for r, s in enumerate(reps)]
    m = torch.tensor(np.concatenate(ms).reshape(-1).astype(np.int32), device=device)
    return px, py, pz, f, ms, m


def _forces(f):
    return np.stack([t.double().cpu().numpy() for t in f], 1)


@pytest.mark.parametrize("neighbor", ["RBL", "Verlet-List"])
def test_batch_matches_fp64_oracle_per_replica(device, neighbor):
    P = 100
    cfg = systems.ljsalt_config(neighbor=neighbor, P=P)
    b, reps = _batch(cfg, SIZES)
    step = 5
    px, py, pz, f, ms, m = _tensors(device, reps, P, step)
    b.force(step, px, py, pz, *f, m_samples=m, energy=True, sync=True)
    F = _forces(f)
    E = b.energies()
    off = b.offsets
    for r, s in enumerate(reps):
        Fo, en, vir, ek, vk = reference_force(s, s.pos.astype(np.float32), cfg, step, ms[r])
        err = rel_rms(F[off[r]:off[r + 1]], Fo)
        print(f"{neighbor} replica {r} N={s.n} L={s.box_len[0]:.2f}: "
              f"rel RMS {err:.2e}")
        assert err < 2e-5
        np.testing.assert_allclose(E[r, 0], en[0], rtol=1e-5)
        np.testing.assert_allclose(E[r, 1], en[1], rtol=1e-5)
        np.testing.assert_allclose(E[r, 2], ek, rtol=1e-5)
        w = np.asarray(vir) + np.asarray(vk)
        np.testing.assert_allclose(E[r, 3:], w, rtol=1e-5, atol=1e-5 * np.abs(w).max())


def test_batch_graph_eager_and_per_replica_agree(device):
    P = 100
    cfg = systems.ljsalt_config(P=P)
    b, reps = _batch(cfg, SIZES, seed=21)
    px, py, pz, f, ms, m = _tensors(device, reps, P, 3)
    out = []
    for kw in ({}, {"eager": True}, {"per_replica": True}, {}):
        for t in f:
            t.zero_()
        b.force(3, px, py, pz, *f, m_samples=m, sync=True, **kw)
        out.append(_forces(f))
    for o in out[1:]:
        assert rel_rms(o, out[0]) < 1e-6
        np.testing.assert_allclose(o, out[0], rtol=0, atol=1e-5 * np.abs(out[0]).max())


def test_batch_wraps_positions_outside_the_box(device):
    """Atoms shifted by whole box lengths: the batch wraps them and matches the
    oracle at the wrapped positions."""
    import torch

    P = 50
    step = 2
    cfg = systems.ljsalt_config(P=P)
    b, reps = _batch(cfg, [900, 1500], seed=31)
    _, _, _, f, ms, m = _tensors(device, reps, P, step)
    rng = np.random.default_rng(1)
    shifted = [(s.pos + rng.integers(-2, 3, size=(s.n, 3)) * s.box_len).astype(np.float32)
               for s in reps]
    p = torch.tensor(np.concatenate(shifted), device=device)
    qx, qy, qz = (p[:, d].contiguous() for d in range(3))
    b.force(step, qx, qy, qz, *f, m_samples=m, sync=True)
    F = _forces(f)
    off = b.offsets
    for r, s in enumerate(reps):
        L = s.box_len
        wrapped = np.mod(shifted[r].astype(np.float64) - s.box_lo, L) + s.box_lo
        Fo = reference_force(s, wrapped.astype(np.float32), cfg, step, ms[r])[0]
        assert rel_rms(F[off[r]:off[r + 1]], Fo) < 2e-5


def test_batch_autotune_cache_roundtrip(device, tmp_path):
    P = 50
    cfg = systems.ljsalt_config(P=P)
    b, reps = _batch(cfg, [700, 1800, 3000], seed=41)
    px, py, pz, f, ms, m = _tensors(device, reps, P, 1)
    cache = tmp_path / "tune.json"
    shape, us = b.autotune(1, px, py, pz, *f, m_samples=m, cache_path=cache)
    assert us > 0
    stored = json.loads(cache.read_text())
    (key, val), = stored.items()
    assert {k: val[k] for k in shape} == shape
    F = _forces(f)
    b2, _ = _batch(cfg, [700, 1800, 3000], seed=41)
    shape2, us2 = b2.autotune(1, px, py, pz, *f, m_samples=m, cache_path=cache)
    assert us2 == 0.0 and shape2 == shape
    # a launch shape changes the geometry and with it the FP32 summation order
    # of a few terms per atom, nothing else
    def same(G):
        assert rel_rms(G, F) < 1e-6
        np.testing.assert_allclose(G, F, rtol=0, atol=1e-5 * np.abs(F).max())

    b2.force(1, px, py, pz, *f, m_samples=m, sync=True)
    same(_forces(f))
    for s in ({"bw": 1, "pair_block": 64, "grid_per_cu": 1, "qlanes": 4},
              {"bw": 2, "pair_block": 256, "grid_per_cu": 8, "qlanes": 16}):
        b2.set_shape(**s)
        b2.force(1, px, py, pz, *f, m_samples=m, sync=True)
        same(_forces(f))


def test_batch_ewald_matches_fp64_oracle_per_replica(device):
    """EWALD k-space in the batch: the kmax grid (|m_d| <= 3, 342 k-vectors)
    weighted exp(-k^2 / 4 alpha) in every replica, against the FP64 oracle of
    each box alone (pair forces plus the weighted k-sum); forces 2e-5, k-space
    energy and virial 1e-5."""
    import oracle

    alpha, kmax = 0.25, (3, 3, 3)
    cfg = systems.ljsalt_config(neighbor="Verlet-List", coulomb="EWALD", kmax=kmax,
                                alpha=alpha)
    b, reps = _batch(cfg, SIZES[:4])
    assert b.num_samples() == 7 ** 3 - 1
    px, py, pz, f, _, _ = _tensors(device, reps, 1, 0)
    b.force(0, px, py, pz, *f, energy=True, sync=True)
    F = _forces(f)
    E = b.energies()
    off = b.offsets
    g = np.arange(-3, 4)
    m = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    m = m[np.any(m != 0, axis=1)]
    for r, s in enumerate(reps):
        pos32 = s.pos.astype(np.float32)
        Fo, en, vir, _, _ = reference_force(s, pos32, cfg, 0, None, with_kspace=False)
        ci = np.nonzero(s.charges)[0]
        k2 = ((2 * np.pi * m / s.box_len) ** 2).sum(1)
        fk, _, ek, vk = oracle.kspace_f64(pos32.astype(np.float64)[ci], s.charges[ci],
                                          s.box_len, m, np.exp(-k2 / (4 * alpha)), 1.0,
                                          alpha)
        Fo[ci] += fk
        V = float(np.prod(s.box_len))
        ek += oracle.kspace_self((s.charges ** 2).sum(), s.charges.sum(), alpha, V, 1.0)
        err = rel_rms(F[off[r]:off[r + 1]], Fo)
        print(f"EWALD replica {r} N={s.n}: rel RMS {err:.2e}")
        assert err < 2e-5
        np.testing.assert_allclose(E[r, 2], ek, rtol=1e-5)
        w = np.asarray(vir) + np.asarray(vk)
        np.testing.assert_allclose(E[r, 3:], w, rtol=1e-5, atol=1e-5 * np.abs(w).max())


def _sog_samples(reps, level, rc, P, seed, step):
    """Per replica the RBSOG k-sample (m (P,3), idx (P,)) and the flat input
    of rbmdx_batch_force: each replica's P k-vectors, then its P sources."""
    from rbmd_amd import sog_sample_host

    ms = [sog_sample_host(level, rc, s.box_len, P, seed + r, step) for r, s in enumerate(reps)]
    flat = np.concatenate([np.concatenate([m.reshape(-1), idx]) for m, idx in ms])
    return ms, flat.astype(np.int32)


def _sog_kspace(s, pos32, m, idx, level, rc, qqr2e):
    """FP64 oracle RBSOG k-space force (N,3), energy with the self term, virial."""
    import oracle

    sset = oracle.sog_setup(level, rc, s.box_len)
    ci = np.nonzero(s.charges)[0]
    fk, ek, vk = oracle.sog_kspace_f64(pos32.astype(np.float64)[ci], s.charges[ci],
                                      s.box_len, m, idx, sset, qqr2e=qqr2e)
    F = np.zeros((s.n, 3))
    F[ci] = fk
    return F, ek + oracle.sog_self((s.charges ** 2).sum(), sset, qqr2e), vk


def test_batch_rbsog_matches_fp64_oracle_per_replica(device):
    """RBSOG k-space (level 1, P = 64 sampled k's and the direct vectors) in the
    batch: every replica of the electrolyte set against the FP64 oracle of that
    box (its own SOG coefficients and normalisation, its own k-sample)."""
    import torch

    P = 64
    cfg = systems.ljsalt_config(P=P)
    cl = cfg["hyper_parameters"]["coulomb"]
    cl.update({"type": "RBSOG", "rbsog_level": 1, "coulomb_sample_num": P})
    cl.pop("energy_rbe_flag", None)
    b, reps = _batch(cfg, SIZES, seed=31)
    assert b.num_samples() == P
    step = 2
    px, py, pz, f, _, _ = _tensors(device, reps, P, step)
    ms, flat = _sog_samples(reps, 1, 3.5, P, 77, step)
    b.force(step, px, py, pz, *f, m_samples=torch.tensor(flat, device=device), energy=True,
            sync=True)
    F = _forces(f)
    E = b.energies()
    off = b.offsets
    for r, s in enumerate(reps):
        pos32 = s.pos.astype(np.float32)
        Fp, en, vir, _, _ = reference_force(s, pos32, cfg, step, None, with_kspace=False)
        Fk, ek, vk = _sog_kspace(s, pos32, ms[r][0], ms[r][1], 1, 3.5, 1.0)
        err = rel_rms(F[off[r]:off[r + 1]], Fp + Fk)
        print(f"RBSOG replica {r} N={s.n}: rel RMS {err:.2e}")
        assert err < 2e-5
        np.testing.assert_allclose(E[r, 0], en[0], rtol=1e-5)
        np.testing.assert_allclose(E[r, 1], en[1], rtol=1e-5)
        np.testing.assert_allclose(E[r, 2], ek, rtol=1e-5)
        w = np.asarray(vir) + np.asarray(vk)
        np.testing.assert_allclose(E[r, 3:], w, rtol=1e-5, atol=1e-5 * np.abs(w).max())


@pytest.mark.parametrize("coulomb", ["RBE", "RBSOG"])
def test_batch_cvff_water_matches_fp64_oracle_per_replica(device, coulomb):
    """CVFF SPC/E water replicas (420 to 1,000 molecules, boxes of 23 to 31 A)
    with special_bonds [0, 0.5, 0.25], RBL and RBE or RBSOG (level 1) k-space:
    every replica against the FP64 CVFF oracle of that box.  Tolerances as the
    single-system CVFF water tests (tests/test_gpu_cvff.py): forces 5e-5,
    pair energies 2e-5; k-space energy 1e-5."""
    import torch

    from tests.refs import cfg_qqr2e

    P = 64
    cfg = systems.cvff_water_config(coulomb=coulomb, alpha=0.0784, P=P,
                                    special_bonds=(0.0, 0.5, 0.25))
    reps = [systems.spce_water(n, seed=40 + r) for r, n in enumerate([420, 600, 800, 1000])]
    b = BatchContext(cfg, device=0)
    b.set_systems(reps)
    b.set_special(reps)
    step = 4
    pos = np.concatenate([s.pos for s in reps]).astype(np.float32)
    p = torch.tensor(pos, device=device)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    f = [torch.zeros(len(pos), dtype=torch.float32, device=device) for _ in range(3)]
    if coulomb == "RBE":
        ms = [rbe_sample_host(0.0784, s.box_len, P, 50 + r, step) for r, s in enumerate(reps)]
        flat = np.concatenate(ms).reshape(-1).astype(np.int32)
    else:
        ms, flat = _sog_samples(reps, 1, 10.0, P, 50, step)
    b.force(step, px, py, pz, *f, m_samples=torch.tensor(flat, device=device), energy=True,
            sync=True)
    F = _forces(f)
    E = b.energies()
    off = b.offsets
    qqr2e = cfg_qqr2e(cfg)
    for r, s in enumerate(reps):
        pos32 = s.pos.astype(np.float32)
        if coulomb == "RBE":
            Fo, en, vir, ek, vk = reference_force(s, pos32, cfg, step, ms[r])
        else:
            Fp, en, vir, _, _ = reference_force(s, pos32, cfg, step, None, with_kspace=False)
            Fk, ek, vk = _sog_kspace(s, pos32, ms[r][0], ms[r][1], 1, 10.0, qqr2e)
            Fo = Fp + Fk
        err = rel_rms(F[off[r]:off[r + 1]], Fo)
        print(f"CVFF water {coulomb} replica {r} N={s.n}: rel RMS {err:.2e}")
        assert err < 5e-5
        np.testing.assert_allclose(E[r, 0], en[0], rtol=2e-5)
        np.testing.assert_allclose(E[r, 1], en[1], rtol=2e-5)
        np.testing.assert_allclose(E[r, 2], ek, rtol=1e-5)
