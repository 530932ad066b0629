"""GPU parity of the CVFF special-bond pair path (force_field.type "CVFF")
against the FP64 oracle (oracle/oracle.cpp pair_force with special bonds,
pinned in tests/test_oracle_special.py).

Systems: rigid SPC/E water in REAL units (every atom charged: all pairs take
the charged-pair path; intramolecular O-H and H-H pairs are special), and the
LJ electrolyte with a synthetic topology (each atom's three nearest
neighbours as its 1-2, 1-3 and 1-4 partners), so that special pairs with a
neutral atom exercise the brick-staged list-A path.

Tolerances: forces relative vector RMS against FP64 <= 5e-5 in REAL units
(the weight-0 O-H pairs subtract a ~120 kcal/mol/A full Coulomb force from
the screened one in FP32) and <= 2e-5 for the electrolyte; energies 1e-5
relative (2e-5 in REAL units).
"""
import numpy as np
import pytest

import oracle
from rbmd_amd import ForceContext, rbe_sample_host, systems
from tests.refs import reference_force, rel_rms, cfg_qqr2e

pytestmark = pytest.mark.gpu


def _ctx(s, cfg, special=True):
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    if special:
        ctx.set_special(s.nspecial, s.special)
    return ctx


def _force(ctx, s, device, step, m=None, num_k=None, energy=True):
    import torch

    p = torch.tensor(s.pos.astype(np.float32), device=device)
    px, py, pz = p[:, 0].contiguous(), p[:, 1].contiguous(), p[:, 2].contiguous()
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    mt = None
    if m is not None:
        mt = torch.tensor(np.asarray(m).reshape(-1).astype(np.int32), device=device)
    ctx.force(step, px, py, pz, *f, m_samples=mt, energy=energy, sync=True, num_k=num_k)
    return np.stack([t.double().cpu().numpy() for t in f], axis=1)


@pytest.fixture(scope="module")
def water():
    return systems.spce_water(1000, seed=4)


@pytest.mark.parametrize("weights", [(0.0, 0.0, 0.0), (0.5, 0.25, 1.0)])
def test_water_rbl_rbe_special_parity(device, water, weights):
    s = water
    cfg = systems.cvff_water_config(alpha=0.0784, P=300, special_bonds=weights)
    ctx = _ctx(s, cfg)
    m = rbe_sample_host(0.0784, s.box_len, 300, 11, 0)
    step = 2
    got = _force(ctx, s, device, step, m)
    ref, en, vir, ek, vk = reference_force(s, s.pos.astype(np.float32), cfg, step, m)
    err = rel_rms(got, ref)
    print(f"CVFF water RBL+RBE special {weights}: force rel RMS vs FP64 = {err:.3e}")
    assert err < 5e-5
    th = ctx.thermo()
    assert th["e_coul"] == pytest.approx(en[1], rel=2e-5)
    assert th["e_vdwl"] == pytest.approx(en[0], rel=2e-5)
    assert th["e_kspace"] == pytest.approx(ek, rel=1e-5)


def test_water_without_topology_equals_lj_cut_coul_long(device, water):
    """No special rows: CVFF is the plain LJ + Coulomb pair path."""
    s = water
    cfg = systems.cvff_water_config(alpha=0.0784, P=300)
    m = rbe_sample_host(0.0784, s.box_len, 300, 11, 0)
    a = _force(_ctx(s, cfg, special=False), s, device, 2, m)
    cfg2 = systems.cvff_water_config(alpha=0.0784, P=300)
    cfg2["hyper_parameters"]["force_field"]["type"] = "LJ/CUT/COUL/LONG"
    del cfg2["hyper_parameters"]["extend"]
    b = _force(_ctx(s, cfg2, special=False), s, device, 2, m)
    np.testing.assert_array_equal(a, b)


def test_water_verlet_ewald_special_parity(device, water):
    s = water
    cfg = systems.cvff_water_config(neighbor="Verlet-List", coulomb="EWALD", alpha=0.0784)
    cfg["hyper_parameters"]["coulomb"]["kmax"] = [6, 6, 6]
    ctx = _ctx(s, cfg)
    got = _force(ctx, s, device, 0)
    qqr2e = cfg_qqr2e(cfg)
    ref, en, vir, _, _ = reference_force(s, s.pos.astype(np.float32), cfg, 0, None,
                                         with_kspace=False)
    # exact Ewald k-sum over the kmax grid, weights exp(-k^2 / 4 alpha)
    r = [np.arange(-6, 7)] * 3
    mg = np.stack(np.meshgrid(*r, indexing="ij"), -1).reshape(-1, 3)
    mg = mg[(mg != 0).any(1)].astype(np.int32)
    k = 2 * np.pi * mg / s.box_len
    w = np.exp(-(k ** 2).sum(1) / (4 * 0.0784))
    fk, _, ek, _ = oracle.kspace_f64(s.pos.astype(np.float32).astype(np.float64), s.charges,
                                     s.box_len, mg, w, qqr2e, 0.0784)
    err = rel_rms(got, ref + fk)
    print(f"CVFF water Verlet+EWALD: force rel RMS vs FP64 = {err:.3e}")
    assert err < 5e-5
    th = ctx.thermo()
    assert th["e_coul"] == pytest.approx(en[1], rel=2e-5)


def _electrolyte_topology(s):
    """Each atom's three nearest neighbours (FP64, minimum image) as its 1-2,
    1-3 and 1-4 partners; ids are 0-based and equal to the index."""
    L = s.box_len
    n = s.n
    cell = 2.0
    nc = np.maximum(1, (L // cell).astype(int))
    c = np.floor(s.pos / (L / nc)).astype(int) % nc
    key = (c[:, 2] * nc[1] + c[:, 1]) * nc[0] + c[:, 0]
    order = np.argsort(key, kind="stable")
    buckets = {}
    for i in order:
        buckets.setdefault(int(key[i]), []).append(int(i))
    sp = np.zeros((n, 3), dtype=np.int32)
    for i in range(n):
        cand = []
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    cc = (c[i] + [dx, dy, dz]) % nc
                    cand += buckets.get(int((cc[2] * nc[1] + cc[1]) * nc[0] + cc[0]), [])
        cand = np.array([j for j in set(cand) if j != i])
        d = s.pos[cand] - s.pos[i]
        d -= np.round(d / L) * L
        near = cand[np.argsort((d ** 2).sum(1), kind="stable")[:3]]
        sp[i] = s.ids[near]
    ns = np.tile([[1, 2, 3]], (n, 1)).astype(np.int32)
    return ns, sp


def test_electrolyte_topology_list_a_and_b(device):
    s = systems.electrolyte(4000, seed=11)
    ns, sp = _electrolyte_topology(s)
    s.nspecial, s.special = ns, sp
    cfg = systems.ljsalt_config(P=100)
    cfg["hyper_parameters"]["force_field"]["type"] = "CVFF"
    cfg["hyper_parameters"]["extend"] = {"special_bonds": [0.0, 0.5, 0.25]}
    ctx = _ctx(s, cfg)
    m = rbe_sample_host(0.09, s.box_len, 100, 7, 0)
    step = 5
    got = _force(ctx, s, device, step, m)
    ref, en, vir, ek, vk = reference_force(s, s.pos.astype(np.float32), cfg, step, m)
    plain = reference_force(s, s.pos.astype(np.float32),
                            systems.ljsalt_config(P=100), step, m)[0]
    # the topology matters at this tolerance
    assert rel_rms(plain, ref) > 1e-3
    err = rel_rms(got, ref)
    print(f"CVFF electrolyte topology: force rel RMS vs FP64 = {err:.3e}")
    assert err < 2e-5
    th = ctx.thermo()
    assert th["e_vdwl"] == pytest.approx(en[0], rel=1e-5)
    assert th["e_coul"] == pytest.approx(en[1], rel=1e-5)


def test_spce_water_c3_rbl_rbsog(device):
    """The 72,981-atom SPC/E water workload: CVFF with special_bonds [0, 0, 0],
    RBL (r_core 6.5 A, 100 samples) and RBSOG k-space (level 1, P = 500, rounded
    up to 512)."""
    s = systems.spce_water(24327, seed=3)
    assert s.n == 72981
    cfg = systems.cvff_water_config(coulomb="RBSOG", alpha=0.0784, P=500, rbsog_level=1)
    ctx = _ctx(s, cfg)
    P = ctx.num_samples()
    assert P == 512
    sset = oracle.sog_setup(1, 10.0, s.box_len)
    m, idx = oracle.sog_sample(31, 0, P, s.box_len, sset)
    step = 1
    got = _force(ctx, s, device, step, np.concatenate([m.reshape(-1), idx]), num_k=P,
                 energy=False)
    pos32 = s.pos.astype(np.float32)
    ref = reference_force(s, pos32, cfg, step, None, with_kspace=False)[0]
    qqr2e = cfg_qqr2e(cfg)
    fk, ek, vk = oracle.sog_kspace_f64(pos32.astype(np.float64), s.charges, s.box_len, m, idx,
                                       sset, qqr2e=qqr2e)
    err = rel_rms(got, ref + fk)
    print(f"C3 SPC/E water RBL+RBSOG: force rel RMS vs FP64 = {err:.3e}")
    assert err < 5e-5
