"""GPU parity of the RBSOG k-space and the VERLET-SOG near field against the
FP64 oracle (oracle/sog.cpp).

Tolerances: forces relative vector RMS against FP64 <= 2e-5 for RBL + RBSOG
(FP32 pair and phase terms, FP64 reductions); <= 1e-4 for VERLET-SOG, whose
near-field force is a 6-term polynomial in r^2 with alternating terms that
FP32 evaluates with more cancellation; energies and virials 1e-5 relative.
The k-samples passed to both sides are identical (oracle sampler), and the
product's own seeded stream is checked against explicit samples bit for bit.
"""
import numpy as np
import pytest

import oracle
from rbmd_amd import ForceContext, systems
from tests.refs import rel_rms

pytestmark = pytest.mark.gpu


def _sog_config(level=1, P=64, neighbor="RBL", exact=False, lj_cut=3.5, seed=20261202,
                near=None):
    cfg = systems.ljsalt_config(lj_cut_off=lj_cut, P=P, seed=seed)
    cl = cfg["hyper_parameters"]["coulomb"]
    cl["type"] = "RBSOG"
    cl["rbsog_level"] = level
    cl.pop("energy_rbe_flag", None)
    if exact:
        cl["rbsog_exact"] = True
    if near is not None:
        cl["rbsog_near_field"] = near
    if neighbor != "RBL":
        nb = cfg["hyper_parameters"]["neighbor"]
        nb["type"] = neighbor
        for key in ("r_core", "neighbor_sample_num", "energy_rbl_flag"):
            nb.pop(key, None)
    return cfg


def _ctx(s, cfg):
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    return ctx


def _run(ctx, s, device, step, m=None, idx=None):
    import torch

    p = torch.tensor(s.pos.astype(np.float32), device=device)
    px, py, pz = p[:, 0].contiguous(), p[:, 1].contiguous(), p[:, 2].contiguous()
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    mt = None
    if m is not None:
        mt = torch.tensor(np.concatenate([np.asarray(m).reshape(-1), idx]).astype(np.int32),
                          device=device)
    ctx.force(step, px, py, pz, *f, m_samples=mt, energy=True, sync=True,
              num_k=None if m is None else len(idx))
    return np.stack([t.double().cpu().numpy() for t in f], axis=1)


def _kspace_ref(s, pos32, m, idx, sset):
    ci = np.nonzero(s.charges)[0]
    fk, ek, vk = oracle.sog_kspace_f64(pos32.astype(np.float64)[ci], s.charges[ci], s.box_len,
                                      m, idx, sset)
    F = np.zeros((s.n, 3))
    F[ci] = fk
    ek += oracle.sog_self((s.charges ** 2).sum(), sset)
    return F, ek, vk


@pytest.fixture(scope="module")
def electrolyte():
    return systems.electrolyte(4000, seed=11)


@pytest.mark.parametrize("level,exact", [(1, False), (0, False), (1, True)])
def test_rbl_rbsog_force_parity(device, electrolyte, level, exact):
    s = electrolyte
    cfg = _sog_config(level=level, P=64, exact=exact)
    ctx = _ctx(s, cfg)
    P = ctx.num_samples()
    assert P == 64
    sset = oracle.sog_setup(level, 3.5, s.box_len, exact=exact)
    m, idx = oracle.sog_sample(99, 0, P, s.box_len, sset)
    step = 3
    got = _run(ctx, s, device, step, m, idx)
    pos32 = s.pos.astype(np.float32)
    nb = cfg["hyper_parameters"]["neighbor"]
    freq = oracle.rbl_frequency(s.n, s.box_len, nb["r_core"], nb["cut_off"],
                                nb["neighbor_sample_num"])
    params = oracle.pair_params(nb["r_core"], 3.5, 3.5, 0.09, 1.0, freq, step, 1, 1, 1)
    fp, en, vir, _ = oracle.pair_force(pos32, s.charges.astype(np.float32), s.types, s.ids,
                                       s.eps, s.sigma, s.box_len, params, prec=0)
    fk, ek, vk = _kspace_ref(s, pos32, m, idx, sset)
    ref = fp + fk
    err = rel_rms(got, ref)
    print(f"RBL+RBSOG level {level} exact {exact}: force rel RMS vs FP64 = {err:.3e}")
    assert err < 2e-5
    th = ctx.thermo()
    assert th["e_kspace"] == pytest.approx(ek, rel=1e-5)
    np.testing.assert_allclose(th["virial_kspace"], vk, rtol=1e-5, atol=1e-6 * abs(vk).max())
    assert th["e_vdwl"] == pytest.approx(en[0], rel=1e-5)
    assert th["e_coul"] == pytest.approx(en[1], rel=1e-5)


def test_rbsog_seeded_stream_matches_explicit_samples(device, electrolyte):
    s = electrolyte
    cfg = _sog_config(level=1, P=48, seed=4242)
    ctx = _ctx(s, cfg)
    ctx2 = _ctx(s, cfg)
    P = ctx.num_samples()
    assert P == 48
    sset = oracle.sog_setup(1, 3.5, s.box_len)
    for call in range(3):
        a = _run(ctx, s, device, call)
        m, idx = oracle.sog_sample(4242, call, P, s.box_len, sset)
        mm, ii = ctx.rbe_sample(call, sog=True)
        np.testing.assert_array_equal(mm, m)
        np.testing.assert_array_equal(ii, idx)
        b = _run(ctx2, s, device, call, m, idx)
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("level", [1, 2])
def test_verlet_sog_force_parity(device, electrolyte, level):
    s = electrolyte
    cfg = _sog_config(level=level, P=64, neighbor="VERLET-SOG", lj_cut=2.5)
    ctx = _ctx(s, cfg)
    sset = oracle.sog_setup(level, 3.5, s.box_len)
    m, idx = oracle.sog_sample(5, 0, ctx.num_samples(), s.box_len, sset)
    got = _run(ctx, s, device, 0, m, idx)
    pos32 = s.pos.astype(np.float32)
    fp, en, vir = oracle.sog_verlet_f64(pos32.astype(np.float64), s.charges, s.types, s.eps,
                                        s.sigma, s.box_len, 3.5, 2.5, sset)
    fk, ek, vk = _kspace_ref(s, pos32, m, idx, sset)
    err = rel_rms(got, fp + fk)
    print(f"VERLET-SOG level {level}: force rel RMS vs FP64 = {err:.3e}")
    assert err < 1e-4
    th = ctx.thermo()
    assert th["e_vdwl"] == pytest.approx(en[0], rel=1e-5)
    assert th["e_coul"] == pytest.approx(en[1], rel=1e-4)
    assert th["e_kspace"] == pytest.approx(ek, rel=1e-5)
    np.testing.assert_allclose(np.array(th["virial"]) - np.array(th["virial_kspace"]), vir,
                               rtol=1e-4, atol=1e-5 * abs(vir).max())


@pytest.mark.parametrize("neighbor", ["RBL", "Verlet-List"])
def test_rbsog_with_sog_near_field(device, electrolyte, neighbor):
    """RBL / Verlet + RBSOG with coulomb.rbsog_near_field "sog": the SOG near
    field in place of RBMD's erfc screening (the far field's own partner)."""
    s = electrolyte
    cfg = _sog_config(level=1, P=64, neighbor=neighbor, lj_cut=2.5, near="sog")
    ctx = _ctx(s, cfg)
    sset = oracle.sog_setup(1, 3.5, s.box_len)
    m, idx = oracle.sog_sample(9, 0, ctx.num_samples(), s.box_len, sset)
    step = 2
    got = _run(ctx, s, device, step, m, idx)
    pos32 = s.pos.astype(np.float32)
    rbl = neighbor == "RBL"
    freq = oracle.rbl_frequency(s.n, s.box_len, 2.5, 3.5, 30) if rbl else 1
    params = oracle.pair_params(2.5, 3.5, 2.5, 0.09, 1.0, freq, step, int(rbl), 1, int(rbl))
    fp, en, vir, _ = oracle.pair_force(pos32, s.charges.astype(np.float32), s.types, s.ids,
                                       s.eps, s.sigma, s.box_len, params, sog_near=sset)
    fk, ek, vk = _kspace_ref(s, pos32, m, idx, sset)
    err = rel_rms(got, fp + fk)
    print(f"{neighbor} + RBSOG, SOG near field: force rel RMS vs FP64 = {err:.3e}")
    assert err < 1e-4
    th = ctx.thermo()
    assert th["e_coul"] == pytest.approx(en[1], rel=1e-4)
