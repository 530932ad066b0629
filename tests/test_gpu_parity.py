"""GPU parity of the HIP force path against the CPU oracle.

Tolerances:
* pair sets (candidates, RBL core/shell selections): bit-exact as sets of
  (gid_i, gid_j, class) -- both sides evaluate the same FP32 distance formula;
* forces: relative vector RMS against the FP64 oracle <= 2e-5 (FP32 pair terms
  and FP32 e^{ik.r} with FP64 S(k) reductions); FP64 LJ path <= 1e-12;
* energies / virials: relative 1e-5 (FP32 path), 1e-10 (FP64 path).
"""
import numpy as np
import pytest

import oracle
from rbmd_amd import ForceContext, rbe_sample_host, systems
from tests.refs import reference_force, rel_rms

pytestmark = pytest.mark.gpu


def _tensors(sys_, device, dtype):
    import torch

    p = torch.tensor(sys_.pos, dtype=dtype, device=device)
    px, py, pz = p[:, 0].contiguous(), p[:, 1].contiguous(), p[:, 2].contiguous()
    f = [torch.zeros(sys_.n, dtype=dtype, device=device) for _ in range(3)]
    return (px, py, pz), f


def _ctx(sys_, cfg):
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(sys_.box_lo, sys_.box_len, sys_.types, sys_.ids, sys_.charges,
                   sys_.eps, sys_.sigma)
    return ctx


def _stack(f):
    return np.stack([t.double().cpu().numpy() for t in f], axis=1)


@pytest.fixture(scope="module")
def small_electrolyte():
    return systems.electrolyte(4000, seed=11)


def _sets(gi, gj, cls):
    return set(zip(gi.tolist(), gj.tolist(), cls.tolist()))


@pytest.mark.parametrize("lj_cut", [3.5, 2.5])
@pytest.mark.parametrize("step", [0, 7])
def test_rbl_pair_sets_bit_exact(device, small_electrolyte, lj_cut, step):
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(lj_cut_off=lj_cut, sample_num=30)
    ctx = _ctx(s, cfg)
    (px, py, pz), _ = _tensors(s, device, torch.float32)
    gi, gj, cls = ctx.export_pairs(step, px, py, pz)
    pos32 = s.pos.astype(np.float32)
    nb = cfg["hyper_parameters"]["neighbor"]
    freq = oracle.rbl_frequency(s.n, s.box_len, nb["r_core"], nb["cut_off"],
                                nb["neighbor_sample_num"])
    assert ctx.thermo()["rbl_frequency"] == freq
    params = oracle.pair_params(nb["r_core"], nb["cut_off"], lj_cut, 0.09, 1.0, freq, step,
                                1, 1, 1)
    ref = oracle.rbl_pairs_f32(pos32, s.ids, s.box_len, params)
    ref_set = set(zip(s.ids[ref[:, 0]].tolist(), s.ids[ref[:, 1]].tolist(),
                      ref[:, 2].tolist()))
    got = _sets(gi, gj, cls)
    if lj_cut == 3.5:
        assert got == ref_set
    else:
        # split LJ cutoff: neutral pairs beyond lj_cut + skin are not stored
        # (they carry no force); every stored pair matches and every pair that
        # can interact is stored
        assert got <= ref_set
        q = s.charges
        inter = {t for t in ref_set if (q[t[0]] != 0 and q[t[1]] != 0)
                 or _d2(pos32, s.box_len, t[0], t[1]) < 2.7 ** 2}
        assert inter <= got


def _d2(pos32, L, i, j):
    d = (pos32[j].astype(np.float64) - pos32[i]).astype(np.float64)
    d -= np.round(d / L) * L
    return float((d ** 2).sum())


def test_candidate_lists_bit_exact(device, small_electrolyte):
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(lj_cut_off=3.5)
    ctx = _ctx(s, cfg)
    (px, py, pz), _ = _tensors(s, device, torch.float32)
    ctx.export_pairs(0, px, py, pz)
    gi, gj, cls = ctx.export_candidates()
    pos32 = s.pos.astype(np.float32)
    R2 = np.float32(np.float32(3.8) * np.float32(3.8))
    ref = oracle.pairs_f32(pos32, s.box_len, float(np.float32(3.8 * 3.8)))
    ref_set = set(zip(ref[:, 0].tolist(), ref[:, 1].tolist()))
    got = set(zip(gi.tolist(), gj.tolist()))
    assert len(got) == len(gi)  # no duplicates
    assert got == ref_set
    del R2


@pytest.mark.parametrize("lj_cut,sample_num", [(3.5, 30), (2.5, 30), (2.5, 1000)])
def test_rbl_rbe_force_parity(device, small_electrolyte, lj_cut, sample_num):
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(lj_cut_off=lj_cut, sample_num=sample_num, P=200)
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    m = rbe_sample_host(0.09, s.box_len, 200, 20261202, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    step = 3
    ctx.force(step, px, py, pz, *f, m_samples=mt, energy=True, sync=True)
    got = _stack(f)
    pos32 = s.pos.astype(np.float32)
    ref, en, vir, ek, vk = reference_force(s, pos32, cfg, step, m)
    err = rel_rms(got, ref)
    print(f"force rel RMS vs FP64 = {err:.3e}")
    assert err < 2e-5
    th = ctx.thermo()
    assert th["e_kspace"] == pytest.approx(ek, rel=1e-5)
    np.testing.assert_allclose(th["virial_kspace"], vk, rtol=1e-5, atol=1e-6 * abs(vk).max())
    assert th["e_vdwl"] == pytest.approx(en[0], rel=1e-5)
    assert th["e_coul"] == pytest.approx(en[1], rel=1e-5)
    np.testing.assert_allclose(np.array(th["virial"]) - np.array(th["virial_kspace"]), vir,
                               rtol=1e-4, atol=1e-5 * abs(vir).max())


def test_seeded_stream_matches_explicit_samples(device, small_electrolyte):
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(P=64)
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    outs = []
    for call in range(3):
        ctx.force(call, px, py, pz, *f)
        outs.append(_stack(f))
    ctx2 = _ctx(s, cfg)
    f2 = [torch.zeros_like(t) for t in f]
    for call in range(3):
        m = ctx2.rbe_sample(call)
        assert (m == rbe_sample_host(0.09, s.box_len, 64, 20261202, call)).all()
        mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
        ctx2.force(call, px, py, pz, *f2, m_samples=mt)
        np.testing.assert_array_equal(_stack(f2), outs[call])


def test_deterministic_and_skin_reuse(device, small_electrolyte):
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(P=100)
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    m = rbe_sample_host(0.09, s.box_len, 100, 5, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    ctx.force(0, px, py, pz, *f, m_samples=mt, sync=True)
    a = _stack(f)
    ctx.force(0, px, py, pz, *f, m_samples=mt, sync=True)
    np.testing.assert_array_equal(_stack(f), a)
    rng = np.random.default_rng(4)
    for amp, step in [(0.04, 1), (0.2, 2)]:  # inside skin/2 (reuse), then rebuild
        pos = np.mod(s.pos + rng.uniform(-amp, amp, size=s.pos.shape), s.box_len)
        pos32 = pos.astype(np.float32)
        px.copy_(torch.tensor(pos32[:, 0], device=device))
        py.copy_(torch.tensor(pos32[:, 1], device=device))
        pz.copy_(torch.tensor(pos32[:, 2], device=device))
        ctx.force(step, px, py, pz, *f, m_samples=mt, sync=True)
        ref = reference_force(s, pos32, cfg, step, m)[0]
        assert rel_rms(_stack(f), ref) < 2e-5
    assert ctx.thermo()["rebuilds"] >= 2


def test_lj_fcc_fp64_verlet(device):
    import torch

    s = systems.lj_fcc(cells=10, seed=1)
    cfg = systems.lj_config(precision="fp64")
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float64)
    ctx.force(0, px, py, pz, *f, energy=True, sync=True)
    ref, en, vir = oracle.verlet_force_f64(s.pos, s.charges, s.types, s.eps, s.sigma,
                                           s.box_len, 2.5)
    assert rel_rms(_stack(f), ref) < 1e-12
    th = ctx.thermo()
    assert th["e_vdwl"] == pytest.approx(en[0], rel=1e-11)
    np.testing.assert_allclose(th["virial"], vir, rtol=1e-10, atol=1e-10)
