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


# neighbor_sample_num 30, 40, 20 give RBL frequencies 4, 3, 5 on this system:
# a power of two and two odd frequencies (the two remainder paths)
@pytest.mark.parametrize("sample_num", [30, 40, 20])
@pytest.mark.parametrize("lj_cut", [3.5, 2.5])
@pytest.mark.parametrize("step", [0, 7])
def test_rbl_pair_sets_bit_exact(device, small_electrolyte, lj_cut, step, sample_num):
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(lj_cut_off=lj_cut, sample_num=sample_num)
    ctx = _ctx(s, cfg)
    (px, py, pz), _ = _tensors(s, device, torch.float32)
    gi, gj, cls = ctx.export_pairs(step, px, py, pz)
    pos32 = s.pos.astype(np.float32)
    nb = cfg["hyper_parameters"]["neighbor"]
    freq = oracle.rbl_frequency(s.n, s.box_len, nb["r_core"], nb["cut_off"],
                                nb["neighbor_sample_num"])
    assert freq == {30: 4, 40: 3, 20: 5}[sample_num]
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


@pytest.mark.parametrize("lj_cut,sample_num", [(3.5, 30), (2.5, 30), (2.5, 1000), (3.5, 40),
                                               (2.5, 20)])
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


def _kgrid(kmax):
    r = [np.arange(-k, k + 1) for k in kmax]
    m = np.stack(np.meshgrid(*r, indexing="ij"), -1).reshape(-1, 3)
    return m[(m != 0).any(1)].astype(np.int32)


def test_ewald_exact_matches_oracle(device):
    import torch

    s = systems.electrolyte(1500, seed=5)
    kmax = (7, 7, 7)
    cfg = systems.ljsalt_config(neighbor="Verlet-List", coulomb="EWALD", kmax=kmax,
                                alpha=0.25)
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    ctx.force(0, px, py, pz, *f, energy=True, sync=True)
    pos32 = s.pos.astype(np.float32)
    ref, en, vir, _, _ = reference_force(s, pos32, cfg, 0, None, with_kspace=False)
    ci = np.nonzero(s.charges)[0]
    m = _kgrid(kmax)
    k2 = ((2 * np.pi * m / s.box_len) ** 2).sum(1)
    fk, _, ek, vk = oracle.kspace_f64(pos32.astype(np.float64)[ci], s.charges[ci], s.box_len,
                                      m, np.exp(-k2 / (4 * 0.25)), 1.0, 0.25)
    ref[ci] += fk
    V = float(np.prod(s.box_len))
    ek += oracle.kspace_self((s.charges ** 2).sum(), s.charges.sum(), 0.25, V, 1.0)
    err = rel_rms(_stack(f), ref)
    print(f"EWALD force rel RMS vs FP64 = {err:.3e}")
    assert err < 2e-5
    th = ctx.thermo()
    assert th["num_k"] == len(m)
    assert th["e_kspace"] == pytest.approx(ek, rel=1e-5)
    np.testing.assert_allclose(th["virial_kspace"], vk, rtol=1e-5, atol=1e-6 * abs(vk).max())


def test_small_box_whole_dimension_stencil(device):
    import torch

    # 2 cells per dimension: every dimension is scanned once, no duplicate pairs
    s = systems.electrolyte(400, seed=8, d_min=0.8)
    cfg = systems.ljsalt_config(cut_off=3.0, lj_cut_off=3.0, r_core=2.0, skin=0.2, P=50)
    assert s.box_len[0] < 8.0
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    m = rbe_sample_host(0.09, s.box_len, 50, 3, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    ctx.force(5, px, py, pz, *f, m_samples=mt, sync=True)
    ref = reference_force(s, s.pos.astype(np.float32), cfg, 5, m)[0]
    assert rel_rms(_stack(f), ref) < 2e-5
    gi, gj, cls = ctx.export_candidates()
    pairs = list(zip(gi.tolist(), gj.tolist()))
    assert len(pairs) == len(set(pairs))


def test_list_overflow_fallback_is_exact(device):
    import torch

    s = systems.electrolyte(4000, seed=13)
    cfg = systems.ljsalt_config(P=64)
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    m = rbe_sample_host(0.09, s.box_len, 64, 9, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    ctx.force(0, px, py, pz, *f, m_samples=mt, sync=True)
    # squeeze a slab of atoms towards a plane: local density (and list counts)
    # jump beyond the planned capacity; the same call must still be exact
    pos = s.pos.copy()
    L = s.box_len[0]
    sel = (pos[:, 2] > 0.4 * L) & (pos[:, 2] < 0.6 * L)
    pos[sel, 2] = 0.5 * L + 0.75 * (pos[sel, 2] - 0.5 * L)
    pos32 = pos.astype(np.float32)
    px.copy_(torch.tensor(pos32[:, 0], device=device))
    py.copy_(torch.tensor(pos32[:, 1], device=device))
    pz.copy_(torch.tensor(pos32[:, 2], device=device))
    ctx.force(1, px, py, pz, *f, m_samples=mt, sync=True)
    ref = reference_force(s, pos32, cfg, 1, m)[0]
    assert rel_rms(_stack(f), ref) < 2e-5
    ctx.force(2, px, py, pz, *f, m_samples=mt, sync=True)  # after growth
    ref = reference_force(s, pos32, cfg, 2, m)[0]
    assert rel_rms(_stack(f), ref) < 2e-5


def test_cpp_force_module_driver(device, tmp_path):
    import json
    import os
    import subprocess

    import rbmd_amd

    drv = os.path.join(os.path.dirname(rbmd_amd.LIB_PATH), "rbmdx_plugin_driver")
    s = systems.lj_fcc(cells=6, seed=4)
    cfg = systems.lj_config(precision="fp64")
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    lines = [f"{s.n} 1", " ".join(map(str, s.box_lo)), " ".join(map(str, s.box_len)),
             "1.0 1.0"]
    lines += [f"{i} 0 0.0 {x!r} {y!r} {z!r}" for i, (x, y, z) in enumerate(s.pos.tolist())]
    (tmp_path / "sys.txt").write_text("\n".join(lines) + "\n")
    out = subprocess.run([drv, str(tmp_path / "cfg.json"), str(tmp_path / "sys.txt"), "LJ/CUT",
                          "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = out.stdout.split("\n")
    e = [float(v) for v in rows[0].split()[1:]]
    fr = np.array([[float(v) for v in r.split()[1:]] for r in rows if r.startswith("f ")])
    ref, en, vir = oracle.verlet_force_f64(s.pos, s.charges, s.types, s.eps, s.sigma,
                                           s.box_len, 2.5)
    assert rel_rms(fr, ref) < 1e-8  # printed with 9 significant digits
    assert e[0] == pytest.approx(en[0], rel=1e-11)


def test_fused_leapfrog_matches_separate_update(device):
    import torch

    s = systems.electrolyte(4000, seed=21)
    cfg = systems.ljsalt_config(P=32)
    ctx_a = _ctx(s, cfg)
    ctx_b = _ctx(s, cfg)
    masses = np.array([2.0, 3.0, 1.5])
    ctx_a.set_masses(masses)
    p0 = torch.tensor(s.pos.T.copy(), dtype=torch.float32, device=device)
    v0 = torch.tensor(s.vel.T.copy(), dtype=torch.float32, device=device)
    pa, va = p0.clone(), v0.clone()
    pb, vb = p0.clone(), v0.clone()
    fa = torch.zeros_like(p0)
    fb = torch.zeros_like(p0)
    inv_m = torch.tensor(1.0 / masses[s.types], dtype=torch.float32, device=device)
    dt = 0.002
    for k in range(3):
        ctx_a.force_leapfrog(k, pa[0], pa[1], pa[2], va[0], va[1], va[2], fa[0], fa[1], fa[2],
                             dt, sync=True)
        ctx_b.force(k, pb[0], pb[1], pb[2], fb[0], fb[1], fb[2], sync=True)
        vb += dt * fb * inv_m
        pb += dt * vb
        if k == 0:
            # same positions: the same kernels give the same forces
            np.testing.assert_array_equal(fa.cpu().numpy(), fb.cpu().numpy())
        else:
            # the two updates round v += dt F / m differently (the kernel
            # contracts to an FMA); close LJ contacts amplify the ulp-level
            # position difference, so forces are compared as a relative RMS
            assert rel_rms(fa.T.double().cpu().numpy(), fb.T.double().cpu().numpy()) < 1e-5
        np.testing.assert_allclose(pa.cpu().numpy(), pb.cpu().numpy(), rtol=1e-6, atol=1e-6)
        np.testing.assert_allclose(va.cpu().numpy(), vb.cpu().numpy(), rtol=1e-4, atol=1e-4)


def test_kspace_stream_schedules_bitwise_equal(device, small_electrolyte, monkeypatch):
    """k-space beside the pair kernel on a second stream and k-space in order
    on one stream run the same kernels with the same reductions: forces,
    energies and virials agree bit for bit; per-phase timing is reported."""
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(P=64)
    m = rbe_sample_host(0.09, s.box_len, 64, 9, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    out = []
    for kst in ("1", "0"):
        monkeypatch.setenv("RBMDX_KSTREAM", kst)
        ctx = _ctx(s, cfg)
        (px, py, pz), f = _tensors(s, device, torch.float32)
        for step in range(3):
            ctx.force(step, px, py, pz, *f, m_samples=mt, energy=True, timing=True)
        kt = ctx.kernel_times()
        assert set(kt) == {"rebuild", "pair", "kspace", "assemble"}
        assert kt["pair"] > 0 and kt["kspace"] > 0
        th = ctx.thermo()
        out.append((_stack(f), th))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for key in ("e_vdwl", "e_coul", "e_kspace", "virial", "mean_force"):
        assert out[0][1][key] == out[1][1][key], key
    ref = reference_force(s, s.pos.astype(np.float32), cfg, 2, m)[0]
    assert rel_rms(out[0][0], ref) < 2e-5


def test_graph_replays_match_eager_launches(device, monkeypatch):
    """Evaluations replayed from the captured graph (the default) agree bit for
    bit with eager launches along a trajectory with rebuilds, a forced
    rebuild, an energy evaluation (a second graph) and torch work plus
    synchronisation between the replays."""
    import torch

    s = systems.electrolyte(4000, seed=31)
    cfg = systems.ljsalt_config(P=32, seed=777)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("RBMDX_GRAPH", mode)
        ctx = _ctx(s, cfg)
        p = torch.tensor(s.pos.T.copy(), dtype=torch.float32, device=device)
        v = torch.tensor(s.vel.T.copy(), dtype=torch.float32, device=device)
        f = torch.zeros_like(p)
        traj = []
        for k in range(14):
            if k == 6:
                ctx.force(k, p[0], p[1], p[2], f[0], f[1], f[2], rebuild=True)
            else:
                ctx.force_leapfrog(k, p[0], p[1], p[2], v[0], v[1], v[2], f[0], f[1], f[2],
                                   0.004, energy=(k % 5 == 4))
            torch.cuda.synchronize()
            scratch = (p * 2.0).sum() + v.abs().max()  # eager work between replays
            assert torch.isfinite(scratch)
            traj.append((p.cpu().numpy().copy(), f.cpu().numpy().copy()))
            if k % 5 == 4:
                th = ctx.thermo()
                traj.append((np.array([th["e_vdwl"], th["e_coul"], th["e_kspace"]]),
                             np.array(th["virial"])))
        th = ctx.thermo()
        out[mode] = (traj, th["rebuilds"])
    assert out["1"][1] == out["0"][1] and out["1"][1] >= 3
    for (a1, b1), (a0, b0) in zip(out["1"][0], out["0"][0]):
        np.testing.assert_array_equal(a1, a0)
        np.testing.assert_array_equal(b1, b0)


def test_edge_shifted_stage_sets_and_reuse_across_wraps(device):
    """Stages shift the cells across the box edge by +-L instead of applying the
    minimum image per pair: pair sets stay bit-exact, forces match the FP64
    oracle, also on a list-reuse step where atoms have wrapped across the box
    edge since the rebuild (9 cells per edge: one-cell bricks at the edge)."""
    import torch

    s = systems.electrolyte(36000, seed=17)  # 9 cells per edge: interior bricks exist
    cfg = systems.ljsalt_config(lj_cut_off=3.5, sample_num=30, P=64)
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    step = 5
    gi, gj, cls = ctx.export_pairs(step, px, py, pz, cap=1 << 24)
    pos32 = s.pos.astype(np.float32)
    nb = cfg["hyper_parameters"]["neighbor"]
    freq = oracle.rbl_frequency(s.n, s.box_len, nb["r_core"], nb["cut_off"],
                                nb["neighbor_sample_num"])
    params = oracle.pair_params(nb["r_core"], nb["cut_off"], 3.5, 0.09, 1.0, freq, step,
                                1, 1, 1)
    ref = oracle.rbl_pairs_f32(pos32, s.ids, s.box_len, params)
    ref_set = set(zip(s.ids[ref[:, 0]].tolist(), s.ids[ref[:, 1]].tolist(),
                      ref[:, 2].tolist()))
    assert _sets(gi, gj, cls) == ref_set
    m = rbe_sample_host(0.09, s.box_len, 64, 3, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    ctx.force(step, px, py, pz, *f, m_samples=mt, sync=True)
    refF = reference_force(s, pos32, cfg, step, m)[0]
    assert rel_rms(_stack(f), refF) < 2e-5
    # lists reused while atoms near the box edges wrap across it: with 9 cells
    # per edge the last brick is one cell wide, so a neighbourhood that does
    # not wrap still holds an edge cell
    rebuilds = ctx.thermo()["rebuilds"]
    rng = np.random.default_rng(23)
    pos = np.mod(s.pos + rng.uniform(-0.05, 0.05, size=s.pos.shape), s.box_len)
    wrapped = (np.abs(pos - s.pos) > 0.5 * s.box_len).any(1).sum()
    assert wrapped > 0
    pos32 = pos.astype(np.float32)
    for q, t in zip((px, py, pz), range(3)):
        q.copy_(torch.tensor(pos32[:, t], device=device))
    ctx.force(step + 1, px, py, pz, *f, m_samples=mt, sync=True)
    assert ctx.thermo()["rebuilds"] == rebuilds  # reuse
    refF = reference_force(s, pos32, cfg, step + 1, m)[0]
    assert rel_rms(_stack(f), refF) < 2e-5


def _ewald_grid_energy(s, alpha, kmax, qqr2e=1.0):
    r = [np.arange(-k, k + 1) for k in kmax]
    mg = np.stack(np.meshgrid(*r, indexing="ij"), -1).reshape(-1, 3)
    mg = mg[(mg != 0).any(1)].astype(np.int32)
    k = 2 * np.pi * mg / s.box_len
    w = np.exp(-(k ** 2).sum(1) / (4 * alpha))
    ci = np.nonzero(s.charges)[0]
    _, _, ek, _ = oracle.kspace_f64(s.pos.astype(np.float32).astype(np.float64)[ci],
                                    s.charges[ci], s.box_len, mg, w, qqr2e, alpha)
    return ek - qqr2e * np.sqrt(alpha / np.pi) * (s.charges ** 2).sum()


def test_energy_rbe_flag_reports_exact_ewald_beside_rbe(device, small_electrolyte):
    """energy_rbe_flag "yes": on energy evaluations the exact Ewald k-space
    energy over the kmax grid (with the self term) is reported as
    e_kspace_exact; the RBE estimate and the forces are unchanged."""
    import torch

    s = small_electrolyte
    cfg = systems.ljsalt_config(P=100, kmax=(5, 5, 5))
    cfg["hyper_parameters"]["coulomb"]["energy_rbe_flag"] = "yes"
    plain = systems.ljsalt_config(P=100)
    m = rbe_sample_host(0.09, s.box_len, 100, 9, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    out = {}
    for name, c in (("diag", cfg), ("plain", plain)):
        ctx = _ctx(s, c)
        (px, py, pz), f = _tensors(s, device, torch.float32)
        ctx.force(4, px, py, pz, *f, m_samples=mt, energy=True, sync=True)
        out[name] = (_stack(f), ctx.thermo())
    np.testing.assert_array_equal(out["diag"][0], out["plain"][0])
    assert out["diag"][1]["e_kspace"] == out["plain"][1]["e_kspace"]
    assert out["plain"][1]["e_kspace_exact"] == 0.0
    want = _ewald_grid_energy(s, 0.09, (5, 5, 5))
    assert out["diag"][1]["e_kspace_exact"] == pytest.approx(want, rel=1e-5)
    # the graph replay (no explicit sample) reports it too
    ctx = _ctx(s, cfg)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    for step in range(3):
        ctx.force(step, px, py, pz, *f, energy=True)
    assert ctx.thermo()["e_kspace_exact"] == pytest.approx(want, rel=1e-5)
    # automatic kmax from the accuracy rule: close to a converged grid
    cfg2 = systems.ljsalt_config(P=100)
    cl = cfg2["hyper_parameters"]["coulomb"]
    del cl["alpha"]
    cl["accuracy"] = 1e-4
    cl["energy_rbe_flag"] = "yes"
    ctx = _ctx(s, cfg2)
    (px, py, pz), f = _tensors(s, device, torch.float32)
    ctx.force(0, px, py, pz, *f, energy=True, sync=True)
    th = ctx.thermo()
    conv = _ewald_grid_energy(s, th["alpha"], (16, 16, 16))
    assert th["e_kspace_exact"] == pytest.approx(conv, rel=1e-3)
