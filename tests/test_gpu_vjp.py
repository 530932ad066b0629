"""GPU parity of the force-matching adjoint (rbmdx_force_vjp and the
torch.autograd.Function over it) against the FP64 oracle adjoint
(oracle/oracle.cpp force_adjoint, pinned by central finite differences in
tests/test_oracle_vjp.py), and a finite difference of the GPU forward itself.

dL/dQ of the neutral solvent type (charge 0) is the oracle's with
neutral_q / neutral_charge_grad, and exactly 0 without (the charged types'
values agree either way).

Tolerances: dL/dr relative vector RMS against FP64 <= 1e-4; per-type
gradients within 1e-4 of the largest of their kind (FP32 pair and phase terms,
FP32 per-lane sums, FP64 block sums); the GPU finite difference of eps within
2e-3 (FP32 forces).
"""
import numpy as np
import pytest

import oracle
from rbmd_amd import ForceContext, rbe_sample_host, systems
from tests.refs import rel_rms

pytestmark = pytest.mark.gpu

EPS = np.array([1.0, 0.8, 1.1])
SIG = np.array([1.0, 0.95, 1.05])


def _system(n=4000, seed=11):
    s = systems.electrolyte(n, seed=seed)
    s.eps = EPS.copy()
    s.sigma = SIG.copy()
    return s


def _ctx(s, cfg):
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    return ctx


def _run(ctx, s, device, step, m, G, neutral_q=True):
    import torch

    p = torch.tensor(s.pos.astype(np.float32), device=device)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    mt = torch.tensor(m.reshape(-1).astype(np.int32), device=device)
    ctx.force(step, px, py, pz, *f, m_samples=mt, sync=True)
    g = torch.tensor(G.astype(np.float32), device=device)
    gx, gy, gz = (g[:, d].contiguous() for d in range(3))
    dr = [torch.zeros_like(px) for _ in range(3)]
    dp = torch.zeros(9, dtype=torch.float64, device=device)
    ctx.force_vjp(step, px, py, pz, gx, gy, gz, *dr, dp, sync=True, neutral_q=neutral_q)
    F = np.stack([t.double().cpu().numpy() for t in f], 1)
    return F, np.stack([t.double().cpu().numpy() for t in dr], 1), dp.cpu().numpy()


def _oracle(s, cfg, step, m, G):
    nb = cfg["hyper_parameters"]["neighbor"]
    freq = oracle.rbl_frequency(s.n, s.box_len, nb["r_core"], nb["cut_off"],
                                nb["neighbor_sample_num"])
    params = oracle.pair_params(nb["r_core"], nb["cut_off"], nb["lj_cut_off"], 0.09, 1.0,
                                freq, step, 1, 1, 1)
    wk = np.full(len(m), oracle.rbe_S(0.09, s.box_len) / len(m))
    Q = np.array([1.0, -1.0, 0.0])
    pos = s.pos.astype(np.float32).astype(np.float64)
    F = oracle.force_f64(pos, s.types, s.ids, Q, s.eps, s.sigma, s.box_len, params, m, wk)
    return F, oracle.force_vjp_f64(pos, s.types, s.ids, Q, s.eps, s.sigma, s.box_len,
                                   params, G, m, wk)


@pytest.mark.parametrize("lj_cut,neutral_q", [(2.5, True), (3.5, True), (2.5, False)])
def test_vjp_matches_fp64_oracle(device, lj_cut, neutral_q):
    s = _system()
    cfg = systems.ljsalt_config(lj_cut_off=lj_cut, P=100)
    ctx = _ctx(s, cfg)
    m = rbe_sample_host(0.09, s.box_len, 100, 13, 0)
    G = np.random.default_rng(2).normal(size=(s.n, 3))
    step = 6
    F, dr, dp = _run(ctx, s, device, step, m, G, neutral_q=neutral_q)
    Fo, (dro, dqo, deo, dso) = _oracle(s, cfg, step, m, G)
    assert rel_rms(F, Fo) < 2e-5
    err = rel_rms(dr, dro)
    print(f"adjoint lj_cut {lj_cut} neutral_q {neutral_q}: dL/dr rel RMS vs FP64 = {err:.3e}")
    print("dQ", dp[:3], dqo, "deps", dp[3:6], deo, "dsig", dp[6:], dso)
    assert err < 1e-4
    tol_q = 1e-4 * np.abs(dqo).max()
    if neutral_q:
        np.testing.assert_allclose(dp[:3], dqo, rtol=0, atol=tol_q)
    else:
        np.testing.assert_allclose(dp[:2], dqo[:2], rtol=0, atol=tol_q)
        assert dp[2] == 0.0
    for got, want in ((dp[3:6], deo), (dp[6:], dso)):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-4 * np.abs(want).max())


def test_vjp_eps_matches_gpu_finite_difference(device):
    import torch

    s = _system(seed=12)
    cfg = systems.ljsalt_config(P=100)
    ctx = _ctx(s, cfg)
    m = rbe_sample_host(0.09, s.box_len, 100, 3, 0)
    G = np.random.default_rng(5).normal(size=(s.n, 3))
    step = 2
    _, _, dp = _run(ctx, s, device, step, m, G)
    p = torch.tensor(s.pos.astype(np.float32), device=device)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    mt = torch.tensor(m.reshape(-1).astype(np.int32), device=device)
    for t in range(3):
        h = 1e-3 * EPS[t]
        vals = []
        for sgn in (1, -1):
            e = EPS.copy()
            e[t] += sgn * h
            ctx.set_lj(e, SIG)
            ctx.force(step, px, py, pz, *f, m_samples=mt, sync=True)
            F = np.stack([x.double().cpu().numpy() for x in f], 1)
            vals.append((G * F).sum())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert dp[3 + t] == pytest.approx(fd, rel=2e-3, abs=1e-3 * abs(dp[3:6]).max())
    ctx.set_lj(EPS, SIG)


def test_autograd_force_matching(device):
    import torch

    from rbmd_amd.autograd import DifferentiableForce

    s = _system(seed=14)
    cfg = systems.ljsalt_config(P=100)
    model = DifferentiableForce(cfg, s, neutral_charge_grad=True)
    target = DifferentiableForce(cfg, s)
    m = rbe_sample_host(0.09, s.box_len, 100, 21, 0)
    mt = torch.tensor(m.reshape(-1).astype(np.int32), device=device)
    pos = torch.tensor(s.pos.astype(np.float32), device=device, requires_grad=True)
    Q = torch.tensor([1.0, -1.0, 0.0], dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(EPS, requires_grad=True)
    sig = torch.tensor(SIG, requires_grad=True)
    rng = np.random.default_rng(7)
    pert = lambda v: torch.tensor(v * (1 + rng.uniform(-0.05, 0.05, len(v))))
    with torch.no_grad():
        Ft = target(pos.detach(), pert(np.array([1.0, -1.0, 0.0])), pert(EPS), pert(SIG),
                    step=4, m_samples=mt)
    F = model(pos, Q, eps, sig, step=4, m_samples=mt)
    loss = ((F - Ft) ** 2).sum()
    loss.backward()
    G = 2 * (F - Ft).detach().double().cpu().numpy()
    _, (dro, dqo, deo, dso) = _oracle(s, cfg, 4, m, G)
    assert rel_rms(pos.grad.double().cpu().numpy(), dro) < 1e-4
    for got, want in ((Q.grad, dqo), (eps.grad, deo), (sig.grad, dso)):
        np.testing.assert_allclose(got.numpy(), want, rtol=0, atol=1e-4 * np.abs(want).max())


@pytest.mark.parametrize("frac_ion", [0.05, 0.3])
def test_vjp_neutral_charge_grad_large_box(device, frac_ion):
    """40,000 atoms: the neutral-charged Coulomb pass stages each brick's
    charged neighbourhood and scans rows of cells (5 % + 5 % ions), or reads
    it from global memory when it exceeds the stage (30 % + 30 % ions)."""
    s = systems.electrolyte(40000, frac_ion=frac_ion, seed=17)
    s.eps = EPS.copy()
    s.sigma = SIG.copy()
    cfg = systems.ljsalt_config(P=50)
    ctx = _ctx(s, cfg)
    m = rbe_sample_host(0.09, s.box_len, 50, 29, 0)
    G = np.random.default_rng(9).normal(size=(s.n, 3))
    step = 3
    F, dr, dp = _run(ctx, s, device, step, m, G, neutral_q=True)
    Fo, (dro, dqo, deo, dso) = _oracle(s, cfg, step, m, G)
    assert rel_rms(F, Fo) < 2e-5
    assert rel_rms(dr, dro) < 1e-4
    print("dQ", dp[:3], dqo)
    for got, want in ((dp[:3], dqo), (dp[3:6], deo), (dp[6:], dso)):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-4 * np.abs(want).max())


def _cvff_water_adjoint_and_fd(device, weights, directions):
    """The adjoint's dL/dr . delta and dL/dparam, and central differences of
    L = sum G . F, for FP64 SPC/E water with CVFF special_bonds `weights`."""
    import torch

    s = systems.spce_water(200, seed=3)
    s.eps = np.array([0.1553, 0.05])
    s.sigma = np.array([3.166, 1.0])
    cfg = systems.cvff_water_config(neighbor="Verlet-List", coulomb="RBE", cut_off=7.0,
                                    skin=1.0, P=100, special_bonds=weights,
                                    precision="fp64")
    ctx = _ctx(s, cfg)
    ctx.set_special(s.nspecial, s.special)
    m = ctx.rbe_sample(0)
    mt = torch.tensor(m.reshape(-1).astype(np.int32), device=device)
    G = np.random.default_rng(5).normal(size=(s.n, 3))
    Q = np.array([-0.8476, 0.4238])

    def loss(pos, eps=s.eps, sig=s.sigma, q=Q):
        ctx.set_lj(eps, sig)
        ctx.set_charges(q[s.types])
        p = torch.tensor(pos, dtype=torch.float64, device=device)
        f = [torch.zeros(s.n, dtype=torch.float64, device=device) for _ in range(3)]
        ctx.force(0, *(p[:, d].contiguous() for d in range(3)), *f, m_samples=mt, sync=True)
        return float(sum((f[d].cpu().numpy() * G[:, d]).sum() for d in range(3)))

    loss(s.pos)
    p = torch.tensor(s.pos, dtype=torch.float64, device=device)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    g = torch.tensor(G, dtype=torch.float64, device=device)
    dr = [torch.zeros_like(px) for _ in range(3)]
    dp = torch.zeros(6, dtype=torch.float64, device=device)
    ctx.force_vjp(0, px, py, pz, *(g[:, d].contiguous() for d in range(3)), *dr, dp,
                  sync=True)
    DR = np.stack([t.cpu().numpy() for t in dr], 1)
    h = 1e-6
    ad = np.array([(DR * dlt).sum() for dlt in directions])
    fd = np.array([(loss(s.pos + h * dlt) - loss(s.pos - h * dlt)) / (2 * h)
                   for dlt in directions])
    fdp = np.zeros(6)
    for t in range(2):
        e = np.zeros(2)
        e[t] = 1.0
        hq, he, hs = 1e-6, 1e-6 * s.eps[t], 1e-6 * s.sigma[t]
        fdp[t] = (loss(s.pos, q=Q + hq * e) - loss(s.pos, q=Q - hq * e)) / (2 * hq)
        fdp[2 + t] = (loss(s.pos, eps=s.eps + he * e)
                      - loss(s.pos, eps=s.eps - he * e)) / (2 * he)
        fdp[4 + t] = (loss(s.pos, sig=s.sigma + hs * e)
                      - loss(s.pos, sig=s.sigma - hs * e)) / (2 * hs)
    return ad, fd, dp.cpu().numpy(), fdp


def test_vjp_cvff_special_bonds_matches_finite_differences(device):
    """CVFF special bonds in the adjoint: SPC/E water with a small H LJ well
    (so the 1-3 H-H pairs carry LJ), special_bonds [0, 0.6, 0.5] against the
    same system with unit weights, Verlet list, RBE with a fixed k-sample,
    FP64.  The forward's k-space force coefficients are FP32, so a central
    difference of L = sum G . F carries their rounding; it is the same in both
    systems (same positions, charges and k-vectors) and cancels in the
    difference, which isolates the special-bond terms: the adjoint's
    special-minus-unit dL/dr along random directions and per-type dL/dQ,
    dL/deps, dL/dsigma equal the finite differences' within 1e-6 relative
    (1e-4 of the largest per-type sum: FP32 per-lane parameter sums).  The
    terms without special bonds are pinned to the FP64 oracle adjoint by the
    tests above; the forward is pinned to the CVFF oracle in
    tests/test_gpu_cvff.py."""
    rng = np.random.default_rng(6)
    dirs = [rng.normal(size=(600, 3)) for _ in range(3)]
    ad0, fd0, dp0, fdp0 = _cvff_water_adjoint_and_fd(device, (0.0, 0.6, 0.5), dirs)
    ad1, fd1, dp1, fdp1 = _cvff_water_adjoint_and_fd(device, (1.0, 1.0, 1.0), dirs)
    print("special - unit, adjoint", ad0 - ad1, "finite differences", fd0 - fd1)
    print("dparam special - unit, adjoint", dp0 - dp1, "finite differences", fdp0 - fdp1)
    np.testing.assert_allclose(ad0 - ad1, fd0 - fd1, rtol=1e-6)
    for k0 in (0, 2, 4):
        dd = (fdp0 - fdp1)[k0:k0 + 2]
        np.testing.assert_allclose((dp0 - dp1)[k0:k0 + 2], dd, atol=1e-4 * np.abs(dd).max())


def _sog_adjoint_and_fd(device, near, directions):
    """Adjoint dL/dr . delta and dL/dQ (charged types) with central
    differences, FP64 Verlet + RBSOG on a 2,000-atom electrolyte with
    coulomb.rbsog_near_field `near`, k-sample fixed."""
    import torch

    s = _system(2000, seed=12)
    cfg = systems.ljsalt_config(P=64)
    cl = cfg["hyper_parameters"]["coulomb"]
    cl.update({"type": "RBSOG", "rbsog_level": 1, "rbsog_near_field": near})
    cl.pop("energy_rbe_flag", None)
    nb = cfg["hyper_parameters"]["neighbor"]
    nb["type"] = "Verlet-List"
    for key in ("r_core", "neighbor_sample_num", "energy_rbl_flag"):
        nb.pop(key, None)
    cfg["hyper_parameters"]["precision"] = "fp64"
    ctx = _ctx(s, cfg)
    m, idx = ctx.rbe_sample(0, sog=True)
    mt = torch.tensor(np.concatenate([m.reshape(-1), idx]).astype(np.int32), device=device)
    G = np.random.default_rng(8).normal(size=(s.n, 3))
    Q = np.array([1.0, -1.0, 0.0])

    def loss(pos, q=Q):
        ctx.set_charges(q[s.types])
        p = torch.tensor(pos, dtype=torch.float64, device=device)
        f = [torch.zeros(s.n, dtype=torch.float64, device=device) for _ in range(3)]
        ctx.force(0, *(p[:, d].contiguous() for d in range(3)), *f, m_samples=mt,
                  num_k=len(idx), sync=True)
        return float(sum((f[d].cpu().numpy() * G[:, d]).sum() for d in range(3)))

    loss(s.pos)
    p = torch.tensor(s.pos, dtype=torch.float64, device=device)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    g = torch.tensor(G, dtype=torch.float64, device=device)
    dr = [torch.zeros_like(px) for _ in range(3)]
    dp = torch.zeros(9, dtype=torch.float64, device=device)
    ctx.force_vjp(0, px, py, pz, *(g[:, d].contiguous() for d in range(3)), *dr, dp,
                  sync=True)
    DR = np.stack([t.cpu().numpy() for t in dr], 1)
    h = 1e-6
    ad = np.array([(DR * dlt).sum() for dlt in directions])
    fd = np.array([(loss(s.pos + h * dlt) - loss(s.pos - h * dlt)) / (2 * h)
                   for dlt in directions])
    fdq = np.zeros(2)
    for t in range(2):
        e = np.zeros(3)
        e[t] = 1e-6
        fdq[t] = (loss(s.pos, q=Q + e) - loss(s.pos, q=Q - e)) / 2e-6
    return ad, fd, dp.cpu().numpy()[:2], fdq


def test_vjp_sog_near_field_matches_finite_differences(device):
    """The SOG near field in the adjoint (VERLET-SOG's and rbsog_near_field
    "sog"): the sog-minus-erfc difference of dL/dr along random directions and
    of dL/dQ per charged type equals that of central differences of the
    forward (pinned to the oracle in tests/test_gpu_sog.py) within 1e-6
    relative (1e-4 of the largest for dL/dQ).  The RBSOG k-space part, whose
    FP32 coefficients make a single central difference noisy, is the same in
    both systems and cancels; the erfc adjoint itself is pinned to the FP64
    oracle adjoint above."""
    rng = np.random.default_rng(9)
    dirs = [rng.normal(size=(2000, 3)) for _ in range(3)]
    ad0, fd0, dq0, fdq0 = _sog_adjoint_and_fd(device, "sog", dirs)
    ad1, fd1, dq1, fdq1 = _sog_adjoint_and_fd(device, "erfc", dirs)
    print("sog - erfc, adjoint", ad0 - ad1, "finite differences", fd0 - fd1)
    print("dQ sog - erfc, adjoint", dq0 - dq1, "finite differences", fdq0 - fdq1)
    assert np.abs(fd0 - fd1).min() > 1e-5 * np.abs(fd0).max()  # the near fields differ
    np.testing.assert_allclose(ad0 - ad1, fd0 - fd1, rtol=1e-6)
    dd = fdq0 - fdq1
    np.testing.assert_allclose(dq0 - dq1, dd, atol=1e-4 * np.abs(dd).max())


def _polar_water(n_mol, seed):
    """SPC/E geometry and topology with a neutral O (type 0) and the two H
    sites charged +0.4238 (type 1) and -0.4238 (type 2): every molecule holds
    charged-neutral special pairs (O-H, 1-2) and a charged 1-3 pair (H-H)."""
    s = systems.spce_water(n_mol, seed=seed)
    s.types = s.types.copy()
    s.types[2::3] = 2
    s.charges = np.tile([0.0, 0.4238, -0.4238], n_mol)
    s.eps = np.array([0.15535, 0.02, 0.02])
    s.sigma = np.array([3.166, 1.0, 1.0])
    return s


def test_vjp_cvff_neutral_charge_grad(device):
    """dL/dQ of a neutral type with CVFF special bonds and
    RBMDX_FLAG_VJP_NEUTRAL_Q: the neutral O of a polar 3-site water carries
    1-2 partners that are charged.  special_bonds [0, 0.6, 0.5] against unit
    weights (RBL, RBE): the adjoint's special-minus-unit dL/dQ_O equals the
    central difference in Q_O at 0 (h = 2^-10) of the FP64 oracle's CVFF pair
    force, L = sum G' . F_pair with G' the mean-shifted cotangent, special
    minus unit.  k-space does not see special bonds and cancels in the
    difference, as do the terms the two systems share, which the FP64
    oracle adjoint pins above.  Tolerance 1e-4 of the difference (FP32
    pair terms and per-lane sums)."""
    from tests.refs import cfg_qqr2e, cfg_special

    s = _polar_water(1000, seed=9)
    G = np.random.default_rng(10).normal(size=(s.n, 3))
    step = 4
    res = {}
    for key, w in (("special", (0.0, 0.6, 0.5)), ("unit", (1.0, 1.0, 1.0))):
        cfg = systems.cvff_water_config(alpha=0.0784, P=100, special_bonds=w)
        ctx = _ctx(s, cfg)
        ctx.set_special(s.nspecial, s.special)
        m = rbe_sample_host(0.0784, s.box_len, 100, 21, 0)
        _, _, dp = _run(ctx, s, device, step, m, G, neutral_q=True)
        # the oracle's FP64 pair force with Q_O = +-h (same RBL epoch and set)
        nb = cfg["hyper_parameters"]["neighbor"]
        freq = ctx.thermo()["rbl_frequency"]
        params = oracle.pair_params(nb["r_core"], nb["cut_off"], nb["cut_off"],
                                    ctx.thermo()["alpha"], cfg_qqr2e(cfg), freq, step, 1, 1, 1)
        pos = s.pos.astype(np.float32)
        Gs = G - G.mean(axis=0)

        def loss(qo):
            q = np.array([qo, 0.4238, -0.4238])[s.types].astype(np.float32)
            f = oracle.pair_force(pos, q, s.types, s.ids, s.eps, s.sigma, s.box_len, params,
                                  special=cfg_special(cfg, s))[0]
            return float((f * Gs).sum())

        # L is quadratic in Q_O, so the central difference is exact up to
        # rounding; h a power of two is exact in the oracle's FP32 charges
        h = 2.0 ** -10
        res[key] = (dp[0], (loss(h) - loss(-h)) / (2 * h))
    ad = res["special"][0] - res["unit"][0]
    fd = res["special"][1] - res["unit"][1]
    print(f"dL/dQ_O special - unit: adjoint {ad:.6e}, oracle central difference {fd:.6e}")
    assert abs(fd) > 1e-3 * abs(res["unit"][1])  # the special terms matter here
    assert ad == pytest.approx(fd, rel=1e-4)
