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
