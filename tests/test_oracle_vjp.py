"""Pins of the oracle's force-matching adjoint (oracle/oracle.cpp,
force_adjoint; FP64).  No reference code differentiates this path; the pins
are mathematical:

* its forward force equals the separately pinned pair oracle plus the k-space
  oracle (same pair set, same k-sample);
* central finite differences of the linear functional L = sum G . F(theta)
  (G fixed) match the vector-Jacobian product for every per-type charge, eps
  and sigma -- including the charge of the neutral type, whose derivative runs
  through Coulomb pairs and structure factors the forward path never
  evaluates -- and for the positions of sampled atoms;
* for the force-matching loss L = sum |F - F_t|^2 with G = 2 (F - F_t), a
  finite difference of L itself matches.
The systems use RBL (core, sampled shell with its directed selection and the
mean-force fix) with list-A shell pairs (LJ cut 3.5 > r_core) and without.
"""
import numpy as np
import pytest

import oracle
from rbmd_amd import rbe_sample_host, systems

ALPHA = 0.09


def _setup(lj_cut=2.5, seed=3, n=800):
    s = systems.electrolyte(n, seed=seed)
    pos = s.pos.astype(np.float32).astype(np.float64)
    Q = np.array([1.0, -1.0, 0.0])
    eps = np.array([1.0, 0.8, 1.1])
    sig = np.array([1.0, 0.95, 1.05])
    freq = oracle.rbl_frequency(s.n, s.box_len, 2.5, 3.5, 30)
    params = oracle.pair_params(2.5, 3.5, lj_cut, ALPHA, 1.0, freq, 7, 1, 1, 1)
    m = rbe_sample_host(ALPHA, s.box_len, 40, 5, 0)
    wk = np.full(len(m), oracle.rbe_S(ALPHA, s.box_len) / len(m))
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(s.n, 3))
    return s, pos, Q, eps, sig, params, m, wk, G


def test_forward_equals_pair_plus_kspace_oracles():
    s, pos, Q, eps, sig, params, m, wk, G = _setup()
    f = oracle.force_f64(pos, s.types, s.ids, Q, eps, sig, s.box_len, params, m, wk)
    q = Q[s.types]
    fp = oracle.pair_force(pos.astype(np.float32), q.astype(np.float32), s.types, s.ids, eps,
                           sig, s.box_len, params, prec=0)[0]
    ci = np.nonzero(q)[0]
    fk = oracle.kspace_f64(pos[ci], q[ci], s.box_len, m, wk, 1.0, ALPHA)[0]
    ref = fp.copy()
    ref[ci] += fk
    np.testing.assert_allclose(f, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def _lin(s, pos, Q, eps, sig, params, m, wk, G):
    return float((G * oracle.force_f64(pos, s.types, s.ids, Q, eps, sig, s.box_len, params,
                                       m, wk)).sum())


@pytest.mark.parametrize("lj_cut", [2.5, 3.5])
def test_parameter_gradients_match_central_differences(lj_cut):
    s, pos, Q, eps, sig, params, m, wk, G = _setup(lj_cut=lj_cut)
    dr, dq, de, ds = oracle.force_vjp_f64(pos, s.types, s.ids, Q, eps, sig, s.box_len, params,
                                          G, m, wk)
    for name, vec, grad in (("Q", Q, dq), ("eps", eps, de), ("sigma", sig, ds)):
        for t in range(3):
            h = 1e-5 * max(1.0, abs(vec[t]))
            vp, vm = vec.copy(), vec.copy()
            vp[t] += h
            vm[t] -= h
            args = {"Q": Q, "eps": eps, "sig": sig}
            key = {"Q": "Q", "eps": "eps", "sigma": "sig"}[name]
            lp = _lin(s, pos, **{**args, key: vp}, params=params, m=m, wk=wk, G=G)
            lm = _lin(s, pos, **{**args, key: vm}, params=params, m=m, wk=wk, G=G)
            fd = (lp - lm) / (2 * h)
            assert grad[t] == pytest.approx(fd, rel=2e-6, abs=1e-6 * (abs(fd) + 1.0)), \
                (name, t, grad[t], fd)
    # the neutral type's charge gradient is live
    assert abs(dq[2]) > 1e-3 * np.abs(dq).max()


def test_position_gradients_match_central_differences():
    s, pos, Q, eps, sig, params, m, wk, G = _setup(lj_cut=3.5, seed=8)
    dr = oracle.force_vjp_f64(pos, s.types, s.ids, Q, eps, sig, s.box_len, params, G, m,
                              wk)[0]
    rng = np.random.default_rng(1)
    charged = np.nonzero(Q[s.types])[0]
    atoms = list(rng.choice(s.n, 4, replace=False)) + list(rng.choice(charged, 2))
    h = 1e-5
    for i in atoms:
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += h
            pm[i, d] -= h
            fd = (_lin(s, pp, Q, eps, sig, params, m, wk, G) -
                  _lin(s, pm, Q, eps, sig, params, m, wk, G)) / (2 * h)
            assert dr[i, d] == pytest.approx(fd, rel=1e-5, abs=1e-5 * (abs(fd) + 1.0)), \
                (i, d, dr[i, d], fd)


def test_force_matching_loss_gradient():
    s, pos, Q, eps, sig, params, m, wk, _ = _setup(seed=4)
    rng = np.random.default_rng(9)
    Qt, et, st = (v * (1 + rng.uniform(-0.05, 0.05, 3)) for v in (Q, eps, sig))
    Ft = oracle.force_f64(pos, s.types, s.ids, Qt, et, st, s.box_len, params, m, wk)

    def loss(Q_, e_, s_):
        F = oracle.force_f64(pos, s.types, s.ids, Q_, e_, s_, s.box_len, params, m, wk)
        return float(((F - Ft) ** 2).sum()), F

    L0, F = loss(Q, eps, sig)
    dr, dq, de, ds = oracle.force_vjp_f64(pos, s.types, s.ids, Q, eps, sig, s.box_len,
                                          params, 2 * (F - Ft), m, wk)
    h = 1e-6
    e1, e2 = eps.copy(), eps.copy()
    e1[1] += h
    e2[1] -= h
    fd = (loss(Q, e1, sig)[0] - loss(Q, e2, sig)[0]) / (2 * h)
    assert de[1] == pytest.approx(fd, rel=1e-4)
    q1, q2 = Q.copy(), Q.copy()
    q1[0] += h
    q2[0] -= h
    fd = (loss(q1, eps, sig)[0] - loss(q2, eps, sig)[0]) / (2 * h)
    assert dq[0] == pytest.approx(fd, rel=1e-4)
