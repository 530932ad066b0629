"""Pins of the CPU oracle against known answers, closed forms, finite
differences and invariants (no GPU needed).

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong sign or factor, a transposed index) fails it.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- generators
def test_mt19937_check_value():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed
    # std::mt19937 (seed 5489) is 4123659995.
    g = golden("known_answers.json")
    assert int(oracle.mt19937(5489, 9999, 1)[0]) == g["mt19937_10000th"]


def test_splitmix64_vectors():
    # MixRblSampleKey(v) is the SplitMix64 output for state v + golden gamma:
    # mix(0) and mix(gamma) are the first two outputs of SplitMix64 seeded 0.
    g = golden("known_answers.json")
    L = oracle.lib()
    gamma = 0x9E3779B97F4A7C15
    assert L.orc_mix64(0) == int(g["splitmix64_seed0"][0], 16)
    assert L.orc_mix64(gamma) == int(g["splitmix64_seed0"][1], 16)
    assert L.orc_mix64(2 * gamma % 2**64) == int(g["splitmix64_seed0"][2], 16)


def test_rbl_selection_rate_and_asymmetry():
    L = oracle.lib()
    f = 4
    n = 20000
    rng = np.random.default_rng(1)
    gi = rng.integers(0, 10**6, n)
    gj = rng.integers(0, 10**6, n)
    sel = np.array([L.orc_rbl_selected(f, 17, int(a), int(b)) for a, b in zip(gi, gj)])
    rate = sel.mean()
    assert abs(rate - 1 / f) < 5 * math.sqrt((1 / f) * (1 - 1 / f) / n)
    swapped = np.array([L.orc_rbl_selected(f, 17, int(b), int(a)) for a, b in zip(gi, gj)])
    assert (sel != swapped).any()  # the directed hash is not symmetric
    assert all(L.orc_rbl_selected(1, 3, int(a), int(b)) == 1 for a, b in zip(gi[:50], gj[:50]))
    # a new epoch resamples
    other = np.array([L.orc_rbl_selected(f, 18, int(a), int(b)) for a, b in zip(gi, gj)])
    assert (sel != other).mean() > 0.2


def test_rbl_frequency_known_values():
    # rho = 0.1002 /A^3 water box: n_rs = 116, n_rc = 420 -> shell 304, p = 100 -> f = 4
    L = (24327 * 3 / 0.1002) ** (1 / 3)
    assert oracle.rbl_frequency(24327 * 3, [L, L, L], 6.5, 10.0, 100) == 4
    # LJ electrolyte rho = 0.8, r_core 2.5, rc 3.5: shell 92 -> p = 30 gives 4, p = 100 gives 1
    L2 = (200000 / 0.8) ** (1 / 3)
    assert oracle.rbl_frequency(200000, [L2] * 3, 2.5, 3.5, 30) == 4
    assert oracle.rbl_frequency(200000, [L2] * 3, 2.5, 3.5, 100) == 1


# ---------------------------------------------------------------- pair sets
def _brute_pairs(pos, L, R):
    d = pos[None, :, :] - pos[:, None, :]
    d -= np.round(d / L) * L
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    return r2


@pytest.mark.parametrize("n,L", [(600, 9.0), (300, 7.3), (150, 5.1)])
def test_pairs_vs_brute_force(n, L):
    rng = np.random.default_rng(n)
    pos = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    R = 2.4
    got = oracle.pairs_f32(pos, [L, L, L], float(np.float32(R * R)))
    r2 = _brute_pairs(pos.astype(np.float64), L, R)
    sure = set(zip(*np.nonzero(r2 < R * R * (1 - 1e-5))))
    maybe = set(zip(*np.nonzero(r2 < R * R * (1 + 1e-5))))
    gs = set(map(tuple, got.tolist()))
    assert len(gs) == len(got)
    assert sure <= gs <= maybe


def test_rbl_classes_partition():
    rng = np.random.default_rng(3)
    n, L = 800, 10.0
    pos = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    gid = rng.permutation(n).astype(np.int32)
    p = oracle.pair_params(2.0, 3.0, 3.0, 0.1, 1.0, 3, 11, 1, 1, 1)
    cls = oracle.rbl_pairs_f32(pos, gid, [L] * 3, p)
    allp = oracle.pairs_f32(pos, [L] * 3, float(np.float32(np.float32(3.0) * np.float32(3.0)) - np.float32(1e-4)))
    core_ref = oracle.pairs_f32(pos, [L] * 3, float(np.float32(np.float32(2.0) * np.float32(2.0)) - np.float32(1e-4)))
    core = {(i, j) for i, j, c in cls.tolist() if c == 1}
    shell = {(i, j) for i, j, c in cls.tolist() if c == 2}
    assert core == set(map(tuple, core_ref.tolist()))
    shell_all = set(map(tuple, allp.tolist())) - core
    assert shell <= shell_all
    Lb = oracle.lib()
    expect = {(i, j) for i, j in shell_all if Lb.orc_rbl_selected(3, 11, int(gid[i]), int(gid[j]))}
    assert shell == expect


# ---------------------------------------------------------------- LJ / Coulomb
def _two_body(r, q=(0.0, 0.0), eps=1.0, sig=1.0):
    pos = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
    return pos, np.array(q), np.zeros(2, dtype=np.int32), np.array([eps]), np.array([sig])


@pytest.mark.parametrize("r", [0.95, 1.12, 1.6, 2.3])
def test_lj_closed_form(r):
    pos, q, t, e, s = _two_body(r, eps=1.7, sig=1.1)
    f, en, vir = oracle.verlet_force_f64(pos, q, t, e, s, [10.0] * 3, 2.5)
    sr6 = (1.1 / r) ** 6
    E = 4 * 1.7 * (sr6 ** 2 - sr6)
    F = 24 * 1.7 * (2 * sr6 ** 2 - sr6) / r  # repulsive magnitude along +x on atom 1
    assert en[0] == pytest.approx(E, rel=1e-12)
    assert f[1, 0] == pytest.approx(F, rel=1e-12)
    assert f[0, 0] == pytest.approx(-F, rel=1e-12)
    assert vir[0] == pytest.approx(F * r, rel=1e-12)  # W_xx = r_ij F_ij
    assert vir[1] == pytest.approx(0.0, abs=1e-14)


@pytest.mark.parametrize("r", [0.9, 1.7, 3.1])
def test_coulomb_erfc_closed_form(r):
    pos, q, t, e, s = _two_body(r, q=(1.3, -0.7))
    alpha = 0.2
    f, en, _ = oracle.verlet_force_f64(pos, q, t, np.array([0.0]), s, [10.0] * 3, 3.5,
                                       alpha=alpha, qqr2e=2.0, coulomb=True)
    g = math.sqrt(alpha)
    E = 2.0 * 1.3 * -0.7 * math.erfc(g * r) / r
    dE = 2.0 * 1.3 * -0.7 * (-math.erfc(g * r) / r ** 2
                              - 2 * g / math.sqrt(math.pi) * math.exp(-alpha * r * r) / r)
    assert en[1] == pytest.approx(E, rel=1e-12)
    assert f[1, 0] == pytest.approx(-dE, rel=1e-10)


def _rand_system(n, L, seed, charged=True, dmin=0.9):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 3)
        if all(np.linalg.norm((p - o) - np.round((p - o) / L) * L) > dmin for o in pts):
            pts.append(p)
    pos = np.array(pts)
    q = rng.choice([-1.0, 1.0], n) if charged else np.zeros(n)
    if charged:
        q[-1] = -q[:-1].sum() if abs(q[:-1].sum()) <= 1 else q[-1]
        q -= q.mean()
    return pos, q


def test_verlet_force_is_minus_energy_gradient_and_newton():
    pos, q = _rand_system(60, 7.0, 5)
    t = np.zeros(60, dtype=np.int32)
    kw = dict(alpha=0.3, qqr2e=1.0, coulomb=True)
    L = [7.0] * 3
    f, en, _ = oracle.verlet_force_f64(pos, q, t, np.array([1.0]), np.array([1.0]), L, 3.0, **kw)
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max()
    h = 1e-6
    for i in (0, 17, 41):
        for a in range(3):
            pp = pos.copy(); pp[i, a] += h
            pm = pos.copy(); pm[i, a] -= h
            ep = sum(oracle.verlet_force_f64(pp, q, t, np.array([1.0]), np.array([1.0]), L, 3.0, **kw)[1])
            em = sum(oracle.verlet_force_f64(pm, q, t, np.array([1.0]), np.array([1.0]), L, 3.0, **kw)[1])
            assert f[i, a] == pytest.approx(-(ep - em) / (2 * h), rel=2e-5, abs=1e-6)


def test_pair_virial_is_strain_derivative():
    pos, q = _rand_system(50, 6.5, 9, charged=False)
    t = np.zeros(50, dtype=np.int32)
    e, s = np.array([1.0]), np.array([1.0])
    _, _, vir = oracle.verlet_force_f64(pos, q, t, e, s, [6.5] * 3, 2.5)
    h = 1e-6
    for a in range(3):
        sc = np.ones(3); sc[a] += h
        Ep = oracle.verlet_force_f64(pos * sc, q, t, e, s, list(6.5 * sc), 2.5)[1][0]
        sc = np.ones(3); sc[a] -= h
        Em = oracle.verlet_force_f64(pos * sc, q, t, e, s, list(6.5 * sc), 2.5)[1][0]
        assert vir[a] == pytest.approx(-(Ep - Em) / (2 * h), rel=1e-5)


def test_rbl_force_equals_full_force_when_unsampled():
    # f = 1 and r_core just below cut_off: RBL keeps every pair (core + full
    # shell), so the pair force equals the full-list force minus the mean.
    pos, q = _rand_system(80, 8.0, 2)
    pos32 = pos.astype(np.float32)
    t = np.zeros(80, dtype=np.int32)
    gid = np.arange(80, dtype=np.int32)
    e, s = np.array([1.0]), np.array([1.0])
    p = oracle.pair_params(2.0, 3.0, 3.0, 0.3, 1.0, 1, 0, 1, 1, 1)
    f, en, vir, mean = oracle.pair_force(pos32, q.astype(np.float32), t, gid, e, s, [8.0] * 3, p)
    fv, env, virv = oracle.verlet_force_f64(pos32.astype(np.float64), q.astype(np.float32).astype(np.float64),
                                            t, e, s, [8.0] * 3, 3.0, alpha=0.3, coulomb=True)
    # pairs in the sliver [r_core^2 - 1e-4, r_core^2] drop out of RBL; none here
    np.testing.assert_allclose(f, fv - fv.mean(0), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(en, env, rtol=1e-12)


# ---------------------------------------------------------------- Ewald / RBE
def _kgrid(kmax):
    r = np.arange(-kmax, kmax + 1)
    m = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    return m[(m != 0).any(1)].astype(np.int32)


def _ewald_total(pos, q, L, alpha, rc, kmax):
    m = _kgrid(kmax)
    k2 = ((2 * np.pi * m / np.asarray(L)) ** 2).sum(1)
    w = np.exp(-k2 / (4 * alpha))
    fk, rho, ek, vk = oracle.kspace_f64(pos, q, L, m, w, 1.0, alpha)
    fr, er = oracle.ewald_real_f64(pos, q, L, alpha, rc)
    V = float(np.prod(L))
    es = oracle.kspace_self((q ** 2).sum(), q.sum(), alpha, V, 1.0)
    return er + ek + es, fr + fk, ek, vk


def test_nacl_madelung_constant():
    g = golden("known_answers.json")
    cells = 4
    base = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 1], [0, 1, 1]], dtype=float)
    na = np.concatenate([base + 2 * np.array(c) for c in np.ndindex(cells, cells, cells)])
    cl = na + np.array([1.0, 0, 0])
    pos = np.concatenate([na, cl]) % (2 * cells)
    q = np.concatenate([np.ones(len(na)), -np.ones(len(cl))])
    L = [2.0 * cells] * 3
    alpha = (4.6 / 3.95) ** 2  # erfc(g rc) ~ 1e-10; exp(-k^2/4 alpha) < 1e-10 at kmax
    E, F, _, _ = _ewald_total(pos, q, L, alpha, 3.95, 16)
    M = -E / (len(pos) / 2)  # energy per ion pair = -M (nearest distance 1)
    assert M == pytest.approx(g["madelung_nacl"], rel=2e-7)
    assert np.abs(F).max() < 1e-6  # perfect lattice: no net force


def test_ewald_total_energy_alpha_independent_and_force_gradient():
    pos, q = _rand_system(40, 8.0, 21)
    L = [8.0] * 3
    # both splittings converged: erfc(g rc) < 1e-9, exp(-k^2/4 alpha) < 1e-9 at kmax
    E1, F1, _, _ = _ewald_total(pos, q, L, 1.3, 3.95, 15)
    E2, F2, _, _ = _ewald_total(pos, q, L, 1.6, 3.95, 16)
    assert E1 == pytest.approx(E2, rel=1e-7, abs=1e-8)
    np.testing.assert_allclose(F1, F2, rtol=1e-5, atol=1e-6)
    h = 1e-5
    for i, a in [(3, 0), (11, 2), (29, 1)]:
        pp = pos.copy(); pp[i, a] += h
        pm = pos.copy(); pm[i, a] -= h
        Ep = _ewald_total(pp, q, L, 1.3, 3.95, 15)[0]
        Em = _ewald_total(pm, q, L, 1.3, 3.95, 15)[0]
        assert F1[i, a] == pytest.approx(-(Ep - Em) / (2 * h), rel=1e-5, abs=1e-7)


def test_kspace_virial_is_strain_derivative():
    pos, q = _rand_system(30, 7.0, 33)
    alpha, kmax = 0.4, 10
    m = _kgrid(kmax)

    def ek(Ls, p):
        k2 = ((2 * np.pi * m / np.asarray(Ls)) ** 2).sum(1)
        return oracle.kspace_f64(p, q, Ls, m, np.exp(-k2 / (4 * alpha)), 1.0, alpha, forces=False)

    L = np.array([7.0] * 3)
    _, _, E0, vir = ek(L, pos)
    h = 1e-6
    for a in range(3):
        sc = np.ones(3); sc[a] += h
        Ep = ek(L * sc, pos * sc)[2]
        sc = np.ones(3); sc[a] -= h
        Em = ek(L * sc, pos * sc)[2]
        assert vir[a] == pytest.approx(-(Ep - Em) / (2 * h), rel=1e-6, abs=1e-9)


def test_kspace_force_sums_to_zero():
    pos, q = _rand_system(50, 9.0, 4)
    m = oracle.rbe_sample(7, 0, 64, 0.3, [9.0] * 3)
    f, _, _, _ = oracle.kspace_f64(pos, q, [9.0] * 3, m, np.ones(64), 1.0, 0.3)
    assert np.abs(f.sum(0)).max() < 1e-10 * np.abs(f).sum()


def test_rbe_normalisation_poisson_dual():
    for alpha, L in [(0.09, [62.996] * 3), (0.3, [12.0, 14.0, 9.0]), (0.02, [30.0] * 3)]:
        S = oracle.rbe_S(alpha, L)
        kmax = int(math.ceil(max(L) * math.sqrt(alpha) * 12 / (2 * math.pi))) + 2
        direct = oracle.gauss_sum_direct(alpha, L, kmax)
        assert S == pytest.approx(direct, rel=1e-10)
        assert oracle.rbe_S(alpha, L, f32=True) == pytest.approx(S, rel=1e-5)


def _kp_real_force_error(g, q2, n, rc, V):
    """Kolafa & Perram (Mol. Sim. 9, 351 (1992)) RMS real-space force error of
    an Ewald sum truncated at rc: 2 q2 exp(-g^2 rc^2) / sqrt(n rc V), with
    q2 = qqr2e * sum q^2."""
    return 2 * q2 * math.exp(-(g * rc) ** 2) / math.sqrt(n * rc * V)


@pytest.mark.parametrize("acc,qqr2e,n,rc,V,sq", [
    (1e-4, 1.0, 2000, 6.0, 4000.0, 500.0),
    (1e-5, 1.0, 200000, 3.5, 62.996 ** 3, 20000.0),
    (1e-4, 332.06371, 72981, 10.0, 89.98 ** 3, 24327 * (0.8476 ** 2 + 2 * 0.4238 ** 2)),
])
def test_alpha_from_accuracy_inverts_kolafa_perram(acc, qqr2e, n, rc, V, sq):
    """DeriveAlphaFromAccuracy (kspace_calculator.cpp:120-157) chooses g so that
    the Kolafa-Perram real-space force error equals the requested accuracy in
    force units (acc * qqr2e): the published estimate at the returned g is
    acc * qqr2e up to the float rounding of g (g rc ~ 3: relative 2 (g rc)^2
    times float epsilon)."""
    q2 = float(np.float32(np.float32(qqr2e) * np.float32(sq)))
    a = oracle.alpha_from_accuracy(acc, qqr2e, n, rc, V, q2)
    g = math.sqrt(a)
    err = _kp_real_force_error(g, q2, n, rc, V)
    assert err == pytest.approx(acc * qqr2e, rel=2e-5)
    # alpha is the float value the reference stores (rbmd::Real _alpha)
    assert a == float(np.float32(a))


def test_alpha_from_accuracy_matches_real_space_truncation_error():
    """An independent check of the rule: the actual RMS force that the
    real-space Ewald sum drops beyond rc, summed over all images for random
    +-1 charges, agrees with the accuracy the returned alpha was derived for
    (the Kolafa-Perram estimate is accurate to some 10-20 % for random
    charges: the ratio is 1.02 and 1.006 here, while g off by 5 % gives 1.8
    and 2.3)."""
    from scipy.special import erfc

    rng = np.random.default_rng(5)
    n, L, rc = 800, 16.0, 5.0
    pos = rng.uniform(0, L, size=(n, 3))
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    V = L ** 3
    for acc in (1e-3, 1e-4):
        a = oracle.alpha_from_accuracy(acc, 1.0, n, rc, V, float((q * q).sum()))
        g = math.sqrt(a)
        d = pos[None, :, :] - pos[:, None, :]
        d -= L * np.round(d / L)
        F = np.zeros((n, 3))
        for s in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
            dd = d + s * L
            r = np.sqrt((dd ** 2).sum(-1))
            far = r > rc
            np.fill_diagonal(far, False)
            rr = np.where(far, r, 1.0)
            fr = (erfc(g * rr) / rr ** 2 + 2 * g / math.sqrt(math.pi)
                  * np.exp(-(g * rr) ** 2) / rr) / rr
            fr = np.where(far, fr * q[None, :] * q[:, None], 0.0)
            F -= (fr[:, :, None] * dd).sum(1)
        rms = math.sqrt((F ** 2).sum() / n)
        assert 0.8 * acc < rms < 1.25 * acc, (acc, rms)


def test_alpha_from_accuracy_large_argument_branch():
    # g >= 1 before the logarithm: g = (1.35 - 0.15 ln(acc qqr2e)) / rc, float log
    a2 = oracle.alpha_from_accuracy(0.1, 1.0, 10 ** 6, 6.0, 10 ** 6, 1.0)
    g = (1.35 - 0.15 * float(np.log(np.float32(0.1)))) / 6.0
    assert a2 == pytest.approx(g * g, rel=1e-7)


def test_sampler_target_distribution_chi2():
    alpha, L, P = 0.09, [20.0, 20.0, 20.0], 200
    ms = np.concatenate([oracle.rbe_sample(99, c, P, alpha, L) for c in range(150)])
    assert not ((ms == 0).all(1)).any()  # (0,0,0) is never emitted
    from scipy.stats import chisquare

    for d in range(3):
        vals = ms[:, d]
        lo, hi = -8, 8
        mm = np.arange(lo, hi + 1)
        p = np.exp(-(2 * np.pi * mm / L[d]) ** 2 / (4 * alpha))
        p /= p.sum()
        obs = np.array([(vals == x).sum() for x in mm])
        ok = obs.sum()
        assert ok > 0.995 * len(vals)
        # MH samples are correlated: test on a thinned subsequence
        thin = vals[::5]
        obs = np.array([(thin == x).sum() for x in mm])
        stat = chisquare(obs, p * obs.sum())
        assert stat.pvalue > 1e-4, (d, stat)


def test_rbe_force_unbiased_against_ewald():
    pos, q = _rand_system(24, 6.0, 12)
    L = [6.0] * 3
    alpha = 0.5
    m = _kgrid(12)
    k2 = ((2 * np.pi * m / 6.0) ** 2).sum(1)
    fe, _, ee, _ = oracle.kspace_f64(pos, q, L, m, np.exp(-k2 / (4 * alpha)), 1.0, alpha)
    S = oracle.rbe_S(alpha, L)
    P, calls = 100, 400
    acc = np.zeros_like(fe)
    eacc = 0.0
    for c in range(calls):
        ms = oracle.rbe_sample(1234, c, P, alpha, L)
        f, _, e, _ = oracle.kspace_f64(pos, q, L, ms, np.full(P, S / P), 1.0, alpha)
        acc += f
        eacc += e
    err = np.sqrt(((acc / calls - fe) ** 2).sum() / (fe ** 2).sum())
    assert err < 0.08
    assert eacc / calls == pytest.approx(ee, rel=0.05)


def test_cpu_force_steps_equals_force_step_and_ignores_thread_count():
    """The CPU baseline program (stored candidate lists, OpenMP over atoms,
    FP32 phases in the k-space) computes the force of the oracle's one-shot
    evaluation (cell search per call, FP64 k-space) at the last step's RBL
    epoch, and gives the same bits on 1 and 4 threads."""
    from rbmd_amd import rbe_sample_host, systems

    s = systems.electrolyte(4000, seed=11)
    L = s.box_len
    pos32 = s.pos.astype(np.float32)
    q32 = s.charges.astype(np.float32)
    freq = oracle.rbl_frequency(s.n, L, 2.5, 3.5, 30)
    m = rbe_sample_host(0.09, L, 200, 1, 0)
    w = oracle.rbe_S(0.09, L) / 200

    def pp(epoch):
        return oracle.pair_params(2.5, 3.5, 2.5, 0.09, 1.0, freq, epoch, 1, 1, 1)

    ref = oracle.force_step(pos32, q32, s.types, s.ids, s.eps, s.sigma, L, pp(7), m, w)
    a = oracle.cpu_force_steps(pos32, q32, s.types, s.ids, s.eps, s.sigma, L, pp(5), 0.3, m,
                               w, 3, 2, 1)
    b = oracle.cpu_force_steps(pos32, q32, s.types, s.ids, s.eps, s.sigma, L, pp(5), 0.3, m,
                               w, 3, 2, 4)
    err = np.sqrt(((a - ref) ** 2).sum() / (ref ** 2).sum())
    assert err < 1e-6
    np.testing.assert_array_equal(a, b)
    # a different last epoch selects other shell pairs
    c = oracle.cpu_force_steps(pos32, q32, s.types, s.ids, s.eps, s.sigma, L, pp(5), 0.3, m,
                               w, 2, 2, 1)
    assert np.sqrt(((c - ref) ** 2).sum() / (ref ** 2).sum()) > 1e-4
