"""Pins of the molecular-dynamics oracle (oracle/md.py) on the CPU.

Each check is fixed by mechanics, not by the oracle's own formulas: exact free
drift with image-flag counting; energy conservation and time reversibility of
velocity Verlet on harmonic oscillators; conservation of the Nose-Hoover
extended energy and relaxation to the target temperature; the 9/10
contraction of the energy-stable rescale; the thermostat's SetUp masses.
"""
import numpy as np
import pytest

from oracle import md


def _free(n, rng, L=10.0):
    x = rng.uniform(0.0, L, (n, 3))
    v = rng.normal(0.0, 1.0, (n, 3))
    return x, v, np.zeros((n, 3), dtype=np.int64), np.ones(n)


def test_free_drift_and_image_flags():
    rng = np.random.default_rng(1)
    L = np.array([10.0, 7.0, 12.0])
    x, v, img, m = _free(64, rng)
    x = x * L / 10.0
    x0, v0 = x.copy(), v.copy()
    zero = lambda x_, step: (np.zeros_like(x_), None)
    run = md.Md(zero, x, v, img, m, [0, 0, 0], L, "NVE", 0.05)
    run.run(200)
    unwrapped = run.x + run.img * L
    np.testing.assert_allclose(unwrapped, x0 + v0 * 0.05 * 200, atol=1e-9)
    assert np.abs(run.img).max() >= 1
    assert (run.x >= 0).all() and (run.x < L).all()


def _harmonic(n, rng, k=None):
    k = rng.uniform(0.5, 2.0, n) if k is None else k
    c = rng.uniform(40.0, 60.0, (n, 3))
    x = c + rng.normal(0.0, 0.5, (n, 3))
    v = rng.normal(0.0, 1.0, (n, 3))
    m = rng.uniform(0.5, 2.0, n)

    def force(x_, step):
        d = x_ - c
        return -k[:, None] * d, float(np.sum(0.5 * k * np.sum(d * d, 1))) / n
    return x, v, np.zeros((n, 3), dtype=np.int64), m, force


def test_vv_energy_and_reversibility():
    rng = np.random.default_rng(2)
    x, v, img, m, force = _harmonic(50, rng)
    x0 = x.copy()
    run = md.Md(force, x, v, img, m, [0, 0, 0], [100.0] * 3, "NVE", 0.01)
    e0 = run.ke + run.pe
    err = []
    for _ in range(100):
        run.run(10)
        err.append(abs(run.ke + run.pe - e0) / e0)
    # velocity Verlet: bounded O(dt^2) energy error, no drift
    assert max(err) < 1e-4
    run.v *= -1.0
    run.F, _ = force(run.x, 0)
    run.run(1000)
    np.testing.assert_allclose(run.x, x0, atol=1e-9)


def _nvt_drift(dt, t_end=2.0):
    rng = np.random.default_rng(3)
    n = 200
    x, v, img, m, force = _harmonic(n, rng)
    v *= np.sqrt(3.0)  # start near T = 3 for a target of 1
    run = md.Md(force, x, v, img, m, [0, 0, 0], [100.0] * 3, "NVT", dt,
                t_control=(1.0, 1.0, 0.1))

    def extended():
        # kinetic energy of the velocities after the chain's last scaling
        return md.temp_sum(run.v, run.m, 1.0) / 2 + n * run.pe + run.chain_energy()
    e0 = extended()
    temps, drift = [], []
    steps = int(round(t_end / dt))
    for _ in range(steps // 10):
        run.run(10)
        temps.append(run.T)
        drift.append(abs(extended() - e0) / abs(e0))
    return max(drift), temps, run


def test_nose_hoover_extended_energy_and_target():
    d1, _, _ = _nvt_drift(0.005)
    d2, temps, run = _nvt_drift(0.0025)
    # the extended energy is conserved up to the integrator's O(dt^2) error
    assert d2 < 5e-4 and 3.0 < d1 / d2 < 5.0
    _, temps, run = _nvt_drift(0.005, t_end=20.0)
    assert abs(np.mean(temps[len(temps) // 2:]) - 1.0) < 0.1
    # the temperature the chain carries is that of the velocities it scaled
    ts = md.temp_sum(run.v, run.m, 1.0)
    assert abs(md.temperature(ts, run.tdof, 1.0) - run.T) < 1e-12 * run.T


def test_nose_hoover_check_catches_a_planted_sign_error():
    """A planted mistake (the velocity factor exp(+dt/2 eta_dot) instead of
    exp(-dt/2 eta_dot)) fails the extended-energy check above."""
    good_half = md.Chain.half

    def bad_half(self, dt, tdof, kb, t_freq):
        return 1.0 / good_half(self, dt, tdof, kb, t_freq)
    md.Chain.half = bad_half
    try:
        with np.errstate(all="ignore"):
            d, _, _ = _nvt_drift(0.0025)
    except (ZeroDivisionError, OverflowError):
        d = np.inf  # the planted run diverged outright
    finally:
        md.Chain.half = good_half
    assert not d <= 1e-2


def test_ramp_and_setup():
    assert md.target_temperature(1.0, 2.0, 100, 0) == 1.0
    assert md.target_temperature(1.0, 2.0, 100, 50) == 1.5
    assert md.target_temperature(1.5, 1.5, 0, 7) == 1.5
    c = md.Chain()
    c.setup(1.2, 2.0, 297.0, 1.0, 10.0)
    # Q_0 = tdof kT / f^2, Q_i = kT / f^2; a chain at rest: eta_dotdot_i = -kT / Q_i
    assert c.eta_mass[0] == pytest.approx(297.0 * 2.0 / 100.0)
    assert c.eta_mass[1] == pytest.approx(0.02)
    assert c.eta_dotdot[1] == pytest.approx(-100.0)
    assert c.eta_dotdot[2] == pytest.approx(-100.0)


def test_energy_stable_contracts_by_nine_tenths():
    rng = np.random.default_rng(4)
    x, v, img, m = _free(30, rng)
    pe = {1: -2.0}
    force = lambda x_, step: (np.zeros_like(x_), pe.get(step, -1.5))
    run = md.Md(force, x, v, img, m, [0, 0, 0], [1e3] * 3, "NVE-ES", 0.01)
    run.run(1)
    assert run.xi == 1.0
    h_init = run.pe_init + run.ke_init
    d = []
    for _ in range(12):
        run.run(1)
        d.append(run.pe + run.ke - h_init)
    assert d[0] == pytest.approx(0.5, rel=1e-12)
    for a, b in zip(d, d[1:]):
        assert b == pytest.approx(0.9 * a, rel=1e-9)
