"""GAUSS velocity creation: the oracle's pins (oracle/gauss.py) and the data
reader's rbmdx_data_gauss_velocities against it, on the CPU (host code).

Pins: the C++ standard's std::mt19937_64 check value; standard-normal draws
(moments and a Kolmogorov-Smirnov test); zero total momentum and the exact
target temperature with 3N - 3 degrees of freedom.  Parity: the library's
velocities equal the oracle's bit for bit (same operations in the same order,
no contraction).  Velocities from a Velocities section are kept; bad seeds,
temperatures, masses and unit styles are errors.
"""
import numpy as np
import pytest

import rbmd_amd
from oracle import gauss
from rbmd_amd import RbmdxError, systems


def test_mt19937_64_check_value():
    g = gauss.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_polar_draws_are_standard_normal():
    from scipy import stats

    g = gauss.MT19937_64(gauss.mix_velocity_seed(12345, 1))
    nd = gauss.NormalPolar()
    x = np.array([nd(g) for _ in range(20000)])
    assert abs(x.mean()) < 0.03 and abs(x.var() - 1.0) < 0.03
    assert abs(stats.kurtosis(x, fisher=False) - 3.0) < 0.15
    assert stats.kstest(x, "norm").pvalue > 1e-3
    # a mix that drops the id leaves every atom with the same stream
    assert gauss.mix_velocity_seed(12345, 1) != gauss.mix_velocity_seed(12345, 2)


def _system(n=300, seed=5):
    rng = np.random.default_rng(seed)
    L = 12.0
    types = rng.integers(0, 2, n)
    ids = rng.permutation(n)
    return systems.System(rng.uniform(0, L, (n, 3)), [0, 0, 0], [L, L, L], types, ids,
                          np.zeros(n), [1.0, 0.8], [1.0, 1.1], mass=[1.0, 2.5])


def test_library_matches_oracle(tmp_path):
    s = _system()
    p = tmp_path / "g.data"
    systems.write_lammps_data(s, p, "atomic")
    r = rbmd_amd.read_data(p, "atomic", velocity_type="GAUSS", temperature=1.3,
                           velocity_seed=777)
    assert r.has_velocities
    ref = gauss.gauss_velocities(r.ids, r.types, r.mass, 1.3, 777)
    np.testing.assert_array_equal(r.vel, ref)
    m = r.mass[r.types]
    np.testing.assert_allclose((m[:, None] * r.vel).sum(0), 0.0, atol=1e-12)
    t = (m * (r.vel ** 2).sum(1)).sum() / (3 * r.n - 3)
    assert t == pytest.approx(1.3, rel=1e-13)
    # REAL units: kB and mvv2e enter the rescale
    r2 = rbmd_amd.read_data(p, "atomic", velocity_type="GAUSS", temperature=300.0,
                            velocity_seed=777, unit="REAL")
    kb, mvv2e = 1.9872067e-3, 1.0 / 4.186e-4
    t2 = mvv2e * (m * (r2.vel ** 2).sum(1)).sum() / ((3 * r2.n - 3) * kb)
    assert t2 == pytest.approx(300.0, rel=1e-9)
    np.testing.assert_allclose(r2.vel / np.linalg.norm(r2.vel),
                               r.vel / np.linalg.norm(r.vel), rtol=1e-12, atol=1e-15)


def test_file_velocities_are_kept_and_errors(tmp_path):
    s = _system(50)
    s.vel = np.arange(150, dtype=float).reshape(50, 3) / 100.0
    p = tmp_path / "v.data"
    systems.write_lammps_data(s, p, "atomic")
    r = rbmd_amd.read_data(p, "atomic", velocity_type="GAUSS", temperature=2.0)
    order = np.argsort(r.ids)
    np.testing.assert_array_equal(r.vel[order], s.vel[np.argsort(s.ids)])
    s.vel = None
    systems.write_lammps_data(s, p, "atomic")
    with pytest.raises(RbmdxError, match="seed"):
        rbmd_amd.read_data(p, "atomic", velocity_type="GAUSS", temperature=1.0,
                           velocity_seed=0)
    with pytest.raises(RbmdxError, match="temperature"):
        rbmd_amd.read_data(p, "atomic", velocity_type="GAUSS", temperature=-1.0)
    with pytest.raises(RbmdxError, match="unit"):
        rbmd_amd.read_data(p, "atomic", velocity_type="GAUSS", temperature=1.0, unit="SI")
    with pytest.raises(RbmdxError, match="velocity_type"):
        rbmd_amd.read_data(p, "atomic", velocity_type="UNIFORM", temperature=1.0)
    s.mass = np.array([1.0, 0.0])
    systems.write_lammps_data(s, p, "atomic")
    with pytest.raises(RbmdxError, match="mass"):
        rbmd_amd.read_data(p, "atomic", velocity_type="GAUSS", temperature=1.0)
