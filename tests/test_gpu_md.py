"""The molecular-dynamics step on the GPU (rbmdx_md_*) against the CPU oracle
(oracle/md.py driving the FP64 LJ oracle force).

* NVT Nose-Hoover and plain NVE (velocity Verlet, wrap with image flags) on a
  500-atom FP64 LJ fluid, 40 steps in runs of 10: image flags equal, wrapped
  positions and velocities within 1e-9 relative RMS, temperature and chain
  variables within 1e-9 relative (the chain's exp() on the GPU and in libm
  may differ in the last bit; FP64 summation order of the temperature).
* NVE with RBL lists: the energy-stable factor of every step equals the
  oracle's formula on the kinetic and potential energies the GPU sampled
  (1e-12 relative), the step-1 factor is exactly 1, and the velocities left
  in the caller's arrays carry the factor.
* Settings the build does not support are rejected at rbmdx_md_init.
"""
import numpy as np
import pytest

import oracle
import rbmd_amd
from oracle import md as omd
from rbmd_amd import ForceContext, RbmdxError, systems

pytestmark = pytest.mark.gpu

DT = 0.005


def _config(ensemble, neighbor="Verlet-List", **extra):
    cfg = systems.lj_config(neighbor=neighbor)
    cfg["execution"] = {"ensemble": ensemble, "timestep": DT, "integration_type": "vv",
                        **extra}
    return cfg


def _setup(device, cfg, s):
    import torch

    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    ctx.set_masses([1.0])
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    x = [t(s.pos[:, d]) for d in range(3)]
    v = [t(s.vel[:, d]) for d in range(3)]
    f = [torch.zeros(s.n, dtype=torch.float64, device=device) for _ in range(3)]
    img = [torch.zeros(s.n, dtype=torch.int32, device=device) for _ in range(3)]
    return ctx, x, v, f, img


def _host(ts):
    return np.stack([a.cpu().numpy() for a in ts], axis=1)


def _oracle_force(s, rc=2.5):
    def force(x, step):
        F, e, _ = oracle.verlet_force_f64(x, s.charges, s.types, s.eps, s.sigma, s.box_len, rc)
        return F, e[0] / len(x)
    return force


def _rel(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


@pytest.mark.parametrize("ensemble", ["NVT", "NVE"])
def test_md_trajectory_matches_oracle(device, ensemble):
    s = systems.lj_fcc(cells=5, temperature=1.5, seed=11)
    # start some atoms just outside the box so the first wrap counts images
    s.pos[:7, 0] += s.box_len[0]
    s.pos[7:12, 2] -= s.box_len[2]
    extra = {"temp_ctrl_type": "NOSE_HOOVER", "temperature": [1.0, 1.0, 0.1]} \
        if ensemble == "NVT" else {}
    cfg = _config(ensemble, **extra)
    ctx, x, v, f, img = _setup(device, cfg, s)
    ctx.md_init(x, v, f, image=img)
    ref = omd.Md(_oracle_force(s), s.pos.astype(np.float64).copy(), s.vel.copy(),
                 np.zeros((s.n, 3), dtype=np.int64), np.ones(s.n), s.box_lo, s.box_len,
                 ensemble, DT, t_control=(1.0, 1.0, 0.1))
    np.testing.assert_array_equal(_host(img), ref.img)
    for chunk in range(4):
        ctx.md_run(10, x, v, f, image=img)
        ref.run(10)
        st = ctx.md_state()
        X, V, I = _host(x), _host(v), _host(img)
        print(f"{ensemble} step {st['step']}: x {_rel(X, ref.x):.2e} v {_rel(V, ref.v):.2e} "
              f"T {st['temperature']:.6f} / {ref.T:.6f}")
        assert st["step"] == ref.step
        np.testing.assert_array_equal(I, ref.img)
        assert _rel(X, ref.x) < 1e-9
        assert _rel(V, ref.v) < 1e-9
        assert st["temperature"] == pytest.approx(ref.T, rel=1e-9)
        if ensemble == "NVT":
            assert st["mode"] == "NVT-Nose-Hoover"
            for k in range(3):
                assert st["eta"][k] == pytest.approx(ref.chain.eta[k], rel=1e-9, abs=1e-12)
                assert st["eta_dot"][k] == pytest.approx(ref.chain.eta_dot[k], rel=1e-9,
                                                         abs=1e-12)
            assert st["thermostat_energy"] == pytest.approx(ref.chain_energy(), rel=1e-9,
                                                            abs=1e-9)
        else:
            assert st["mode"] == "NVE"
    assert np.abs(ref.img).max() >= 1


def test_energy_stable_rescale(device):
    import torch

    s = systems.lj_fcc(cells=8, temperature=1.2, seed=12)
    cfg = _config("NVE", neighbor="RBL")
    nb = cfg["hyper_parameters"]["neighbor"]
    nb.update({"cut_off": 3.5, "r_core": 2.0, "neighbor_sample_num": 50,
               "energy_rbl_flag": "yes"})
    ctx, x, v, f, img = _setup(device, cfg, s)
    ctx.md_init(x, v, f, image=img)
    states = []
    for _ in range(8):
        ctx.md_run(1, x, v, f, image=img)
        st = ctx.md_state()
        states.append(st)
        th = ctx.thermo()
        pe = (th["e_vdwl"] + th["e_coul"] + th["e_kspace"]) / s.n
        assert st["pe"] == pytest.approx(pe, rel=1e-12)
        V = torch.stack(v, 1)
        t_now = float((V * V).sum()) / (3 * s.n - 3)
        # the caller's velocities carry the factor: T(v) = xi^2 T(sampled)
        assert t_now == pytest.approx(st["xi"] ** 2 * st["temperature"], rel=1e-11)
    assert states[0]["mode"] == "NVE-energy-stable"
    assert states[0]["xi"] == 1.0
    ke0, pe0 = states[0]["ke"], states[0]["pe"]
    for st in states[1:]:
        xi = omd.energy_stable_xi(st["ke"], st["pe"], ke0, pe0, DT)
        assert st["xi"] == pytest.approx(xi, rel=1e-12)
    assert any(abs(st["xi"] - 1.0) > 1e-9 for st in states[1:])


@pytest.mark.parametrize("extra, match", [
    ({"ensemble": "NVT", "temp_ctrl_type": "BERENDSEN", "temperature": [1, 1, 0.1]},
     "temp_ctrl_type"),
    ({"ensemble": "NPT"}, "ensemble"),
    ({"ensemble": "NVE", "integration_type": "bm"}, "integration_type"),
    ({"ensemble": "NVT", "temp_ctrl_type": "NOSE_HOOVER", "temperature": [1, 1]},
     "temperature"),
])
def test_unsupported_settings_rejected(device, extra, match):
    s = systems.lj_fcc(cells=4, seed=1)
    cfg = _config("NVE")
    cfg["execution"].update(extra)
    ctx, x, v, f, img = _setup(device, cfg, s)
    with pytest.raises(RbmdxError, match=match):
        ctx.md_init(x, v, f, image=img)
    with pytest.raises(RbmdxError, match="md_init"):
        ctx.md_run(1, x, v, f, image=img)


def test_driver_md_loop_matches_python(device, tmp_path):
    """The C++ driver's simulate loop (--md: data file, GAUSS velocities, NVT,
    thermo lines and rbmd.trj frames) against the same run through Python."""
    import json
    import os
    import subprocess

    import torch

    s = systems.lj_fcc(cells=5, seed=21)
    s.vel = None
    systems.write_lammps_data(s, tmp_path / "lj.data", "atomic")
    cfg = _config("NVT", temp_ctrl_type="NOSE_HOOVER", temperature=[1.2, 1.2, 0.1],
                  num_steps=20)
    cfg["init_configuration"]["read_data"].update(
        {"file": "lj.data", "velocity_type": "GAUSS", "velocity_seed": 99})
    cfg["outputs"] = {"thermo_out": {"interval": 10}, "trajectory_out": {"interval": 10}}
    (tmp_path / "rbmd.json").write_text(json.dumps(cfg))
    drv = os.path.join(os.path.dirname(rbmd_amd.LIB_PATH), "rbmdx_plugin_driver")
    out = subprocess.run([drv, "--md", str(tmp_path / "rbmd.json")], capture_output=True,
                         text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stderr
    thermo = [ln.split() for ln in out.stdout.splitlines() if ln.startswith("thermo ")]
    assert [int(t[1]) for t in thermo] == [0, 10, 20]
    text = (tmp_path / "rbmd.trj").read_text().splitlines()
    starts = [k for k, ln in enumerate(text) if ln == "ITEM: TIMESTEP"]
    assert [int(text[k + 1]) for k in starts] == [0, 10, 20]
    last = np.array([[float(v) for v in ln.split()] for ln in text[starts[-1] + 9:]])

    r = rbmd_amd.read_data(tmp_path / "lj.data", "atomic", velocity_type="GAUSS",
                           temperature=1.2, velocity_seed=99)
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(r.box_lo, r.box_len, r.types, r.ids, r.charges, r.eps, r.sigma)
    ctx.set_masses(r.mass)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    x = [t(r.pos[:, d]) for d in range(3)]
    v = [t(r.vel[:, d]) for d in range(3)]
    f = [torch.zeros(r.n, dtype=torch.float64, device=device) for _ in range(3)]
    img = [torch.tensor(r.image[:, d].copy(), device=device) for d in range(3)]
    ctx.md_init(x, v, f, image=img, energy=True)
    temps = [ctx.md_state()["temperature"]]
    for _ in range(2):
        ctx.md_run(10, x, v, f, image=img, energy=True)
        temps.append(ctx.md_state()["temperature"])
    np.testing.assert_allclose([float(t[2]) for t in thermo], temps, rtol=1e-9)
    o = np.argsort(r.ids)
    X, V, I = _host(x)[o], _host(v)[o], _host(img)[o]
    np.testing.assert_array_equal(last[:, 0], r.ids[o] + 1)
    assert _rel(last[:, 3:6], X) < 1e-12
    assert _rel(last[:, 6:9], X + I * r.box_len) < 1e-12
    assert _rel(last[:, 9:12], V) < 1e-12


def test_real_units_argon_nvt(device, tmp_path):
    """REAL units end to end: GAUSS velocities at 120 K from the data reader,
    then 200 Nose-Hoover steps of FP64 LJ argon (eps 0.238 kcal/mol, sigma
    3.405 A, m 39.948 g/mol, dt 2 fs).  The sampled temperature at step 0 is
    the GAUSS target exactly (same kB, mvv2e and 3N - 3); the extended energy
    N (ke + pe) + chain energy stays within 2 % of the kinetic energy (a
    wrong ftm2v, mvv2e or kB breaks it by orders of magnitude)."""
    import torch

    s = systems.lj_fcc(cells=6, jitter=0.05, seed=31)
    sigma, eps, mass = 3.405, 0.238, 39.948
    s.pos *= sigma
    s.box_len = s.box_len * sigma
    s.eps, s.sigma, s.mass, s.vel = np.array([eps]), np.array([sigma]), np.array([mass]), None
    systems.write_lammps_data(s, tmp_path / "ar.data", "atomic")
    r = rbmd_amd.read_data(tmp_path / "ar.data", "atomic", velocity_type="GAUSS",
                           temperature=120.0, velocity_seed=7, unit="REAL")
    cfg = _config("NVT", temp_ctrl_type="NOSE_HOOVER", temperature=[120.0, 120.0, 100.0])
    cfg["execution"]["timestep"] = 2.0
    cfg["init_configuration"]["read_data"]["unit"] = "REAL"
    cfg["hyper_parameters"]["neighbor"].update({"cut_off": 2.5 * sigma, "skin": 1.0})
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(r.box_lo, r.box_len, r.types, r.ids, r.charges, r.eps, r.sigma)
    ctx.set_masses(r.mass)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    x = [t(r.pos[:, d]) for d in range(3)]
    v = [t(r.vel[:, d]) for d in range(3)]
    f = [torch.zeros(r.n, dtype=torch.float64, device=device) for _ in range(3)]
    ctx.md_init(x, v, f, energy=True)
    st0 = ctx.md_state()
    assert st0["temperature"] == pytest.approx(120.0, rel=1e-9)
    h0 = r.n * (st0["ke"] + st0["pe"]) + st0["thermostat_energy"]
    for _ in range(4):
        ctx.md_run(50, x, v, f, energy=True)
        st = ctx.md_state()
        h = r.n * (st["ke"] + st["pe"]) + st["thermostat_energy"]
        print(f"step {st['step']}: T {st['temperature']:.2f} K, H_ext drift "
              f"{(h - h0) / (r.n * st0['ke']):.2e}")
        assert np.isfinite(h) and 60.0 < st["temperature"] < 240.0
        assert abs(h - h0) < 0.02 * r.n * st0["ke"]
