"""RBMD-style input end to end on the GPU: a LAMMPS data file (full style,
bonds and angles) read by rbmdx_data_read, its special-bond topology built
from the bonds and angles, and the CVFF force against the FP64 oracle, both
through the Python binding and through the C++ force-module driver reading
the RBMD JSON and the data file it names.  Tolerance: force relative vector
RMS <= 2e-5 (the driver prints 9 significant digits).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import rbmd_amd
from rbmd_amd import ForceContext, rbe_sample_host, systems
from tests.refs import reference_force, rel_rms

pytestmark = pytest.mark.gpu

ALPHA, P, SEED = 0.0784, 200, 20261202


def _water_file(tmp_path, n_mol=1000, seed=5):
    s = systems.spce_water(n_mol, seed=seed)
    bonds, angles = systems.water_topology(n_mol)
    path = tmp_path / "water.data"
    systems.write_lammps_data(s, path, "full", bonds=bonds, angles=angles)
    return s, path


def _cfg(neighbor):
    return systems.cvff_water_config(neighbor=neighbor, coulomb="RBE", alpha=ALPHA, P=P,
                                     seed=SEED)


@pytest.mark.parametrize("neighbor", ["RBL", "Verlet-List"])
def test_data_file_cvff_parity(device, tmp_path, neighbor):
    import torch

    s, path = _water_file(tmp_path)
    r = rbmd_amd.read_data(path, "full")
    assert np.array_equal(r.special, s.special)
    cfg = _cfg(neighbor)
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(r.box_lo, r.box_len, r.types, r.ids, r.charges, r.eps, r.sigma)
    ctx.set_special(r.nspecial, r.special)
    step = 3
    m = rbe_sample_host(ALPHA, r.box_len, P, 77, step)
    pos32 = r.pos.astype(np.float32)
    p = torch.tensor(pos32, device=device)
    px, py, pz = (p[:, d].contiguous() for d in range(3))
    f = [torch.zeros(r.n, dtype=torch.float32, device=device) for _ in range(3)]
    mt = torch.tensor(m.reshape(-1).astype(np.int32), device=device)
    ctx.force(step, px, py, pz, *f, m_samples=mt, sync=True)
    F = np.stack([t.double().cpu().numpy() for t in f], 1)
    ref = reference_force(r, pos32, cfg, step, m)[0]
    err = rel_rms(F, ref)
    print(f"{neighbor}: data-file water force rel RMS vs FP64 = {err:.3e}")
    assert err < 2e-5


def test_plugin_driver_reads_rbmd_json_and_data(device, tmp_path):
    s, path = _water_file(tmp_path)
    cfg = _cfg("Verlet-List")
    cfg["init_configuration"]["read_data"]["file"] = "water.data"
    (tmp_path / "rbmd.json").write_text(json.dumps(cfg))
    drv = os.path.join(os.path.dirname(rbmd_amd.LIB_PATH), "rbmdx_plugin_driver")
    out = subprocess.run([drv, "--rbmd", str(tmp_path / "rbmd.json"), "1"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = out.stdout.split("\n")
    assert rows[0].split()[1] == str(s.n)
    F = np.array([[float(v) for v in ln.split()[1:]] for ln in rows if ln.startswith("f ")])
    # the first evaluation of a context draws call 0 of the seeded k-sample stream
    m = rbe_sample_host(ALPHA, s.box_len, P, SEED, 0)
    r = rbmd_amd.read_data(path, "full")
    ref = reference_force(r, r.pos.astype(np.float32), cfg, 0, m)[0]
    assert rel_rms(F, ref) < 2e-5
