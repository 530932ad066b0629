"""Host-side checks of the C ABI (no GPU): exported symbols, configuration
validation with the reference's error behaviour, and the k-sampler stream."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import rbmd_amd
from rbmd_amd import systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rbmdx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbmdx_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 14
    lib = ctypes.CDLL(rbmd_amd.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding binds exactly the declared set
    assert set(rbmd_amd.SIGNATURES) == set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", rbmd_amd.LIB_PATH],
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (rbmdx_\w+)", out))
    assert set(names) <= exported


def test_binding_loads_and_binds():
    lib = rbmd_amd.load_library()
    assert lib.rbmdx_create.restype == ctypes.c_int


def test_product_never_imports_the_oracle():
    for dirpath, _, files in os.walk(os.path.join(ROOT, "rbmd_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "liboracle" not in src, f
    for f in ("rbmdx.h",):
        assert "oracle" not in open(os.path.join(ROOT, "include", f)).read()


GOOD = systems.ljsalt_config()


def _cfg(**edits):
    import copy

    c = copy.deepcopy(GOOD)
    for path, val in edits.items():
        node = c
        keys = path.split(".")
        for k in keys[:-1]:
            node = node[k]
        if val is None:
            node.pop(keys[-1], None)
        else:
            node[keys[-1]] = val
    return c


@pytest.mark.parametrize("edit,msg", [
    ({}, None),
    ({"hyper_parameters.neighbor.energy_rbl_flag": None}, "energy_rbl_flag"),
    ({"hyper_parameters.neighbor.r_core": 3.5}, "r_core"),
    ({"hyper_parameters.neighbor.neighbor_sample_num": 0}, "neighbor_sample_num"),
    ({"hyper_parameters.coulomb.energy_rbe_flag": None}, "energy_rbe_flag"),
    ({"hyper_parameters.coulomb.coulomb_sample_num": 0}, "sample number"),
    ({"hyper_parameters.coulomb.alpha": -1.0}, "alpha"),
    ({"hyper_parameters.coulomb.alpha": None}, "accuracy"),
    ({"hyper_parameters.coulomb.kmax": [2, 2]}, "kmax"),
    ({"hyper_parameters.coulomb.random_seed": 2 ** 33}, "UINT_MAX"),
    ({"hyper_parameters.coulomb.type": "PPPM"}, "coulomb.type"),
    ({"hyper_parameters.force_field.type": "EAM"}, "force_field.type"),
    ({"hyper_parameters.neighbor.type": "HALF"}, "neighbor.type"),
    ({"hyper_parameters.neighbor.type": "VERLET-SOG"}, "rbsog_level"),
    ({"hyper_parameters.coulomb.type": "RBSOG"}, "rbsog_level"),
    ({"hyper_parameters.coulomb.type": "RBSOG", "hyper_parameters.coulomb.rbsog_level": 7},
     "[0, 4]"),
    ({"hyper_parameters.coulomb.type": "RBSOG", "hyper_parameters.coulomb.rbsog_level": 1,
      "hyper_parameters.coulomb.rbsog_b": 2.0}, "Legacy"),
    ({"hyper_parameters.coulomb.type": "RBSOG", "hyper_parameters.coulomb.rbsog_level": 1,
      "hyper_parameters.coulomb.accuracy": None, "hyper_parameters.coulomb.alpha": None},
     None),
    ({"hyper_parameters.neighbor.type": "VERLET-SOG",
      "hyper_parameters.coulomb.type": "RBSOG", "hyper_parameters.coulomb.rbsog_level": 2},
     None),
    ({"hyper_parameters.coulomb.type": "RBSOG", "hyper_parameters.coulomb.rbsog_level": 1,
      "hyper_parameters.coulomb.rbsog_near_field": "sog"}, None),
    ({"hyper_parameters.coulomb.type": "RBSOG", "hyper_parameters.coulomb.rbsog_level": 1,
      "hyper_parameters.coulomb.rbsog_near_field": "gauss"}, "rbsog_near_field"),
    ({"hyper_parameters.coulomb.rbsog_near_field": "sog"}, "RBSOG only"),
    ({"hyper_parameters.neighbor.skin": -0.1}, "skin"),
    ({"init_configuration.read_data.unit": "SI"}, "unit"),
    ({"hyper_parameters.force_field.type": "CVFF"}, None),
    ({"hyper_parameters.force_field.type": "CVFF",
      "hyper_parameters.extend": {"special_bonds": [0, 0.5, 1]}}, None),
    ({"hyper_parameters.force_field.type": "CVFF",
      "hyper_parameters.extend": {"special_bonds": [0, 0, 0, 0]}}, "special_bonds"),
    ({"hyper_parameters.force_field.type": "CVFF",
      "hyper_parameters.extend": {"special_bonds": ["x"]}}, "special_bonds"),
    ({"hyper_parameters.force_field.type": "CVFF", "hyper_parameters.coulomb": None},
     "CVFF needs hyper_parameters.coulomb"),
])
def test_config_validation(edit, msg):
    rc, err = rbmd_amd.check_config(_cfg(**edit))
    if msg is None:
        assert rc == 0, err
    else:
        assert rc == 1 and msg in err, (rc, err)


def test_config_accepts_comments_and_bare_hyper_parameters():
    text = """{
      // RBMD allows comments
      "force_field": {"type": "LJ/CUT"}, /* block comment */
      "neighbor": {"type": "Verlet-List", "cut_off": 2.5, "skin": 0.3}
    }"""
    assert rbmd_amd.check_config(text)[0] == 0
    assert rbmd_amd.check_config("{ bad json")[0] == 1


@pytest.mark.parametrize("seed", [0, 1, 12345, 20261202, 4294967295])
def test_sampler_stream_equals_oracle(seed):
    L = [62.99605249474366] * 3
    for call in (0, 1, 2, 17, 999):
        a = rbmd_amd.rbe_sample_host(0.09, L, 200, seed, call)
        b = oracle.rbe_sample(seed, call, 200, 0.09, L)
        np.testing.assert_array_equal(a, b)


def test_sampler_other_boxes():
    for alpha, L, P in [(0.3, [12.0, 14.5, 9.25], 64), (0.01, [90.0] * 3, 500),
                        (1.5, [5.0] * 3, 17)]:
        a = rbmd_amd.rbe_sample_host(alpha, L, P, 77, 3)
        b = oracle.rbe_sample(77, 3, P, alpha, L)
        np.testing.assert_array_equal(a, b)
        assert not (a == 0).all(1).any()


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("exact", [False, True])
def test_sog_setup_equals_oracle(level, exact):
    L = [62.99605249474366, 60.0, 61.5]
    a = rbmd_amd.sog_setup_host(level, 3.5, L, exact=exact)
    b = oracle.sog_setup(level, 3.5, L, exact=exact)
    assert a["mmax"] == b["mmax"]
    for key in ("b", "sigma", "w0", "S", "S_npt"):
        if exact:
            assert a[key] == pytest.approx(b[key], rel=1e-12), key
        else:
            assert a[key] == b[key], key  # the same float arithmetic
    np.testing.assert_allclose(a["coef"], b["coef"], rtol=0 if not exact else 1e-13)
    np.testing.assert_allclose(a["coef_npt"], b["coef_npt"], rtol=0 if not exact else 1e-13)


@pytest.mark.parametrize("level,P,seed", [(0, 64, 20261202), (1, 512, 7), (3, 48, 0)])
def test_sog_sampler_stream_equals_oracle(level, P, seed):
    L = [29.24, 31.0, 27.5]
    for exact in (False, True):
        s = oracle.sog_setup(level, 3.5, L, exact=exact)
        for call in (0, 1, 5):
            m, idx = rbmd_amd.sog_sample_host(level, 3.5, L, P, seed, call, exact=exact)
            mo, io = oracle.sog_sample(seed, call, P, L, s)
            np.testing.assert_array_equal(m, mo)
            np.testing.assert_array_equal(idx, io)


@pytest.mark.parametrize("edit,msg", [
    (lambda hp: hp.update(precision="fp64"), "FP32"),
    (lambda hp: (hp["neighbor"].update(type="VERLET-SOG"), hp["coulomb"].update(rbsog_level=1)),
     "RBL and Verlet-List"),
    (lambda hp: hp["coulomb"].update(type="RBSOG", rbsog_level=1, rbsog_near_field="sog"),
     "erfc near field"),
    (lambda hp: hp["coulomb"].update(type="RBSOG", rbsog_level=1, coulomb_sample_num=510),
     "506"),
    (lambda hp: hp["coulomb"].update(type="EWALD"), "kmax"),
    (lambda hp: hp["coulomb"].update(type="EWALD", kmax=[5, 5, 5]), "512 k-vectors"),
    (lambda hp: (hp["coulomb"].pop("alpha"), hp["coulomb"].update(accuracy=1e-4)),
     "coulomb.alpha"),
    (lambda hp: hp["coulomb"].update(coulomb_sample_num=600), "sample_num"),
])
def test_batch_config_validation(edit, msg):
    """The replica batch rejects what it does not implement, before touching
    a device."""
    cfg = systems.ljsalt_config(P=100)
    edit(cfg["hyper_parameters"])
    lib = rbmd_amd.load_library()
    h = ctypes.c_void_p()
    import json

    rc = lib.rbmdx_batch_create(json.dumps(cfg).encode(), 0, ctypes.byref(h))
    try:
        assert rc == 1, rc  # RBMDX_E_CONFIG
        assert msg in lib.rbmdx_batch_last_error(h).decode()
    finally:
        lib.rbmdx_batch_destroy(h)


def test_replica_set_sizes_are_log_uniform_and_seeded():
    a = systems.replica_set(6, seed=3)
    b = systems.replica_set(6, seed=3)
    assert [s.n for s in a] == [s.n for s in b]
    assert all(500 <= s.n <= 8000 for s in a)
    assert np.array_equal(a[2].pos, b[2].pos)
