"""rbmd_amd: MI355X-native RBMD force path (cell list + RBL pair + RBE k-space).

The compute lives in ``librbmdx.so`` (hand-written HIP kernels for gfx950 and a
C++ host engine) behind the C ABI declared in ``include/rbmdx.h``.  This module
is a thin ctypes binding over that ABI; device arrays are passed as raw
pointers (torch tensors are used only for device memory and streams).

There is no CPU fallback: if the library is missing, :func:`load_library`
raises.
"""
import ctypes
import json
import os

import numpy as np

__all__ = ["load_library", "ForceContext", "RbmdxError", "LIB_PATH"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librbmdx.so")

RBMDX_FLAG_ENERGY = 1
RBMDX_FLAG_SYNC = 2
RBMDX_FLAG_REBUILD = 4
RBMDX_FLAG_TIMING = 8
RBMDX_FLAG_EAGER = 16
RBMDX_FLAG_VJP_NEUTRAL_Q = 32
RBMDX_FLAG_PER_REPLICA = 64

_STATUS = {
    0: "RBMDX_OK",
    1: "RBMDX_E_CONFIG",
    2: "RBMDX_E_CAPACITY",
    3: "RBMDX_E_HIP",
    4: "RBMDX_E_NONFINITE",
    5: "RBMDX_E_ARG",
    6: "RBMDX_E_STATE",
}


class RbmdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class Thermo(ctypes.Structure):
    _fields_ = [
        ("e_vdwl", ctypes.c_double),
        ("e_coul", ctypes.c_double),
        ("e_kspace", ctypes.c_double),
        ("virial", ctypes.c_double * 6),
        ("virial_kspace", ctypes.c_double * 6),
        ("mean_force", ctypes.c_double * 3),
        ("steps", ctypes.c_int64),
        ("rebuilds", ctypes.c_int64),
        ("rbl_frequency", ctypes.c_int32),
        ("num_k", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("rbe_S", ctypes.c_double),
        ("pairs_a", ctypes.c_int64),
        ("pairs_b", ctypes.c_int64),
        ("atoms", ctypes.c_int64),
        ("charged", ctypes.c_int64),
        ("e_kspace_exact", ctypes.c_double),
    ]

    def as_dict(self):
        return {
            "e_vdwl": self.e_vdwl,
            "e_coul": self.e_coul,
            "e_kspace": self.e_kspace,
            "virial": list(self.virial),
            "virial_kspace": list(self.virial_kspace),
            "mean_force": list(self.mean_force),
            "steps": self.steps,
            "rebuilds": self.rebuilds,
            "rbl_frequency": self.rbl_frequency,
            "num_k": self.num_k,
            "alpha": self.alpha,
            "rbe_S": self.rbe_S,
            "pairs_a": self.pairs_a,
            "pairs_b": self.pairs_b,
            "atoms": self.atoms,
            "charged": self.charged,
            "e_kspace_exact": self.e_kspace_exact,
        }


class MdState(ctypes.Structure):
    _fields_ = [
        ("step", ctypes.c_int64),
        ("mode", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("temperature", ctypes.c_double),
        ("ke", ctypes.c_double),
        ("pe", ctypes.c_double),
        ("t_target", ctypes.c_double),
        ("xi", ctypes.c_double),
        ("eta", ctypes.c_double * 3),
        ("eta_dot", ctypes.c_double * 3),
        ("eta_mass", ctypes.c_double * 3),
        ("thermostat_energy", ctypes.c_double),
    ]

    def as_dict(self):
        out = {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}
        for f in ("eta", "eta_dot", "eta_mass"):
            out[f] = list(out[f])
        out["mode"] = ("NVE", "NVE-energy-stable", "NVT-Nose-Hoover")[self.mode]
        return out


# Every symbol include/rbmdx.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
SIGNATURES = {
    "rbmdx_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_P)]),
    "rbmdx_destroy": (None, [_P]),
    "rbmdx_last_error": (ctypes.c_char_p, [_P]),
    "rbmdx_set_system": (ctypes.c_int, [_P, _I64, _DP, _DP, _IP, _IP, _DP, _I32, _DP, _DP]),
    "rbmdx_set_charges": (ctypes.c_int, [_P, _DP]),
    "rbmdx_set_special": (ctypes.c_int, [_P, _I32, _IP, _IP]),
    "rbmdx_set_lj": (ctypes.c_int, [_P, _DP, _DP]),
    "rbmdx_force_vjp": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                       ctypes.c_uint32, _P]),
    "rbmdx_force": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _I32,
                                   ctypes.c_uint32, _P]),
    "rbmdx_force_noalloc": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _I32,
                                           ctypes.c_uint32, _P]),
    "rbmdx_thermo": (ctypes.c_int, [_P, ctypes.POINTER(Thermo)]),
    "rbmdx_rbe_sample": (ctypes.c_int, [_P, _I64, _IP]),
    "rbmdx_num_samples": (_I32, [_P]),
    "rbmdx_export_pairs": (_I64, [_P, _I64, _P, _P, _P, _IP, _IP, _IP, _I64]),
    "rbmdx_export_candidates": (_I64, [_P, _IP, _IP, _IP, _I64]),
    "rbmdx_kernel_times": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_char_p), _DP,
                                          ctypes.c_int]),
    "rbmdx_set_masses": (ctypes.c_int, [_P, _DP]),
    "rbmdx_force_leapfrog": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                            ctypes.c_double, ctypes.c_uint32, _P]),
    "rbmdx_md_init": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                     ctypes.c_uint32, _P]),
    "rbmdx_md_run": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _P, ctypes.c_uint32, _P]),
    "rbmdx_md_state": (ctypes.c_int, [_P, ctypes.POINTER(MdState)]),
    "rbmdx_rbe_sample_host": (ctypes.c_int, [ctypes.c_double, _DP, _I32, ctypes.c_uint32,
                                             _I64, _IP]),
    "rbmdx_check_config": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _I32]),
    "rbmdx_sog_setup_host": (ctypes.c_int, [_I32, ctypes.c_double, _DP, _I32, _DP, _DP,
                                            _DP]),
    "rbmdx_sog_sample_host": (ctypes.c_int, [_I32, ctypes.c_double, _DP, _I32, _I32,
                                             ctypes.c_uint32, _I64, _IP, _IP]),
    "rbmdx_batch_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_P)]),
    "rbmdx_batch_destroy": (None, [_P]),
    "rbmdx_batch_last_error": (ctypes.c_char_p, [_P]),
    "rbmdx_batch_set_systems": (ctypes.c_int, [_P, _I32, _IP, _DP, _DP, _IP, _IP, _DP, _I32,
                                               _DP, _DP]),
    "rbmdx_batch_set_special": (ctypes.c_int, [_P, _I32, _IP, _IP]),
    "rbmdx_batch_force": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P,
                                         ctypes.c_uint32, _P]),
    "rbmdx_batch_energies": (ctypes.c_int, [_P, _DP]),
    "rbmdx_batch_autotune": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _P,
                                            ctypes.c_char_p, _P, _DP]),
    "rbmdx_batch_set_shape": (ctypes.c_int, [_P, _IP]),
    "rbmdx_batch_get_shape": (ctypes.c_int, [_P, _IP]),
    "rbmdx_batch_num_samples": (_I32, [_P]),
    "rbmdx_data_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _IP,
                                       ctypes.POINTER(_P)]),
    "rbmdx_write_dump": (ctypes.c_int, [ctypes.c_char_p, _I32, _I64, _I64, _DP, _DP, _IP,
                                        _IP, _DP, _DP, _IP, _DP]),
    "rbmdx_write_force_samples": (ctypes.c_int, [ctypes.c_char_p, _I32, _I64, _I64, _IP,
                                                 _DP, _DP]),
    "rbmdx_data_destroy": (None, [_P]),
    "rbmdx_data_replicate": (ctypes.c_int, [_P, _IP, _I32]),
    "rbmdx_data_gauss_velocities": (ctypes.c_int, [_P, ctypes.c_double, _I32,
                                                   ctypes.c_char_p]),
    "rbmdx_data_last_error": (ctypes.c_char_p, [_P]),
    "rbmdx_data_info": (ctypes.c_int, [_P, ctypes.POINTER(_I64), _IP, _IP, _IP]),
    "rbmdx_data_get": (ctypes.c_int, [_P, _DP, _DP, _IP, _IP, _DP, _DP, _DP, _IP, _DP, _DP,
                                      _DP]),
    "rbmdx_data_special": (ctypes.c_int, [_P, _IP, _IP]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load librbmdx.so and bind every ABI symbol (raises if anything is missing)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: run __graft_entry__.build() first")
    # torch bundles its own HIP runtime with the same soname as ROCm's
    # (libamdhip64.so.7).  Loading torch first makes librbmdx.so bind to that
    # already-loaded runtime; otherwise the process would hold two HIP runtimes
    # and torch's handles would be invalid in ours (and vice versa).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is absent
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def rbe_sample_host(alpha, L, P, base_seed, call_index):
    """Host k-sample of evaluation `call_index` (P x 3 int32), no device needed."""
    lib = load_library()
    Ld = np.ascontiguousarray(L, dtype=np.float64)
    m = np.zeros(3 * P, dtype=np.int32)
    rc = lib.rbmdx_rbe_sample_host(float(alpha), _dptr(Ld), int(P), int(base_seed) & 0xFFFFFFFF,
                                   int(call_index), _iptr(m))
    if rc != 0:
        raise RbmdxError(rc, "rbmdx_rbe_sample_host failed")
    return m.reshape(P, 3)


def sog_setup_host(level, cut_off, L, exact=False):
    """RBSOG set-up: dict(b, sigma, w0, S, S_npt, mmax, coef, coef_npt)."""
    lib = load_library()
    Ld = np.ascontiguousarray(L, dtype=np.float64)
    scal = np.zeros(6)
    coef = np.zeros(128)
    coefn = np.zeros(128)
    rc = lib.rbmdx_sog_setup_host(int(level), float(cut_off), _dptr(Ld), 1 if exact else 0,
                                  _dptr(scal), _dptr(coef), _dptr(coefn))
    if rc != 0:
        raise RbmdxError(rc, "rbmdx_sog_setup_host failed")
    M = int(scal[5])
    return {"b": scal[0], "sigma": scal[1], "w0": scal[2], "S": scal[3], "S_npt": scal[4],
            "mmax": M, "coef": coef[:M].copy(), "coef_npt": coefn[:M].copy()}


def sog_sample_host(level, cut_off, L, P, base_seed, call_index, exact=False):
    """RBSOG k-sample of evaluation call_index: (m (P,3), idx (P,))."""
    lib = load_library()
    Ld = np.ascontiguousarray(L, dtype=np.float64)
    m = np.zeros(3 * P, dtype=np.int32)
    idx = np.zeros(P, dtype=np.int32)
    rc = lib.rbmdx_sog_sample_host(int(level), float(cut_off), _dptr(Ld), 1 if exact else 0,
                                   int(P), int(base_seed) & 0xFFFFFFFF, int(call_index),
                                   _iptr(m), _iptr(idx))
    if rc != 0:
        raise RbmdxError(rc, "rbmdx_sog_sample_host failed")
    return m.reshape(P, 3), idx


def check_config(config):
    """Validate a JSON configuration without a device: (status, message)."""
    lib = load_library()
    text = config if isinstance(config, str) else json.dumps(config)
    buf = ctypes.create_string_buffer(512)
    rc = lib.rbmdx_check_config(text.encode(), buf, 512)
    return rc, buf.value.decode()


def _dptr(a):
    return a.ctypes.data_as(_DP)


def _iptr(a):
    return a.ctypes.data_as(_IP)


def _tensor_ptr(t):
    if t is None:
        return None
    return t.data_ptr()


def _current_stream(device):
    """The raw HIP stream of torch's current stream on `device` (the int that
    ``torch.cuda.current_stream(device).cuda_stream`` gives, without building
    the Stream object on every call)."""
    import torch

    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return raw(device.index if device.index is not None else torch.cuda.current_device())
    return torch.cuda.current_stream(device).cuda_stream


class ForceContext:
    """One periodic system on one GPU (wraps an ``rbmdx_ctx``)."""

    def __init__(self, config, device=0):
        self.lib = load_library()
        text = config if isinstance(config, str) else json.dumps(config)
        h = _P()
        rc = self.lib.rbmdx_create(text.encode(), int(device), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.rbmdx_last_error(h).decode() if h else ""
            self.lib.rbmdx_destroy(h)
            self._h = None
            raise RbmdxError(rc, msg)
        self.device = device
        self.config = json.loads(text)
        hp = self.config.get("hyper_parameters", self.config)
        self.fp64 = hp.get("precision", "fp32") == "fp64"
        self.n = 0
        self.calls = 0

    def close(self):
        if getattr(self, "_h", None):
            self.lib.rbmdx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RbmdxError(rc, self.lib.rbmdx_last_error(self._h).decode())

    def set_system(self, box_lo, box_len, types, ids, charges, eps, sigma):
        types = np.ascontiguousarray(types, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        charges = np.ascontiguousarray(charges, dtype=np.float64)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        lo = np.ascontiguousarray(box_lo, dtype=np.float64)
        L = np.ascontiguousarray(box_len, dtype=np.float64)
        self.n = len(types)
        self._check(self.lib.rbmdx_set_system(
            self._h, self.n, _dptr(lo), _dptr(L), _iptr(types), _iptr(ids),
            _dptr(charges), len(eps), _dptr(eps), _dptr(sigma)))

    def set_charges(self, charges):
        q = np.ascontiguousarray(charges, dtype=np.float64)
        self._check(self.lib.rbmdx_set_charges(self._h, _dptr(q)))

    def set_special(self, nspecial, special):
        """CVFF special-bond topology: nspecial (N, 3) cumulative 1-2/1-3/1-4
        counts, special (N, maxspecial) partner global ids (caller order)."""
        ns = np.ascontiguousarray(nspecial, dtype=np.int32).reshape(-1, 3)
        sp = np.ascontiguousarray(special, dtype=np.int32)
        ms = sp.shape[1] if sp.ndim == 2 else 0
        sp = np.ascontiguousarray(sp.reshape(-1))
        self._check(self.lib.rbmdx_set_special(self._h, ms, _iptr(ns), _iptr(sp)))

    def force(self, step, px, py, pz, fx, fy, fz, m_samples=None, energy=False,
              sync=False, rebuild=False, timing=False, stream=None, num_k=None):
        """Total force into fx/fy/fz (device tensors, caller order).  m_samples:
        explicit k-sample (int32 device tensor) of num_k vectors (default
        numel // 3; RBSOG: num_k = P m-vectors followed by P chain indices)."""
        flags = ((RBMDX_FLAG_ENERGY if energy else 0) | (RBMDX_FLAG_SYNC if sync else 0)
                 | (RBMDX_FLAG_REBUILD if rebuild else 0)
                 | (RBMDX_FLAG_TIMING if timing else 0))
        nk = 0 if m_samples is None else int(num_k or m_samples.numel() // 3)
        if stream is None:
            stream = _current_stream(px.device)
        self._check(self.lib.rbmdx_force(
            self._h, int(step), _tensor_ptr(px), _tensor_ptr(py), _tensor_ptr(pz),
            _tensor_ptr(fx), _tensor_ptr(fy), _tensor_ptr(fz),
            _tensor_ptr(m_samples), nk, flags, ctypes.c_void_p(stream)))
        self.calls += 1

    def force_noalloc(self, step, px, py, pz, fx, fy, fz, m_samples=None, energy=False,
                      sync=False, rebuild=False, stream=None, num_k=None):
        """force() through rbmdx_force_noalloc: no device allocation in the
        call, the per-evaluation scratch in the library's static per-device
        array (after one force() call has sized the lists)."""
        flags = ((RBMDX_FLAG_ENERGY if energy else 0) | (RBMDX_FLAG_SYNC if sync else 0)
                 | (RBMDX_FLAG_REBUILD if rebuild else 0))
        nk = 0 if m_samples is None else int(num_k or m_samples.numel() // 3)
        if stream is None:
            stream = _current_stream(px.device)
        self._check(self.lib.rbmdx_force_noalloc(
            self._h, int(step), _tensor_ptr(px), _tensor_ptr(py), _tensor_ptr(pz),
            _tensor_ptr(fx), _tensor_ptr(fy), _tensor_ptr(fz),
            _tensor_ptr(m_samples), nk, flags, ctypes.c_void_p(stream)))
        self.calls += 1

    def set_masses(self, masses):
        m = np.ascontiguousarray(masses, dtype=np.float64)
        self._check(self.lib.rbmdx_set_masses(self._h, _dptr(m)))

    def set_lj(self, eps, sigma):
        e = np.ascontiguousarray(eps, dtype=np.float64)
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        self._check(self.lib.rbmdx_set_lj(self._h, _dptr(e), _dptr(s)))

    def force_vjp(self, step, px, py, pz, gx, gy, gz, dx, dy, dz, dparam, sync=False,
                  neutral_q=False, stream=None):
        """Vector-Jacobian product of the last evaluation's force with the
        cotangent (gx, gy, gz): dL/dr into (dx, dy, dz) and [dL/dQ | dL/deps |
        dL/dsigma] per type into dparam (float64 device tensor of 3 ntypes).
        dL/dQ shifts the charged atoms of a type; neutral_q=True shifts every
        atom of the type (neutral ones included)."""
        flags = ((RBMDX_FLAG_SYNC if sync else 0)
                 | (RBMDX_FLAG_VJP_NEUTRAL_Q if neutral_q else 0))
        if stream is None:
            stream = _current_stream(px.device)
        args = [_tensor_ptr(t) for t in (px, py, pz, gx, gy, gz, dx, dy, dz)]
        self._check(self.lib.rbmdx_force_vjp(
            self._h, int(step), *args, ctypes.cast(_tensor_ptr(dparam),
                                                   ctypes.POINTER(ctypes.c_double)),
            flags, ctypes.c_void_p(stream)))
        self.calls += 1

    def force_leapfrog(self, step, px, py, pz, vx, vy, vz, fx, fy, fz, dt, energy=False,
                       sync=False, timing=False, eager=False, stream=None):
        """Force evaluation fused with v += dt F/m, x += dt v (in place)."""
        flags = ((RBMDX_FLAG_ENERGY if energy else 0) | (RBMDX_FLAG_SYNC if sync else 0)
                 | (RBMDX_FLAG_TIMING if timing else 0) | (RBMDX_FLAG_EAGER if eager else 0))
        if stream is None:
            stream = _current_stream(px.device)
        args = [_tensor_ptr(t) for t in (px, py, pz, vx, vy, vz, fx, fy, fz)]
        self._check(self.lib.rbmdx_force_leapfrog(self._h, int(step), *args, float(dt), flags,
                                                  ctypes.c_void_p(stream)))
        self.calls += 1

    def _md_args(self, x, v, img, f):
        ptrs = [_tensor_ptr(t) for t in (*x, *v)]
        ptrs += [None if img is None else _tensor_ptr(t) for t in (img if img else (None,) * 3)]
        ptrs += [_tensor_ptr(t) for t in f]
        return ptrs

    def md_init(self, x, v, f, image=None, energy=False, stream=None):
        """Molecular-dynamics start: wrap x = (px, py, pz) into the box (image
        = (ix, iy, iz) int32 tensors updated, or None), evaluate the force of
        step 0 into f and set the thermostat up.  Settings come from the
        configuration's execution section; masses from set_masses."""
        if stream is None:
            stream = _current_stream(x[0].device)
        flags = RBMDX_FLAG_ENERGY if energy else 0
        self._check(self.lib.rbmdx_md_init(self._h, *self._md_args(x, v, image, f), flags,
                                           ctypes.c_void_p(stream)))

    def md_run(self, nsteps, x, v, f, image=None, energy=False, stream=None):
        """nsteps velocity-Verlet steps with the configured thermostat (in place)."""
        if stream is None:
            stream = _current_stream(x[0].device)
        flags = RBMDX_FLAG_ENERGY if energy else 0
        self._check(self.lib.rbmdx_md_run(self._h, int(nsteps),
                                          *self._md_args(x, v, image, f), flags,
                                          ctypes.c_void_p(stream)))
        self.calls += int(nsteps)

    def md_state(self):
        s = MdState()
        self._check(self.lib.rbmdx_md_state(self._h, ctypes.byref(s)))
        return s.as_dict()

    def thermo(self):
        t = Thermo()
        self._check(self.lib.rbmdx_thermo(self._h, ctypes.byref(t)))
        return t.as_dict()

    def num_samples(self):
        return int(self.lib.rbmdx_num_samples(self._h))

    def rbe_sample(self, call_index, sog=False):
        """k-sample of evaluation call_index: (P,3) m-vectors; with sog=True
        (an RBSOG context) the pair (m (P,3), virial-chain indices (P,))."""
        P = self.num_samples()
        m = np.zeros((4 if sog else 3) * P, dtype=np.int32)
        self._check(self.lib.rbmdx_rbe_sample(self._h, int(call_index), _iptr(m)))
        if sog:
            return m[:3 * P].reshape(P, 3), m[3 * P:].copy()
        return m.reshape(P, 3)

    def _export(self, fn, *args, cap):
        gi = np.zeros(cap, dtype=np.int32)
        gj = np.zeros(cap, dtype=np.int32)
        cls = np.zeros(cap, dtype=np.int32)
        n = fn(self._h, *args, _iptr(gi), _iptr(gj), _iptr(cls), cap)
        if n < 0:
            self._check(int(-n))
        if n > cap:
            return self._export(fn, *args, cap=int(n))
        return gi[:n], gj[:n], cls[:n]

    def export_pairs(self, step, px, py, pz, cap=1 << 20):
        """(gid_i, gid_j, cls) of every directed pair the RBL classification keeps."""
        return self._export(self.lib.rbmdx_export_pairs, int(step), _tensor_ptr(px),
                            _tensor_ptr(py), _tensor_ptr(pz), cap=cap)

    def export_candidates(self, cap=1 << 20):
        return self._export(self.lib.rbmdx_export_candidates, cap=cap)

    def kernel_times(self):
        names = (ctypes.c_char_p * 8)()
        us = np.zeros(8, dtype=np.float64)
        n = self.lib.rbmdx_kernel_times(self._h, names, _dptr(us), 8)
        if n < 0:
            self._check(-n)
        return {names[k].decode(): float(us[k]) for k in range(n)}


class BatchContext:
    """Many independent periodic boxes of one force field on one GPU (wraps an
    ``rbmdx_batch``).  Positions and forces are the replicas' atoms
    concatenated in replica order."""

    def __init__(self, config, device=0):
        self.lib = load_library()
        text = config if isinstance(config, str) else json.dumps(config)
        h = _P()
        rc = self.lib.rbmdx_batch_create(text.encode(), int(device), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = self.lib.rbmdx_batch_last_error(h).decode() if h else ""
            self.lib.rbmdx_batch_destroy(h)
            self._h = None
            raise RbmdxError(rc, msg)
        self.device = device
        self.nrep = 0
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            self.lib.rbmdx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RbmdxError(rc, self.lib.rbmdx_batch_last_error(self._h).decode())

    def set_systems(self, systems, eps=None, sigma=None):
        """systems: objects with box_lo, box_len, types, ids, charges (and eps,
        sigma per type, taken from the first unless given)."""
        nat = np.array([len(s.types) for s in systems], dtype=np.int32)
        lo = np.ascontiguousarray(np.concatenate([np.asarray(s.box_lo, dtype=np.float64)
                                                  for s in systems]))
        L = np.ascontiguousarray(np.concatenate([np.asarray(s.box_len, dtype=np.float64)
                                                 for s in systems]))
        types = np.ascontiguousarray(np.concatenate([s.types for s in systems]), dtype=np.int32)
        ids = np.ascontiguousarray(np.concatenate([s.ids for s in systems]), dtype=np.int32)
        q = np.ascontiguousarray(np.concatenate([s.charges for s in systems]), dtype=np.float64)
        e = np.ascontiguousarray(systems[0].eps if eps is None else eps, dtype=np.float64)
        sg = np.ascontiguousarray(systems[0].sigma if sigma is None else sigma,
                                  dtype=np.float64)
        self._check(self.lib.rbmdx_batch_set_systems(
            self._h, len(systems), _iptr(nat), _dptr(lo), _dptr(L), _iptr(types), _iptr(ids),
            _dptr(q), len(e), _dptr(e), _dptr(sg)))
        self.nrep = len(systems)
        self.n = int(nat.sum())
        self.offsets = np.concatenate([[0], np.cumsum(nat)])

    def set_special(self, systems):
        """CVFF special bonds of every replica from the systems' nspecial (N, 3)
        and special (N, maxspecial) arrays (partner ids within the replica)."""
        ms = max(np.asarray(s.special).shape[1] for s in systems)
        ns = np.ascontiguousarray(np.concatenate([s.nspecial for s in systems]), dtype=np.int32)
        sp = np.zeros((len(ns), ms), dtype=np.int32)
        o = 0
        for s in systems:
            a = np.asarray(s.special, dtype=np.int32)
            sp[o:o + len(a), :a.shape[1]] = a
            o += len(a)
        self._check(self.lib.rbmdx_batch_set_special(self._h, int(ms), _iptr(ns),
                                                     _iptr(np.ascontiguousarray(sp))))

    def num_samples(self):
        return int(self.lib.rbmdx_batch_num_samples(self._h))

    def force(self, step, px, py, pz, fx, fy, fz, m_samples=None, energy=False, sync=False,
              eager=False, per_replica=False, stream=None):
        flags = ((RBMDX_FLAG_ENERGY if energy else 0) | (RBMDX_FLAG_SYNC if sync else 0)
                 | (RBMDX_FLAG_EAGER if eager else 0)
                 | (RBMDX_FLAG_PER_REPLICA if per_replica else 0))
        if stream is None:
            stream = _current_stream(px.device)
        self._check(self.lib.rbmdx_batch_force(
            self._h, int(step), _tensor_ptr(px), _tensor_ptr(py), _tensor_ptr(pz),
            _tensor_ptr(fx), _tensor_ptr(fy), _tensor_ptr(fz), _tensor_ptr(m_samples), flags,
            ctypes.c_void_p(stream)))

    def energies(self):
        """(nrep, 9): E_vdwl, E_coul, E_kspace, virial xx yy zz xy xz yz."""
        out = np.zeros(9 * self.nrep, dtype=np.float64)
        self._check(self.lib.rbmdx_batch_energies(self._h, _dptr(out)))
        return out.reshape(self.nrep, 9)

    def autotune(self, step, px, py, pz, fx, fy, fz, m_samples=None, cache_path=None,
                 stream=None):
        """Measure and keep the fastest launch shape; returns (shape, us)."""
        if stream is None:
            stream = _current_stream(px.device)
        us = ctypes.c_double(0.0)
        self._check(self.lib.rbmdx_batch_autotune(
            self._h, int(step), _tensor_ptr(px), _tensor_ptr(py), _tensor_ptr(pz),
            _tensor_ptr(fx), _tensor_ptr(fy), _tensor_ptr(fz), _tensor_ptr(m_samples),
            None if cache_path is None else str(cache_path).encode(), ctypes.c_void_p(stream),
            ctypes.byref(us)))
        return self.shape(), us.value

    def shape(self):
        s = np.zeros(4, dtype=np.int32)
        self._check(self.lib.rbmdx_batch_get_shape(self._h, _iptr(s)))
        return {"bw": int(s[0]), "pair_block": int(s[1]), "grid_per_cu": int(s[2]),
                "qlanes": int(s[3])}

    def set_shape(self, bw, pair_block, grid_per_cu, qlanes):
        s = np.array([bw, pair_block, grid_per_cu, qlanes], dtype=np.int32)
        self._check(self.lib.rbmdx_batch_set_shape(self._h, _iptr(s)))


def read_data(path, atom_style="full", replicate=None, velocity_type=None,
              temperature=None, velocity_seed=12345, unit="LJ", bond_periodic=True):
    """A LAMMPS data file as RBMD reads it (rbmdx_data_read): a
    ``systems.System`` with 0-based ids and types, per-type eps / sigma from
    Pair Coeffs and masses, velocities and image flags, and the special-bond
    topology (``nspecial`` (N, 3), ``special`` (N, maxspecial)) built from the
    bonds, angles and dihedrals.  ``velocity_type="GAUSS"`` creates velocities
    at ``temperature`` from ``velocity_seed`` when the file has none
    (rbmdx_data_gauss_velocities).  ``replicate`` tiles the box; with
    ``bond_periodic`` (RBMD's default) bonded partners come from the tile of
    their nearest image, otherwise from their own tile after unwrapping by
    image flags (rbmdx_data_replicate)."""
    from . import systems

    lib = load_library()
    h = _P()
    rep = None if replicate is None else np.ascontiguousarray(replicate, dtype=np.int32)
    rc = lib.rbmdx_data_read(str(path).encode(), atom_style.encode(), None, ctypes.byref(h))
    try:
        if rc != 0:
            raise RbmdxError(rc, lib.rbmdx_data_last_error(h).decode() if h else "")
        if rep is not None:
            rc = lib.rbmdx_data_replicate(h, _iptr(rep), int(bool(bond_periodic)))
            if rc != 0:
                raise RbmdxError(rc, lib.rbmdx_data_last_error(h).decode())
        if velocity_type is not None and velocity_type.upper() not in ("NONE", ""):
            if velocity_type.upper() not in ("GAUSS", "GAUSSIAN"):
                raise RbmdxError(1, "Unsupported velocity_type: " + velocity_type)
            rc = lib.rbmdx_data_gauss_velocities(
                h, float(temperature if temperature is not None else float("nan")),
                int(velocity_seed), unit.encode())
            if rc != 0:
                raise RbmdxError(rc, lib.rbmdx_data_last_error(h).decode())
        n, nt, ms, fl = ctypes.c_int64(0), ctypes.c_int32(0), ctypes.c_int32(0), \
            ctypes.c_int32(0)
        lib.rbmdx_data_info(h, ctypes.byref(n), ctypes.byref(nt), ctypes.byref(ms),
                            ctypes.byref(fl))
        n, nt, ms = n.value, nt.value, ms.value
        lo, L = np.zeros(3), np.zeros(3)
        types, ids = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        q, pos, vel = np.zeros(n), np.zeros(3 * n), np.zeros(3 * n)
        img = np.zeros(3 * n, dtype=np.int32)
        mass, eps, sig = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        lib.rbmdx_data_get(h, _dptr(lo), _dptr(L), _iptr(types), _iptr(ids), _dptr(q),
                           _dptr(pos), _dptr(vel), _iptr(img), _dptr(mass), _dptr(eps),
                           _dptr(sig))
        nsp = np.zeros(3 * n, dtype=np.int32)
        sp = np.zeros(n * ms, dtype=np.int32)
        lib.rbmdx_data_special(h, _iptr(nsp), _iptr(sp))
    finally:
        lib.rbmdx_data_destroy(h)
    s = systems.Water(pos.reshape(n, 3), lo, L, types, ids, q, eps, sig,
                      vel=vel.reshape(n, 3), mass=mass, name=os.path.basename(str(path)),
                      nspecial=nsp.reshape(n, 3), special=sp.reshape(n, ms))
    s.image = img.reshape(n, 3)
    s.has_pair_coeffs = bool(fl.value & 1)
    s.has_velocities = bool(fl.value & 2)
    s.has_bonds = bool(fl.value & 4)
    return s


def write_dump(path, timestep, box_lo, box_len, ids, types, pos, charges=None, image=None,
               vel=None, append=False):
    """Append (or write) one LAMMPS-style trajectory frame (rbmdx_write_dump)."""
    lib = load_library()
    n = len(ids)
    c = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t)
    ids_, types_ = c(ids, np.int32), c(types, np.int32)
    pos_, q_, img_, vel_ = c(pos, np.float64), c(charges, np.float64), c(image, np.int32), \
        c(vel, np.float64)
    lo, L = c(box_lo, np.float64), c(box_len, np.float64)
    p = lambda a: None if a is None else (_iptr(a) if a.dtype == np.int32 else _dptr(a))
    rc = lib.rbmdx_write_dump(str(path).encode(), int(bool(append)), int(timestep), n, p(lo),
                              p(L), p(ids_), p(types_), p(q_), p(pos_), p(img_), p(vel_))
    if rc != 0:
        raise RbmdxError(rc, f"cannot write {path}")


def write_force_samples(path, sample, ids, pos, force, append=False):
    """Rows sample,id,x,y,z,fx,fy,fz of one sample (rbmdx_write_force_samples)."""
    lib = load_library()
    ids_ = np.ascontiguousarray(ids, dtype=np.int32)
    pos_ = np.ascontiguousarray(pos, dtype=np.float64)
    f_ = np.ascontiguousarray(force, dtype=np.float64)
    rc = lib.rbmdx_write_force_samples(str(path).encode(), int(bool(append)), int(sample),
                                       len(ids_), _iptr(ids_), _dptr(pos_), _dptr(f_))
    if rc != 0:
        raise RbmdxError(rc, f"cannot write {path}")
