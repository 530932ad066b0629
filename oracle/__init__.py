"""CPU oracle for the RBMD force path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package.  It wraps liboracle.so (oracle/oracle.cpp, built by the Makefile
in this directory), a plain CPU restatement of the reference algorithm; see the
header of oracle.cpp for what each function restates and how it is pinned.
Two pure-Python modules sit beside it: oracle.md (the molecular-dynamics step
around the force) and oracle.gauss (GAUSS velocity creation), each with its
own header of what it restates and how it is pinned.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None
_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)
_L = ctypes.POINTER(ctypes.c_long)
_U32 = ctypes.POINTER(ctypes.c_uint32)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_pairs_f32.restype = ctypes.c_long
        L.orc_pairs_f32.argtypes = [ctypes.c_int, _F, _F, _F, _D, ctypes.c_float, _L, _I]
        L.orc_rbl_pairs_f32.restype = ctypes.c_long
        L.orc_rbl_pairs_f32.argtypes = [ctypes.c_int, _F, _F, _F, _I, _D, _D, _L, _I, _I]
        L.orc_mix64.restype = ctypes.c_uint64
        L.orc_mix64.argtypes = [ctypes.c_uint64]
        L.orc_rbl_selected.restype = ctypes.c_int
        L.orc_rbl_selected.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                       ctypes.c_int]
        L.orc_rbl_frequency.restype = ctypes.c_int
        L.orc_rbl_frequency.argtypes = [ctypes.c_long, _D, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int]
        L.orc_pair_force.restype = None
        L.orc_pair_force.argtypes = [ctypes.c_int, ctypes.c_int, _F, _F, _F, _F, _I, _I,
                                     ctypes.c_int, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
        L.orc_force_f64.restype = None
        L.orc_force_f64.argtypes = [ctypes.c_int, _D, _D, _D, _I, _I, ctypes.c_int, _D, _D,
                                    _D, _D, _D, ctypes.c_int, _I, _D, _D, _D, _D]
        L.orc_force_vjp_f64.restype = None
        L.orc_force_vjp_f64.argtypes = [ctypes.c_int, _D, _D, _D, _I, _I, ctypes.c_int, _D,
                                        _D, _D, _D, _D, ctypes.c_int, _I, _D, _D, _D, _D,
                                        _D, _D]
        L.orc_cpu_force_steps.restype = None
        L.orc_cpu_force_steps.argtypes = [ctypes.c_int, _F, _F, _F, _F, _I, _I, ctypes.c_int,
                                          _D, _D, _D, _D, ctypes.c_double, ctypes.c_int, _I,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _D, _D, _D]
        L.orc_pair_force_special.restype = None
        L.orc_pair_force_special.argtypes = [ctypes.c_int, ctypes.c_int, _F, _F, _F, _F, _I,
                                             _I, ctypes.c_int, _D, _D, _D, _D, ctypes.c_int,
                                             _I, _I, _D, _D, _D, _D, _D, _D, _D]
        L.orc_pair_force_sog.restype = None
        L.orc_pair_force_sog.argtypes = [ctypes.c_int, ctypes.c_int, _F, _F, _F, _F, _I, _I,
                                         ctypes.c_int, _D, _D, _D, _D, ctypes.c_int, _I, _I,
                                         _D, _D, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_double, _D, _D, _D, _D, _D,
                                         _D]
        L.orc_verlet_force_f64.restype = None
        L.orc_verlet_force_f64.argtypes = [ctypes.c_int, _D, _D, _D, _D, _I, ctypes.c_int,
                                           _D, _D, _D, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           _D, _D, _D, _D, _D]
        L.orc_kspace_f64.restype = None
        L.orc_kspace_f64.argtypes = [ctypes.c_int, _D, _D, _D, _D, _D, ctypes.c_int, _I,
                                     _D, ctypes.c_double, ctypes.c_double, _D, _D, _D, _D,
                                     _D, _D]
        L.orc_kspace_self.restype = ctypes.c_double
        L.orc_kspace_self.argtypes = [ctypes.c_double] * 5
        L.orc_rbe_S.restype = ctypes.c_double
        L.orc_rbe_S.argtypes = [ctypes.c_int, ctypes.c_double, _D]
        L.orc_gauss_sum_direct.restype = ctypes.c_double
        L.orc_gauss_sum_direct.argtypes = [ctypes.c_double, _D, ctypes.c_int]
        L.orc_ewald_real_f64.restype = None
        L.orc_ewald_real_f64.argtypes = [ctypes.c_int, _D, _D, _D, _D, _D, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, _D, _D, _D, _D]
        L.orc_alpha_from_accuracy.restype = ctypes.c_double
        L.orc_alpha_from_accuracy.argtypes = [ctypes.c_double] * 6
        L.orc_rbe_sample.restype = None
        L.orc_rbe_sample.argtypes = [ctypes.c_uint32, ctypes.c_long, ctypes.c_int,
                                     ctypes.c_double, _D, _I]
        L.orc_mt19937.restype = None
        L.orc_mt19937.argtypes = [ctypes.c_uint32, ctypes.c_long, ctypes.c_int, _U32]
        L.orc_force_step.restype = None
        L.orc_force_step.argtypes = [ctypes.c_int, _F, _F, _F, _F, _I, _I, ctypes.c_int,
                                     _D, _D, _D, _D, ctypes.c_int, _I, ctypes.c_double,
                                     _D, _D, _D]
        c_int, c_dbl = ctypes.c_int, ctypes.c_double
        L.orc_sog_preset.restype = None
        L.orc_sog_preset.argtypes = [c_int, _D]
        L.orc_sog_setup.restype = None
        L.orc_sog_setup.argtypes = [c_int, c_dbl, _D, c_int, c_int, _D, _D, _D]
        L.orc_sog_sample.restype = None
        L.orc_sog_sample.argtypes = [ctypes.c_uint32, ctypes.c_long, c_int, _D, c_dbl,
                                     c_dbl, c_int, _D, _D, c_int, _I, _I]
        L.orc_sog_f.restype = c_dbl
        L.orc_sog_f.argtypes = [c_dbl, c_dbl, c_dbl, c_int, _D]
        L.orc_sog_kspace_f64.restype = None
        L.orc_sog_kspace_f64.argtypes = [c_int, _D, _D, _D, _D, _D, c_int, _I, _I, c_dbl,
                                         c_dbl, c_dbl, c_dbl, c_int, _D, _D, c_int, c_dbl,
                                         _D, _D, _D, _D, _D]
        L.orc_sog_self.restype = c_dbl
        L.orc_sog_self.argtypes = [c_dbl, c_dbl, c_dbl, c_int, c_dbl, c_dbl]
        L.orc_sog_far_r.restype = c_dbl
        L.orc_sog_far_r.argtypes = [c_dbl, c_dbl, c_dbl, c_int, c_dbl]
        L.orc_sog_taylor.restype = None
        L.orc_sog_taylor.argtypes = [c_dbl, c_dbl, c_int, c_dbl, _D]
        L.orc_sog_verlet_f64.restype = None
        L.orc_sog_verlet_f64.argtypes = [c_int, _D, _D, _D, _D, _I, c_int, _D, _D, _D,
                                         c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_int, c_dbl,
                                         _D, _D, _D, _D, _D, _D]
        _lib = L
    return _lib


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_F)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I)


def pairs_f32(pos, L, R2):
    """Set of directed pairs (i, j) with FP32-mirror d^2 < R2, as (N,2) array."""
    n = len(pos)
    x, xp = _f(pos[:, 0])
    y, yp = _f(pos[:, 1])
    z, zp = _f(pos[:, 2])
    Ld, Lp = _d(L)
    off = np.zeros(n + 1, dtype=np.int64)
    tot = lib().orc_pairs_f32(n, xp, yp, zp, Lp, ctypes.c_float(R2),
                              off.ctypes.data_as(_L), None)
    js = np.zeros(max(tot, 1), dtype=np.int32)
    lib().orc_pairs_f32(n, xp, yp, zp, Lp, ctypes.c_float(R2), off.ctypes.data_as(_L),
                        js.ctypes.data_as(_I))
    i = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    return np.stack([i, js[:tot]], axis=1)


def pair_params(r_core, rc, rc_lj, alpha, qqr2e, freq, epoch, rbl, coulomb, mean_fix):
    return np.array([r_core, rc, rc_lj, alpha, qqr2e, freq, epoch, rbl, coulomb, mean_fix],
                    dtype=np.float64)


def rbl_pairs_f32(pos, gid, L, params):
    """(i, j, cls) of every directed pair the RBL classification keeps."""
    n = len(pos)
    x, xp = _f(pos[:, 0])
    y, yp = _f(pos[:, 1])
    z, zp = _f(pos[:, 2])
    g, gp = _i(gid)
    Ld, Lp = _d(L)
    pp, ppp = _d(params)
    off = np.zeros(n + 1, dtype=np.int64)
    tot = lib().orc_rbl_pairs_f32(n, xp, yp, zp, gp, Lp, ppp, off.ctypes.data_as(_L),
                                  None, None)
    js = np.zeros(max(tot, 1), dtype=np.int32)
    cl = np.zeros(max(tot, 1), dtype=np.int32)
    lib().orc_rbl_pairs_f32(n, xp, yp, zp, gp, Lp, ppp, off.ctypes.data_as(_L),
                            js.ctypes.data_as(_I), cl.ctypes.data_as(_I))
    i = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    return np.stack([i, js[:tot], cl[:tot]], axis=1)


def rbl_frequency(n, L, r_core, rc, p):
    Ld, Lp = _d(L)
    return lib().orc_rbl_frequency(n, Lp, r_core, rc, p)


def pair_force(pos, q, types, gid, eps, sig, L, params, prec=0, special=None,
               sog_near=None):
    """Pair forces (N,3), energies (vdwl, coul), virial[6], mean[3] (FP64 out).

    special: optional CVFF topology (nspecial (N,3), special (N,maxspecial)
    global ids, (w12, w13, w14)).  sog_near: optional SOG set (sog_setup) whose
    near field replaces erfc for Coulomb pairs."""
    n = len(pos)
    x, xp = _f(pos[:, 0])
    y, yp = _f(pos[:, 1])
    z, zp = _f(pos[:, 2])
    qq, qp = _f(q)
    t, tp = _i(types)
    g, gp = _i(gid)
    e, ep = _d(eps)
    s, sp = _d(sig)
    Ld, Lp = _d(L)
    pp, ppp = _d(params)
    fx = np.zeros(n)
    fy = np.zeros(n)
    fz = np.zeros(n)
    en = np.zeros(2)
    vir = np.zeros(6)
    mean = np.zeros(3)
    outs = (fx.ctypes.data_as(_D), fy.ctypes.data_as(_D), fz.ctypes.data_as(_D),
            en.ctypes.data_as(_D), vir.ctypes.data_as(_D), mean.ctypes.data_as(_D))
    if sog_near is not None:
        s_ = sog_near
        T, Tp = _d(sog_taylor(s_))
        if special is None:
            ns = np.zeros((n, 3), dtype=np.int32)
            ids = np.zeros((n, 1), dtype=np.int32)
            w = (1.0, 1.0, 1.0)
        else:
            ns, ids, w = special
        nsa, nsp = _i(np.asarray(ns).reshape(-1))
        ida, idp = _i(np.asarray(ids).reshape(-1))
        wa, wp = _d(w)
        ms = np.asarray(ids).shape[1] if special is not None else 0
        lib().orc_pair_force_sog(prec, n, xp, yp, zp, qp, tp, gp, len(eps), ep, sp, Lp, ppp,
                                 ms, nsp, idp, wp, Tp, s_["sigma"], s_["b"], s_["mmax"],
                                 s_["w0"], *outs)
    elif special is None:
        lib().orc_pair_force(prec, n, xp, yp, zp, qp, tp, gp, len(eps), ep, sp, Lp, ppp,
                             *outs)
    else:
        ns, ids, w = special
        nsa, nsp = _i(np.asarray(ns).reshape(-1))
        ida, idp = _i(np.asarray(ids).reshape(-1))
        wa, wp = _d(w)
        ms = np.asarray(ids).shape[1]
        lib().orc_pair_force_special(prec, n, xp, yp, zp, qp, tp, gp, len(eps), ep, sp, Lp,
                                     ppp, ms, nsp, idp, wp, *outs)
    return np.stack([fx, fy, fz], axis=1), en, vir, mean


def verlet_force_f64(pos, q, types, eps, sig, L, rc, rc_lj=None, alpha=0.0, qqr2e=1.0,
                     coulomb=False):
    n = len(pos)
    x, xp = _d(pos[:, 0])
    y, yp = _d(pos[:, 1])
    z, zp = _d(pos[:, 2])
    qq, qp = _d(q)
    t, tp = _i(types)
    e, ep = _d(eps)
    s, sp = _d(sig)
    Ld, Lp = _d(L)
    f = [np.zeros(n) for _ in range(3)]
    en = np.zeros(2)
    vir = np.zeros(6)
    lib().orc_verlet_force_f64(n, xp, yp, zp, qp, tp, len(eps), ep, sp, Lp, rc,
                               rc if rc_lj is None else rc_lj, alpha, qqr2e, int(coulomb),
                               *[a.ctypes.data_as(_D) for a in f], en.ctypes.data_as(_D),
                               vir.ctypes.data_as(_D))
    return np.stack(f, axis=1), en, vir


def kspace_f64(pos, q, L, m, w, qqr2e, alpha, forces=True):
    """FP64 weighted k-sum: forces (N,3) or None, rho (K,2), energy, virial[6]."""
    n = len(pos)
    x, xp = _d(pos[:, 0])
    y, yp = _d(pos[:, 1])
    z, zp = _d(pos[:, 2])
    qq, qp = _d(q)
    Ld, Lp = _d(L)
    mm, mp = _i(np.asarray(m).reshape(-1))
    K = len(mm) // 3
    ww, wp = _d(np.broadcast_to(np.asarray(w, dtype=np.float64), (K,)))
    f = [np.zeros(n) for _ in range(3)]
    rho = np.zeros(2 * K)
    en = np.zeros(1)
    vir = np.zeros(6)
    fp = [a.ctypes.data_as(_D) for a in f] if forces else [None, None, None]
    lib().orc_kspace_f64(n, xp, yp, zp, qp, Lp, K, mp, wp, qqr2e, alpha, *fp,
                         rho.ctypes.data_as(_D), en.ctypes.data_as(_D),
                         vir.ctypes.data_as(_D))
    return (np.stack(f, axis=1) if forces else None), rho.reshape(K, 2), en[0], vir


def kspace_self(sum_q2, sum_q, alpha, V, qqr2e):
    return lib().orc_kspace_self(sum_q2, sum_q, alpha, V, qqr2e)


def rbe_S(alpha, L, f32=False):
    Ld, Lp = _d(L)
    return lib().orc_rbe_S(int(f32), alpha, Lp)


def gauss_sum_direct(alpha, L, kmax):
    Ld, Lp = _d(L)
    return lib().orc_gauss_sum_direct(alpha, Lp, kmax)


def ewald_real_f64(pos, q, L, alpha, rc, qqr2e=1.0):
    n = len(pos)
    x, xp = _d(pos[:, 0])
    y, yp = _d(pos[:, 1])
    z, zp = _d(pos[:, 2])
    qq, qp = _d(q)
    Ld, Lp = _d(L)
    f = [np.zeros(n) for _ in range(3)]
    en = np.zeros(1)
    lib().orc_ewald_real_f64(n, xp, yp, zp, qp, Lp, alpha, rc, qqr2e,
                             *[a.ctypes.data_as(_D) for a in f], en.ctypes.data_as(_D))
    return np.stack(f, axis=1), en[0]


def alpha_from_accuracy(acc, qqr2e, n, rc, V, q2):
    return lib().orc_alpha_from_accuracy(acc, qqr2e, n, rc, V, q2)


def rbe_sample(base_seed, call_index, P, alpha, L):
    Ld, Lp = _d(L)
    m = np.zeros(3 * P, dtype=np.int32)
    lib().orc_rbe_sample(base_seed & 0xFFFFFFFF, call_index, P, alpha, Lp,
                         m.ctypes.data_as(_I))
    return m.reshape(P, 3)


def mt19937(seed, skip, count):
    out = np.zeros(count, dtype=np.uint32)
    lib().orc_mt19937(seed, skip, count, out.ctypes.data_as(_U32))
    return out


def force_step(pos, q, types, gid, eps, sig, L, params, m, rbe_weight):
    """One full RBL + RBE force evaluation (CPU baseline workload); forces (N,3)."""
    n = len(pos)
    x, xp = _f(pos[:, 0])
    y, yp = _f(pos[:, 1])
    z, zp = _f(pos[:, 2])
    qq, qp = _f(q)
    t, tp = _i(types)
    g, gp = _i(gid)
    e, ep = _d(eps)
    s, sp = _d(sig)
    Ld, Lp = _d(L)
    pp, ppp = _d(params)
    mm, mp = _i(np.asarray(m).reshape(-1))
    f = [np.zeros(n) for _ in range(3)]
    lib().orc_force_step(n, xp, yp, zp, qp, tp, gp, len(eps), ep, sp, Lp, ppp,
                         len(mm) // 3, mp, rbe_weight, *[a.ctypes.data_as(_D) for a in f])
    return np.stack(f, axis=1)


def cpu_force_steps(pos, q, types, gid, eps, sig, L, params, skin, m, rbe_weight, nsteps,
                    rebuild_every, threads):
    """The CPU baseline program: nsteps force evaluations at fixed positions
    over a candidate list (rc + skin) rebuilt every rebuild_every steps, RBL
    epoch advancing per step, OpenMP over atoms with `threads` threads.
    Returns the forces (N,3) of the last step."""
    n = len(pos)
    x, xp = _f(pos[:, 0])
    y, yp = _f(pos[:, 1])
    z, zp = _f(pos[:, 2])
    qq, qp = _f(q)
    t, tp = _i(types)
    g, gp = _i(gid)
    e, ep = _d(eps)
    s, sp = _d(sig)
    Ld, Lp = _d(L)
    pp, ppp = _d(params)
    mm, mp = _i(np.asarray(m).reshape(-1))
    f = [np.zeros(n) for _ in range(3)]
    lib().orc_cpu_force_steps(n, xp, yp, zp, qp, tp, gp, len(eps), ep, sp, Lp, ppp, skin,
                              len(mm) // 3, mp, rbe_weight, nsteps, rebuild_every, threads,
                              *[a.ctypes.data_as(_D) for a in f])
    return np.stack(f, axis=1)


# ---- RBSOG / VERLET-SOG (oracle/sog.cpp) -----------------------------------
def sog_preset(level):
    out = np.zeros(4)
    lib().orc_sog_preset(level, out.ctypes.data_as(_D))
    return {"b": out[0], "sigma_rc10": out[1], "accuracy": out[2], "mmax": int(out[3])}


def sog_setup(level, cut_off, L, kcut=1, exact=False):
    """Coefficients and normalisation: dict(b, sigma, w0, S, S_npt, mmax, coef,
    coef_npt).  exact=False: the reference's float set-up (with its defects);
    exact=True: the corrected double set-up."""
    Ld, Lp = _d(L)
    scal = np.zeros(6)
    coef = np.zeros(128)
    coefn = np.zeros(128)
    lib().orc_sog_setup(level, float(cut_off), Lp, kcut, 1 if exact else 0,
                        scal.ctypes.data_as(_D), coef.ctypes.data_as(_D),
                        coefn.ctypes.data_as(_D))
    M = int(scal[5])
    return {"b": scal[0], "sigma": scal[1], "w0": scal[2], "S": scal[3], "S_npt": scal[4],
            "mmax": M, "coef": coef[:M].copy(), "coef_npt": coefn[:M].copy()}


def sog_sample(base_seed, call_index, P, L, s, kcut=1):
    """(m (P,3), idx (P,)) of force evaluation call_index for setup dict s."""
    Ld, Lp = _d(L)
    c, cp = _d(s["coef"])
    cn, cnp = _d(s["coef_npt"])
    m = np.zeros(3 * P, dtype=np.int32)
    idx = np.zeros(P, dtype=np.int32)
    lib().orc_sog_sample(base_seed & 0xFFFFFFFF, call_index, P, Lp, s["sigma"], s["b"],
                         s["mmax"], cp, cnp, kcut, m.ctypes.data_as(_I),
                         idx.ctypes.data_as(_I))
    return m.reshape(P, 3), idx


def sog_f(k2, s, npt=False):
    c, cp = _d(s["coef_npt"] if npt else s["coef"])
    return lib().orc_sog_f(float(k2), s["sigma"], s["b"], s["mmax"], cp)


def sog_kspace_f64(pos, q, L, m, idx, s, qqr2e=1.0, kcut=1):
    """(F (N,3), energy without self term, virial[6]) of the RBSOG k-space."""
    n = len(pos)
    x, xp = _d(pos[:, 0])
    y, yp = _d(pos[:, 1])
    z, zp = _d(pos[:, 2])
    qq, qp = _d(q)
    Ld, Lp = _d(L)
    mm, mp = _i(np.asarray(m).reshape(-1))
    ii, ip = _i(idx)
    c, cp = _d(s["coef"])
    cn, cnp = _d(s["coef_npt"])
    f = [np.zeros(n) for _ in range(3)]
    e = np.zeros(1)
    v = np.zeros(6)
    lib().orc_sog_kspace_f64(n, xp, yp, zp, qp, Lp, len(ii), mp, ip, s["S"], s["S_npt"],
                             s["sigma"], s["b"], s["mmax"], cp, cnp, kcut, qqr2e,
                             *[a.ctypes.data_as(_D) for a in f], e.ctypes.data_as(_D),
                             v.ctypes.data_as(_D))
    return np.stack(f, axis=1), e[0], v


def sog_self(sum_q2, s, qqr2e=1.0):
    return lib().orc_sog_self(sum_q2, s["sigma"], s["b"], s["mmax"], s["w0"], qqr2e)


def sog_far_r(r, s):
    return lib().orc_sog_far_r(float(r), s["sigma"], s["b"], s["mmax"], s["w0"])


def sog_taylor(s):
    T = np.zeros(6)
    lib().orc_sog_taylor(s["sigma"], s["b"], s["mmax"], s["w0"], T.ctypes.data_as(_D))
    return T


def sog_verlet_f64(pos, q, types, eps, sig, L, rc, rc_lj, s, qqr2e=1.0):
    """(F (N,3), (E_lj, E_coul), virial[6]) of VERLET-SOG pairs, O(N^2)."""
    n = len(pos)
    x, xp = _d(pos[:, 0])
    y, yp = _d(pos[:, 1])
    z, zp = _d(pos[:, 2])
    qq, qp = _d(q)
    t, tp = _i(types)
    e_, ep = _d(eps)
    s_, sp = _d(sig)
    Ld, Lp = _d(L)
    T, Tp = _d(sog_taylor(s))
    f = [np.zeros(n) for _ in range(3)]
    en = np.zeros(2)
    v = np.zeros(6)
    lib().orc_sog_verlet_f64(n, xp, yp, zp, qp, tp, len(eps), ep, sp, Lp, rc, rc_lj, qqr2e,
                             s["sigma"], s["b"], s["mmax"], s["w0"], Tp,
                             *[a.ctypes.data_as(_D) for a in f], en.ctypes.data_as(_D),
                             v.ctypes.data_as(_D))
    return np.stack(f, axis=1), en, v


def _adj_inputs(pos, types, gid, Q, eps, sig, L, params, m, wk):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    x, xp = _d(pos[:, 0])
    y, yp = _d(pos[:, 1])
    z, zp = _d(pos[:, 2])
    t, tp = _i(types)
    g, gp = _i(gid)
    qa, qp = _d(Q)
    e, ep = _d(eps)
    s, sp = _d(sig)
    Ld, Lp = _d(L)
    pp, ppp = _d(params)
    K = 0 if m is None else len(m)
    ma, mp = _i(np.zeros(3, np.int32) if m is None else np.asarray(m).reshape(-1))
    wa, wp = _d(np.zeros(1) if wk is None else wk)
    keep = (x, y, z, t, g, qa, e, s, Ld, pp, ma, wa)
    return keep, (len(pos), xp, yp, zp, tp, gp, len(Q), qp, ep, sp, Lp, ppp, K, mp, wp)


def force_f64(pos, types, gid, Q, eps, sig, L, params, m=None, wk=None):
    """FP64 force (N,3) of one evaluation (RBL/Verlet pairs, mean-force fix,
    RBE k-space over the atoms of nonzero charge) for per-type Q, eps, sigma."""
    keep, args = _adj_inputs(pos, types, gid, Q, eps, sig, L, params, m, wk)
    n = args[0]
    f = [np.zeros(n) for _ in range(3)]
    lib().orc_force_f64(*args, *(a.ctypes.data_as(_D) for a in f))
    del keep
    return np.stack(f, axis=1)


def force_vjp_f64(pos, types, gid, Q, eps, sig, L, params, G, m=None, wk=None):
    """(dL/dr (N,3), dL/dQ, dL/deps, dL/dsigma) for L = sum G . F."""
    keep, args = _adj_inputs(pos, types, gid, Q, eps, sig, L, params, m, wk)
    n, nt = args[0], args[6]
    Ga, Gp = _d(np.asarray(G, dtype=np.float64).reshape(-1))
    dr = np.zeros(3 * n)
    dq, de, ds = np.zeros(nt), np.zeros(nt), np.zeros(nt)
    lib().orc_force_vjp_f64(*args, Gp, dr.ctypes.data_as(_D), dq.ctypes.data_as(_D),
                            de.ctypes.data_as(_D), ds.ctypes.data_as(_D))
    del keep, Ga
    return dr.reshape(n, 3), dq, de, ds
