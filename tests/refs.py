"""Oracle-side helpers shared by the tests: full reference force evaluations."""
import numpy as np

import oracle


def rel_rms(f, ref):
    """Relative vector RMS: sqrt(sum |f - ref|^2 / sum |ref|^2)."""
    f = np.asarray(f, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    return float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))


QQR2E = {"LJ": 1.0, "REAL": 332.06371, "METAL": 14.399645}


def cfg_qqr2e(cfg):
    unit = cfg.get("init_configuration", {}).get("read_data", {}).get("unit", "LJ")
    return QQR2E[unit]


def cfg_special(cfg, sys_):
    """(nspecial, special, (w12, w13, w14)) for a CVFF configuration of a
    system with topology, else None."""
    hp = cfg["hyper_parameters"]
    if hp["force_field"]["type"] != "CVFF" or getattr(sys_, "nspecial", None) is None:
        return None
    w = [1.0, 1.0, 1.0]
    for k, v in enumerate(hp.get("extend", {}).get("special_bonds", [])[:3]):
        w[k] = float(v)
    return sys_.nspecial, sys_.special, tuple(w)


def cfg_values(cfg):
    hp = cfg["hyper_parameters"]
    nb = hp["neighbor"]
    rc = nb["cut_off"]
    rclj = nb.get("lj_cut_off", rc)
    rbl = nb["type"] == "RBL"
    cl = hp.get("coulomb")
    return hp, nb, rc, rclj, rbl, cl


def reference_force(sys_, pos32, cfg, step, m, prec=0, with_kspace=True):
    """Oracle total force of one evaluation (pair + k-space over charged atoms).

    pos32: float32 positions (N,3) exactly as handed to the GPU.
    m: (P,3) integer k-vectors of this evaluation (RBE), or None.
    Returns (F (N,3), pair energies (vdwl, coul), pair virial, kspace energy,
    kspace virial).
    """
    hp, nb, rc, rclj, rbl, cl = cfg_values(cfg)
    coulomb = cl is not None
    alpha = cl["alpha"] if coulomb else 0.0
    qqr2e = cfg_qqr2e(cfg)
    n = len(pos32)
    freq = oracle.rbl_frequency(n, sys_.box_len, nb.get("r_core", 0.0), rc,
                                nb.get("neighbor_sample_num", 1)) if rbl else 1
    params = oracle.pair_params(nb.get("r_core", 0.0), rc, rclj, alpha, qqr2e, freq, step,
                                1 if rbl else 0, 1 if coulomb else 0, 1 if rbl else 0)
    q32 = sys_.charges.astype(np.float32)
    f, en, vir, mean = oracle.pair_force(pos32, q32, sys_.types, sys_.ids, sys_.eps,
                                         sys_.sigma, sys_.box_len, params, prec=prec,
                                         special=cfg_special(cfg, sys_))
    ek, vk = 0.0, np.zeros(6)
    if coulomb and with_kspace:
        ci = np.nonzero(sys_.charges)[0]
        p64 = pos32.astype(np.float64)[ci]
        if cl["type"] == "RBE":
            S = oracle.rbe_S(alpha, sys_.box_len)
            w = np.full(len(m), S / len(m))
        else:
            raise ValueError("reference_force: EWALD weights are built by the caller")
        fk, rho, ek, vk = oracle.kspace_f64(p64, sys_.charges[ci], sys_.box_len, m, w, qqr2e,
                                            alpha)
        f[ci] += fk
        V = float(np.prod(sys_.box_len))
        ek += oracle.kspace_self((sys_.charges ** 2).sum(), sys_.charges.sum(), alpha, V,
                                 qqr2e)
    return f, en, vir, ek, vk


def coul_abs_scale(sys_, rc, alpha, qqr2e):
    """Sum over charged pairs within rc of |E_pair| = 1/2 |q_i q_j| erfc(g r) / r
    (both directions, g = sqrt(alpha)): the size of the terms the real-space
    Coulomb energy adds up.  Their signs cancel in the total, so an FP32
    pair evaluation is bounded relative to this sum, not to the total."""
    from scipy.special import erfc

    ci = np.nonzero(sys_.charges)[0]
    p = sys_.pos[ci].astype(np.float64)
    q = sys_.charges[ci].astype(np.float64)
    L = np.asarray(sys_.box_len, dtype=np.float64)
    tot = 0.0
    for a in range(0, len(ci), 512):
        d = p[None, :, :] - p[a:a + 512, None, :]
        d -= L * np.round(d / L)
        r = np.sqrt((d * d).sum(-1))
        qq = np.abs(q[a:a + 512, None] * q[None, :])
        m = (r > 0) & (r < rc)
        tot += 0.5 * (qq[m] * erfc(np.sqrt(alpha) * r[m]) / r[m]).sum()
    return qqr2e * tot
