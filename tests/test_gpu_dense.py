"""GPU parity of the mostly charged path (more than half the atoms charged:
the per-cell charged-pair list build and the one-wave-per-atom charged-pair
kernel) against the CPU oracle, on systems whose grids take both periodic
branches of the list build: 4 cells per dimension (the smallest grid where a
stencil cell across the box edge is shifted by L) and 7, and the all-charged
and the 60 % charged compositions (the latter with neutral pairs in list A).

Tolerances as in test_gpu_parity.py: candidate and RBL pair sets bit-exact as
sets of (gid_i, gid_j[, class]); forces relative vector RMS <= 2e-5 against
the FP64 oracle (5e-5 for CVFF water, as in test_gpu_cvff.py); energies
relative 1e-5, the pair virial 1e-4.
"""
import numpy as np
import pytest

import oracle
from rbmd_amd import ForceContext, rbe_sample_host, systems
from tests.refs import reference_force, rel_rms

pytestmark = pytest.mark.gpu

# (atoms, fraction of cations = fraction of anions): 4000 atoms at density
# 0.8 give a box of 17.1 sigma, 4 cells of >= 3.8 per dimension; 20000 give 7
CASES = [(4000, 0.5), (20000, 0.3)]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"n{c[0]}_ion{c[1]}")
def dense_sys(request):
    n, frac = request.param
    return systems.electrolyte(n, frac_ion=frac, seed=17)


def _ctx(s, cfg):
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    return ctx


def _pos(s, device):
    import torch

    p = torch.tensor(s.pos, dtype=torch.float32, device=device)
    return tuple(p[:, k].contiguous() for k in range(3))


def test_dense_candidate_lists_bit_exact(device, dense_sys):
    s = dense_sys
    assert 2 * int((s.charges != 0).sum()) > s.n  # the mostly charged path
    ctx = _ctx(s, systems.ljsalt_config(lj_cut_off=3.5))
    px, py, pz = _pos(s, device)
    ctx.export_pairs(0, px, py, pz, cap=1 << 23)
    gi, gj, _ = ctx.export_candidates(cap=1 << 23)
    pos32 = s.pos.astype(np.float32)
    ref = oracle.pairs_f32(pos32, s.box_len, float(np.float32(3.8 * 3.8)))
    got = set(zip(gi.tolist(), gj.tolist()))
    assert len(got) == len(gi)  # no duplicates
    assert got == set(zip(ref[:, 0].tolist(), ref[:, 1].tolist()))


@pytest.mark.parametrize("step", [0, 5])
def test_dense_rbl_pair_sets_bit_exact(device, dense_sys, step):
    s = dense_sys
    cfg = systems.ljsalt_config(lj_cut_off=3.5, sample_num=30)
    ctx = _ctx(s, cfg)
    px, py, pz = _pos(s, device)
    gi, gj, cls = ctx.export_pairs(step, px, py, pz, cap=1 << 23)
    nb = cfg["hyper_parameters"]["neighbor"]
    freq = ctx.thermo()["rbl_frequency"]
    assert freq == oracle.rbl_frequency(s.n, s.box_len, nb["r_core"], nb["cut_off"],
                                        nb["neighbor_sample_num"])
    params = oracle.pair_params(nb["r_core"], nb["cut_off"], 3.5, 0.09, 1.0, freq, step,
                                1, 1, 1)
    ref = oracle.rbl_pairs_f32(s.pos.astype(np.float32), s.ids, s.box_len, params)
    ref_set = set(zip(s.ids[ref[:, 0]].tolist(), s.ids[ref[:, 1]].tolist(),
                      ref[:, 2].tolist()))
    assert set(zip(gi.tolist(), gj.tolist(), cls.tolist())) == ref_set


@pytest.mark.parametrize("energy", [False, True])
def test_dense_rbl_rbe_force_parity(device, dense_sys, energy):
    import torch

    s = dense_sys
    cfg = systems.ljsalt_config(lj_cut_off=2.5, sample_num=30, P=200)
    ctx = _ctx(s, cfg)
    px, py, pz = _pos(s, device)
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    m = rbe_sample_host(0.09, s.box_len, 200, 20261202, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    step = 3
    ctx.force(step, px, py, pz, *f, m_samples=mt, energy=energy, sync=True)
    got = np.stack([t.double().cpu().numpy() for t in f], axis=1)
    ref, en, vir, ek, vk = reference_force(s, s.pos.astype(np.float32), cfg, step, m)
    err = rel_rms(got, ref)
    print(f"dense n={s.n}: force rel RMS vs FP64 = {err:.3e}")
    assert err < 2e-5
    if energy:
        th = ctx.thermo()
        assert th["e_vdwl"] == pytest.approx(en[0], rel=1e-5)
        assert th["e_coul"] == pytest.approx(en[1], rel=1e-5)
        # the total virial less its k-space part, to test_gpu_parity.py's bound
        np.testing.assert_allclose(np.array(th["virial"]) - np.array(th["virial_kspace"]),
                                   vir, rtol=1e-4, atol=1e-5 * abs(vir).max())


def test_dense_water_four_cells_special(device):
    """SPC/E water of 5,000 molecules (box 53 A, 4 cells of 13.3 A per
    dimension): CVFF special bonds [0, 0, 0], RBL + RBE."""
    import torch

    s = systems.spce_water(5000, seed=5)
    cfg = systems.cvff_water_config(alpha=0.0784, P=300)
    ctx = _ctx(s, cfg)
    ctx.set_special(s.nspecial, s.special)
    px, py, pz = _pos(s, device)
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    m = rbe_sample_host(0.0784, s.box_len, 300, 20261202, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    ctx.force(2, px, py, pz, *f, m_samples=mt, sync=True)
    got = np.stack([t.double().cpu().numpy() for t in f], axis=1)
    ref = reference_force(s, s.pos.astype(np.float32), cfg, 2, m)[0]
    err = rel_rms(got, ref)
    print(f"water 4-cell grid: force rel RMS vs FP64 = {err:.3e}")
    assert err < 5e-5  # the CVFF water bound of test_gpu_cvff.py


@pytest.mark.parametrize("nmol,cut,lo,hi", [(7600, 13.0, 257, 512), (5000, 13.5, 513, 1 << 20)],
                         ids=["cells_257_512", "cells_over_512"])
def test_dense_water_large_cells(device, nmol, cut, lo, hi):
    """SPC/E water whose cells hold more than 256 atoms: 7,600 molecules (box
    61 A) at rc 13 A + skin 2 A give 4 cells of 15.25 A per dimension, about
    360 atoms each (the cell sort ranks eight groups of 64 per wave); 5,000
    molecules (box 53 A) at rc 13.5 A give 3 cells of 17.7 A, about 560 atoms
    each (sorted by one lane).  CVFF special bonds [0, 0, 0], RBL + RBE."""
    import torch

    s = systems.spce_water(nmol, seed=8)
    rb = cut + 2.0  # the cell side bound, rc + skin (the LJ and Coulomb cutoffs agree)
    nc = np.floor(np.asarray(s.box_len) / (rb * (1.0 + 1e-5))).astype(int)
    cell = np.floor((s.pos - np.asarray(s.box_lo)) / (np.asarray(s.box_len) / nc)).astype(int)
    cell = np.clip(cell, 0, nc - 1)
    occ = np.bincount((cell[:, 2] * nc[1] + cell[:, 1]) * nc[0] + cell[:, 0])
    assert lo <= occ.max() <= hi, (nc, occ.max())
    cfg = systems.cvff_water_config(cut_off=cut, r_core=6.5, alpha=0.0784, P=300)
    ctx = _ctx(s, cfg)
    ctx.set_special(s.nspecial, s.special)
    px, py, pz = _pos(s, device)
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    m = rbe_sample_host(0.0784, s.box_len, 300, 20261203, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    ctx.force(1, px, py, pz, *f, m_samples=mt, sync=True)
    got = np.stack([t.double().cpu().numpy() for t in f], axis=1)
    ref = reference_force(s, s.pos.astype(np.float32), cfg, 1, m)[0]
    err = rel_rms(got, ref)
    print(f"water {nmol} molecules, rc {cut}: largest cell {occ.max()}, "
          f"force rel RMS vs FP64 = {err:.3e}")
    assert err < 5e-5  # the CVFF water bound of test_gpu_cvff.py


def _slab():
    """The all-charged 4,000-atom electrolyte in a box twice as tall (vacuum
    above the atoms): 4 x 4 x 8 cells with a mean of 31 atoms, so each cell
    has one block of the list build, while 19 filled cells hold more than the
    64 atoms of one pass and take a second."""
    s = systems.electrolyte(4000, frac_ion=0.5, seed=17)
    box = np.array(s.box_len, dtype=np.float64)
    box[2] *= 2.0
    return systems.System(s.pos, s.box_lo, box, s.types, s.ids, s.charges, s.eps, s.sigma,
                          name="electrolyte_4000_slab")


def test_dense_slab_lists_and_forces(device):
    import torch

    s = _slab()
    nc = np.floor(s.box_len / (3.8 * (1.0 + 1e-5))).astype(int)
    cell = np.clip(np.floor((s.pos - s.box_lo) / (s.box_len / nc)).astype(int), 0, nc - 1)
    occ = np.bincount((cell[:, 2] * nc[1] + cell[:, 1]) * nc[0] + cell[:, 0])
    assert 1.25 * s.n / nc.prod() <= 64 and occ.max() > 64  # one block per cell, two passes
    # candidate lists (rc + skin = 3.8, the cell side bound)
    ctx = _ctx(s, systems.ljsalt_config(lj_cut_off=3.5))
    px, py, pz = _pos(s, device)
    ctx.export_pairs(0, px, py, pz, cap=1 << 23)
    gi, gj, _ = ctx.export_candidates(cap=1 << 23)
    ref = oracle.pairs_f32(s.pos.astype(np.float32), s.box_len, float(np.float32(3.8 * 3.8)))
    got = set(zip(gi.tolist(), gj.tolist()))
    assert len(got) == len(gi)
    assert got == set(zip(ref[:, 0].tolist(), ref[:, 1].tolist()))
    # forces, RBL + RBE, against the FP64 oracle
    cfg = systems.ljsalt_config(lj_cut_off=3.5, sample_num=30, P=200)
    ctx = _ctx(s, cfg)
    f = [torch.zeros(s.n, dtype=torch.float32, device=device) for _ in range(3)]
    m = rbe_sample_host(0.09, s.box_len, 200, 20261203, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    ctx.force(3, px, py, pz, *f, m_samples=mt, sync=True)
    got_f = np.stack([t.double().cpu().numpy() for t in f], axis=1)
    ref_f = reference_force(s, s.pos.astype(np.float32), cfg, 3, m)[0]
    err = rel_rms(got_f, ref_f)
    print(f"slab: largest cell {occ.max()}, force rel RMS vs FP64 = {err:.3e}")
    assert err < 2e-5
