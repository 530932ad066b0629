"""rbmdx_force_noalloc (the allocation-free variant of rbmdx_force, its
scratch in the library's static per-device array) against rbmdx_force on the
same context and inputs: two calls in a row, the second on a smaller system,
so the second runs over the first's leftover scratch contents.  The kernels
are the same and deterministic, so forces, energies and virials must agree
bit for bit.  Also the errors: before the context's first rbmdx_force (no
lists yet) and for a system whose scratch exceeds the static array."""
import numpy as np
import pytest

from rbmd_amd import ForceContext, RbmdxError, rbe_sample_host, systems

pytestmark = pytest.mark.gpu

RBMDX_E_CAPACITY = 2
RBMDX_E_STATE = 6


def _setup(n, seed, P, device):
    import torch

    s = systems.electrolyte(n, seed=seed)
    cfg = systems.ljsalt_config(P=P)
    ctx = ForceContext(cfg, device=0)
    ctx.set_system(s.box_lo, s.box_len, s.types, s.ids, s.charges, s.eps, s.sigma)
    p = torch.tensor(s.pos.astype(np.float32), device=device)
    pos = [p[:, d].contiguous() for d in range(3)]
    m = rbe_sample_host(0.09, s.box_len, P, 1234 + seed, 0)
    mt = torch.tensor(m.reshape(-1), dtype=torch.int32, device=device)
    return s, ctx, pos, mt


def _forces(n, device):
    import torch

    return [torch.zeros(n, dtype=torch.float32, device=device) for _ in range(3)]


def _thermo_vec(th):
    return np.array([th["e_vdwl"], th["e_coul"], th["e_kspace"]] + list(th["virial"]))


def test_noalloc_equals_force_on_two_calls_second_smaller(device):
    import torch

    step = 7
    cases = [_setup(8000, 31, 200, device), _setup(2000, 32, 200, device)]
    want = []
    for s, ctx, pos, mt in cases:
        f = _forces(s.n, device)
        ctx.force(step, *pos, *f, m_samples=mt, energy=True, sync=True)
        want.append(([t.clone() for t in f], _thermo_vec(ctx.thermo())))
    # the two allocation-free calls back to back: the larger system first,
    # then the smaller one over the scratch the first left behind
    got = []
    for s, ctx, pos, mt in cases:
        f = _forces(s.n, device)
        print(f"force_noalloc N={s.n}", flush=True)
        ctx.force_noalloc(step, *pos, *f, m_samples=mt, energy=True)
        got.append((f, ctx))
    torch.cuda.synchronize()
    for (f, ctx), (fw, thw) in zip(got, want):
        for a, b in zip(f, fw):
            assert torch.equal(a, b)
        assert np.array_equal(_thermo_vec(ctx.thermo()), thw)
    # and once more on the larger system, after the smaller one
    s, ctx, pos, mt = cases[0]
    f = _forces(s.n, device)
    ctx.force_noalloc(step, *pos, *f, m_samples=mt, energy=True, sync=True)
    for a, b in zip(f, want[0][0]):
        assert torch.equal(a, b)


def test_noalloc_errors(device):
    # no lists before the context's first rbmdx_force
    s, ctx, pos, mt = _setup(2000, 33, 64, device)
    f = _forces(s.n, device)
    with pytest.raises(RbmdxError) as e:
        ctx.force_noalloc(0, *pos, *f, m_samples=mt, sync=True)
    assert e.value.code == RBMDX_E_STATE
    # a sample count whose per-k scratch (at least 64 bytes per k) exceeds
    # the 16 MiB array
    s, ctx, pos, mt = _setup(2000, 34, 300000, device)
    with pytest.raises(RbmdxError) as e:
        ctx.force_noalloc(0, *pos, *_forces(s.n, device), m_samples=mt, sync=True)
    assert e.value.code == RBMDX_E_CAPACITY
