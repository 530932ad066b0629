"""Differentiable force evaluation: a ``torch.autograd.Function`` over the C ABI.

The forward pass is one ``rbmdx_force`` evaluation; the backward pass is
``rbmdx_force_vjp`` with the incoming gradient of the forces as the cotangent:
hand-written adjoint kernels give dL/d(positions) and, per atom type,
dL/dQ (the type's charge), dL/deps and dL/dsigma.  The structure factor is
recomputed in the backward pass, not stored per atom and k.

Per-type charges are applied as per-atom charges Q[type]; an update may not
turn a neutral type charged or back (the charged/neutral partition of the pair
work is fixed when the system is set).  dL/dQ of a neutral type is therefore 0
unless ``neutral_charge_grad=True``, which computes it as the derivative at
Q = 0 (the Coulomb pairs of charged and neutral atoms and the neutral atoms'
structure sums enter the backward pass).  The pair set (RBL sampling epoch
``step``) and the k-sample are those of the forward evaluation; the backward
pass must run before the next forward evaluation of the same context.
"""
import numpy as np
import torch

from . import ForceContext


class _ForceFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pos, Q, eps, sigma, model, step, m_samples, num_k):
        model._apply(Q, eps, sigma)
        pos = pos.contiguous()
        px, py, pz = (pos[:, d].contiguous() for d in range(3))
        f = [torch.empty_like(px) for _ in range(3)]
        model.ctx.force(step, px, py, pz, *f, m_samples=m_samples, num_k=num_k)
        fctx.save_for_backward(px, py, pz)
        fctx.model = model
        fctx.step = step
        fctx.pdev = (Q.device, eps.device, sigma.device)
        fctx.pdtype = (Q.dtype, eps.dtype, sigma.dtype)
        return torch.stack(f, dim=1)

    @staticmethod
    def backward(fctx, gF):
        px, py, pz = fctx.saved_tensors
        model = fctx.model
        gF = gF.contiguous().to(px.dtype)
        gx, gy, gz = (gF[:, d].contiguous() for d in range(3))
        d = [torch.empty_like(px) for _ in range(3)]
        nt = model.ntypes
        dparam = torch.empty(3 * nt, dtype=torch.float64, device=px.device)
        model.ctx.force_vjp(fctx.step, px, py, pz, gx, gy, gz, *d, dparam,
                            neutral_q=model.neutral_charge_grad)
        dpos = torch.stack(d, dim=1).to(gF.dtype)
        grads = [dparam[k * nt:(k + 1) * nt].to(device=dev, dtype=dt)
                 for k, (dev, dt) in enumerate(zip(fctx.pdev, fctx.pdtype))]
        return (dpos, *grads, None, None, None, None)


class DifferentiableForce:
    """Force field of one system as a differentiable function of positions
    and per-type (Q, eps, sigma).

    >>> model = DifferentiableForce(cfg, system)
    >>> F = model(pos, Q, eps, sigma, step=0, m_samples=m)
    >>> ((F - F_target) ** 2).sum().backward()
    """

    def __init__(self, cfg, system, device=0, neutral_charge_grad=False):
        self.neutral_charge_grad = bool(neutral_charge_grad)
        self.ctx = ForceContext(cfg, device=device)
        self.types = np.asarray(system.types, dtype=np.int32)
        self.ntypes = len(system.eps)
        self.ctx.set_system(system.box_lo, system.box_len, system.types, system.ids,
                            system.charges, system.eps, system.sigma)
        self._state = None

    def _apply(self, Q, eps, sigma):
        q = np.asarray(Q.detach().cpu(), dtype=np.float64)
        e = np.asarray(eps.detach().cpu(), dtype=np.float64)
        s = np.asarray(sigma.detach().cpu(), dtype=np.float64)
        if self._state is not None:
            q0, e0, s0 = self._state
            same_q = np.array_equal(q, q0)
            same_lj = np.array_equal(e, e0) and np.array_equal(s, s0)
        else:
            same_q = same_lj = False
        if not same_q:
            self.ctx.set_charges(q[self.types])
        if not same_lj:
            self.ctx.set_lj(e, s)
        self._state = (q, e, s)

    def __call__(self, pos, Q, eps, sigma, step=0, m_samples=None, num_k=None):
        return _ForceFn.apply(pos, Q, eps, sigma, self, int(step), m_samples, num_k)
