"""torch.autograd.Function over the CGFDM forward run and its hand-written adjoint.

`seismograms(vs, vp, rho, rate, model, nsteps)` returns the station velocity
[ns, nsteps, 3] of a CGFDM run whose medium is (vs, vp, rho) per padded cell and
whose finite-fault source is the moment-rate table `rate` [npts, nt, 6] (the
table `Solver.set_source` takes). Its backward pass runs the adjoint kernels
(`cgfdm_forward_record` keeps checkpoints, `cgfdm_adjoint` integrates back) and
turns the gradients with respect to (mu, lambda, 1/rho) into gradients with
respect to (vs, vp, rho):
    mu = rho vs^2, lambda = rho (vp^2 - 2 vs^2), b = 1 / rho.
The HIP library is the computing backend; tensors are moved to host memory
at the boundary (the C ABI takes host arrays).
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from .model import Model


def _hip_backend(model: Model, precision: int):
    from .solver import Solver
    return Solver(model, precision=precision)


class _Seismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vs, vp, rho, rate, model, nsteps, precision, checkpoint_every, backend):
        m = copy.copy(model)
        med = torch.stack([vs.detach(), vp.detach(), rho.detach()]).to("cpu", torch.float32).numpy()
        m.medium = np.ascontiguousarray(med)
        if rate is not None:
            m.src_rate = np.ascontiguousarray(rate.detach().to("cpu", torch.float32).numpy())
        solver = (backend or _hip_backend)(m, precision)
        rec = solver.forward_record(int(nsteps), int(checkpoint_every))
        ctx.solver = solver
        ctx.has_rate = rate is not None
        ctx.save_for_backward(vs, vp, rho)
        return torch.as_tensor(rec, dtype=torch.float64, device=vs.device)

    @staticmethod
    def backward(ctx, grad_rec):
        vs, vp, rho = ctx.saved_tensors
        g = ctx.solver.adjoint(grad_rec.detach().to("cpu", torch.float64).numpy())
        dev = vs.device
        gmu, glam, gb = (torch.as_tensor(g[k], dtype=torch.float64, device=dev) for k in ("gmu", "glam", "gb"))
        vs64, vp64, rho64 = (t.detach().to(torch.float32).to(torch.float64) for t in (vs, vp, rho))
        g_vs = gmu * 2 * rho64 * vs64 - glam * 4 * rho64 * vs64
        g_vp = glam * 2 * rho64 * vp64
        g_rho = gmu * vs64 ** 2 + glam * (vp64 ** 2 - 2 * vs64 ** 2) - gb / rho64 ** 2
        g_rate = torch.as_tensor(g["grate"], dtype=torch.float64, device=dev) if ctx.has_rate else None
        ctx.solver = None
        return (g_vs.to(vs.dtype), g_vp.to(vp.dtype), g_rho.to(rho.dtype),
                g_rate, None, None, None, None, None)


def seismograms(vs, vp, rho, rate, model: Model, nsteps: int, precision: int = 4, checkpoint_every: int = 100,
                backend=None):
    """Differentiable station velocity [ns, nsteps, 3] (float64 tensor).

    vs, vp, rho: tensors of the padded grid shape (NZ, NY, NX); rate: tensor
    [npts, nt, 6] or None (the model's own table, not differentiated).
    backend(model, precision) may replace the HIP solver by any object with
    forward_record(nsteps, checkpoint_every) and adjoint(src)."""
    return _Seismograms.apply(vs, vp, rho, rate, model, nsteps, precision, checkpoint_every, backend)


def slip_rate_gradient(grate, moment_tensor_area):
    """Gradient with respect to the slip rate per subfault and time sample from the
    moment-rate table gradient: rate[p, t, :] = M6[p, :] * area[p] * slip_rate[p, t],
    so d/d slip_rate = sum_q grate[p, t, q] M6[p, q] area[p] (moment_tensor_area[p, q])."""
    return (np.asarray(grate) * np.asarray(moment_tensor_area)[:, None, :]).sum(-1)
