"""The torch.autograd wrapper: chain rule from (mu, lambda, 1/rho) to (vs, vp, rho)
and the source-table gradient, checked on the CPU with the FP64 oracle as the
backend against central finite differences of the same oracle."""
import numpy as np
import pytest
import torch

from cgfdm_amd import model as M
from cgfdm_amd import autograd as A


class OracleBackend:
    """forward_record / adjoint implemented with the FP64 oracle (test double of the HIP solver)."""

    def __init__(self, model, precision, oracle_mod):
        self.m = model
        self.om = oracle_mod

    def forward_record(self, nsteps, ck):
        o = self.om.Oracle(self.m, prec=8, threads=1)
        st = self.m.stations
        rec = np.zeros((len(st), nsteps, 3))
        for n in range(nsteps):
            o.step(1)
            for c in range(3):
                f = o.field(c)
                rec[:, n, c] = f[st[:, 2], st[:, 1], st[:, 0]]
        self.rec = rec
        self.nsteps = nsteps
        return rec

    def adjoint(self, src):
        o = self.om.Oracle(self.m, prec=8, threads=1)
        r = self.om.adjoint_misfit(o, self.nsteps, self.rec - src, npts=len(self.m.src_cells),
                                   nt_src=self.m.src_rate.shape[1])
        return r


def test_wrapper_gradients_match_finite_differences(oracle_mod):
    nst = 8
    m = M.small_config(nx=12, ny=11, nz=10, npml=3, hills=2, free_surface=True, nt=nst)
    backend = lambda mm, prec: OracleBackend(mm, prec, oracle_mod)
    vs = torch.tensor(m.medium[0], dtype=torch.float64, requires_grad=True)
    vp = torch.tensor(m.medium[1], dtype=torch.float64, requires_grad=True)
    rho = torch.tensor(m.medium[2], dtype=torch.float64, requires_grad=True)
    rate = torch.tensor(m.src_rate, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(np.random.default_rng(3).standard_normal((len(m.stations), nst, 3)))
    loss = (A.seismograms(vs, vp, rho, rate, m, nst, precision=8, backend=backend) * w).sum()
    loss.backward()

    def loss_at(vs_, vp_, rho_, rate_):
        mm = M.small_config(nx=12, ny=11, nz=10, npml=3, hills=2, free_surface=True, nt=nst)
        mm.medium = np.stack([vs_, vp_, rho_]).astype(np.float32)
        mm.src_rate = rate_.astype(np.float32)
        rec = OracleBackend(mm, 8, oracle_mod).forward_record(nst, 1)
        return float((rec * w.numpy()).sum())

    rng = np.random.default_rng(7)
    base = [m.medium[0].astype(np.float64), m.medium[1].astype(np.float64), m.medium[2].astype(np.float64),
            m.src_rate.astype(np.float64)]
    grads = [vs.grad.numpy(), vp.grad.numpy(), rho.grad.numpy(), rate.grad.numpy()]
    for k in range(4):
        d = base[k] * rng.uniform(-0.5, 0.5, base[k].shape)
        eps = 1e-3
        vals, used = [], []
        for sg in (1, -1):
            args = [b.copy() for b in base]
            args[k] = (args[k] + sg * eps * d).astype(np.float32).astype(np.float64)
            used.append(args[k])
            vals.append(loss_at(*args))
        fd = (vals[0] - vals[1]) / (2 * eps)
        an = float((grads[k] * (used[0] - used[1]) / (2 * eps)).sum())
        assert abs(fd - an) <= 1e-5 * abs(fd), (k, fd, an)


def test_slip_rate_gradient_contracts_the_tensor():
    g = np.arange(2 * 3 * 6, dtype=float).reshape(2, 3, 6)
    mt = np.array([[1, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 2.0]])
    out = A.slip_rate_gradient(g, mt)
    np.testing.assert_array_equal(out[0], g[0, :, 0])
    np.testing.assert_array_equal(out[1], 2 * g[1, :, 5])
