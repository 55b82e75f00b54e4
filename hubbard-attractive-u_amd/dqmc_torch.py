"""PyTorch autograd over the engine's fixed-field forward and its hand-written adjoint.

At a fixed set of Ising-field configurations (one per chain) the engine computes
    logdet_c   = log|det(1 + B_L ... B_1)|            (one spin species)
    gdiag_c    = G_c(t)(i, i) for every slice t and site i
and, through `dqmc_adjoint`, their derivatives with respect to theta = (nnt, mu, u), the
[Model.HubbardAttractiveU] keys of config.toml (the interaction is U = -u).  `greens()` wraps the
two calls in a torch.autograd.Function, so any torch expression of logdet and the diagonals
(the reference's equal-time estimators, include/observable_methods.hpp:39-74) can be
differentiated with respect to theta.  Torch only carries the autograd graph and the small
per-chain arrays; the Green's functions and the adjoint run in the HIP engine.

Weight convention: the reference's Metropolis ratio (src/hubbard.cpp:123-130) is that of
    W(s) = [e^{-dt u/4} / 2]^{N L} e^{-alpha sum s} det(1 + B_L ... B_1)^2,
the HS decomposition of H_U = -u sum_i (n_up - 1/2)(n_dn - 1/2) with cosh(alpha) = e^{dt u/2}
(src/hubbard.cpp:35).  The field-independent prefactor drops out of every derivative estimator.
"""
import math

import numpy as np
import torch

import dqmc_amd as gpu


class FixedFields:
    """Field configurations [chains, nt, ns] (+1/-1) of an nl x nl lattice with nt slices at
    inverse temperature beta.  Engines are built per parameter point and cached."""

    def __init__(self, nl, nt, beta, fields, pace=10, cache_size=8):
        self.nl, self.nt, self.beta, self.pace = nl, nt, float(beta), pace
        self.ns = nl * nl
        self.fields = np.ascontiguousarray(fields, dtype=np.int8).reshape(-1, nt, self.ns)
        self.chains = self.fields.shape[0]
        self.dt = self.beta / nt
        self._cache = {}
        self._order = []
        self._cache_size = cache_size

    def engine(self, nnt, u, mus):
        key = (float(nnt), float(u), tuple(float(m) for m in mus))
        e = self._cache.get(key)
        if e is None:
            e = gpu.Engine(self.nl, self.nt, self.beta, float(u), mu=float(mus[0]), nnt=float(nnt), pace=self.pace,
                           n_chains=self.chains, precision=gpu.FP64, delay_k=1)
            for c in range(self.chains):
                e.set_chain_mu(c, float(mus[c]))
                e.set_fields(c, self.fields[c])
            e.init()
            self._cache[key] = e
            self._order.append(key)
            while len(self._order) > self._cache_size:
                self._cache.pop(self._order.pop(0)).close()
        return e

    def close(self):
        for e in self._cache.values():
            e.close()
        self._cache.clear()
        self._order.clear()

    def field_sums(self):
        return self.fields.reshape(self.chains, -1).sum(axis=1).astype(np.float64)


class _FixedFieldGreens(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta_c, ff):
        th = theta_c.detach().to("cpu", torch.float64).numpy()
        if not (np.all(th[:, 0] == th[0, 0]) and np.all(th[:, 2] == th[0, 2])):
            raise ValueError("nnt and u must be the same for every chain (mu may differ)")
        eng = ff.engine(th[0, 0], th[0, 2], th[:, 1])
        logdet, gdiag = eng.forward_fixed()
        ctx.eng = eng
        ctx.set_materialize_grads(False)
        ctx.dev, ctx.dtype = theta_c.device, theta_c.dtype
        return (torch.from_numpy(logdet).to(theta_c.device, theta_c.dtype),
                torch.from_numpy(gdiag).to(theta_c.device, theta_c.dtype))

    @staticmethod
    def backward(ctx, g_logdet, g_gdiag):
        if g_logdet is None and g_gdiag is None:
            return None, None
        gl = None if g_logdet is None else g_logdet.detach().to("cpu", torch.float64).numpy()
        gg = None if g_gdiag is None else g_gdiag.detach().to("cpu", torch.float64).numpy()
        grad = ctx.eng.adjoint(gl, gg)                                  # [chains, 3]
        return torch.from_numpy(grad).to(ctx.dev, ctx.dtype), None


def greens(theta, ff):
    """(logdet [chains], gdiag [chains, nt, ns]) at the fields of `ff`, differentiable in theta.

    theta: tensor (nnt, mu, u) of shape [3] (shared by all chains) or [chains, 3] (mu may differ
    per chain; the backward then yields the per-chain gradient rows)."""
    theta_c = theta.expand(ff.chains, 3) if theta.dim() == 1 else theta
    return _FixedFieldGreens.apply(theta_c, ff)


def log_weight(theta, ff, logdet=None):
    """log W(s) per chain (weight convention in the module docstring)."""
    if logdet is None:
        logdet, _ = greens(theta, ff)
    u = theta[..., 2]
    alpha = torch.acosh(torch.exp(0.5 * ff.dt * u))
    ssum = torch.from_numpy(ff.field_sums()).to(logdet)
    return 2.0 * logdet - alpha * ssum + ff.ns * ff.nt * (-math.log(2.0) - ff.dt * u / 4.0)


def double_occupation(gdiag):
    """Per-chain estimator of include/observable_methods.hpp:57-74: mean_{t,i} (1 - G_ii)^2."""
    return ((1.0 - gdiag) ** 2).mean(dim=(1, 2))


def filling(gdiag):
    """Per-chain estimator of include/observable_methods.hpp:39-52: mean_t (2 - 2 tr G(t) / N)."""
    return (2.0 - 2.0 * gdiag.mean(dim=2)).mean(dim=1)


def derivative_estimator(obs, dobs, dlogw):
    """d<O>/dp from per-configuration samples: <dO/dp> + <O dlogW/dp> - <O><dlogW/dp>."""
    obs, dobs, dlogw = (torch.as_tensor(x, dtype=torch.float64) for x in (obs, dobs, dlogw))
    return dobs.mean() + (obs * dlogw).mean() - obs.mean() * dlogw.mean()
