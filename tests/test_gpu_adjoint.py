"""Fixed-field forward and parameter adjoint (dqmc_forward_fixed / dqmc_adjoint) on the GPU.

Pinned by: high-precision brute force of log det(1 + B_L..B_1) and G(t) (mpmath); the U=0
closed forms (G = (1 + e^{-beta K})^{-1}); the atomic limit nnt=0 (K diagonal: every site is a
two-level problem with x_i = alpha sum_t s_ti + beta mu); and central finite differences
(h = 1e-5) of the engine's own forward on the 8x8, L=40 workload with 256 field configurations.
"""
import math
import sys

import numpy as np
import pytest
import torch

import brute
import dqmc_amd as gpu
import dqmc_torch as dt_

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_fields(seed, chains, nt, ns):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((chains, nt, ns)) < 0.5, 1, -1).astype(np.int8)


def test_forward_fixed_matches_brute_force_and_keeps_chain_state():
    nl, nt, beta, u, mu, pace = 4, 20, 2.0, 4.0, -0.3, 5
    say("case: forward_fixed vs mpmath, nl=4 nt=20 beta=2")
    f = random_fields(11, 3, nt, nl * nl)
    dt = beta / nt
    alpha = math.acosh(math.exp(0.5 * dt * u))
    e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=3, delay_k=8)
    twin = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=3, delay_k=8)
    for c in range(3):
        for x in (e, twin):
            x.set_seed(c, 70 + c)
            x.set_fields(c, f[c])
    e.init()
    twin.init()
    ld, gd = e.forward_fixed()
    expk = brute.fourier_expk(nl, dt, 1.0, mu)
    for c in range(3):
        want = brute.logdet_mp(expk, alpha, f[c].astype(int), dps=60)
        say("  chain %d logdet %.12f brute %.12f" % (c, ld[c], want))
        assert abs(ld[c] - want) < 1e-10 * max(1.0, abs(want))
        for t in (0, 1, 7, nt - 1):
            g = brute.greens_mp(expk, alpha, f[c].astype(int), t=t if t else nt, dps=60)
            assert np.abs(gd[c, t] - np.diag(g)).max() < 1e-10, (c, t)
    # the chain continues exactly as if the forward had not run
    for x in (e, twin):
        x.sweep_up()
    for c in range(3):
        assert np.array_equal(e.accept_mask(c), twin.accept_mask(c))
        assert np.abs(e.gf(c) - twin.gf(c)).max() < 1e-11
    e.close()
    twin.close()


def test_adjoint_u0_closed_form():
    # U = 0: B_t = e^{-dt K} for every t, G(t) = (1 + e^{-beta K})^{-1}; with f_k the Fermi factor
    # of eps_k = -nnt a_k - mu:  d logdet/dmu = beta sum f_k,  d logdet/dnnt = beta sum a_k f_k,
    # D = n^2 with n = mean f_k.  (d/du diverges at u = 0 and is not checked.)
    nl, nt, beta, nnt, mu = 4, 16, 2.0, 0.8, -0.4
    say("case: adjoint at U=0 against closed forms")
    ns = nl * nl
    f = random_fields(5, 2, nt, ns)
    e = gpu.Engine(nl, nt, beta, 0.0, mu=mu, nnt=nnt, pace=4, n_chains=2, delay_k=1)
    for c in range(2):
        e.set_fields(c, f[c])
    e.init()
    ld, gd = e.forward_fixed()
    a = brute.square_lattice_bands(nl)
    eps = -nnt * a - mu
    fk = 1.0 / (1.0 + np.exp(beta * eps))
    assert np.abs(ld - np.sum(np.log1p(np.exp(-beta * eps)))).max() < 1e-11
    g1 = e.adjoint(np.ones(2), None)
    say("  d logdet/d(nnt, mu):", g1[0, :2], "closed form", beta * np.sum(a * fk), beta * np.sum(fk))
    assert np.abs(g1[:, 0] - beta * np.sum(a * fk)).max() < 1e-9
    assert np.abs(g1[:, 1] - beta * np.sum(fk)).max() < 1e-9
    n = fk.mean()
    wD = -2.0 * (1.0 - gd) / (nt * ns)                     # d D / d G(t)_ii
    g2 = e.adjoint(None, wD)
    dn_dmu = np.mean(beta * fk * (1 - fk))
    dn_dnnt = np.mean(beta * a * fk * (1 - fk))
    say("  dD/d(nnt, mu):", g2[0, :2], "closed form", 2 * n * dn_dnnt, 2 * n * dn_dmu)
    assert np.abs(g2[:, 0] - 2 * n * dn_dnnt).max() < 1e-10
    assert np.abs(g2[:, 1] - 2 * n * dn_dmu).max() < 1e-10
    e.close()


def test_adjoint_atomic_limit_closed_form():
    # nnt = 0: site i sees x_i = alpha m_i + beta mu with m_i = sum_t s_ti; f_i = 1/(1+e^{-x_i}),
    # logdet = sum log(1 + e^{x_i}), G(t)_ii = 1 - f_i; the hopping derivative of the diagonal
    # vanishes (bipartite lattice) and so does that of logdet (G diagonal, A off-diagonal).
    nl, nt, beta, u, mu = 4, 24, 3.0, 5.0, 0.3
    say("case: adjoint in the atomic limit against closed forms")
    ns = nl * nl
    f = random_fields(9, 3, nt, ns)
    e = gpu.Engine(nl, nt, beta, u, mu=mu, nnt=0.0, pace=6, n_chains=3, delay_k=1)
    for c in range(3):
        e.set_fields(c, f[c])
    e.init()
    ld, gd = e.forward_fixed()
    dt = beta / nt
    alpha = math.acosh(math.exp(0.5 * dt * u))
    dalpha = dt * math.exp(0.5 * dt * u) / (2.0 * math.sinh(alpha))
    m = f.astype(float).sum(axis=1)                         # [chains, ns]
    x = alpha * m + beta * mu
    fi = 1.0 / (1.0 + np.exp(-x))
    assert np.abs(ld - np.log1p(np.exp(x)).sum(axis=1)).max() < 1e-10
    assert np.abs(gd - (1.0 - fi)[:, None, :]).max() < 1e-12
    g1 = e.adjoint(np.ones(3), None)
    assert np.abs(g1[:, 0]).max() < 1e-10
    assert np.abs(g1[:, 1] - beta * fi.sum(axis=1)).max() < 1e-9
    assert np.abs(g1[:, 2] - dalpha * (m * fi).sum(axis=1)).max() < 1e-9
    wD = -2.0 * (1.0 - gd) / (nt * ns)
    g2 = e.adjoint(None, wD)
    dD_dx = 2.0 * fi * fi * (1.0 - fi) / ns
    say("  dD/du", g2[:, 2], "closed form", (dD_dx * dalpha * m).sum(axis=1))
    assert np.abs(g2[:, 0]).max() < 1e-10
    assert np.abs(g2[:, 1] - (dD_dx * beta).sum(axis=1)).max() < 1e-10
    assert np.abs(g2[:, 2] - (dD_dx * dalpha * m).sum(axis=1)).max() < 1e-10
    e.close()


def test_adjoint_matches_central_differences_8x8_256_configs():
    # 8x8, U=-4, mu=-0.3, beta=4, dtau=0.1 (L=40), 256 field configurations from seed; gradient
    # of log W and of the double occupancy with respect to (nnt, mu, u) through the autograd
    # Function, against central differences of the engine's own forward with h = 1e-5.
    nl, nt, beta = 8, 40, 4.0
    theta0 = np.array([1.0, -0.3, 4.0])
    chains, h = 256, 1e-5
    say("case: adjoint vs central differences, nl=8 nt=40, %d configs" % chains)
    ff = dt_.FixedFields(nl, nt, beta, random_fields(20240520, chains, nt, nl * nl), pace=10)
    rng = np.random.default_rng(3)
    wa = torch.from_numpy(rng.standard_normal(chains))
    wb = torch.from_numpy(rng.standard_normal(chains))

    def evaluate(theta):
        logdet, gd = dt_.greens(theta, ff)
        return dt_.log_weight(theta, ff, logdet), dt_.double_occupation(gd)

    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    lw, dd = evaluate(theta)
    ((wa * lw).sum() + (wb * dd).sum()).backward()
    adj = theta.grad.numpy().copy()
    # per-chain rows of d logW / d theta from a [chains, 3] leaf
    theta_c = torch.tensor(np.tile(theta0, (chains, 1)), dtype=torch.float64, requires_grad=True)
    lw_c, _ = evaluate(theta_c)
    lw_c.sum().backward()
    rows = theta_c.grad.numpy()
    fd = np.zeros(3)
    fd_rows = np.zeros((chains, 3))
    with torch.no_grad():
        for p in range(3):
            vals = []
            for sgn in (+1, -1):
                th = theta0.copy()
                th[p] += sgn * h
                lwp, ddp = evaluate(torch.tensor(th, dtype=torch.float64))
                vals.append(((wa * lwp).sum() + (wb * ddp).sum()).item())
                fd_rows[:, p] += sgn * lwp.numpy() / (2 * h)
            fd[p] = (vals[0] - vals[1]) / (2 * h)
    for p, name in enumerate(("nnt", "mu", "u")):
        say("  d/d%-3s adjoint % .10e  central difference % .10e  rel %.2e" %
            (name, adj[p], fd[p], abs(adj[p] - fd[p]) / abs(fd[p])))
    assert np.all(np.abs(adj - fd) <= 1e-6 * np.abs(fd) + 1e-6)
    err_rows = np.abs(rows - fd_rows) / (np.abs(fd_rows) + 1.0)
    say("  per-chain d logW rows: max rel diff %.2e" % err_rows.max())
    assert err_rows.max() < 1e-6
    ff.close()
