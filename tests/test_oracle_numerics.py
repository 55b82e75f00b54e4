"""Pins the oracle's numerics against mathematics (closed forms, brute force, invariants)."""
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla

import brute
import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_nn_hoppings_accumulate():
    # square_lattice.cpp:48-61: nl=2 -> 2 on the single x and y neighbours, nl=1 -> 4 on the diagonal.
    assert orc.nn_hoppings(1)[0, 0] == 4.0
    a2 = orc.nn_hoppings(2)
    assert a2[0, 1] == 2.0 and a2[0, 2] == 2.0 and a2[0, 3] == 0.0
    a4 = orc.nn_hoppings(4)
    assert np.allclose(a4, a4.T) and np.all(a4.sum(0) == 4)


@pytest.mark.parametrize("nl,dt,nnt,mu", [(1, 0.1, 1.0, 0.0), (2, 0.1, 1.0, -0.5), (4, 0.05, 1.0, 0.0), (6, 0.1, 0.7, 0.3)])
def test_expk_closed_form(nl, dt, nnt, mu):
    e, ie = orc.expk(nl, dt, nnt, mu)
    n = nl * nl
    K = -nnt * orc.nn_hoppings(nl) - mu * np.eye(n)
    assert np.abs(e - sla.expm(-dt * K)).max() < 1e-13
    assert np.abs(ie - sla.expm(dt * K)).max() < 1e-13
    assert np.abs(e - brute.fourier_expk(nl, dt, nnt, mu, -1.0)).max() < 1e-13
    assert np.abs(e @ ie - np.eye(n)).max() < 1e-13


def test_svd_reconstructs_graded_matrix():
    rng = np.random.default_rng(0)
    n = 24
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.exp(np.linspace(30, -30, n))
    a = (q1 * d) @ q2.T @ np.diag(np.exp(rng.uniform(-20, 20, n)))
    rc, u, s, v = orc.svd(a)
    assert rc == 0
    assert np.all(np.diff(s) <= 0)
    assert np.abs(u.T @ u - np.eye(n)).max() < 1e-13
    assert np.abs(v.T @ v - np.eye(n)).max() < 1e-13
    assert np.abs((u * s) @ v.T - a).max() / np.abs(a).max() < 1e-13
    # singular values against numpy for a well-conditioned case
    b = rng.standard_normal((n, n))
    assert np.allclose(orc.svd(b)[2], np.linalg.svd(b, compute_uv=False), rtol=1e-12)


def test_inverse_and_fullpiv_solve():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((30, 30))
    rc, ai = orc.inverse(a)
    assert rc == 0 and np.abs(ai @ a - np.eye(30)).max() < 1e-12
    b = rng.standard_normal((30, 30))
    x = orc.fullpiv_qr_solve(a, b)
    assert np.abs(a @ x - b).max() < 1e-11


def _chain(nl, nt, beta, u, mu, pace, seed, eq=False, dyn=False):
    c = orc.Chain(nl, nt, beta, u, mu, pace=pace, eqtime=eq, dynamic=dyn)
    c.set_seed(seed)
    c.fields_random()
    return c


@pytest.mark.parametrize("op", range(5))
def test_bops_against_numpy(op):
    c = _chain(3, 8, 1.0, 4.0, 0.2, 4, 7)
    e, ie = c.expk()
    rng = np.random.default_rng(op)
    g = rng.standard_normal((9, 9))
    t = 5
    d = np.diag(np.exp(c.alpha * c.fields[t]))
    di = np.diag(np.exp(-c.alpha * c.fields[t]))
    want = [d @ e @ g, g @ d @ e, ie @ di @ g, g @ ie @ di, e @ d @ g][op]
    assert np.abs(c.apply_bop(op, t, g) - want).max() < 1e-13


@pytest.mark.parametrize("t", [0, 10, 20, 37, 40])
def test_equaltime_gf_vs_mpmath(t):
    # nl=4, beta=4, L=40, pace 10: general branch (0<t<L) and the two QR branches (t=0, t=L).
    c = _chain(4, 40, 4.0, 4.0, -0.3, 10, 11)
    rc, g = c.equaltime_gf_at(t)
    assert rc == 0
    e, _ = c.expk()
    gm = brute.greens_mp(e, c.alpha, c.fields, t=t % 40, dps=80)
    assert np.abs(g - gm).max() < 1e-10


def test_dynamic_gf_vs_mpmath():
    # G(t,0) = B(t)..B(1) G(0) and G(0,t) = -(1 - G(0)) (B(t)..B(1))^{-1} for every stored slice,
    # including the tau -> 0 conventions at index 0, from a dynamic sweep with 3 stabilizations.
    nl, nt = 3, 12
    c = orc.Chain(nl, nt, 1.5, 4.0, 0.2, pace=5, eqtime=True, dynamic=True)
    c.set_seed(23)
    c.fields_random()
    assert c.init() == 0
    c.set_thermalization(False)
    assert c.sweep_dynamic() == 0
    e, _ = c.expk()
    gt0, g0t = brute.displaced_mp(e, c.alpha, c.fields, dps=60)
    for t in range(nt):
        assert np.abs(c.gf(2, t) - gt0[t]).max() < 1e-11, t
        assert np.abs(c.gf(3, t) - g0t[t]).max() < 1e-11, t


def test_init_gf_example_fields_vs_mpmath():
    # The reference's own field fixture (example/ising.fields: nl=4, nt=160, beta=8, u=4, mu=0).
    rows = np.loadtxt(os.path.join(GOLD, "example", "ising.fields"), skiprows=1, dtype=int)
    f = np.zeros((160, 16), dtype=np.int32)
    f[rows[:, 0], rows[:, 1]] = rows[:, 2]
    c = orc.Chain(4, 160, 8.0, 4.0, 0.0, pace=10)
    c.fields = f
    assert c.init() == 0
    e, _ = c.expk()
    gm = brute.greens_mp(e, c.alpha, f, t=0, dps=140)
    assert np.abs(c.gf(0) - gm).max() < 1e-10


def test_ratio_and_sherman_morrison_invariant():
    # One up sweep at tiny size: after the sweep G must equal (1+B_L..B_1)^{-1} of the
    # final fields (so every accepted SM update and flip was consistent), and the
    # acceptance probability min(1,r) matches the brute-force determinant ratio.
    c = _chain(2, 8, 1.0, 4.0, 0.1, 4, 3)
    assert c.init() == 0
    e, _ = c.expk()
    g0 = brute.greens_np(e, c.alpha, c.fields, 0)
    assert np.abs(c.gf(0) - g0).max() < 1e-12
    # determinant ratio for flipping (t=1, i=2) from G(1,1)
    f = c.fields
    g1 = brute.greens_np(e, c.alpha, f, 1)
    a = c.alpha
    delta = math.exp(-2 * a * f[1, 2]) - 1
    r = (1 + (1 - g1[2, 2]) * delta) ** 2 / (1 + delta)
    f2 = f.copy(); f2[1, 2] *= -1
    def detw(ff):
        M = np.eye(4)
        for l in range(1, 9):
            M = brute.bmat(e, a, ff[l % 8]) @ M
        return np.linalg.det(np.eye(4) + M) ** 2 * math.exp(-a * ff.sum())
    assert abs(r - detw(f2) / detw(f)) < 1e-12 * max(1, r)
    assert c.sweep_up() == 0
    gf = brute.greens_np(e, c.alpha, c.fields, 0)
    assert np.abs(c.gf(0) - gf).max() < 1e-11
    st = c.stats()
    assert st["visits"] == 8 * 4 and 0 < st["accepts"] <= st["visits"]


def test_u0_closed_form():
    # U=0: alpha=0, every move accepted, G = (1+exp(-beta K))^{-1} exactly.
    c = _chain(4, 20, 2.0, 0.0, -0.4, 5, 5, eq=True)
    c.set_thermalization(False)
    assert c.init() == 0 and c.alpha == 0.0
    c.sweep_up()
    c.sweep_down()
    st = c.stats()
    assert st["accepts"] == st["visits"]
    K = -orc.nn_hoppings(4) + 0.4 * np.eye(16)
    g = np.linalg.inv(np.eye(16) + sla.expm(-2.0 * K))
    assert np.abs(c.gf(0) - g).max() < 1e-12
    for t in (0, 7, 19):
        assert np.abs(c.gf(1, t) - g).max() < 1e-12


def test_atomic_limit_double_occupancy():
    # nnt=0: K = -mu I is diagonal; sites decouple and the Trotter error vanishes.
    # <n_up n_dn> per site = e^{beta U/4} / (2 e^{beta U/4} + 2 e^{-beta U/4}) at mu=0 (attractive U).
    nl, nt, beta, u = 2, 20, 2.0, 2.0
    c = orc.Chain(nl, nt, beta, u, 0.0, nnt=0.0, pace=5, eqtime=True)
    c.set_seed(42)
    c.fields_random()
    c.init()
    c.set_thermalization(True)
    for _ in range(20):
        c.sweep_up(); c.sweep_down()
    c.set_thermalization(False)
    acc = []
    for _ in range(400):
        c.sweep_up()
        acc.append(c.measure_equaltime()[1])
        c.sweep_down()
        acc.append(c.measure_equaltime()[1])
    acc = np.array(acc)
    want = math.exp(beta * u / 4) / (2 * math.exp(beta * u / 4) + 2 * math.exp(-beta * u / 4))
    err = acc.std() / math.sqrt(len(acc) / 20)   # generous autocorrelation allowance
    assert abs(acc.mean() - want) < 5 * err + 1e-3


def test_observables_u0_closed_form():
    # U = 0: G(t) = (1 + e^{-beta K})^{-1} for every slice and G(t,0) = e^{-tau K} G, so with the
    # Fermi factor f_k of eps_k = -2 nnt (cos kx + cos ky) - mu every estimator of
    # include/observable_methods.hpp has a closed form:
    #   filling 2 mean f, double occupancy (mean f)^2, s-wave pairing mean(f^2 + (1-f)^2),
    #   G(k, tau) = e^{-tau eps_k} (1 - f_k), LDOS mean_q e^{-tau eps_q} (1 - f_q),
    #   transverse spin mean_q(e^{tau eps} f) * mean_q(e^{-tau eps} (1 - f)).
    nl, nt, beta, nnt, mu = 4, 16, 2.0, 1.0, -0.4
    c = orc.Chain(nl, nt, beta, 0.0, mu, nnt=nnt, pace=4, eqtime=True, dynamic=True)
    c.set_seed(3)
    c.fields_random()
    assert c.init() == 0
    c.set_thermalization(False)
    assert c.sweep_dynamic() == 0
    ks = [2 * math.pi * m / nl for m in range(nl)]
    eps = np.array([-2 * nnt * (math.cos(kx) + math.cos(ky)) - mu for kx in ks for ky in ks])
    f = 1.0 / (1.0 + np.exp(beta * eps))
    fill, docc, swave = c.measure_equaltime()
    assert abs(fill - 2 * f.mean()) < 1e-12
    assert abs(docc - f.mean() ** 2) < 1e-12
    assert abs(swave - np.mean(f ** 2 + (1 - f) ** 2)) < 1e-12
    kst = np.array([(i / nl * 2 * math.pi - math.pi, j / nl * 2 * math.pi - math.pi)
                    for i in range(math.ceil(nl / 2), nl + 1) for j in range(math.ceil(nl / 2), nl + 1)])
    gk, ldos, spin = c.measure_dynamic(kst)
    tau = np.arange(nt) * beta / nt
    for q, (kx, ky) in enumerate(kst):
        ek = -2 * nnt * (math.cos(kx) + math.cos(ky)) - mu
        fk = 1.0 / (1.0 + math.exp(beta * ek))
        assert np.abs(gk[q] - np.exp(-tau * ek) * (1 - fk)).max() < 1e-11
    want_ldos = np.array([np.mean(np.exp(-t * eps) * (1 - f)) for t in tau])
    want_spin = np.array([np.mean(np.exp(t * eps) * f) * np.mean(np.exp(-t * eps) * (1 - f)) for t in tau])
    assert np.abs(ldos - want_ldos).max() < 1e-11
    assert np.abs(spin - want_spin).max() < 1e-11


def test_compare_error_is_signed_max():
    """StableNumerics::matrix_compare_error (include/utils/stable_numerics.hpp:38-48) is the signed
    max of (a - b), not the max-abs: a pair whose largest |a - b| entry is negative returns the
    largest positive difference, and an all-negative difference returns a negative number."""
    rng = np.random.default_rng(3)
    b = rng.standard_normal((6, 6))
    a = b.copy()
    a[1, 2] -= 1e-3              # largest |difference|, negative
    a[4, 0] += 2e-4              # largest positive difference
    assert orc.compare_error(a, b) == (a - b).max()
    assert abs(orc.compare_error(a, b) - 2e-4) < 1e-15
    c = b - 0.5
    assert orc.compare_error(c, b) == (c - b).max() and (c - b).max() < -0.49


def _planted_chain(plant):
    """U = 0, nnt = 0, mu = 0: B = I, G = I/2 exactly and every flip is accepted with gamma = 0,
    so a matrix planted into G after init reaches the only stabilization of a sweep (nt = pace)
    unchanged, and the wrap error is exactly the signed max of (I/2 - (I/2 + P)) = max(-P),
    taken into a running max that starts at 0 (src/dqmc_core.cpp:211-212)."""
    nl, nt = 4, 4
    o = orc.Chain(nl, nt, 1.0, 0.0, 0.0, nnt=0.0, pace=nt)
    o.set_seed(5)
    o.fields_random()
    assert o.init() == 0
    g0 = o.gf(0)
    assert np.array_equal(g0, 0.5 * np.eye(nl * nl))
    o.set_gftt(g0 + plant)
    assert o.sweep_up() == 0
    assert np.array_equal(o.gf(0), g0)
    st = o.stats()
    o.close()
    return st["wrap_eq"]


def test_wrap_error_running_signed_max_planted():
    n = 16
    p = np.zeros((n, n))
    p[1, 2] = 1e-3               # recompute - working = -1e-3: the largest |difference|
    p[3, 0] = -2e-4              # recompute - working = +2e-4: the signed max
    assert _planted_chain(p) == 2e-4
    # every difference negative: the running max stays at its initial 0, neither -eps nor +eps
    assert _planted_chain(np.full((n, n), 1e-3)) == 0.0
