"""The is_fft products (src/hubbard.cpp:50-81, 159-246): the oracle's momentum-space restatement,
pinned against numpy's FFT, scipy's matrix exponential and the dense products.

The reference applies (inv)expK = F^-1 diag(e^{-+dt E(k)}) F column by column with a 2-D DFT F over
the site index i = x + nl*y (src/hubbard.cpp:172-189) and from the right through a transpose
(:218-230).  Its eigenvalues E(k) are the diagonal of F K F^-1 (:63-76).
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle as orc


def hopping(nl, nnt, mu):
    return -nnt * orc.nn_hoppings(nl) - mu * np.eye(nl * nl)


@pytest.mark.parametrize("nl,nnt,mu", [(1, 1.0, 0.0), (2, 1.0, -0.5), (3, 0.8, 0.2), (4, 1.0, 0.0), (5, 1.0, -0.3),
                                       (8, 1.2, 0.4)])
def test_fft_eigens_vs_numpy_and_closed_form(nl, nnt, mu):
    dt = 0.1
    lf, li = orc.fft_eigens(nl, dt, nnt, mu)
    # numpy's 2-D FFT of K's column 0 laid out as [y][x]: E[ky][kx]
    K = hopping(nl, nnt, mu)
    E = np.fft.fft2(K[:, 0].reshape(nl, nl)).real.ravel()
    assert np.abs(lf - np.exp(-dt * E)).max() < 1e-14
    assert np.abs(li - np.exp(dt * E)).max() < 1e-13
    # they are the spectrum of K (eigenvalues with multiplicity)
    assert np.abs(np.sort(np.log(lf) / -dt) - np.linalg.eigvalsh(K)).max() < 1e-12
    if nl >= 3:   # single bonds: E(k) = -2 nnt (cos kx + cos ky) - mu
        k = 2 * np.pi * np.arange(nl) / nl
        closed = (-2 * nnt * (np.cos(k)[None, :] + np.cos(k)[:, None]) - mu).ravel()
        assert np.abs(E - closed).max() < 1e-13


def numpy_fft_left(g, nl, lam):
    """(inv)expK g column by column with numpy's FFT (an independent DFT)."""
    out = np.empty_like(g)
    for j in range(g.shape[1]):
        v = g[:, j].reshape(nl, nl)
        out[:, j] = np.fft.ifft2(np.fft.fft2(v) * lam.reshape(nl, nl)).real.ravel()
    return out


@pytest.mark.parametrize("nl", [1, 2, 3, 4, 6])
def test_fft_bops_match_numpy_and_dense(nl):
    nt, beta, u, mu = 6, 1.2, 4.0, -0.2
    n = nl * nl
    dense = orc.Chain(nl, nt, beta, u, mu, pace=3)
    fft = orc.Chain(nl, nt, beta, u, mu, pace=3, is_fft=True)
    dense.set_seed(7)
    dense.fields_random()
    fft.fields = dense.fields
    rng = np.random.default_rng(nl)
    g = rng.standard_normal((n, n))
    dt = beta / nt
    lf, li = orc.fft_eigens(nl, dt, 1.0, mu)
    K = hopping(nl, 1.0, mu)
    e, ie = sla.expm(-dt * K), sla.expm(dt * K)
    alpha = fft.alpha
    s = fft.fields[2]
    D = np.exp(alpha * s)
    expected = {
        0: D[:, None] * (e @ g),                 # B(t) g            (src/hubbard.cpp:293-299)
        1: (g * D[None, :]) @ e,                 # g B(t)            (:301-307)
        2: ie @ (g / D[:, None]),                # B(t)^-1 g         (:309-315)
        3: (g @ ie) / D[None, :],                # g B(t)^-1         (:317-323)
        4: e @ (D[:, None] * g),                 # B(t)^T g          (:333-341)
    }
    assert np.abs(numpy_fft_left(g, nl, lf) - e @ g).max() < 1e-12
    for op in range(5):
        a = fft.apply_bop(op, 2, g)
        b = dense.apply_bop(op, 2, g)
        scale = np.abs(expected[op]).max()
        assert np.abs(a - expected[op]).max() < 1e-13 * scale, op
        assert np.abs(a - b).max() < 1e-13 * scale, op
    if nl >= 2:
        ng = numpy_fft_left(D[:, None] * g, nl, lf)
        assert np.abs(fft.apply_bop(4, 2, g) - ng).max() < 1e-13 * np.abs(ng).max()


def test_fft_sweeps_follow_dense_trajectory():
    nl, nt, beta, u, mu, pace = 4, 20, 2.0, 4.0, -0.3, 5
    chains = []
    for fft in (False, True):
        c = orc.Chain(nl, nt, beta, u, mu, pace=pace, is_fft=fft)
        c.set_seed(20240520)
        c.fields_random()
        assert c.init() == 0
        chains.append(c)
    d, f = chains
    assert np.abs(d.gf(0) - f.gf(0)).max() < 1e-12
    for step in ("sweep_up", "sweep_down", "sweep_up"):
        assert getattr(d, step)() == 0 and getattr(f, step)() == 0
        assert np.array_equal(d.accmask(), f.accmask())
        assert np.array_equal(d.fields, f.fields)
        assert np.abs(d.gf(0) - f.gf(0)).max() < 1e-10


@pytest.mark.parametrize("nl", [1, 2, 3, 4, 5, 6])
def test_packed_real_transform_holds_too_few_frequencies(nl):
    """Why the restatement keeps the full spectrum.  A real 2-D transform in the conjugate-even packed
    layout the reference requests (fft_solver.hpp:60-61) stores nl x (nl/2+1) frequencies, the form
    numpy's rfft2 / irfft2 also use.  The dispersion step of src/hubbard.cpp:63-76 run on that layout
    (cache zero-initialised, only the packed part written; the second batch reads a row of site-indexed
    values as a packed spectrum) reproduces E(k) for nl <= 2, where nothing is dropped, and not beyond."""
    n = nl * nl
    K = hopping(nl, 1.0, 0.3)
    P = nl * (nl // 2 + 1)
    disp = np.zeros((n, n), complex)
    for i in range(n):
        col = np.zeros(n, complex)
        col[:P] = np.fft.rfft2(K[:, i].reshape(nl, nl)).ravel()
        disp[:, i] = col
    for j in range(n):
        disp[j, :] = np.fft.irfft2(disp[j, :P].reshape(nl, nl // 2 + 1), s=(nl, nl)).ravel()
    packed_diag = disp.diagonal().real
    full = np.fft.fft2(K[:, 0].reshape(nl, nl)).real.ravel()
    # the packed frequency index p = k1 (nl/2+1) + k2 and the site-ordered full spectrum
    k1, k2 = np.divmod(np.arange(P), nl // 2 + 1)
    want = full.reshape(nl, nl)[k1, k2]
    if nl <= 2:
        assert np.abs(packed_diag[:P] - want).max() < 1e-13
    else:
        assert np.abs(packed_diag[:P] - want).max() > 0.5
