"""Brute-force references for the DQMC numerics, from the mathematics only.

B(t) = diag(exp(alpha*s_t)) @ expK with expK = exp(-dt*K), K = -nnt*A - mu*I,
G(tau_t) = (1 + B(t)...B(1) B(L)...B(t+1))^{-1}.  Products and the inverse are done
in mpmath at high precision on the FP64 input matrices, so the result is the exact
inverse of those inputs; no stabilization is involved.
"""
import math

import mpmath as mp
import numpy as np


def fourier_expk(nl, dt, nnt, mu, sign=-1.0):
    """exp(sign*dt*K) from the plane-wave eigenbasis of the periodic square lattice:
    eps_k = -2 nnt (cos kx + cos ky) - mu, counting periodic images like the adjacency does."""
    n = nl * nl
    ks = [2 * math.pi * m / nl for m in range(nl)]
    out = np.zeros((n, n))
    x = np.arange(n) % nl
    y = np.arange(n) // nl
    for kx in ks:
        for ky in ks:
            eps = -2.0 * nnt * (math.cos(kx) + math.cos(ky)) - mu
            ph = np.exp(1j * (kx * x + ky * y))
            out += (np.outer(ph, ph.conj()) * math.exp(sign * dt * eps)).real
    return out / n


def bmat(expk, alpha, s_t):
    return np.exp(alpha * np.asarray(s_t, dtype=float))[:, None] * expk


def _to_mp(a):
    return mp.matrix(a.tolist())


def _to_np(m):
    return np.array(m.tolist(), dtype=float)


def greens_mp(expk, alpha, fields, t=0, dps=120):
    """(1 + B(t)...B(1) B(L)...B(t+1))^{-1} at `dps` digits; fields is [L][N]."""
    L, n = fields.shape
    with mp.workdps(dps):
        E = _to_mp(expk)
        M = mp.eye(n)
        order = list(range(t + 1, L + 1)) + list(range(1, t + 1))   # rightmost first
        for l in order:
            d = [mp.e ** (mp.mpf(alpha) * int(fields[l % L][i])) for i in range(n)]
            EM = E * M
            for i in range(n):
                for j in range(n):
                    EM[i, j] *= d[i]
            M = EM
        G = (mp.eye(n) + M) ** -1
        return _to_np(G)


def greens_np(expk, alpha, fields, t=0):
    """Plain FP64 version, adequate only for small beta."""
    L, n = fields.shape
    M = np.eye(n)
    for l in list(range(t + 1, L + 1)) + list(range(1, t + 1)):
        M = bmat(expk, alpha, fields[l % L]) @ M
    return np.linalg.inv(np.eye(n) + M)


def logdet_mp(expk, alpha, fields, dps=120):
    """log|det(1 + B(L)...B(1))| at `dps` digits on the FP64 inputs; fields is [L][N]."""
    L, n = fields.shape
    with mp.workdps(dps):
        E = _to_mp(expk)
        M = mp.eye(n)
        for l in list(range(1, L + 1)):
            d = [mp.e ** (mp.mpf(alpha) * int(fields[l % L][i])) for i in range(n)]
            EM = E * M
            for i in range(n):
                for j in range(n):
                    EM[i, j] *= d[i]
            M = EM
        return float(mp.log(abs(mp.det(mp.eye(n) + M))))


def square_lattice_bands(nl):
    """a_k = 2 (cos kx + cos ky): eigenvalues of the adjacency (periodic images counted)."""
    ks = [2 * math.pi * m / nl for m in range(nl)]
    return np.array([2.0 * (math.cos(kx) + math.cos(ky)) for kx in ks for ky in ks])


def displaced_mp(expk, alpha, fields, dps=100):
    """Time-displaced Green's functions at `dps` digits on the FP64 inputs; fields is [L][N].

    Returns (gt0, g0t), each [L][N][N] in the reference's storage order (src/dqmc_core.cpp:343-382):
    index t >= 1 holds G(t,0) = B(t)...B(1) G(0) and G(0,t) = -(1 - G(0)) (B(t)...B(1))^{-1};
    index 0 holds G(0,0) and G(0,0) - 1 (the tau -> 0 conventions)."""
    L, n = fields.shape
    with mp.workdps(dps):
        E = _to_mp(expk)
        bs = []
        for l in range(1, L + 1):
            d = [mp.e ** (mp.mpf(alpha) * int(fields[l % L][i])) for i in range(n)]
            Bl = E.copy()
            for i in range(n):
                for j in range(n):
                    Bl[i, j] *= d[i]
            bs.append(Bl)
        M = mp.eye(n)
        for Bl in bs:
            M = Bl * M
        G0 = (mp.eye(n) + M) ** -1
        gt0 = [_to_np(G0)]
        g0t = [_to_np(G0 - mp.eye(n))]
        P = mp.eye(n)
        for t in range(1, L):
            P = bs[t - 1] * P
            gt0.append(_to_np(P * G0))
            g0t.append(_to_np(-(mp.eye(n) - G0) * P ** -1))
        return np.array(gt0), np.array(g0t)
