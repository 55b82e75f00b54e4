"""Batched chains and graph replay on the GPU.

* 64 chains of the 12x12 workload (N=144, U=-6, beta=6, dtau=0.05, L=120, mu in {-1,-0.5,0}
  round-robin) in one context: every chain follows exactly the Markov chain of a single-chain
  context with the same seed and mu (the reference runs each as its own MPI rank,
  src/main.cpp:111), and one chain follows the oracle.
* Sweeps replayed from captured hipGraphs give the same chains as eager sweeps, on every replay,
  with eager work between the replays.
"""
import sys

import numpy as np
import pytest

import dqmc_amd as gpu
import oracle as orc

pytestmark = pytest.mark.gpu

TOL_G = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_batched_64_chains_12x12_match_single_chains_and_oracle():
    nl, nt, beta, u, pace, k = 12, 120, 6.0, 6.0, 10, 32
    nch, base = 64, 9000
    mus = [[-1.0, -0.5, 0.0][c % 3] for c in range(nch)]
    say("case: batched %d chains nl=%d nt=%d graphs on" % (nch, nl, nt))
    e = gpu.Engine(nl, nt, beta, u, mu=0.0, pace=pace, n_chains=nch, delay_k=k)
    for c in range(nch):
        e.set_chain_mu(c, mus[c])
        e.set_seed(c, base + c)
        e.fields_random(c)
    e.init()
    e.use_graphs(True)
    say("  autotuned tiles", e.tuning())
    picks = [0, 1, 2, 37, 63]
    singles = []
    for c in picks:
        s = gpu.Engine(nl, nt, beta, u, mu=mus[c], pace=pace, n_chains=1, delay_k=k)
        s.set_seed(0, base + c)
        s.fields_random(0)
        s.init()
        singles.append(s)
    o = orc.Chain(nl, nt, beta, u, mus[1], pace=pace)
    o.set_seed(base + 1)
    o.fields_random()
    assert o.init() == 0
    for sweep in range(4):
        up = sweep % 2 == 0
        e.sweep_up() if up else e.sweep_down()
        for c, s in zip(picks, singles):
            s.sweep_up() if up else s.sweep_down()
            assert np.array_equal(e.accept_mask(c), s.accept_mask(0)), "chain %d sweep %d" % (c, sweep)
            d = np.abs(e.gf(c) - s.gf(0)).max()
            assert d < TOL_G, "chain %d sweep %d: %g" % (c, sweep, d)
        if sweep < 2:
            assert (o.sweep_up() if up else o.sweep_down()) == 0
            assert np.array_equal(e.accept_mask(1), o.accmask())
            # At U=-6, beta=6 the FP64 wrap error itself is ~1e-9, and two oracle builds that
            # differ only in their SVD routine (one-sided Jacobi vs LAPACK dgesvd) already
            # disagree by 2e-10 after one sweep; the bar is therefore the oracle's own wrap
            # error (never looser than 1e-8), with identical accept/reject decisions.
            d = np.abs(e.gf(1) - o.gf(0)).max()
            tol = min(1e-8, max(TOL_G, o.stats()["wrap_eq"]))
            say("  sweep %d chain 1 vs oracle max|dG| = %.3e (oracle wrap error %.3e)" % (sweep, d, tol))
            assert d < tol
    acc = sum(e.stats(c)["accepts"] for c in range(nch)) / sum(e.stats(c)["visits"] for c in range(nch))
    wrap = max(e.stats(c)["wrap_eq"] for c in range(nch))
    say("  acceptance %.3f, max wrap error %.3e" % (acc, wrap))
    assert 0.0 < acc < 1.0 and wrap < 1e-6
    for s in singles:
        s.close()
    e.close()


@pytest.mark.parametrize("prec", [gpu.FP64, gpu.MIXED])
def test_graph_replay_matches_eager(prec):
    nl, nt, beta, u, mu, pace, nch = 6, 40, 4.0, 4.0, -0.2, 8, 4
    say("case: graph replay vs eager, precision %d" % prec)
    eng = []
    for graphs in (False, True):
        e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=nch, precision=prec, delay_k=16,
                       eqtime=True)
        for c in range(nch):
            e.set_seed(c, 500 + c)
            e.fields_random(c)
        e.init()
        if eng:                             # same GEMM tiles, so the two differ only in graphs
            t = eng[0].tuning()
            e.set_tuning(t["fp64"], t["fp32"])
        e.use_graphs(graphs)
        eng.append(e)
    for sweep in range(8):
        th = sweep < 4                      # both graph keys: thermalization on, then collecting G(t,t)
        for e in eng:
            e.set_thermalization(th)
            e.sweep_up() if sweep % 2 == 0 else e.sweep_down()
        for c in range(nch):                # eager work between replays
            assert np.array_equal(eng[0].accept_mask(c), eng[1].accept_mask(c)), "sweep %d chain %d" % (sweep, c)
            d = np.abs(eng[0].gf(c) - eng[1].gf(c)).max()
            assert d < (TOL_G if prec == gpu.FP64 else 1e-5), "sweep %d chain %d: %g" % (sweep, c, d)
            if not th:
                for t in (0, nt // 2, nt - 1):
                    d = np.abs(eng[0].gf(c, gpu.GF_TT, t) - eng[1].gf(c, gpu.GF_TT, t)).max()
                    assert d < (TOL_G if prec == gpu.FP64 else 1e-5)
    for e in eng:
        e.close()
