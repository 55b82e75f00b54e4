"""GPU parity: the HIP engine (through the C ABI) against the FP64 CPU oracle on the same
seeds and fields.  Integer/decision data (accept masks, fields, RNG state) must be identical;
FP64 Green's functions must agree within 1e-10 (max abs)."""
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


def make_pair(nl, nt, beta, u, mu, pace, seed, k=1, prec=gpu.FP64, eq=False, dyn=False, chains=1):
    o = orc.Chain(nl, nt, beta, u, mu, pace=pace, eqtime=eq, dynamic=dyn)
    o.set_seed(seed)
    o.fields_random()
    e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=chains, precision=prec, delay_k=k, eqtime=eq,
                   dynamic=dyn)
    e.set_seed(0, seed)
    e.fields_random(0)
    return o, e


def test_fields_random_and_rng_match_oracle():
    say("case: fields_random nl=4 nt=40 seed=20240520")
    o, e = make_pair(4, 40, 4.0, 4.0, 0.0, 10, 20240520)
    assert np.array_equal(e.fields(0).astype(np.int32), o.fields)
    assert e.rng_state(0) == o.rng_state
    e.close()


@pytest.mark.parametrize("op", range(5))
def test_bops_against_oracle(op):
    say("case: apply_bop op=%d nl=6" % op)
    o, e = make_pair(6, 12, 1.2, 4.0, -0.3, 4, 99)
    e.init()
    rng = np.random.default_rng(op)
    g = rng.standard_normal((36, 36))
    for t in (0, 5, 11):
        want = o.apply_bop(op, t, g)
        got = e.apply_bop(0, op, t, g)
        assert np.abs(got - want).max() < 1e-13 * max(1.0, np.abs(want).max())
    e.close()


@pytest.mark.parametrize("nl,nt,beta,mu", [(4, 40, 4.0, 0.0), (8, 80, 8.0, -0.5), (3, 10, 1.0, 0.2)])
def test_init_green_function(nl, nt, beta, mu):
    say("case: init G nl=%d nt=%d beta=%g" % (nl, nt, beta))
    o, e = make_pair(nl, nt, beta, 4.0, mu, 10, 7 + nl)
    assert o.init() == 0
    e.init()
    d = np.abs(e.gf(0) - o.gf(0)).max()
    say("  max|dG| =", d)
    assert d < TOL_G
    e.close()


def _run_parity(nl, nt, beta, u, mu, pace, seed, k, sweeps):
    o, e = make_pair(nl, nt, beta, u, mu, pace, seed, k=k)
    assert o.init() == 0
    e.init()
    worst = 0.0
    for s in range(sweeps):
        if s % 2 == 0:
            assert o.sweep_up() == 0
            e.sweep_up()
        else:
            assert o.sweep_down() == 0
            e.sweep_down()
        assert np.array_equal(e.accept_mask(0), o.accmask()), "accept trace differs in sweep %d" % s
        assert np.array_equal(e.fields(0).astype(np.int32), o.fields)
        assert e.rng_state(0) == o.rng_state
        d = np.abs(e.gf(0) - o.gf(0)).max()
        worst = max(worst, d)
        assert d < TOL_G, "sweep %d max|dG| %g" % (s, d)
    so, sg = o.stats(), e.stats(0)
    assert so["accepts"] == sg["accepts"]
    say("  accepts", sg["accepts"], "worst |dG|", worst, "wrap", sg["wrap_eq"], so["wrap_eq"])
    e.close()


def test_sweep_parity_fp64_rank1_c1():
    # 4x4, U=-4 (u=4), mu=0, beta=4, dt=0.1 (L=40), pace 10: the CPU-runnable configuration.
    say("case: sweep parity FP64 k=1 nl=4 nt=40")
    _run_parity(4, 40, 4.0, 4.0, 0.0, 10, 12345, 1, 6)


@pytest.mark.parametrize("k", [8, 32])
def test_sweep_parity_fp64_delayed(k):
    say("case: sweep parity FP64 delayed k=%d nl=4 nt=40" % k)
    _run_parity(4, 40, 4.0, 4.0, 0.0, 10, 777, k, 4)


def test_sweep_parity_fp64_8x8_c2_shape():
    # 8x8, mu=-0.5, beta=8, L=80, pace 10 (the single-chain FP64 configuration), delayed k=16.
    say("case: sweep parity FP64 k=16 nl=8 nt=80")
    _run_parity(8, 80, 8.0, 4.0, -0.5, 10, 4242, 16, 2)


def test_batched_chains_match_single_chains():
    say("case: batched 3 chains, per-chain mu")
    nl, nt, beta, u, pace = 4, 20, 2.0, 6.0, 5
    mus = [-1.0, -0.5, 0.0]
    seeds = [101, 202, 303]
    e = gpu.Engine(nl, nt, beta, u, mu=0.0, pace=pace, n_chains=3, delay_k=8)
    os_ = []
    for c in range(3):
        e.set_chain_mu(c, mus[c])
        e.set_seed(c, seeds[c])
        e.fields_random(c)
        o = orc.Chain(nl, nt, beta, u, mus[c], pace=pace)
        o.set_seed(seeds[c])
        o.fields_random()
        assert o.init() == 0
        os_.append(o)
    e.init()
    for s in range(4):
        e.sweep_up() if s % 2 == 0 else e.sweep_down()
        for c in range(3):
            o = os_[c]
            assert (o.sweep_up() if s % 2 == 0 else o.sweep_down()) == 0
            assert np.array_equal(e.accept_mask(c), o.accmask())
            assert np.abs(e.gf(c) - o.gf(0)).max() < TOL_G
    e.close()


def test_dynamic_sweep_parity():
    say("case: dynamic sweep nl=4 nt=40 beta=4")
    nl, nt, beta = 4, 40, 4.0
    o, e = make_pair(nl, nt, beta, 4.0, -0.3, 10, 55, k=8, eq=True, dyn=True)
    assert o.init() == 0
    e.init()
    o.set_thermalization(False)
    e.set_thermalization(False)
    for rnd in range(2):
        assert o.sweep_dynamic() == 0
        e.sweep_dynamic()
        for t in range(nt):
            for kind in (1, 2, 3):
                d = np.abs(e.gf(0, kind, t) - o.gf(kind, t)).max()
                assert d < TOL_G, "kind %d t %d: %g" % (kind, t, d)
        assert o.sweep_down() == 0
        e.sweep_down()
        assert np.array_equal(e.accept_mask(0), o.accmask())
    so, sg = o.stats(), e.stats(0)
    say("  wrap eq/dyn gpu", sg["wrap_eq"], sg["wrap_dyn"], "cpu", so["wrap_eq"], so["wrap_dyn"])
    e.close()


def test_mixed_precision_stabilization_restores_fp64():
    # Mixed mode (FP32 wraps and flushes, FP64 stabilization): every sweep ends on a
    # stabilization, so the working G must then equal the FP64 G of the chain's own fields
    # (oracle: UDV stacks from scratch) up to the FP32 rounding of the working copy.  The
    # reported drift max|dG| is finite and at least that rounding.
    say("case: mixed precision nl=8 nt=40 beta=4")
    nl, nt, beta, u, mu, pace = 8, 40, 4.0, 4.0, -0.3, 10
    e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=2, precision=gpu.MIXED, delay_k=32)
    for c in range(2):
        e.set_seed(c, 31 + c)
        e.fields_random(c)
    e.init()
    for s in range(4):
        e.sweep_up() if s % 2 == 0 else e.sweep_down()
        for c in range(2):
            o = orc.Chain(nl, nt, beta, u, mu, pace=pace)
            o.fields = e.fields(c)
            assert o.init() == 0
            d = np.abs(e.gf(c) - o.gf(0)).max()
            say("  sweep %d chain %d |G_mixed - G_FP64(fields)| = %.3e" % (s, c, d))
            assert d < 1e-6
            o.close()
    for c in range(2):
        st = e.stats(c)
        say("  chain %d reported max|dG| = %.3e" % (c, st["max_abs_dG"]))
        assert np.isfinite(st["max_abs_dG"]) and st["max_abs_dG"] > 1e-9
    e.close()
