"""GPU parity on lattices above 16 x 16 (N up to 1024) and with is_fft, through the C ABI, against the
FP64 CPU oracle on the same seeds and fields.

Above 16 x 16 the engine runs the slice products factored on a 32 x 32 zero-padded ring factor (the
right product through two batched transposes), the wide QR panel kernel (8 waves, up to two
rows per thread) with the three-GEMM trailing update, the Householder QR recompose solve and a delay
capped by LDS (src/hubbard.cpp:25-48 and :293-341; src/dqmc_core.cpp:167-284; the reference accepts any
nl >= 1, include/dqmc_parser.hpp:125).  With is_fft the reference forms exp(-+dt K) through the lattice
dispersion (src/hubbard.cpp:50-81, 159-246); the oracle restates that form, and the engine applies the
same operator either way.

Decisions (accept masks, fields, RNG state) must be identical; G must agree within
min(1e-6, max(1e-9, oracle wrap error)).
"""
import sys

import numpy as np
import pytest

import dqmc_amd as gpu
import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def oracle_mkl():
    say("oracle MKL:", orc.set_mkl(1) == 0)
    yield
    orc.set_mkl(0)


def check_chain(e, c, o, what, floor=1e-9):
    assert np.array_equal(e.accept_mask(c), o.accmask()), "%s: accept trace differs" % what
    assert np.array_equal(e.fields(c).astype(np.int32), o.fields), "%s: fields differ" % what
    assert e.rng_state(c) == o.rng_state, "%s: RNG state differs" % what
    st = o.stats()
    tol = min(1e-6, max(floor, st["wrap_eq"], st["wrap_dyn"]))
    d = float(np.abs(e.gf(c) - o.gf(0)).max())
    say("  %s: decisions identical, max|dG| = %.3e (bar %.3e)" % (what, d, tol))
    assert d < tol, "%s: max|dG| %g above %g" % (what, d, tol)


@pytest.mark.parametrize("nl,nt,beta,pace,k,nch", [(17, 12, 1.2, 4, 32, 2), (18, 10, 1.5, 5, 32, 3),
                                                    (20, 8, 1.0, 3, 16, 2), (24, 6, 0.9, 3, 32, 1)])
def test_large_lattice_oracle_parity(nl, nt, beta, pace, k, nch):
    u, mu, base = 4.0, -0.2, 4242
    say("case: nl=%d (N=%d) nt=%d pace=%d k=%d chains=%d" % (nl, nl * nl, nt, pace, k, nch))
    e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=nch, delay_k=k)
    for c in range(nch):
        e.set_seed(c, base + c)
        e.fields_random(c)
    e.init()
    pick = nch - 1
    o = orc.Chain(nl, nt, beta, u, mu, pace=pace)
    o.set_seed(base + pick)
    o.fields_random()
    assert o.init() == 0
    d = float(np.abs(e.gf(pick) - o.gf(0)).max())
    say("  init max|dG| = %.3e" % d)
    assert d < 1e-9
    # the five slice operators B, B^-1, left and right, and B^T (src/hubbard.cpp:293-341)
    g = np.random.default_rng(nl).standard_normal((nl * nl, nl * nl))
    for op in range(5):
        want = o.apply_bop(op, nt - 2, g)
        err = float(np.abs(e.apply_bop(pick, op, nt - 2, g) - want).max() / np.abs(want).max())
        say("  B-op %d: rel err %.2e" % (op, err))
        assert err < 1e-13
    for sweep in range(2):
        up = sweep % 2 == 0
        e.sweep_up() if up else e.sweep_down()
        assert (o.sweep_up() if up else o.sweep_down()) == 0
        check_chain(e, pick, o, "sweep %d" % sweep)
    o.close()
    e.close()


def test_largest_lattice_32x32():
    nl, nt, beta, u, mu, pace = 32, 4, 0.4, 4.0, 0.1, 2
    if orc.set_mkl(1) != 0:
        pytest.skip("the oracle needs MKL's dgesvd / dsyev to finish an N = 1024 chain in seconds")
    say("case: nl=32 (N=1024) nt=%d pace=%d" % (nt, pace))
    e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=1, delay_k=32)
    e.set_seed(0, 99)
    e.fields_random(0)
    e.init()
    o = orc.Chain(nl, nt, beta, u, mu, pace=pace)
    o.set_seed(99)
    o.fields_random()
    assert o.init() == 0
    d = float(np.abs(e.gf(0) - o.gf(0)).max())
    say("  init max|dG| = %.3e" % d)
    assert d < 1e-9
    e.sweep_up()
    assert o.sweep_up() == 0
    check_chain(e, 0, o, "sweep 0")
    o.close()
    e.close()


@pytest.mark.parametrize("nl,k", [(6, 8), (16, 32), (20, 16)])
def test_is_fft_matches_oracle_fft_products(nl, k):
    nt, beta, u, mu, pace, seed = 20, 2.0, 4.0, -0.4, 5, 515
    say("case: is_fft nl=%d k=%d" % (nl, k))
    ef = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=1, delay_k=k, is_fft=True)
    ed = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=1, delay_k=k)
    o = orc.Chain(nl, nt, beta, u, mu, pace=pace, is_fft=True)
    for e in (ef, ed):
        e.set_seed(0, seed)
        e.fields_random(0)
        e.init()
    o.set_seed(seed)
    o.fields_random()
    assert o.init() == 0
    g = np.random.default_rng(nl).standard_normal((nl * nl, nl * nl))
    for op in range(5):
        want = o.apply_bop(op, 3, g)
        assert np.abs(ef.apply_bop(0, op, 3, g) - want).max() < 1e-12 * np.abs(want).max()
    for sweep in range(2):
        up = sweep % 2 == 0
        for e in (ef, ed):
            e.sweep_up() if up else e.sweep_down()
        assert (o.sweep_up() if up else o.sweep_down()) == 0
        check_chain(ef, 0, o, "sweep %d" % sweep, floor=1e-10)
        # the flag selects how the reference builds the products, not what they are
        assert np.array_equal(ef.gf(0), ed.gf(0))
    o.close()
    ef.close()
    ed.close()
