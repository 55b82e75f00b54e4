"""Alternative kernel paths selected by environment switches, on the GPU.

Every switch changes how a step is computed, never what it computes, so a context built under
it must follow exactly the Markov chains of a default context with the same seeds: identical
accept/reject decisions, and Green's functions equal to rounding (the bar 1e-9 is the wrap error
of this workload's FP64 chains, never reached by rounding differences between two correct
evaluation orders).  Workload: 16x16 (N = 256), U = -4, mu = -0.2, beta = 4, L = 40, pace 8,
k = 32, 8 chains (a multiple of 8, so the chain-major block order of the batched kernels is the
one exercised).

  DQMC_GEMM_TILE=1|3|4   every GEMM with the 32x32 tile, the 64x64 tile with double-buffered
                         LDS and two-step global prefetch, or the FP64 128x128 tile, instead of
                         the autotuner's choice
  DQMC_SOLVE=qr          recompose solve by Householder QR instead of partial-pivot LU
  DQMC_QR_PANEL=1024     QR panel by 16 waves (four threads per row) instead of 4 waves
  DQMC_QR_UPDATE=gemm    QR trailing update by three GEMM launches instead of one fused kernel
  DQMC_WRAP=dense        slice products by dense GEMMs with exp(-dt K) instead of Kronecker MFMA
  DQMC_TRSM=v1           column-parallel back-substitution instead of the blocked one
  DQMC_TRSM_BLOCK=32|128 blocked back-substitution with 32- or 128-row diagonal blocks instead of 64
  DQMC_RECOMPOSE_FUSE=0  the recompose system by two GEMMs and a right-hand-side copy, instead of one
                         launch over both products that also stores the right-hand side, and the
                         stabilization's compare-and-commit as its own pass
  DQMC_NORM_FUSE=0       the push's pre-pivot reading the B-chain for its column norms, instead of
                         taking them from the B-chain kernel's last pass
  DQMC_RHS_IDENTITY=0    the push's pre-pivot writing the identity right-hand side of its QR, instead of
                         the QR's first update taking those columns as the identity unread
  DQMC_QR=v1             one-workgroup unblocked Householder QR instead of the blocked one
  DQMC_METRO_THREADS     Metropolis workgroup of 256 or 512 threads instead of 1024
  DQMC_LOOKAHEAD=0       blocked QR and LU panels and updates in order on one stream, instead of
                         each next panel beside the rest of the previous update on a second stream
  DQMC_WRAP_FUSE=0       every wrap's left product by its own launch, instead of by the Metropolis
                         workgroups of the slice (up sweeps: the wrap to the next slice unless a
                         stabilization comes between; down sweeps: the slice's own wrap, its left
                         product before its right one)
  DQMC_METRO_SPLIT=1|4   every Metropolis launch over all chains, or over four groups of the chains
                         on four streams, instead of two halves on two streams between the points
                         where all chains meet again
  DQMC_PANEL_FUSE=0      a separate launch for the look-ahead panel's own columns of the previous
                         trailing update, instead of the panel applying it first
  DQMC_CHAIN=slice       each B-chain factor applied when its slice is done, instead of all factors
                         of a stack block in one pass at its push (also checked at pace 17, where
                         a block takes two passes of at most 16 factors)
"""
import sys

import numpy as np
import pytest

import dqmc_amd as gpu

pytestmark = pytest.mark.gpu

NL, NT, BETA, U, MU, PACE, K, NCH = 16, 40, 4.0, 4.0, -0.2, 8, 32, 8
TOL_G = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_chains(n_sweeps, pace=PACE, beta=BETA):
    e = gpu.Engine(NL, NT, beta, U, mu=MU, pace=pace, n_chains=NCH, delay_k=K)
    for c in range(NCH):
        e.set_seed(c, 4100 + c)
        e.fields_random(c)
    e.init()
    out = []
    for sweep in range(n_sweeps):
        e.sweep_up() if sweep % 2 == 0 else e.sweep_down()
        out.append(([e.accept_mask(c).copy() for c in range(NCH)], [e.gf(c).copy() for c in range(NCH)]))
    wrap = max(e.stats(c)["wrap_eq"] for c in range(NCH))
    e.close()
    return out, wrap


@pytest.fixture(scope="module")
def default_run():
    say("case: default kernels, 16x16 L=40, 8 chains, 2 sweeps")
    return run_chains(2)


@pytest.mark.parametrize("var,val", [("DQMC_GEMM_TILE", "1"), ("DQMC_GEMM_TILE", "3"), ("DQMC_GEMM_TILE", "4"), ("DQMC_SOLVE", "qr"), ("DQMC_QR_PANEL", "1024"), ("DQMC_QR_UPDATE", "gemm"), ("DQMC_WRAP", "dense"), ("DQMC_TRSM", "v1"),
                                     ("DQMC_QR", "v1"), ("DQMC_METRO_THREADS", "256"),
                                     ("DQMC_METRO_THREADS", "512"), ("DQMC_LOOKAHEAD", "0"),
                                     ("DQMC_CHAIN", "slice"), ("DQMC_WRAP_FUSE", "0"), ("DQMC_METRO_SPLIT", "1"),
                                     ("DQMC_METRO_SPLIT", "4"), ("DQMC_PANEL_FUSE", "0"),
                                     ("DQMC_TRSM_BLOCK", "32"), ("DQMC_TRSM_BLOCK", "128"),
                                     ("DQMC_RECOMPOSE_FUSE", "0"), ("DQMC_NORM_FUSE", "0"),
                                     ("DQMC_RHS_IDENTITY", "0")])
def test_switch_follows_default_chains(default_run, monkeypatch, var, val):
    say("case: %s=%s, 16x16 L=40, 8 chains, 2 sweeps" % (var, val))
    monkeypatch.setenv(var, val)
    alt, wrap = run_chains(2)
    compare_runs(alt, wrap, *default_run)


def test_chain_block_longer_than_one_pass(monkeypatch):
    # blocks of 17, 17 and 6 slices at beta = 1 (dt = 0.025), where wrap intervals of 17 slices keep
    # the FP64 wrap error as small as pace 8 does at beta = 4
    say("case: pace 17 (B-chain blocks of 17 factors), beta 1, fused against DQMC_CHAIN=slice, 16x16 L=40, 8 chains")
    fused, wf = run_chains(2, pace=17, beta=1.0)
    monkeypatch.setenv("DQMC_CHAIN", "slice")
    per_slice, ws = run_chains(2, pace=17, beta=1.0)
    say("  wrap errors: fused %.2e, per slice %.2e" % (wf, ws))
    compare_runs(fused, wf, per_slice, ws)


def compare_runs(alt, wrap, ref, ref_wrap, wrap_max=1e-8, tol=TOL_G):
    assert ref_wrap < wrap_max and wrap < wrap_max
    for sweep, ((ma, ga), (mr, gr)) in enumerate(zip(alt, ref)):
        for c in range(NCH):
            assert np.array_equal(ma[c], mr[c]), "chain %d sweep %d: accept decisions differ" % (c, sweep)
            d = np.abs(ga[c] - gr[c]).max()
            assert d < tol, "chain %d sweep %d: max|dG| %g" % (c, sweep, d)
        say("  sweep %d: decisions identical, max|dG| = %.2e" %
            (sweep, max(np.abs(ga[c] - gr[c]).max() for c in range(NCH))))


def run_stored(n_sweeps):
    """Measuring sweeps (equal-time G of every slice stored) on the 16x16 workload."""
    e = gpu.Engine(NL, NT, BETA, U, mu=MU, pace=PACE, n_chains=NCH, delay_k=K)
    for c in range(NCH):
        e.set_seed(c, 4300 + c)
        e.fields_random(c)
    e.set_collect(True, False)
    e.init()
    e.set_thermalization(False)
    out = []
    for sweep in range(n_sweeps):
        e.sweep_up() if sweep % 2 == 0 else e.sweep_down()
        out.append(([e.accept_mask(c).copy() for c in range(NCH)],
                    [np.stack([e.gf(c, gpu.GF_TT, t) for t in range(NT)]) for c in range(NCH)]))
    e.close()
    return out


def test_stored_gf_with_fused_wraps(monkeypatch):
    # The slices' wraps run inside the Metropolis launches; the G(t) stored for measurements must
    # still be the one of each slice after its update (up sweeps) or after its wrap (down sweeps), as
    # with separate wraps.
    say("case: stored G(t), fused wraps against DQMC_WRAP_FUSE=0, 16x16 L=40, 8 chains")
    fused = run_stored(2)
    monkeypatch.setenv("DQMC_WRAP_FUSE", "0")
    plain = run_stored(2)
    for sweep, ((ma, ga), (mr, gr)) in enumerate(zip(fused, plain)):
        for c in range(NCH):
            assert np.array_equal(ma[c], mr[c]), "chain %d sweep %d: accept decisions differ" % (c, sweep)
            d = np.abs(ga[c] - gr[c]).max()
            assert d < TOL_G, "chain %d sweep %d: max|dG(t)| %g" % (c, sweep, d)
        say("  sweep %d: decisions identical, max|dG(t)| = %.2e" %
            (sweep, max(np.abs(ga[c] - gr[c]).max() for c in range(NCH))))
