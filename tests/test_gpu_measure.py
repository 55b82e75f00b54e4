"""Observables measured on the device (dqmc_measure_*) against the oracle's restatement of the
reference's estimators (include/observable_methods.hpp:39-195) on the same Markov chains, bin by
bin, including the bin normalisation (include/observable.hpp:160-207).

Workload: 8x8, U=-4 (u=4), mu=-0.5, beta=8, L=80, pace 10, k=16 -- the single-chain configuration
that measures equal-time and time-displaced Green's functions -- with 2 chains, 2 bins of 2
samples each, following the reference's measuring schedule (src/dqmc_handle.cpp:114-151).  The
estimators are sums of products of G entries, so the device values equal the oracle's within the
G agreement (the oracle's own wrap error) times the sum's condition; the bar is 1e-8 relative.
"""
import math
import sys

import numpy as np
import pytest

import dqmc_amd as gpu
import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def kstars(nl):
    ks = []
    for i in range(math.ceil(nl / 2), nl + 1):
        for j in range(math.ceil(nl / 2), nl + 1):
            ks.append((i / nl * 2 * math.pi - math.pi, j / nl * 2 * math.pi - math.pi))
    return np.array(ks)


def test_device_observables_match_oracle_bins():
    nl, nt, beta, u, mu, pace, k, nch, nbins, cap = 8, 80, 8.0, 4.0, -0.5, 10, 16, 2, 2, 2
    say("case: device observables 8x8 L=80 beta=8, %d chains, %d bins x %d samples" % (nch, nbins, cap))
    kxy = kstars(nl)
    e = gpu.Engine(nl, nt, beta, u, mu=mu, pace=pace, n_chains=nch, delay_k=k, eqtime=True, dynamic=True)
    orcs = []
    for c in range(nch):
        e.set_seed(c, 61 + c)
        e.fields_random(c)
        o = orc.Chain(nl, nt, beta, u, mu, pace=pace, eqtime=True, dynamic=True)
        o.set_seed(61 + c)
        o.fields_random()
        assert o.init() == 0
        o.set_thermalization(False)
        orcs.append(o)
    e.init()
    e.set_thermalization(False)
    e.measure_setup(63, nbins, kxy)
    want = [{n: [] for n in ("fill", "docc", "sw", "gf", "ldos", "spin")} for _ in range(nch)]
    for b in range(nbins):
        acc = [{n: 0.0 for n in want[0]} for _ in range(nch)]
        for s in range(cap):
            e.sweep_dynamic()
            e.measure_dynamic()
            for c, o in enumerate(orcs):
                assert o.sweep_dynamic() == 0
                gf, ldos, spin = o.measure_dynamic(kxy)
                acc[c]["gf"] = acc[c]["gf"] + gf
                acc[c]["ldos"] = acc[c]["ldos"] + ldos
                acc[c]["spin"] = acc[c]["spin"] + spin
            e.sweep_down()
            e.measure_equaltime()
            for c, o in enumerate(orcs):
                assert o.sweep_down() == 0
                assert np.array_equal(e.accept_mask(c), o.accmask())
                f, d, w = o.measure_equaltime()
                acc[c]["fill"] += f
                acc[c]["docc"] += d
                acc[c]["sw"] += w
        e.push_bin(b)
        for c in range(nch):
            for n in want[c]:
                want[c][n].append(acc[c][n] / cap)
    names = {"fill": gpu.OBS_FILLING, "docc": gpu.OBS_DOUBLE_OCCU, "sw": gpu.OBS_SWAVE, "gf": gpu.OBS_GREEN_KT,
             "ldos": gpu.OBS_LDOS, "spin": gpu.OBS_SPIN}
    for c in range(nch):
        for n, obs in names.items():
            got = e.bins(c, obs)
            ref = np.array(want[c][n]).reshape(got.shape)
            rel = float(np.abs(got - ref).max() / max(1e-300, np.abs(ref).max()))
            say("  chain %d %-4s max rel diff %.2e" % (c, n, rel))
            assert rel < 1e-8, (c, n, rel)
    # an empty cache cannot be pushed (the reference throws from normalize_cache)
    with pytest.raises(gpu.DqmcError):
        e.push_bin(0)
    for o in orcs:
        o.close()
    e.close()
