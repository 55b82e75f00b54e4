"""Oracle RNG against libstdc++ <random> known answers (tests/golden/gen_rng_golden.cpp)."""
import json
import os

import numpy as np
import pytest

import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_kat_10000th():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed minstd_rand0.
    d = _load("rng_libstdcxx_nofma.json")
    assert d["kat_10000th"] == 1043618065
    assert int(orc.rng_raw(1, 10000)[-1]) == 1043618065


@pytest.mark.parametrize("fname,fma", [("rng_libstdcxx_nofma.json", False), ("rng_libstdcxx_fma.json", True)])
def test_raw_and_canonical(fname, fma):
    d = _load(fname)
    for s in d["seeds"]:
        raw = orc.rng_raw(s["seed"], len(s["raw"]))
        assert [int(x) for x in raw] == s["raw"], s["seed"]
        can = orc.rng_canonical(s["seed"], len(s["canonical_hex"]), fma=fma)
        want = np.array([float.fromhex(h) for h in s["canonical_hex"]])
        assert np.array_equal(can, want), s["seed"]


def test_bernoulli_consumes_two_draws():
    # bernoulli(p) draws one canonical (= 2 raw draws) for every p, including p = 0 and p = 1.
    d = _load("rng_libstdcxx_fma.json")
    for s in d["seeds"]:
        n = len(s["bernoulli"])
        can = orc.rng_canonical(s["seed"], n, fma=True)
        for (p, b), u in zip(s["bernoulli"], can):
            assert int(u < p) == b
        # state after n bernoulli calls = raw draw number 2n+1
        raw = orc.rng_raw(s["seed"], 2 * n + 1)
        assert int(raw[-1]) == s["state_after_bernoulli"]
