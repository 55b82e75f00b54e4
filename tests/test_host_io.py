"""Host-side drop-in pieces (config, fields, output formats, statistics) against the
reference's own example files (tests/golden/example, from the reference's example/ run)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "example")


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "host_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-o", exe, os.path.join(HERE, "cpp", "host_check.cpp")])
    return exe


def run(exe, *args):
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True)


def test_parse_example_config(host_check):
    r = run(host_check, "parse", os.path.join(EX, "config.toml"), 4)
    assert r.returncode == 0, r.stderr
    # 20 bins split over 4 ranks (include/dqmc_parser.hpp:94), dt = beta / nt
    assert "nl=4 ns=16 nt=160 beta=8 dt=0.050000000000000003 nnt=1 u=4 mu=0 pace=10 warmup=1000 bins=5 cap=100 between=10 fft=0 nobs=6" in r.stdout
    assert "momentum=AllKstars" in r.stdout


@pytest.mark.parametrize("body,msg", [
    ("[Model.HubbardAttractiveU]\nu = -1.0\n[Measurement]\nobservables = [\"none\"]\n", "'u' out of range"),
    ("[Lattice.SquareLattice]\nnl = 0\n[Measurement]\nobservables = [\"none\"]\n", "'nl' out of range"),
    ("[Measurement]\nobservables = [\"all\", \"none\"]\n", "conflict observable options"),
    ("[Measurement]\nbin_num = 3\n", "invalid input observables"),
])
def test_parse_errors(host_check, tmp_path, body, msg):
    f = tmp_path / "c.toml"
    f.write_text(body)
    r = run(host_check, "parse", f, 1)
    assert r.returncode == 3 and msg in r.stderr


def test_fields_round_trip_is_byte_exact(host_check, tmp_path):
    out = tmp_path / "ising.fields"
    r = run(host_check, "fields", os.path.join(EX, "ising.fields"), out, 160, 16)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == open(os.path.join(EX, "ising.fields"), "rb").read()


def test_fields_dimension_mismatch(host_check, tmp_path):
    r = run(host_check, "fields", os.path.join(EX, "ising.fields"), tmp_path / "x", 80, 16)
    assert r.returncode == 3 and "contradictory dimensions" in r.stderr


@pytest.mark.parametrize("stem,dims", [("filling", []), ("double_occu", []), ("swave_corr", []), ("ldos", [160]),
                                       ("spin_corr", [160])])
def test_statistics_and_out_format_match_example(host_check, tmp_path, stem, dims):
    # mean, ddof=0 stddev and stderr of the example's own bins reproduce its *.out byte for byte,
    # including the trailing-zero quirk of scalar observables (example/double_occu.out).
    out = tmp_path / (stem + ".out")
    r = run(host_check, "stats", os.path.join(EX, stem + ".bins.out"), out, len(dims), *dims)
    assert r.returncode == 0, r.stderr
    got = out.read_text().splitlines()
    want = open(os.path.join(EX, stem + ".out")).read().splitlines()
    if not dims:
        assert got == want            # byte-exact, including the 0-d printer's trailing blanks
        return
    # vectors: identical layout; values equal up to the last printed digit (the example's bins
    # are printed to 15 decimals and the reference's mean sums in another order)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert len(g) == len(w)
        gv, wv = [float(x) for x in g.split()], [float(x) for x in w.split()]
        assert gv[:len(dims)] == wv[:len(dims)]
        assert max(abs(a - b) for a, b in zip(gv, wv)) <= 5e-15


def test_momenta_and_tgrids_match_example(host_check, tmp_path):
    m = tmp_path / "momenta.out"
    assert run(host_check, "momenta", 4, m).returncode == 0
    assert m.read_text() == open(os.path.join(EX, "momenta.out")).read()
    t = tmp_path / "tgrids.out"
    assert run(host_check, "tgrids", 160, 8.0, t).returncode == 0
    assert t.read_text() == open(os.path.join(EX, "tgrids.out")).read()
