#!/usr/bin/env python3
"""Benchmark of the MI355X DQMC sweep engine on the north-star workload.

Workload (one "step" = one directional updating sweep of every chain of the batch):
  16x16 square lattice (N=256), nnt=1, u=4 (U=-4), mu=0, beta=8, nt=80 (dt=0.1),
  stabilization every 8 slices (the largest pace whose FP64 recompute error stays below
  1e-6 over all chains at this beta*|K|; at pace 10 the FP64 wrap error of the worst chain
  reaches ~3e-6), fields Bernoulli(0.5) from seeds base+c, FP64 G, delayed rank-k updates,
  B independent chains batched on one GPU.
Metric: aggregate sweeps/s = B * steps / wall seconds of the timed region (inputs resident).
Also measured in the same run:
  * the whole sweep against the FP64 peak with the canonical algorithmic flop count,
  * the drop-in application's sweeps/s at 8x8, beta=8, L=80 with all six observables,
  * the time of one fixed-field forward and one parameter adjoint over 256 configurations at 8x8, L=40,
  * the plain FP64 GPU form (1 chain, rank-1 Sherman-Morrison updates),
  * max |G - G_FP64| over the run (FP64 recompute at every stabilization),
  * the mixed FP32 form's |dG| on a short run (reported, not the value),
  * the CPU baseline: the FP64 CPU restatement (oracle) on the host cores,
  * live per-launch HIP-event timing of the dominant kernel for the roofline entry.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hubbard-attractive-u_amd"))

NL, NT, BETA, U, MU, PACE = 16, 80, 8.0, 4.0, 0.0, 8
BASE_SEED = 20240520
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


# ----------------------------------------------------------------------------- CPU baseline
def _cpu_chain(args):
    idx, seconds, use_mkl = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # test infrastructure: CPU baseline leg only
    mkl = orc.set_mkl(1) == 0 if use_mkl else False
    c = orc.Chain(NL, NT, BETA, U, MU, pace=PACE)
    c.set_seed(BASE_SEED + idx)
    c.fields_random()
    c.init()
    c.set_thermalization(True)
    c.sweep_up()                       # untimed warm-up sweep
    t0 = time.time()
    n = 0
    while True:
        c.sweep_down() if n % 2 == 0 else c.sweep_up()
        n += 1
        if time.time() - t0 >= seconds and n >= 2:
            break
    return n, time.time() - t0, mkl


def usable_cores():
    """CPUs this process may run on (its affinity mask), not capped."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(seconds):
    cores = usable_cores()
    ctx = mp.get_context("spawn")
    t0 = time.time()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_chain, [(i, seconds, True) for i in range(cores)])
    wall = time.time() - t0
    sweeps = sum(r[0] for r in res)
    rate = sum(r[0] / r[1] for r in res)
    return {"value": round(rate, 4), "unit": "sweeps/s", "cores": cores, "kind": "port",
            "sample": "%d single-threaded FP64 oracle chains (one per core, %s GEMM/SVD), N=256 L=80, "
                      "%d sweeps in %.1f s after 1 warm-up sweep each" %
                      (cores, "MKL" if res[0][2] else "in-repo C", sweeps, wall)}


def profile_value(name, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def measured_traffic():
    """HBM bytes per launch of the Metropolis kernel on this workload, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE x fetch scale + WRITE_SIZE x write scale, the scales
    from a device copy of known size), and its MFMA busy fraction from the MFMA counter run."""
    out = {}
    t = profile_value("metropolis_kernel_n256_b256_pmc.md", "traffic_bytes_per_launch")
    if t is not None:
        out["traffic"] = int(t)
        out["traffic_source"] = "profiles/metropolis_kernel_n256_b256_pmc.md"
    u = profile_value("mfma_f64_sweep_n256_b256_pmc.md", "metropolis_mfma_busy_frac")
    if u is not None:
        out["mfma_busy_frac"] = u
        out["mfma_source"] = "profiles/mfma_f64_sweep_n256_b256_pmc.md"
    return out


def mixed_table(gpu, paces, chains, sweeps):
    """Markdown table of the mixed-precision error max|G_mixed - G_FP64| against the stabilization
    pace on the benchmark lattice (FP32 wraps and flushes, FP64 stabilization)."""
    rows = ["| pace | stabilizations per sweep | max over chains and sweeps of max|G_mixed - G_FP64| |",
            "|---|---|---|"]
    for pace in paces:
        e = gpu.Engine(NL, NT, BETA, U, mu=MU, pace=pace, n_chains=chains, precision=gpu.MIXED, delay_k=32)
        for c in range(chains):
            e.set_seed(c, BASE_SEED + c)
            e.fields_random(c)
        e.init()
        run_sweeps(e, sweeps, True)
        dG = max(e.stats(c)["max_abs_dG"] for c in range(chains))
        e.close()
        rows.append("| %d | %d | %.3e |" % (pace, -(-NT // pace), dG))
        print("[bench] mixed pace %d: max|dG| %.3e" % (pace, dG), file=sys.stderr, flush=True)
    return "\n".join(rows)


def mixed_batched_12x12(gpu, sweeps=2):
    """Mixed precision on the batched workload: 64 chains of 12x12, U=-6, beta=6, L=120 (dt=0.05),
    pace 10, mu in {-1, -0.5, 0} round-robin over the chains, FP32 wraps and flushes, k=32: the
    largest max|G_mixed - G_FP64| over the chains and stabilizations of a short run."""
    chains = 64
    e = gpu.Engine(12, 120, 6.0, 6.0, mu=0.0, pace=10, n_chains=chains, precision=gpu.MIXED, delay_k=32)
    for c in range(chains):
        e.set_chain_mu(c, (-1.0, -0.5, 0.0)[c % 3])
        e.set_seed(c, BASE_SEED + c)
        e.fields_random(c)
    e.init()
    run_sweeps(e, sweeps, True)
    dG = max(e.stats(c)["max_abs_dG"] for c in range(chains))
    e.close()
    return dG


def app_throughput():
    """The drop-in application at the 8x8, U=-4, mu=-0.5, beta=8, L=80 (pace 10) configuration with
    all six observables measured on the device: directional sweeps per second of its simulation
    (warm-up, dynamic and down sweeps, measurements and binning), from the time it prints."""
    import re
    import subprocess
    import tempfile
    app = os.path.join(ROOT, "hubbard-attractive-u_amd", "hubbard-attractive-u")
    if not os.path.exists(app):
        return None
    warm, bins, cap = 20, 2, 10
    with tempfile.TemporaryDirectory() as d:
        cfg = os.path.join(d, "c2.toml")
        with open(cfg, "w") as f:
            f.write("""[Model.HubbardAttractiveU]
    nnt = 1.0
    u = 4.0
    mu = -0.5
[Lattice.SquareLattice]
    nl = 8
    momentum_list = "AllKstars"
[MonteCarlo]
    beta = 8.0
    nt = 80
    stabilization_pace = 10
[Measurement]
    sweeps_warmup = %d
    bin_num = %d
    bin_capacity = %d
    sweeps_between_bins = 0
    observables = [ "FillingNumber", "DoubleOccupation", "StaticSWavePairingCorrelation",
                    "DynamicGreenFunctions", "LocalDensityOfStates", "LocalDynamicTransverseSpinCorrelation" ]
[Engine]
    chains = 1
    seed = %d
    delay = 16
""" % (warm, bins, cap, BASE_SEED))
        r = subprocess.run([app, "-c", cfg, "-o", os.path.join(d, "out")], capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            return {"error": r.stderr[-300:]}
        m = re.search(r"finished in (?:(\d+) m )?([0-9.]+) s", r.stdout)
        if not m:
            return None
        secs = (int(m.group(1)) * 60 if m.group(1) else 0) + float(m.group(2))
    sweeps = warm + bins * cap * 2   # updating warm-up, then per sample one dynamic and one down sweep
    return {"sweeps_per_s": round(sweeps / secs, 2), "sweeps": sweeps, "seconds": secs,
            "workload": "app: 8x8 U=-4 mu=-0.5 beta=8 L=80 pace 10, 1 chain, all six observables on the device"}


def adjoint_c5(gpu):
    """The parameter adjoint on its workload (8x8, U=-4, mu=-0.3, beta=4, L=40, 256 field
    configurations from seed): one fixed-field forward (log|det|, G(t)_ii of every configuration)
    and one adjoint (the cotangents contracted with d/d(nnt, mu, u)), each timed after a warm-up call.
    The adjoint runs L stabilized time-displaced passes per gradient."""
    import numpy as np
    nl, nt, beta, chains = 8, 40, 4.0, 256
    e = gpu.Engine(nl, nt, beta, 4.0, mu=-0.3, nnt=1.0, pace=10, n_chains=chains, precision=gpu.FP64, delay_k=1)
    for c in range(chains):
        e.set_seed(c, BASE_SEED + c)
        e.fields_random(c)
    e.init()
    gl = np.ones(chains)
    gg = np.full((chains, nt, nl * nl), 1.0 / (nt * nl * nl))
    e.forward_fixed()
    e.adjoint(gl, gg)
    e.synchronize()
    t0 = time.perf_counter()
    e.forward_fixed()
    t1 = time.perf_counter()
    e.adjoint(gl, gg)
    t2 = time.perf_counter()
    e.close()
    return {"workload": "adjoint: 8x8 U=-4 mu=-0.3 beta=4 L=40, 256 fixed configurations, FP64",
            "forward_ms": round((t1 - t0) * 1e3, 3), "adjoint_ms": round((t2 - t1) * 1e3, 3),
            "gradients_per_s": round(chains / (t2 - t0), 2)}


# ----------------------------------------------------------------------------- GPU legs
def make_engine(gpu, chains, precision, k):
    e = gpu.Engine(NL, NT, BETA, U, mu=MU, pace=PACE, n_chains=chains, precision=precision, delay_k=k)
    for c in range(chains):
        e.set_seed(c, BASE_SEED + c)
        e.fields_random(c)
    e.init()
    e.set_thermalization(True)
    return e


def run_sweeps(e, n, start_up):
    for s in range(n):
        (e.sweep_up if (s % 2 == 0) == start_up else e.sweep_down)()


def timed(e, chains, steps, warmup, graphs):
    e.use_graphs(graphs)
    run_sweeps(e, warmup, True)
    e.synchronize()
    start_up = warmup % 2 == 0
    t0 = time.perf_counter()
    run_sweeps(e, steps, start_up)
    e.synchronize()
    dt = time.perf_counter() - t0
    return chains * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--chains", type=int, default=256)
    ap.add_argument("--delay", type=int, default=32)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 (1 chain, k=1) leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision check leg")
    ap.add_argument("--no-app", action="store_true", help="skip the drop-in application leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the parameter-adjoint timing leg")
    ap.add_argument("--mixed-table", action="store_true",
                    help="only print the table of mixed-precision |dG| against the stabilization pace")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one device copy of 256 MiB (the reference kernel for rocprofv3 --pmc scales)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported "
              "(chains are batched inside one device instead)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.mixed_table:
        import dqmc_amd as gpu
        print(mixed_table(gpu, (1, 2, 4, 8), 8, 2))
        print("\nbatched 12x12 (64 chains, U=-6, beta=6, L=120, pace 10, mu in {-1,-0.5,0}): max|G_mixed - G_FP64| "
              "= %.3e" % mixed_batched_12x12(gpu))
        return

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)

    import dqmc_amd as gpu
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.calibrate_copy:
        import torch
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: %d bytes read, %d bytes written" % (src.numel() * 8, dst.numel() * 8))
        del src, dst

    # headline: FP64 batched, delayed updates
    log("[bench] batched FP64: chains=%d k=%d" % (args.chains, args.delay))
    e = make_engine(gpu, args.chains, gpu.FP64, args.delay)
    value, dt = timed(e, args.chains, args.steps, args.warmup, not args.no_graphs)
    dG = max(e.stats(c)["max_abs_dG"] for c in range(args.chains))
    st0 = e.stats(0)
    acc = sum(e.stats(c)["accepts"] for c in range(args.chains)) / float(sum(e.stats(c)["visits"] for c in range(args.chains)))
    log("[bench] value %.2f sweeps/s (%.3f s for %d steps)" % (value, dt, args.steps))

    # roofline: live per-launch timing (HIP events on the engine's stream) of the Metropolis
    # kernel, the path's largest single kernel, over two extra sweeps; the other timer classes
    # give the kernel-time shares
    qr_name = ("qr_v1_kernel" if os.environ.get("DQMC_QR", "") == "v1"
               else "blocked QR and LU (panel + update kernels per 32 columns)")
    trsm_name = ("trsm_v1_kernel" if os.environ.get("DQMC_TRSM", "") == "v1"
                 else "blocked trsm (trsm_diag_kernel + gemm_kernel<double>)")
    kernel_names = {0: "metropolis_kernel<double, 1024>", 1: qr_name, 2: "GEMMs (stack/recompose)",
                    3: trsm_name, 4: "kron_left/right_kernel<double>"}
    prof = {}
    for kid in (0, 1, 2, 3, 4):
        e.kernel_timer_begin(kid, 4096)
        a0 = sum(e.stats(c)["accepts"] for c in range(args.chains))
        run_sweeps(e, 2, (args.warmup + args.steps) % 2 == 0)
        ms, n = e.kernel_timer_end()
        a1 = sum(e.stats(c)["accepts"] for c in range(args.chains))
        prof[kid] = (ms, n, a1 - a0)
    N = NL * NL
    ms, n, accs = prof[0]
    avg_s = ms / n / 1e3
    # Algorithmic work of one launch (one slice of every chain): one read and one write of each
    # chain's N x N FP64 G, and the accepted flips' rank-1 Sherman-Morrison updates, 2 N^2 flops
    # each (src/hubbard.cpp:132-139).  The launches also apply the slice wrap that follows, in an up
    # sweep for the slices not followed by a stabilization (NT - S of the NT) and in a down sweep for
    # every slice: its left product (G read, the N x N product written) and its right product (that
    # product read, G written), each 2 (16^3 + 16^3) flops per lattice vector with the Kronecker
    # factors.  The two timed sweeps are one up and one down sweep.  The kernel moves several times
    # the algorithmic bytes (the flushes of its site blocks), so HBM is what bounds it.
    S = -(-NT // PACE)
    fused = (2 * NT - S) if os.environ.get("DQMC_WRAP_FUSE", "") != "0" else 0
    ffrac = fused / float(n)
    abytes = 2.0 * N * N * 8 * args.chains * (1.0 + 2.0 * ffrac)
    flops = 2.0 * N * N * accs / n + ffrac * 2.0 * 4.0 * 16 ** 3 * N * args.chains
    roof = {"bound": "hbm", "achieved": round(abytes / avg_s / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 5)
    roof["traffic"] = None
    roof.update(measured_traffic())
    roof["kernel"] = kernel_names[0]
    roof["avg_launch_us"] = round(avg_s * 1e6, 2)
    roof["algorithmic_bytes_per_launch"] = round(abytes)
    roof["flops_per_launch"] = round(flops)
    roof["launches_with_wrap"] = "%d of %d" % (fused, n)
    roof["achieved_tflops"] = round(flops / avg_s / 1e12, 3)
    if roof.get("traffic"):
        roof["traffic_gbs"] = round(roof["traffic"] / avg_s / 1e9, 1)
    roof["kernel_time_ms_2_sweeps"] = {kernel_names[k]: round(prof[k][0], 2) for k in prof}
    # Whole sweep against the FP64 peak with the canonical algorithmic flop count per chain and
    # sweep, F = (6 + 2a) L N^3 + (44/3) S N^3 (a = acceptance, S = stabilizations per sweep).
    fsweep = ((6.0 + 2.0 * acc) * NT + 44.0 / 3.0 * S) * float(N) ** 3 * args.chains
    sweep_roof = {"flops_per_step": round(fsweep), "achieved_tflops": round(fsweep / (dt / args.steps) / 1e12, 3),
                  "peak": PEAK_FP64_TFLOPS}
    sweep_roof["frac"] = round(sweep_roof["achieved_tflops"] / PEAK_FP64_TFLOPS, 4)
    e.close()

    # denominator: plain FP64 GPU form, 1 chain, rank-1 updates
    plain = plain_dG = None
    if not args.no_plain:
        log("[bench] plain FP64: 1 chain k=1")
        p = make_engine(gpu, 1, gpu.FP64, 1)
        plain, _ = timed(p, 1, args.plain_steps, 1, not args.no_graphs)
        plain_dG = p.stats(0)["max_abs_dG"]
        p.close()

    # mixed FP32 wraps / flushes: |dG| on a short run (reported; fails the 1e-6 bound here)
    mixed_rate = mixed_dG = None
    if not args.no_mixed:
        log("[bench] mixed FP32 check: 8 chains")
        m = make_engine(gpu, 8, gpu.MIXED, args.delay)
        mixed_rate, _ = timed(m, 8, 2, 0, False)
        mixed_dG = max(m.stats(c)["max_abs_dG"] for c in range(8))
        m.close()

    adj = None
    if not args.no_adjoint:
        log("[bench] adjoint, 8x8 L=40, 256 configurations")
        adj = adjoint_c5(gpu)

    app = None
    if not args.no_app:
        log("[bench] drop-in app, 8x8 beta=8 L=80, all observables")
        app = app_throughput()

    out = {
        "metric": "aggregate sweeps/s at N=256, L=80: mixed+batched vs FP64 GPU, |dG|<=1e-6",
        "value": round(value, 3),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random Ising fields from fixed seeds)",
        "config": {"workload": "dqmc-sweep 16x16 U=-4 mu=0 beta=8 L=80 pace=8", "chains": args.chains,
                   "delay_k": args.delay, "precision": "fp64", "graphs": not args.no_graphs},
        "max_abs_dG": dG,
        "dG_ok": dG <= 1e-6,
        "acceptance": round(acc, 4),
        "plain_fp64_sweeps_per_s": round(plain, 4) if plain else None,
        "plain_fp64_max_abs_dG": plain_dG,
        "speedup_vs_plain_fp64": round(value / plain, 2) if plain else None,
        "mixed_fp32_max_abs_dG": mixed_dG,
        "mixed_fp32_sweeps_per_s_8chains": round(mixed_rate, 3) if mixed_rate else None,
        "roofline": roof,
        "sweep_roofline": sweep_roof,
        "app_c2": app,
        "adjoint_c5": adj,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
