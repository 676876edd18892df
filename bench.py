"""Benchmark of the NRG hot path on one MI355X.

Workload: the keep = 10000 two-channel Anderson solve (QS symmetry, U = 0.5,
eps_d = -0.25, Gamma = 0.02 per channel, Lambda = 3, Nmax = 40, FDM spectral
function at T = 1e-10).  Its iterations hold ~146 000 eigenpairs in ~66
invariant subspaces of up to 13 136 rows.

A step is one NRG iteration of that solve: assembly of all subspace
Hamiltonians, the batched FP64 eigensolve, truncation, and the recalculation
of the f and singlet operators, with every iteration's eigenvectors kept for
the FDM pass.  The solve runs in process through libnrg_engine.so: `warmup`
iterations (the first three grow the basis to steady state) and then exactly
`steps` timed ones, each ended by a drained stream (--iter-log); the run then
stops through the reference's own `forcestop` control.  value = eigenpairs of
the timed iterations / their wall time.

Also reported: the roofline of the dominant kernel (HIP events on its launch
stream, algorithmic work counted by the library), the same iterations on the
plain per-block GPU path (one eigensolver call per subspace), a CPU baseline
(LAPACK through numpy on the host cores, on subspace sizes of the same
iteration) and the serving stream (64 seeded Poisson SIAM requests, four
z-members each, keep 1000-8000).
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "nrgljubljana-old_amd")
sys.path.insert(0, PKG)

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and MFMA)
PEAK_HBM_GBS = 8000.0
# What bounds each kernel, and which of its counted quantities is algorithmic.
# k_eig_tri_blk: the band reduction of a block of dimension d reads the
# trailing matrix once for Y = A V and reads and writes its lower triangle
# for the rank-64 update per panel of 32 columns (d^3 / 6 bytes; the
# one-stage reduction below the two-stage size, sum_j 4 (d-j)^2); the streaming one
# (k_eig_tri_coop) the whole trailing matrix per column; binning / assembly
# / transfers are byte streams; the contractions run on FP64 MFMA, the
# bulge-chasing back-transformation (k_sbr_bt, 2 d^3) on the FP64 FMA units
# (the same 78.6 TFLOP/s peak); the bulge chasing itself (k_sbr_chase) is
# bound by the latency of its sequential tasks and is not priced here.
FP64_VECTOR = {"k_sbr_bt"}  # compute-priced kernels that run on the FP64 FMA units, not MFMA
KERNEL_BOUND = {"k_eig_tri_blk": "hbm", "k_eig_tri_coop": "hbm", "k_spec_bin": "hbm", "k_assemble": "hbm",
                "k_gather_scatter": "hbm", "k_zero": "hbm", "k_eig_copy": "hbm", "k_gemm": "mfma",
                "k_gemm_dc": "mfma", "k_gemm_bt": "mfma", "k_sbr_bt": "mfma"}

CONFIGS = {
    # SIAM QS, keep <= 500 states (Ecut = 8 omega_N), Nmax = 30, DM-NRG spectrum at T=1e-8
    "siam_qs_keep500_dmnrg": dict(model="siam", sym="QS", nch=1, Lambda=2.0, Nmax=30,
                                  model_kw=dict(U=1.0, eps_d=-0.5, Gamma=0.04),
                                  param=dict(keep=500, keepenergy=8.0, T=1e-8, dmnrg=True, specd="A_d-A_d",
                                             bins=50, savebins=True)),
    # SIAM QS, keep <= 4000 states (Ecut = 10 omega_N), Nmax = 50, FDM spectrum + thermodynamics
    "siam_qs_keep4000_fdm": dict(model="siam", sym="QS", nch=1, Lambda=2.0, Nmax=50,
                                 model_kw=dict(U=1.0, eps_d=-0.5, Gamma=0.04),
                                 param=dict(keep=4000, keepenergy=10.0, T=1e-10, fdm=True, specd="A_d-A_d",
                                            bins=100, savebins=True)),
    # two-channel Anderson QS, keep 10000, Lambda = 3, Nmax = 40, FDM spectrum + susceptibility (td)
    "anderson2ch_qs_keep10000_fdm": dict(model="siam", sym="QS", nch=2, Lambda=3.0, Nmax=40,
                                         model_kw=dict(U=0.5, eps_d=-0.25, Gamma=0.02),
                                         param=dict(keep=10000, keepenergy=100.0, T=1e-10, fdm=True,
                                                    specd="A_d-A_d", bins=100, savebins=True)),
}
DEFAULT = "anderson2ch_qs_keep10000_fdm"
# serving stream: seeds of the 64 requests and their Poisson rate (requests/s),
# about 1.5x the throughput the stream reaches when all requests arrive at once
SERVE_SEED = 1000
SERVE_RATE = 10.0


def make_inputs(wd, cfg, **extra):
    from nrgamd import datagen
    m = datagen.siam(sym=cfg["sym"], nch=cfg["nch"], Lambda=cfg["Lambda"], **cfg["model_kw"])
    m.write(os.path.join(wd, "data"), cfg["Lambda"], 1.0, cfg["Nmax"])
    datagen.write_param(os.path.join(wd, "param"), Lambda=cfg["Lambda"], **dict(cfg["param"], **extra))


def ensure_tables():
    tab = os.path.join(PKG, "tables", "qs-1ch.tab")
    if not os.path.exists(tab):
        import subprocess
        subprocess.check_call([sys.executable, "-m", "nrgamd.symtab", os.path.join(PKG, "tables")], cwd=PKG)


def run_engine(lib, wd, args):
    """One solve in-process: chdir into the work directory and call nrg_main
    (its log, a line per iteration, goes to stderr, so stdout keeps the one
    JSON line)."""
    argv_l = [b"nrg", b"--tables", os.path.join(PKG, "tables").encode()] + [a.encode() for a in args]
    argv = (ctypes.c_char_p * len(argv_l))(*argv_l)
    cwd = os.getcwd()
    os.chdir(wd)
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        rc = lib.nrg_main(len(argv_l), argv)
    finally:
        libc = ctypes.CDLL(None)
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.chdir(cwd)
    return rc


def iterations(lib, cfg, warmup, steps, plain=False, budget_gb=None):
    """Runs the solve for warmup + steps iterations; returns the timed
    iterations' eigenpairs, wall time and the library's kernel table."""
    import nrgamd.capi as capi
    with tempfile.TemporaryDirectory() as wd:
        make_inputs(wd, cfg, forcestop=warmup + steps)
        log = os.path.join(wd, "iter.log")
        args = ["--iter-log", log, "--stats", os.path.join(wd, "stats.json"), "--stats-light"]
        if plain:
            args.append("--plain-diag")
        if budget_gb:
            args += ["--hbm-budget", "%g" % budget_gb]
        rc = run_engine(lib, wd, args)
        rows = [l.split() for l in open(log)] if os.path.exists(log) else []
        if len(rows) != warmup + steps:
            raise RuntimeError("solve ended after %d of %d iterations (status %d)" % (len(rows), warmup + steps, rc))
        t = {int(r[0]): float(r[1]) for r in rows}
        ep = {int(r[0]): int(r[2]) for r in rows}
        timed = range(warmup, warmup + steps)
        t0 = t[warmup - 1] if warmup > 0 else 0.0
        wall = t[warmup + steps - 1] - t0
        pairs = sum(ep[n] for n in timed)
        return pairs, wall, capi.kernel_stats(lib)


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration), or
    None when no such measurement exists."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t[workload][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(budget_s=20.0):
    """LAPACK (numpy eigh: dsyevd, all eigenpairs, as dsyev) on the host cores,
    on random symmetric blocks with the dimensions of one steady-state
    iteration of the workload, smallest first, until the budget is spent: the
    eigensolver is the reference's dominant cost (README:36-38)."""
    import numpy as np
    try:
        from threadpoolctl import threadpool_info
        cores = max([p.get("num_threads", 1) for p in threadpool_info()] + [1])
    except Exception:
        cores = os.cpu_count() or 1
    sizes = sorted([12680, 11450, 11450, 9272, 8873, 8873, 7413, 7413, 6290, 6290, 5421, 5421, 4815, 4815, 3648,
                    3648, 2934, 2934, 2396, 2076, 2076, 1870, 1870, 1523, 1523, 1107, 1107, 818, 818, 731])
    rng = np.random.default_rng(0)
    done, pairs, dt = [], 0, 0.0
    for n in sizes:
        a = rng.standard_normal((n, n))
        a = (a + a.T) / 2
        t0 = time.perf_counter()
        np.linalg.eigh(a)
        dt += time.perf_counter() - t0
        pairs += n
        done.append(n)
        if dt > budget_s:
            break
    return {"value": pairs / dt, "unit": "eigenpairs/s", "cores": cores, "kind": "port",
            "sample": "LAPACK eigh (numpy, %d threads) on %d random symmetric blocks of the workload's subspace "
                      "dimensions %d..%d: %d eigenpairs in %.1f s" % (cores, len(done), done[0], done[-1], pairs, dt)}


def serve_measure(nreq, rate, keep_max):
    """The serving stream: seeded SIAM QS requests with Poisson arrivals,
    four z-members each, keep drawn from [1000, keep_max]; latency p50/p99
    and throughput."""
    from nrgamd import serve
    with tempfile.TemporaryDirectory() as out:
        reqs = serve.make_requests(out, nreq=nreq, seed=SERVE_SEED, rate=rate, keep_range=(1000, keep_max))
        rep = serve.run(out, reqs, slots=16)
    keys = ["solves", "failed", "makespan_s", "solves_per_s", "eigenpairs_per_s", "latency_p50_s", "latency_p99_s",
            "request_latency_p50_s", "request_latency_p99_s", "requests_per_s", "shared_launches",
            "merged_iterations", "slots"]
    res = {k: rep[k] for k in keys}
    res.update(requests=nreq, rate_per_s=rate, seed=SERVE_SEED, keep_range=[1000, keep_max])
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--config", default=DEFAULT, choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the unbatched per-block comparison")
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--serve-requests", type=int, default=64, help="0 skips the serving stream")
    ap.add_argument("--serve-rate", type=float, default=SERVE_RATE)
    ap.add_argument("--serve-keep-max", type=int, default=8000)
    ap.add_argument("--hbm-budget", type=float, default=230.0,
                    help="GB of HBM for the eigenvectors kept for the FDM pass (beyond: pinned host memory)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % a.gpus)
        sys.exit(2)
    if a.warmup < 3:
        ap.error("--warmup must be at least 3: the first three iterations grow the basis to steady state")
    cfg = CONFIGS[a.config]
    ensure_tables()
    import nrgamd.capi as capi
    gpulib = capi.load()
    eng = ctypes.CDLL(os.path.join(PKG, "lib", "libnrg_engine.so"))
    eng.nrg_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    print("solve: %d warmup + %d timed iterations of %s" % (a.warmup, a.steps, a.config), file=sys.stderr, flush=True)
    pairs, wall, kst = iterations(eng, cfg, a.warmup, a.steps, budget_gb=a.hbm_budget)
    value = pairs / wall
    timed = {k: v for k, v in kst.items() if k in KERNEL_BOUND and v["ms"] > 0}
    roof = None
    if timed:
        name, k = max(timed.items(), key=lambda kv: kv[1]["ms"])
        avg_ms = k["ms"] / max(1, k["launches"])
        if KERNEL_BOUND[name] == "mfma":
            ach = k["flops"] / (k["ms"] * 1e-3) / 1e12
            roof = {"kernel": name, "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / PEAK_FP64_TFLOPS}
            if name in FP64_VECTOR:
                roof["compute"] = ("FP64 vector FMA, no MFMA instructions; the peak is MI355X's FP64 rate, "
                                   "78.6 TFLOP/s for vector and matrix alike")
        else:
            ach = k["bytes"] / (k["ms"] * 1e-3) / 1e9
            roof = {"kernel": name, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": ach / PEAK_HBM_GBS}
        roof["traffic"] = measured_traffic(a.config, name)
        roof["algorithmic_per_launch"] = (k["flops"] if KERNEL_BOUND[name] == "mfma" else k["bytes"]) / max(1, k["launches"])
        roof["avg_launch_ms"] = avg_ms
        roof["launches"] = k["launches"]
    plain = None
    if not a.no_plain:
        print("plain per-block path: %d + %d iterations" % (3, a.plain_steps), file=sys.stderr, flush=True)
        pp, pw, _ = iterations(eng, cfg, 3, a.plain_steps, plain=True, budget_gb=a.hbm_budget)
        plain = {"value": pp / pw, "iterations_timed": [3, 3 + a.plain_steps - 1], "speedup": value / (pp / pw)}
    cpu = None if a.no_cpu_baseline else cpu_baseline()
    serving = None
    if a.serve_requests > 0:
        print("serving %d requests" % a.serve_requests, file=sys.stderr, flush=True)
        serving = serve_measure(a.serve_requests, a.serve_rate if a.serve_rate > 0 else None, a.serve_keep_max)
    out = {
        "metric": "NRG eigenpairs/s, keep=10000 two-channel QS solve (steady-state iterations)",
        "value": value,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated two-channel Anderson data file)",
        "config": {"workload": a.config, "step": "one NRG iteration (assembly, batched eigensolve, truncation, "
                   "f/singlet recalculation; eigenvectors kept for the FDM pass)",
                   "iterations_timed": [a.warmup, a.warmup + a.steps - 1], "eigenpairs_per_step": pairs / a.steps,
                   "kernels": {k: {"launches": v["launches"], "ms": v["ms"]} for k, v in kst.items() if v["launches"]},
                   "profile": "kernels: HIP events on every launch of the run (warmup included)",
                   "full_solve": "the whole solve (pass 1, FDM sweep, spectra) is timed by bin/nrg, not here: "
                                 "profiles/full_solve_anderson2ch_keep10000_fdm.md"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_per_block": plain,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
