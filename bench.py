"""Benchmark of the NRG hot path on one MI355X.

A step is one complete NRG solve of the configured workload through the
drop-in engine (libnrg_engine.so, called in-process): NRG pass with batched
assembly + eigensolve + recalculation, DM/FDM backward sweeps, and the DMNRG
pass with operator recalculation and spectral binning.  The input `data` file
is generated before the timed region.

Prints one JSON line: value = eigenpairs/s (sum over iterations and
subspaces of the eigenpairs produced, divided by the run time), plus the
roofline of the dominant kernel (HIP events on the launch stream, algorithmic
work counted by the library) and a CPU baseline (the numpy oracle on a bounded
sample of the same workload, i.e. LAPACK dsyevd + BLAS on the host cores).
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
# What bounds each kernel, and which of its counted quantities is algorithmic:
# the streaming tridiagonalisation (blocks above 576) reads the trailing
# matrix once per column (sum_j 8 (d-j)^2 bytes per block); the register-
# resident one (blocks up to 576) keeps its rows on chip, so its ceiling is
# the FP64 rate of its 4/3 d^3 flops; binning / assembly / transfers are byte
# streams, the contractions and back-transformations are FP64 arithmetic.
KERNEL_BOUND = {"k_eig_tri_coop": "hbm", "k_eig_tri_one": "mfma", "k_eig_tri_reg": "mfma", "k_spec_bin": "hbm",
                "k_assemble": "hbm", "k_gather_scatter": "hbm",
                "k_zero": "hbm", "k_eig_copy": "hbm", "k_gemm": "mfma", "k_gemm_dc": "mfma", "k_gemm_bt": "mfma",
                "k_eig_bt": "mfma"}

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
    # two-channel Anderson QS, keep 10000, Lambda = 3, Nmax = 40, FDM spectrum + susceptibility
    "anderson2ch_qs_keep10000_fdm": dict(model="siam", sym="QS", nch=2, Lambda=3.0, Nmax=40,
                                         model_kw=dict(U=0.5, eps_d=-0.25, Gamma=0.02),
                                         param=dict(keep=10000, keepenergy=100.0, T=1e-10, fdm=True,
                                                    specd="A_d-A_d", bins=100, savebins=True)),
}
DEFAULT = "siam_qs_keep4000_fdm"


def make_inputs(wd, cfg, Nmax=None):
    from nrgamd import datagen
    N = Nmax or cfg["Nmax"]
    m = datagen.siam(sym=cfg["sym"], nch=cfg["nch"], Lambda=cfg["Lambda"], **cfg["model_kw"])
    m.write(os.path.join(wd, "data"), cfg["Lambda"], 1.0, N)
    datagen.write_param(os.path.join(wd, "param"), Lambda=cfg["Lambda"], **cfg["param"])


def ensure_tables():
    tab = os.path.join(PKG, "tables", "qs-1ch.tab")
    if not os.path.exists(tab):
        import subprocess
        subprocess.check_call([sys.executable, "-m", "nrgamd.symtab", os.path.join(PKG, "tables")], cwd=PKG)


def run_engine(lib, wd, stats, plain=False, light=False):
    """One solve in-process: chdir into the work directory and call nrg_main
    (light: no stage timing, so no stream syncs between stages)."""
    args = [b"nrg", b"--quiet", b"--stats", stats.encode(), b"--tables", os.path.join(PKG, "tables").encode()]
    if plain:
        args.append(b"--plain-diag")
    if light:
        args.append(b"--stats-light")
    argv = (ctypes.c_char_p * len(args))(*args)
    cwd = os.getcwd()
    os.chdir(wd)
    # the engine logs to stdout; keep bench's stdout for the one JSON line
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        rc = lib.nrg_main(len(args), argv)
    finally:
        libc = ctypes.CDLL(None)
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.chdir(cwd)
    if rc != 0:
        raise RuntimeError("nrg_main returned %d" % rc)
    with open(stats) as f:
        return json.load(f)


def serve_measure(nreq, keep_max):
    """Queued whole solves through the serving loop (continuous batching on
    this GPU): seeded SIAM QS requests, 4 z-members each, all arriving at t=0,
    keep drawn from [1000, keep_max]; latency p50/p99 and throughput."""
    from nrgamd import serve
    with tempfile.TemporaryDirectory() as out:
        reqs = serve.make_requests(out, nreq=nreq, seed=1000, rate=None, keep_range=(1000, keep_max))
        rep = serve.run(out, reqs, slots=16)
    keys = ["solves", "failed", "makespan_s", "solves_per_s", "eigenpairs_per_s", "latency_p50_s", "latency_p99_s",
            "request_latency_p50_s", "request_latency_p99_s", "requests_per_s", "shared_launches",
            "merged_iterations", "slots"]
    res = {k: rep[k] for k in keys}
    res["requests"] = nreq
    res["keep_range"] = [1000, keep_max]
    return res


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


def cpu_baseline(cfg, budget_s=20.0):
    """Oracle (numpy/LAPACK) on the first shells of the same workload."""
    sys.path.insert(0, ROOT)
    from oracle import nrg_oracle
    try:
        from threadpoolctl import threadpool_info
        cores = max([p.get("num_threads", 1) for p in threadpool_info()] + [1])
    except Exception:
        cores = os.cpu_count() or 1
    N = 4
    best = None
    while True:
        with tempfile.TemporaryDirectory() as wd:
            make_inputs(wd, cfg, Nmax=N)
            t0 = time.time()
            o = nrg_oracle.Oracle(wd).run()
            dt = time.time() - t0
        pairs = sum(len(v) for it in o.iters for v, k in it.values())
        best = (pairs / dt, N, dt, pairs)
        if dt > budget_s / 3 or N >= cfg["Nmax"]:
            break
        N = min(cfg["Nmax"], N + max(2, N // 2))
    rate, N, dt, pairs = best
    return {"value": rate, "unit": "eigenpairs/s", "cores": cores, "kind": "port",
            "sample": "numpy oracle (LAPACK eigh + BLAS), both passes, first %d of %d shells: %d eigenpairs in %.1f s"
                      % (N, cfg["Nmax"], pairs, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # ten timed solves (about 1.2 s): a single solve varies by a few percent
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=DEFAULT, choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the unbatched per-block comparison solve")
    ap.add_argument("--serve-requests", type=int, default=8,
                    help="requests (x4 z-members) of the serving measurement; 0 skips it, 64 is the full stream")
    ap.add_argument("--serve-keep-max", type=int, default=2000)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % a.gpus)
        sys.exit(2)
    cfg = CONFIGS[a.config]
    ensure_tables()
    import nrgamd.capi as capi
    gpulib = capi.load()
    eng = ctypes.CDLL(os.path.join(PKG, "lib", "libnrg_engine.so"))
    eng.nrg_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    with tempfile.TemporaryDirectory() as wd:
        make_inputs(wd, cfg)
        stats = os.path.join(wd, "stats.json")
        # warmup: the last warmup solve runs fully instrumented (HIP events
        # around every launch, a stream sync at every stage boundary) for the
        # stage breakdown and the kernel table, which names the dominant
        # kernel; the timed solves then carry events for that kernel alone
        profile = None
        for w in range(a.warmup):
            print("warmup %d" % w, file=sys.stderr, flush=True)
            profile = run_engine(eng, wd, stats)
        dom_name = None
        if profile:
            cand = {k: v for k, v in profile["kernels"].items() if k in KERNEL_BOUND and v["ms"] > 0}
            if cand:
                dom_name = max(cand.items(), key=lambda kv: kv[1]["ms"])[0]
        gpulib.nrg_timing_filter(dom_name.encode() if dom_name else None)
        t0 = time.perf_counter()
        pairs = 0
        stage_sum = {}
        kst = {}
        for s in range(a.steps):
            print("step %d" % s, file=sys.stderr, flush=True)
            st = run_engine(eng, wd, stats, light=dom_name is not None)
            pairs += st["eigenpairs"]
            for k, v in st["stages"].items():
                stage_sum[k] = stage_sum.get(k, 0.0) + v["s"]
            for k, v in capi.kernel_stats(gpulib).items():  # the engine resets them per solve
                acc = kst.setdefault(k, {"launches": 0, "ms": 0.0, "flops": 0.0, "bytes": 0.0})
                for f in acc:
                    acc[f] += v[f]
        t1 = time.perf_counter()
        gpulib.nrg_timing_filter(None)
        gpulib.nrg_timing_enable(0)
        if profile:  # breakdown from the instrumented warmup solve
            stage_sum = {k: v["s"] * a.steps for k, v in profile["stages"].items()}
        # the unbatched GPU path on the same workload (one eigensolver call per
        # block), outside the timed region
        plain = None
        if not a.no_plain:
            print("plain per-block path", file=sys.stderr, flush=True)
            tp0 = time.perf_counter()
            stp = run_engine(eng, wd, stats, plain=True)
            tp = time.perf_counter() - tp0
            plain = {"value": stp["eigenpairs"] / tp, "ms_per_step": 1e3 * tp, "diag_s": stp["stages"]["diag"]["s"]}
    total = t1 - t0
    value = pairs / total
    if plain:
        plain["speedup"] = value / plain["value"]
        plain["diag_speedup"] = plain["diag_s"] / (stage_sum["diag"] / a.steps)
    timed = {k: v for k, v in kst.items() if k in KERNEL_BOUND and v["ms"] > 0}
    dom = max(timed.items(), key=lambda kv: kv[1]["ms"]) if timed else None
    roof = None
    if dom:
        name, k = dom
        avg_ms = k["ms"] / max(1, k["launches"])
        if KERNEL_BOUND[name] == "mfma":
            ach = k["flops"] / (k["ms"] * 1e-3) / 1e12
            roof = {"kernel": name, "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / PEAK_FP64_TFLOPS}
        else:
            ach = k["bytes"] / (k["ms"] * 1e-3) / 1e9
            roof = {"kernel": name, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": ach / PEAK_HBM_GBS}
        roof["traffic"] = measured_traffic(a.config, name)
        roof["algorithmic_per_launch"] = (k["flops"] if KERNEL_BOUND[name] == "mfma" else k["bytes"]) / max(1, k["launches"])
        roof["avg_launch_ms"] = avg_ms
    cpu = None if a.no_cpu_baseline else cpu_baseline(cfg)
    serving = None
    if a.serve_requests > 0:
        print("serving %d requests" % a.serve_requests, file=sys.stderr, flush=True)
        serving = serve_measure(a.serve_requests, a.serve_keep_max)
    out = {
        "metric": "NRG eigenpairs/s (full two-pass solve)",
        "value": value,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * total / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated SIAM data file)",
        "config": {"workload": a.config, "run_time_s": total / a.steps, "eigenpairs_per_solve": pairs // a.steps,
                   "stages_s": {k: v / a.steps for k, v in stage_sum.items()},
                   "kernels": {k: {"launches": v["launches"], "ms": v["ms"]}
                               for k, v in (profile["kernels"] if profile else kst).items()},
                   "profile": ("stages and kernels: the last warmup solve, fully instrumented; timed solves: "
                               "events on the dominant kernel only") if profile else "timed solves, fully instrumented"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain_per_block": plain,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
