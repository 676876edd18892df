"""Measures on the GPU the size thresholds of libnrg_gpu.so and writes them to
tables/tune.txt, which bin/nrg, nrg-serve and the in-process engine read at
start (environment variables of the same names take precedence):

NRG_GEMM_MFMA_MIN_TILES  grouped GEMM launches with at least this many 64 x 64
                         output tiles run on the FP64-MFMA kernel, smaller ones
                         on the split 32 x 32 FMA kernel;
NRG_BLK_MIN_N            blocks from this dimension up take the two-stage
                         eigensolver (band reduction, bulge chasing), smaller
                         ones the one-stage team reductions.

Each threshold is the smallest candidate from which the faster variant wins at
every larger tested size: single-term launches of 16..1024 tiles with k = 128
and 512 (the recalculation and back-transformation shapes), and batches of
eight random symmetric matrices of 640..1536 rows.  Times are the library's
own HIP-event kernel times (nrg_timing_*), best of three.

    python -m nrgamd.autotune [--out DIR/tune.txt]
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
sys.path.insert(0, PKG)

from nrgamd import capi  # noqa: E402

GEMM_TILES = [16, 32, 64, 128, 256, 512, 1024]
GEMM_K = [128, 512]
BLK_SIZES = [640, 768, 896, 1024, 1280, 1536]


def _ctx(env):
    """A context created, and through its first eigensolve, with `env` set
    (the library reads these variables at those two points)."""
    saved = {k: os.environ.get(k) for k in env}
    for k, v in env.items():
        os.environ[k] = str(v)
    c = capi.Ctx(0)
    c.diag_batch([np.eye(3)])
    for k, v in saved.items():
        if v is None:
            del os.environ[k]
        else:
            os.environ[k] = v
    return c


def _kernel_ms(lib, names=None):
    ks = capi.kernel_stats(lib)
    return sum(v["ms"] for k, v in ks.items() if names is None or k in names)


def time_gemm(ctx, tiles, k, reps=3):
    """Kernel time of one launch with `tiles` 64 x 64 outputs of inner
    dimension k (outputs of 256 x 256, or one narrower output)."""
    rng = np.random.default_rng(tiles * 1000 + k)
    outs, terms = [], []
    left = tiles
    while left > 0:
        t = min(16, left)
        m, n = 256, 64 * t // 4 if t < 16 else 256
        o = capi.GemmOut()
        o.C = ctx.alloc(m * n)
        o.m, o.n, o.ldc, o.beta = m, n, n, 0.0
        o.term_begin, o.term_count = len(terms), 1
        tm = capi.GemmTerm()
        tm.A, tm.B = ctx.put(rng.standard_normal((m, k))), ctx.put(rng.standard_normal((k, n)))
        tm.s = None
        tm.alpha, tm.k, tm.lda, tm.ldb, tm.transA, tm.transB = 1.0, k, k, n, 0, 0
        outs.append(o)
        terms.append(tm)
        left -= t
    best = None
    for _ in range(reps + 1):
        ctx.lib.nrg_timing_reset()
        ctx.gemm_grouped(outs, terms)
        ctx.sync()
        ms = _kernel_ms(ctx.lib, {"k_gemm"})
        best = ms if best is None else min(best, ms)
    return best


def time_eig(ctx, n, count=8, reps=3):
    rng = np.random.default_rng(n)
    mats = []
    for _ in range(count):
        a = rng.standard_normal((n, n))
        mats.append((a + a.T) / 2)
    best = None
    for _ in range(reps + 1):
        ctx.lib.nrg_timing_reset()
        ctx.diag_batch(mats)
        ms = _kernel_ms(ctx.lib)
        best = ms if best is None else min(best, ms)
    return best


def crossover(cands, faster):
    """The smallest candidate c with faster[x] for every candidate x >= c
    (one past the largest when the last one is not faster)."""
    thr = None
    for c in reversed(cands):
        if not faster[c]:
            break
        thr = c
    return thr if thr is not None else cands[-1] + 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(PKG, "tables", "tune.txt"))
    a = ap.parse_args()
    lib = capi.load()
    lib.nrg_timing_enable(1)
    lines = []

    fma, mfma = _ctx({"NRG_GEMM_MFMA_MIN_TILES": 1 << 30}), _ctx({"NRG_GEMM_MFMA_MIN_TILES": 1})
    faster = {}
    for t in GEMM_TILES:
        wins = True
        for k in GEMM_K:
            tf, tm = time_gemm(fma, t, k), time_gemm(mfma, t, k)
            wins = wins and tm < tf
            lines.append("# gemm tiles %5d k %4d: FMA %8.3f ms  MFMA %8.3f ms" % (t, k, tf, tm))
            print(lines[-1], flush=True)
        faster[t] = wins
    fma.close()
    mfma.close()
    gemm_thr = crossover(GEMM_TILES, faster)

    one, two = _ctx({"NRG_BLK_MIN_N": 1 << 30}), _ctx({"NRG_BLK_MIN_N": min(BLK_SIZES)})
    faster = {}
    for n in BLK_SIZES:
        t1, t2 = time_eig(one, n), time_eig(two, n)
        faster[n] = t2 < t1
        lines.append("# eigensolve 8 x %5d: one-stage %8.2f ms  two-stage %8.2f ms" % (n, t1, t2))
        print(lines[-1], flush=True)
    one.close()
    two.close()
    blk_thr = crossover(BLK_SIZES, faster)

    with open(a.out, "w") as f:
        f.write("# size thresholds measured on this device by: python -m nrgamd.autotune"
                " (from nrgljubljana-old_amd/)\n")
        for line in lines:
            f.write(line + "\n")
        f.write("NRG_GEMM_MFMA_MIN_TILES %d\n" % gemm_thr)
        f.write("NRG_BLK_MIN_N %d\n" % blk_thr)
    print("NRG_GEMM_MFMA_MIN_TILES %d\nNRG_BLK_MIN_N %d\nwritten to %s" % (gemm_thr, blk_thr, a.out), flush=True)


if __name__ == "__main__":
    main()
