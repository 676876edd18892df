"""nrg_gemm_grouped_noalloc (fixed device scratch array, no allocation)
against nrg_gemm_grouped on the same inputs, through the C ABI: the same
bits, on two calls in a row (the second smaller), for a launch that splits
its tiles (partial tiles in the scratch array) and one on the MFMA kernel;
a launch beyond the array's size is refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_nrgamd

pytestmark = pytest.mark.gpu


def _case(ctx, capi, rng, shapes, kmax):
    """Outputs (with their initial C, beta) and terms; returns (outs, terms,
    handles) where each output gets two copies of C (one per entry point)."""
    outs_a, outs_b, terms, handles = [], [], [], []
    for m, n in shapes:
        Cinit = rng.standard_normal((m, n))
        beta = float(rng.choice([0.0, 1.0, -0.5]))
        pa, pb = ctx.put(Cinit), ctx.put(Cinit)
        oa, ob = capi.GemmOut(), capi.GemmOut()
        for o, p in ((oa, pa), (ob, pb)):
            o.C, o.m, o.n, o.ldc, o.beta = p, m, n, n, beta
            o.term_begin = len(terms)
            o.term_count = 2
        for _ in range(2):
            k = int(rng.integers(1, kmax))
            A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
            t = capi.GemmTerm()
            t.A, t.B, t.s = ctx.put(A), ctx.put(B), None
            t.alpha, t.k, t.lda, t.ldb, t.transA, t.transB = float(rng.standard_normal()), k, k, n, 0, 0
            terms.append(t)
        outs_a.append(oa)
        outs_b.append(ob)
        handles.append((pa, pb, m, n))
    return outs_a, outs_b, terms, handles


@pytest.mark.parametrize("shapes,kmax", [
    ([(40, 50), (33, 17)], 3000),          # few tiles, long k: split, partial tiles in the scratch array
    ([(1100, 1030), (300, 257)], 200),     # enough 64 x 64 tiles: the MFMA kernel
])
def test_noalloc_matches_grouped_twice(shapes, kmax):
    import nrgamd.capi as capi
    load_nrgamd()
    ctx = capi.Ctx(0)
    rng = np.random.default_rng(31)
    # two calls in a row, the second smaller
    for call, scale in ((0, 1.0), (1, 0.5)):
        sh = [(max(1, int(m * scale)), max(1, int(n * scale))) for m, n in shapes]
        outs_a, outs_b, terms, handles = _case(ctx, capi, rng, sh, kmax)
        ctx.gemm_grouped(outs_a, terms)
        st = ctx.gemm_grouped_noalloc(outs_b, terms)
        assert st == 0, ctx.lib.nrg_last_error(ctx.ctx)
        ctx.sync()
        for pa, pb, m, n in handles:
            np.testing.assert_array_equal(ctx.get(pb, (m, n)), ctx.get(pa, (m, n)))
    ctx.close()


def test_noalloc_refuses_beyond_its_array():
    import nrgamd.capi as capi
    load_nrgamd()
    ctx = capi.Ctx(0)
    # one output with 320 000 terms: its descriptors alone exceed the 16 MiB
    # array; refused before any device work
    a = ctx.put(np.ones((8, 1)))
    b = ctx.put(np.ones((1, 8)))
    cbuf = ctx.put(np.zeros((8, 8)))
    o = capi.GemmOut()
    o.C, o.m, o.n, o.ldc, o.beta, o.term_begin, o.term_count = cbuf, 8, 8, 8, 0.0, 0, 320000
    nt = 320000
    terms = (capi.GemmTerm * nt)()
    for t in terms:
        t.A, t.B, t.s, t.alpha, t.k, t.lda, t.ldb, t.transA, t.transB = a, b, None, 1.0, 1, 1, 8, 0, 0
    st = ctx.lib.nrg_gemm_grouped_noalloc(ctx.ctx, (capi.GemmOut * 1)(o), 1, terms, nt)
    msg = ctx.lib.nrg_last_error(ctx.ctx)
    ctx.close()
    assert st != 0 and b"scratch" in msg
