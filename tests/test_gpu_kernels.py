"""Each C-ABI entry point on the GPU against an independent CPU computation.

eigensolver: numpy/LAPACK eigh (eigenvalues) and the defining identities
(A u = lambda u, orthonormal rows) on random, degenerate and structured
matrices; grouped GEMM / assembly: numpy products; spectral binning: the same
peaks binned by the oracle's Bins (src/bins.h:102-131).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import load_nrgamd, load_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    capi = load_nrgamd().capi if hasattr(load_nrgamd(), "capi") else __import__("nrgamd.capi").capi
    c = capi.Ctx(0)
    yield c
    c.close()


def _capi():
    import nrgamd.capi as capi
    return capi


def _sym(rng, n):
    a = rng.standard_normal((n, n))
    return (a + a.T) / 2


def _check_eig(A, W, U, tol=1e-12):
    n = A.shape[0]
    nrm = max(1.0, np.abs(A).max() * n)
    ref = np.linalg.eigvalsh(A)
    np.testing.assert_allclose(W, ref, atol=tol * nrm, rtol=0)
    assert np.all(np.diff(W) >= 0)
    # rows of U are eigenvectors: A U^T = U^T diag(W); U U^T = 1
    np.testing.assert_allclose(U @ U.T, np.eye(n), atol=1e-12 * max(1, n ** 0.5))
    res = A @ U.T - U.T * W[None, :]
    assert np.abs(res).max() <= tol * nrm


def test_diag_batch_random_sizes(ctx):
    rng = np.random.default_rng(1)
    # boundaries of the tridiagonalisation variants: one workgroup up to 128
    # (here drained by the team launch), register teams of 4 to 9 column
    # registers (one per 64 columns; two rows per wavefront up to 512, three
    # above), streaming teams above 576
    sizes = [1, 2, 3, 4, 7, 16, 33, 64, 65, 127, 128, 129, 256, 257, 300, 320, 321, 384, 385, 448, 449, 511, 512,
             513, 576, 577]
    mats = [_sym(rng, n) for n in sizes]
    out = ctx.diag_batch(mats)
    for A, (W, U) in zip(mats, out):
        _check_eig(A, W, U)


def test_diag_batch_degenerate_and_structured(ctx):
    rng = np.random.default_rng(2)
    mats = []
    # exact degeneracies (as with channel symmetry in two-channel models)
    Q, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    lam = np.repeat(np.arange(10, dtype=float), 4)
    mats.append(Q @ np.diag(lam) @ Q.T)
    # block diagonal with identical blocks
    b = _sym(rng, 12)
    mats.append(np.kron(np.eye(3), b))
    # already diagonal, zero matrix, rank one, tridiagonal with tiny couplings
    mats.append(np.diag(rng.standard_normal(20)))
    mats.append(np.zeros((9, 9)))
    v = rng.standard_normal(15)
    mats.append(np.outer(v, v))
    t = np.diag(np.linspace(0, 1, 30)) + np.diag(np.full(29, 1e-14), 1) + np.diag(np.full(29, 1e-14), -1)
    mats.append(t)
    # NRG-like scale spread: large diagonal, small hopping
    d = np.sort(rng.uniform(0, 50, 80))
    h = np.diag(d) + 0.3 * _sym(rng, 80)
    mats.append(h)
    # multiplets in a register-team size: deflation at every merge level
    Q, _ = np.linalg.qr(rng.standard_normal((400, 400)))
    lam = np.repeat(np.round(rng.uniform(0, 3, 100), 2), 4)
    mats.append((Q * lam) @ Q.T)
    out = ctx.diag_batch(mats)
    for A, (W, U) in zip(mats, out):
        _check_eig(A, W, U)


def test_diag_batch_large(ctx):
    # 700: the streaming team tridiagonalisation; 1100, 1345, 2100: the
    # blocked one (from 1024 rows); the divide-and-conquer merges of all four
    # deflate in LDS (the global-memory deflation above 5000 rows is covered
    # by the 7000-row case of the blocked tests)
    rng = np.random.default_rng(3)
    mats = [_sym(rng, 700), _sym(rng, 1100), _sym(rng, 1345), _sym(rng, 2100)]
    out = ctx.diag_batch(mats)
    for A, (W, U) in zip(mats, out):
        _check_eig(A, W, U, tol=2e-13)
    # the workspace returned to the device and allocated again: the same bits
    ctx.release_workspace()
    out2 = ctx.diag_batch(mats)
    for (W, U), (W2, U2) in zip(out, out2):
        assert np.array_equal(W, W2) and np.array_equal(U, U2)


def test_diag_batch_streaming_teams_with_small_blocks(ctx):
    # no register team: the one-workgroup matrices run in their own launch
    # beside the streaming teams
    rng = np.random.default_rng(8)
    mats = [_sym(rng, n) for n in (700, 1, 50, 100, 128)]
    out = ctx.diag_batch(mats)
    for A, (W, U) in zip(mats, out):
        _check_eig(A, W, U, tol=2e-13)


def test_diag_batch_beyond_register_limit(ctx):
    # 2500: a team spread over XCDs; 4200: back-transformation with the
    # eigenvector partly in LDS
    rng = np.random.default_rng(7)
    mats = [_sym(rng, 4200), _sym(rng, 2500), _sym(rng, 300)]
    out = ctx.diag_batch(mats)
    for A, (W, U) in zip(mats, out):
        _check_eig(A, W, U, tol=2e-13)


def test_diag_batch_rejects_oversize(ctx):
    # the size check happens before any device access: a small buffer suffices
    capi = _capi()
    t = (capi.EigTask * 1)()
    t[0].A, t[0].W, t[0].n = ctx.alloc(16), ctx.alloc(16), 32769  # above the documented limit of 32768
    st = ctx.lib.nrg_diag_batch(ctx.ctx, t, 1)
    assert st != 0 and b"dimension" in ctx.lib.nrg_last_error(ctx.ctx)


def test_diag_values_early_in_task_order(ctx):
    # nrg_diag_values returns the batch's eigenvalues, concatenated in the
    # order the tasks were passed (the eigensolver sorts them by size), the
    # same bits as reading W back once everything has finished
    capi = _capi()
    rng = np.random.default_rng(12)
    sizes = [5, 130, 1, 300, 64, 2, 129]
    tasks = (capi.EigTask * len(sizes))()
    ws = []
    for t, n in zip(tasks, sizes):
        m = rng.standard_normal((n, n))
        t.A, t.W, t.n = ctx.put(m + m.T), ctx.alloc(n), n
        ws.append((t.W, n))
    assert ctx.lib.nrg_diag_batch(ctx.ctx, tasks, len(sizes)) == 0
    early = ctx.diag_values(sum(sizes))
    late = np.concatenate([ctx.get(w, (n,)) for w, n in ws])
    np.testing.assert_array_equal(early, late)
    out = np.empty(4)
    # nothing queued any more, and a count that is not the batch's
    assert ctx.lib.nrg_diag_values(ctx.ctx, out.ctypes.data_as(C.c_void_p), C.c_size_t(4)) != 0
    assert ctx.lib.nrg_diag_batch(ctx.ctx, tasks, 2) == 0
    assert ctx.lib.nrg_diag_values(ctx.ctx, out.ctypes.data_as(C.c_void_p), C.c_size_t(4)) != 0


def test_diag_batch_many_small_with_shared_spectrum(ctx):
    # many blocks of one iteration: identical spectra in different blocks and
    # clusters inside blocks (SU(2) multiplets of neighbouring subspaces)
    rng = np.random.default_rng(6)
    mats = []
    for k in range(60):
        n = int(rng.integers(1, 90))
        lam = np.round(rng.uniform(0, 5, n), 1)
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        mats.append(Q @ np.diag(lam) @ Q.T)
    out = ctx.diag_batch(mats)
    for A, (W, U) in zip(mats, out):
        _check_eig(A, W, U)


def _gemm_cases(ctx, rng, shapes):
    capi = _capi()
    outs, terms, refs, handles = [], [], [], []
    for case, (m, n) in enumerate(shapes):
        ldc = n + (case % 3)
        Cinit = rng.standard_normal((m, ldc))
        beta = [0.0, 1.0, -0.5][case % 3]
        ref = beta * Cinit[:, :n] if beta != 0.0 else np.zeros((m, n))
        o = capi.GemmOut()
        o.C = ctx.put(Cinit)
        o.m, o.n, o.ldc, o.beta = m, n, ldc, beta
        o.term_begin = len(terms)
        nt = 1 + case % 3
        for t_ in range(nt):
            k = int(rng.integers(1, 90 if m * n < 40000 else 700))
            tA, tB = bool(rng.integers(0, 2)), bool(rng.integers(0, 2))
            lda_extra, ldb_extra = int(rng.integers(0, 3)), int(rng.integers(0, 3))
            A = rng.standard_normal((k, m + lda_extra) if tA else (m, k + lda_extra))
            B = rng.standard_normal((n, k + ldb_extra) if tB else (k, n + ldb_extra))
            opA = A[:, :m].T if tA else A[:, :k]
            opB = B[:, :k].T if tB else B[:, :n]
            s = rng.standard_normal(k) if (t_ % 2) else None
            alpha = float(rng.standard_normal())
            t = capi.GemmTerm()
            t.A, t.B = ctx.put(A), ctx.put(B)
            t.s = ctx.put(s) if s is not None else None
            t.alpha, t.k = alpha, k
            t.lda, t.ldb = A.shape[1], B.shape[1]
            t.transA, t.transB = int(tA), int(tB)
            terms.append(t)
            ref = ref + alpha * (opA * (s[None, :] if s is not None else 1.0)) @ opB
        o.term_count = nt
        outs.append(o)
        refs.append((ref, Cinit))
        handles.append((o.C, m, ldc, n))
    ctx.gemm_grouped(outs, terms)
    for (ref, Cinit), (p, m, ldc, n) in zip(refs, handles):
        got = ctx.get(p, (m, ldc))
        np.testing.assert_allclose(got[:, :n], ref, rtol=1e-13, atol=1e-12)
        np.testing.assert_array_equal(got[:, n:], Cinit[:, n:])  # outside the view untouched


def test_gemm_grouped_all_transposes_and_scales(ctx):
    # few tiles: the 32 x 32 FMA kernel
    rng = np.random.default_rng(4)
    _gemm_cases(ctx, rng, [tuple(rng.integers(1, 150, 2)) for _ in range(12)])


def test_gemm_grouped_mfma_tiles(ctx):
    # at least 256 tiles of 64 x 64 in the launch: the FP64-MFMA kernel, with
    # ragged edges (sizes not multiples of 64 or 16), every transpose, scale
    # vectors, several terms and beta
    rng = np.random.default_rng(14)
    shapes = [(1100, 1030), (577, 700), (64, 4096), (130, 65), (1, 900), (333, 1), (1023, 1025)]
    _gemm_cases(ctx, rng, shapes)


def test_gemm_grouped_split_long_k(ctx):
    # few output tiles with long k: the launch splits every tile's (term, k)
    # sequence into slices whose partial tiles are reduced in slice order;
    # a k = 0 term between others must be skipped
    capi = _capi()
    rng = np.random.default_rng(11)
    outs, terms, refs, handles = [], [], [], []
    for case, (m, n) in enumerate([(70, 130), (64, 64), (1, 200), (129, 3)]):
        ldc = n + case
        Cinit = rng.standard_normal((m, ldc))
        beta = [0.0, 1.0, -0.5, 1.0][case]
        ref = beta * Cinit[:, :n] if beta != 0.0 else np.zeros((m, n))
        o = capi.GemmOut()
        o.C = ctx.put(Cinit)
        o.m, o.n, o.ldc, o.beta = m, n, ldc, beta
        o.term_begin = len(terms)
        ks = [700, 0, 1500, 33]
        for t_, k in enumerate(ks):
            tA, tB = bool(t_ & 1), bool(case & 1)
            A = rng.standard_normal((max(k, 1), m) if tA else (m, max(k, 1)))
            B = rng.standard_normal((n, max(k, 1)) if tB else (max(k, 1), n))
            opA = A[:k, :].T if tA else A[:, :k]
            opB = B[:, :k].T if tB else B[:k, :]
            s = rng.standard_normal(k) if t_ == 2 else None
            alpha = float(rng.standard_normal())
            t = capi.GemmTerm()
            t.A, t.B = ctx.put(A), ctx.put(B)
            t.s = ctx.put(s) if s is not None else None
            t.alpha, t.k = alpha, k
            t.lda, t.ldb = A.shape[1], B.shape[1]
            t.transA, t.transB = int(tA), int(tB)
            terms.append(t)
            if k:
                ref = ref + alpha * (opA * (s[None, :] if s is not None else 1.0)) @ opB
        o.term_count = len(ks)
        outs.append(o)
        refs.append((ref, Cinit))
        handles.append((o.C, m, ldc, n))
    ctx.gemm_grouped(outs, terms)
    for (ref, Cinit), (p, m, ldc, n) in zip(refs, handles):
        got = ctx.get(p, (m, ldc))
        np.testing.assert_allclose(got[:, :n], ref, rtol=1e-12, atol=1e-11)
        np.testing.assert_array_equal(got[:, n:], Cinit[:, n:])  # beyond n untouched


def test_assemble_batch(ctx):
    capi = _capi()
    rng = np.random.default_rng(5)
    d = 20
    H = ctx.alloc(d * d)
    e1 = rng.standard_normal(6)
    e2 = rng.standard_normal(14)
    F = rng.standard_normal((6, 16))  # ld 16, used 6 x 14
    terms = []
    for kind, src, scale, r0, c0, rows, cols, ld in [
            (0, ctx.put(e1), 1.4, 0, 0, 6, 0, 0), (0, ctx.put(e2), 1.4, 6, 0, 14, 0, 0),
            (1, ctx.put(F), -0.7, 0, 6, 6, 14, 16), (2, None, 0.25, 6, 0, 14, 0, 0)]:
        t = capi.AsmTerm()
        t.src, t.scale, t.kind, t.r0, t.c0, t.rows, t.cols, t.ld = src, scale, kind, r0, c0, rows, cols, ld
        terms.append(t)
    b = capi.AsmBlock()
    b.H, b.d, b.term_begin, b.term_count = H, d, 0, len(terms)
    ctx.assemble([b], terms)
    ref = np.zeros((d, d))
    ref[np.arange(6), np.arange(6)] += 1.4 * e1
    ref[6 + np.arange(14), 6 + np.arange(14)] += 1.4 * e2 + 0.25
    ref[:6, 6:] += -0.7 * F[:, :14]
    ref[6:, :6] += -0.7 * F[:, :14].T
    np.testing.assert_allclose(ctx.get(H, (d, d)), ref, rtol=0, atol=1e-15)


def test_assemble_batch_bands(ctx):
    # blocks of several sizes, each cut into three parts: diagonal vectors and
    # shifts on every part, block terms (padded leading dimension) between
    # every pair of parts, so that terms and their transposed halves cross the
    # row bands the assembly is split into
    capi = _capi()
    rng = np.random.default_rng(11)

    def term(kind, src, scale, r0, c0, rows, cols, ld):
        t = capi.AsmTerm()
        t.src, t.scale, t.kind, t.r0, t.c0, t.rows, t.cols, t.ld = src, scale, kind, r0, c0, rows, cols, ld
        return t

    blocks, terms, refs = [], [], []
    for d in (1, 17, 100, 257):
        H = ctx.alloc(d * d)
        c = sorted(int(x) for x in rng.integers(0, d + 1, size=2))
        parts = [(0, c[0]), (c[0], c[1]), (c[1], d)]
        ref = np.zeros((d, d))
        tb = len(terms)
        for a0, a1 in parts:
            if a1 > a0:
                e = rng.standard_normal(a1 - a0)
                idx = np.arange(a0, a1)
                terms.append(term(0, ctx.put(e), 0.9, a0, 0, a1 - a0, 0, 0))
                ref[idx, idx] += 0.9 * e
                terms.append(term(2, None, -0.3, a0, 0, a1 - a0, 0, 0))
                ref[idx, idx] += -0.3
        for p_, q_ in ((0, 1), (0, 2), (1, 2)):
            (r0, r1), (c0, c1) = parts[p_], parts[q_]
            if r1 > r0 and c1 > c0:
                ld = c1 - c0 + 3
                F = rng.standard_normal((r1 - r0, ld))
                sc = float(rng.standard_normal())
                terms.append(term(1, ctx.put(F), sc, r0, c0, r1 - r0, c1 - c0, ld))
                ref[r0:r1, c0:c1] += sc * F[:, :c1 - c0]
                ref[c0:c1, r0:r1] += sc * F[:, :c1 - c0].T
        b = capi.AsmBlock()
        b.H, b.d, b.term_begin, b.term_count = H, d, tb, len(terms) - tb
        blocks.append(b)
        refs.append((H, d, ref))
    ctx.assemble(blocks, terms)
    for H, d, ref in refs:
        np.testing.assert_allclose(ctx.get(H, (d, d)), ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize("mode", ["FDM", "DMNRG"])
def test_spec_bin_matches_cpu_binning(ctx, mode):
    capi = _capi()
    orc = load_oracle()
    rng = np.random.default_rng(6)
    p = dict(orc.DEFAULTS)
    p["bins"] = 50
    scale0, scalel = 1.06, 1e-6
    nb = orc.Bins(p, scale0, scalel).nr
    pos_ref, neg_ref = orc.Bins(p, scale0, scalel), orc.Bins(p, scale0, scalel)
    pairs = []
    T, wnf, escale = 0.003, 0.7, 0.01
    for _ in range(5):
        alli, allj = rng.integers(1, 70, 2)
        reti, retj = (rng.integers(0, alli + 1), rng.integers(0, allj + 1)) if mode == "FDM" else (alli, allj)
        op1 = rng.standard_normal((allj, alli))
        op2 = rng.standard_normal((allj, alli))
        M1 = rng.standard_normal((allj, max(reti, 1)))
        M2 = rng.standard_normal((max(retj, 1), alli))
        if mode == "DMNRG":
            M1 = rng.standard_normal((allj, alli))
            M2 = rng.standard_normal((allj, alli))
        Ei = np.sort(rng.uniform(0, 0.05, alli))
        Ej = np.sort(rng.uniform(0, 0.05, allj))
        Ei[0] = Ej[0]  # a zero-energy peak
        sf = float(rng.uniform(0.5, 2))
        q = capi.SpecPair()
        q.op1, q.op2, q.M1, q.M2 = ctx.put(op1), ctx.put(op2), ctx.put(M1), ctx.put(M2)
        q.Ei, q.Ej = ctx.put(Ei), ctx.put(Ej)
        q.alli, q.allj, q.reti, q.retj = int(alli), int(allj), int(reti), int(retj)
        q.ld1, q.ld2, q.ldm1, q.ldm2 = int(alli), int(alli), int(M1.shape[1]), int(M2.shape[1])
        q.sf = sf
        pairs.append(q)
        E = (Ej[:, None] - Ei[None, :]) * (escale if mode == "DMNRG" else 1.0)
        ws, es = [], []
        if mode == "FDM":
            Di = np.arange(alli)[None, :] >= reti
            Dj = np.arange(allj)[:, None] >= retj
            Di, Dj = np.broadcast_to(Di, E.shape), np.broadcast_to(Dj, E.shape)
            P = sf * op1 * op2
            A = np.where(Di, P * wnf * np.exp(-Ei[None, :] / T), 0.0)
            m1 = np.zeros(E.shape)
            if reti:
                m1[:, :reti] = sf * op1[:, :reti] * M1[:, :reti]
            A1 = np.where(~Di & Dj, m1, 0.0)
            B = np.where(Dj, P * wnf * np.exp(-Ej[:, None] / T), 0.0)
            m2 = np.zeros(E.shape)
            if retj:
                m2[:retj, :] = sf * op1[:retj, :] * M2[:retj, :]
            B2 = np.where(~Dj & Di, m2, 0.0)
            for mask, W in ((Di, A), (~Di & Dj, A1), (Dj, B), (~Dj & Di, B2)):
                es.append(E[mask])
                ws.append(W[mask])
        else:
            aE = np.abs(Ej[:, None] - Ei[None, :])
            ok = ~((aE < 2.0 * 0) | (aE > 1e9))
            W = sf * (M1 * op1 + M2 * op2)
            es.append(E[ok])
            ws.append(W[ok])
        for e, w in zip(es, ws):
            pos = e >= 0
            pos_ref.add_many(e[pos], w[pos])
            neg_ref.add_many(-e[~pos], w[~pos])
    prm = capi.SpecParams()
    prm.mode = 0 if mode == "FDM" else 1
    prm.nbins = nb
    prm.log10emin = pos_ref.l10min
    prm.bins_per_decade = 50
    prm.discard = 1e-16
    prm.T, prm.wnf, prm.escale = T, wnf, (escale if mode == "DMNRG" else 1.0)
    prm.emin_w, prm.emax_w = 0.0, 1e9
    dbins = ctx.alloc(2 * nb)
    ctx.lib.nrg_zero(ctx.ctx, C.c_void_p(dbins), C.c_size_t(2 * nb))
    ctx.spec_bin(prm, pairs, dbins)
    got = ctx.get(dbins, (2 * nb,))
    scale = np.abs(np.concatenate([pos_ref.w, neg_ref.w])).max()
    np.testing.assert_allclose(got[:nb], pos_ref.w, rtol=1e-11, atol=1e-13 * scale)
    np.testing.assert_allclose(got[nb:], neg_ref.w, rtol=1e-11, atol=1e-13 * scale)
