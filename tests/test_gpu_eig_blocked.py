"""The blocked tridiagonalisation (panels of Householder reflectors, a
half-matrix product per column, an FP64-MFMA rank-2 panel update) through
nrg_diag_batch, against LAPACK and the defining identities.

Matrices from the dimension NRG_BLK_MIN_N up take the blocked path (default
1024).  A context created with a lower threshold runs it on small matrices
too, so every corner (tau = 0 columns, partial first and last 64-blocks, one
and many panels, teams inside one XCD and across XCDs) is exercised at sizes
LAPACK checks in seconds.  The largest cases are the block sizes of the
two-channel keep-10000 workload (8820, 12964, 13136 rows): there the check is
the residual |A u - lambda u| and the orthonormality of the rows, which fix
every eigenpair to the same tolerance without a CPU eigensolver.
"""
import os

import numpy as np
import pytest

from conftest import load_nrgamd

pytestmark = pytest.mark.gpu


def _ctx(blk_min_n=None):
    import nrgamd.capi as capi
    old = os.environ.get("NRG_BLK_MIN_N")
    if blk_min_n is not None:
        os.environ["NRG_BLK_MIN_N"] = str(blk_min_n)
    try:
        c = capi.Ctx(0)
        c.diag_batch([np.eye(3)])  # the threshold is read at the context's first eigensolve
    finally:
        if blk_min_n is not None:
            if old is None:
                del os.environ["NRG_BLK_MIN_N"]
            else:
                os.environ["NRG_BLK_MIN_N"] = old
    return c


@pytest.fixture(scope="module")
def ctx_small():
    load_nrgamd()
    c = _ctx(blk_min_n=200)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx():
    load_nrgamd()
    c = _ctx()
    yield c
    c.close()


def _sym(rng, n):
    a = rng.standard_normal((n, n))
    return (a + a.T) / 2


def _check_pairs(A, W, U, tol=1e-12):
    """Rows of U orthonormal, A u_r = W_r u_r, W ascending."""
    n = A.shape[0]
    nrm = max(1.0, np.abs(A).max() * n)
    assert np.all(np.isfinite(W)) and np.all(np.diff(W) >= 0)
    G = U @ U.T
    G[np.diag_indices(n)] -= 1.0
    assert np.abs(G).max() <= 1e-12 * max(1.0, n ** 0.5), np.abs(G).max()
    res = A @ U.T - U.T * W[None, :]
    assert np.abs(res).max() <= tol * nrm, (np.abs(res).max(), tol * nrm)


def _check_eig(A, W, U, tol=1e-12):
    n = A.shape[0]
    nrm = max(1.0, np.abs(A).max() * n)
    np.testing.assert_allclose(W, np.linalg.eigvalsh(A), atol=tol * nrm, rtol=0)
    _check_pairs(A, W, U, tol)


def test_blocked_small_sizes(ctx_small):
    # partial first/last 64-blocks, one panel (200), several panels, teams of
    # 2 to 10 workgroups inside one XCD
    rng = np.random.default_rng(21)
    sizes = [200, 201, 255, 256, 257, 319, 320, 577, 640, 700, 1100]
    mats = [_sym(rng, n) for n in sizes]
    for A, (W, U) in zip(mats, ctx_small.diag_batch(mats)):
        _check_eig(A, W, U)


def test_blocked_structured(ctx_small):
    rng = np.random.default_rng(22)
    mats = []
    # already diagonal: every column has sigma = 0 (tau = 0, no reflector)
    mats.append(np.diag(rng.standard_normal(260)))
    # block diagonal: tau = 0 columns between nonzero ones
    b = _sym(rng, 70)
    mats.append(np.kron(np.eye(4), b))
    # multiplets (exact degeneracies) and a rank-one matrix
    Q, _ = np.linalg.qr(rng.standard_normal((400, 400)))
    lam = np.repeat(np.round(rng.uniform(0, 3, 100), 2), 4)
    mats.append((Q * lam) @ Q.T)
    v = rng.standard_normal(300)
    mats.append(np.outer(v, v))
    # NRG-like: large ordered diagonal, weak couplings between blocks
    d = np.sort(rng.uniform(0, 60, 500))
    h = np.diag(d)
    h[:250, 250:] = 0.05 * rng.standard_normal((250, 250))
    h[250:, :250] = h[:250, 250:].T
    mats.append(h)
    # tridiagonal input with tiny couplings
    t = np.diag(np.linspace(0, 1, 230)) + np.diag(np.full(229, 1e-14), 1) + np.diag(np.full(229, 1e-14), -1)
    mats.append(t)
    for A, (W, U) in zip(mats, ctx_small.diag_batch(mats)):
        _check_eig(A, W, U)


def test_blocked_default_threshold_local_and_global_teams(ctx):
    # 1024 / 1500 / 2600: teams inside one XCD; 4200 and 5000: teams across
    # XCDs (more than 32 workgroups)
    rng = np.random.default_rng(23)
    sizes = [1024, 1500, 2600, 4200, 5000]
    mats = [_sym(rng, n) for n in sizes]
    for A, (W, U) in zip(mats, ctx.diag_batch(mats)):
        _check_eig(A, W, U, tol=2e-13)


def test_blocked_result_independent_of_batch(ctx):
    # the same matrix alone and among others.  The reduction's sums run in an
    # order fixed by the matrix alone; the back-transformation GEMMs split
    # their k range by the launch's tile count, which changes the last bits
    # (1e-13 relative is the stated bound of "alone vs in the mix")
    rng = np.random.default_rng(24)
    A = _sym(rng, 1800)
    (W1, U1), = ctx.diag_batch([A])
    out = ctx.diag_batch([_sym(rng, 3000), A, _sym(rng, 1200), _sym(rng, 100)])
    W2, U2 = out[1]
    np.testing.assert_allclose(W2, W1, rtol=0, atol=1e-13 * np.abs(W1).max())
    np.testing.assert_allclose(U2, U1, rtol=0, atol=1e-12)


def test_blocked_deflation_heavy_above_lds_sizes(ctx):
    # above the deflation kernel's LDS size (5000) and the streaming team's
    # LDS size (6800): clustered spectrum, so rotation chains and deflation
    # happen at the top merges
    rng = np.random.default_rng(25)
    n = 7000
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.repeat(np.round(rng.uniform(-2, 2, n // 4), 3), 4)
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    (W, U), = ctx.diag_batch([A])
    np.testing.assert_allclose(W, np.sort(lam), atol=1e-12 * n * np.abs(A).max(), rtol=0)
    _check_pairs(A, W, U, tol=2e-13)


@pytest.mark.parametrize("n", [8820, 12964, 13136])
def test_blocked_largest_blocks_of_keep10000(ctx, n):
    rng = np.random.default_rng(n)
    A = _sym(rng, n)
    (W, U), = ctx.diag_batch([A])
    _check_pairs(A, W, U, tol=2e-13)
    # trace and Frobenius norm are invariants of the spectrum
    assert abs(W.sum() - np.trace(A)) <= 1e-11 * n * np.abs(A).max()
    assert abs(np.dot(W, W) - np.sum(A * A)) <= 1e-10 * np.sum(A * A)
