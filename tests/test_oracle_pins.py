"""Closed-form and exact-diagonalisation pins of the oracle's binning,
broadening and DM-NRG weights.

Bins (src/bins.h:50-131): a delta peak between two representative energies
is shared linearly in log10(E) between them, so a peak at
10^(l10min + (i + f)/bpd) puts (1 - f) w into bin i and f w into bin i + 1,
and the log-energy centre of mass is conserved; below the grid everything
lands in the first bin, above it in the last.  A swapped rem / 1 - rem, an
index off by one or a wrong grid exponent fails here.

Broadening (src/spectral.h:56-102): above omega0 the kernel BR_L(e, e') is the
log-normal density of e with ln e ~ N(ln e' + alpha^2/4, alpha^2/2) (scipy's
lognorm), symmetric in its arguments; well below omega0 it is the normal
density of width omega0/sqrt(2) around e' (scipy's norm); it integrates to one.

DM-NRG (src/spec_DMNRG.cc:13-52): on the last shell, with nothing truncated,
the density matrix is the whole chain's Boltzmann matrix, so the raw weights
the oracle bins there are the exact finite-temperature Lehmann weights
|<j|d^dag|m>|^2 (rho_m + rho_j) of the chain inside the window
goodE <= |E_j - E_m| / omega_N <= goodE Lambda^2.
"""
import math
import os

import numpy as np
import pytest

from conftest import load_nrgamd, load_oracle
import ed_chain


def _bins(bpd=10, emin=2e-6, emax=30.0):
    orc = load_oracle()
    p = dict(orc.DEFAULTS, bins=bpd, emin=emin, emax=emax)
    return orc.Bins(p, 1.0, 1e-6)


def test_bin_grid_closed_form():
    b = _bins(bpd=10, emin=2e-6, emax=30.0)
    # floor(log10 2e-6) = -6, ceil(log10 30) = 2: 8 decades, 10 per decade
    assert (b.l10min, b.l10max, b.nr) == (-6, 2, 81)
    np.testing.assert_allclose(b.e, 10.0 ** (-6.0 + np.arange(81) / 10.0), rtol=1e-15)


@pytest.mark.parametrize("i,f", [(0, 0.25), (7, 0.5), (33, 0.9), (79, 0.1)])
def test_bin_split_is_linear_in_log_energy(i, f):
    b = _bins()
    E = 10.0 ** (b.l10min + (i + f) / b.bpd)
    w = 0.75
    b.add_many(np.array([E]), np.array([w]))
    expect = np.zeros(b.nr)
    expect[i], expect[i + 1] = (1.0 - f) * w, f * w
    np.testing.assert_allclose(b.w, expect, atol=1e-12)
    # the log-energy centre of mass of the shared weight is the peak's
    assert abs((b.w * np.log10(b.e)).sum() / b.w.sum() - math.log10(E)) < 1e-12


def test_bin_clamps_and_discard():
    b = _bins()
    # below the grid (and E = 0): the first bin; at or above the last
    # representative energy: the last bin
    b.add_many(np.array([1e-9, 0.0, 10.0 ** b.l10max, 1e3]), np.array([1.0, 2.0, 4.0, 8.0]))
    assert b.w[0] == 3.0 and b.w[-1] == 12.0 and b.w[1:-1].sum() == 0.0
    # |w| < discard_immediately * E is dropped before binning
    b2 = _bins()
    b2.add_many(np.array([1.0, 1e-5]), np.array([1e-20, 1e-20]))
    assert b2.w.sum() == pytest.approx(1e-20, rel=1e-12)
    # many peaks: weight is conserved
    rng = np.random.default_rng(3)
    b3 = _bins()
    E = 10.0 ** rng.uniform(-7, 3, 2000)
    W = rng.uniform(0.1, 1.0, 2000)
    b3.add_many(E, W)
    assert abs(b3.w.sum() - W.sum()) < 1e-10


def _broadener(T=1e-6, **kw):
    orc = load_oracle()
    o = orc.Oracle.__new__(orc.Oracle)
    o.p = dict(orc.DEFAULTS, T=T, **kw)
    return o


def test_broadening_kernel_is_lognormal_above_omega0():
    from scipy import stats
    o = _broadener(T=1e-9)
    a = o.p["alpha"]
    ept = 3e-3
    e = ept * np.exp(np.linspace(-6 * a, 6 * a, 101))
    ref = stats.lognorm.pdf(e, s=a / math.sqrt(2.0), scale=ept * math.exp(a * a / 4.0))
    np.testing.assert_allclose(o.br_new(e, ept), ref, rtol=1e-12)
    # negative side mirrors, opposite signs give nothing
    np.testing.assert_allclose(o.br_new(-e, -ept), ref, rtol=1e-12)
    assert np.all(o.br_new(-e, ept) == 0.0)
    # symmetric in its two arguments
    rng = np.random.default_rng(5)
    x, y = 10.0 ** rng.uniform(-5, 0, 50), 10.0 ** rng.uniform(-5, 0, 50)
    np.testing.assert_allclose(o.br_new(x, y), o.br_new(y, x), rtol=1e-12)


def test_broadening_kernel_normalisation_and_log_mean():
    o = _broadener(T=1e-12)
    a = o.p["alpha"]
    ept = 0.01
    u = np.linspace(math.log(ept) - 12 * a, math.log(ept) + 12 * a, 20001)
    E = np.exp(u)
    f = o.br_new(E, ept) * E  # density in ln E
    norm = np.trapezoid(f, u)
    assert abs(norm - 1.0) < 1e-10
    assert abs(np.trapezoid(f * u, u) - (math.log(ept) + a * a / 4.0)) < 1e-10


def test_broadening_kernel_below_omega0_is_gaussian():
    from scipy import stats
    T = 1e-4
    o = _broadener(T=T)  # omega0 = omega0_ratio * T = T
    w0 = T
    e = w0 * np.array([1e-6, 1e-5, 1e-4])
    ept = 0.4 * w0
    ref = stats.norm.pdf(e, loc=ept, scale=w0 / math.sqrt(2.0))
    np.testing.assert_allclose(o.br_new(e, ept), ref, rtol=1e-12)
    # continuous where the blend starts (|e| = omega0)
    lo, hi = o.br_new(w0 * (1 - 1e-12), ept), o.br_new(w0 * (1 + 1e-12), ept)
    assert abs(lo - hi) <= 1e-9 * abs(hi)


def test_broadening_mesh():
    o = _broadener(broaden_min=1e-4)
    o.Nmax = 10
    m = o.broadening_mesh()
    assert m[0] == o.p["broaden_max"]
    np.testing.assert_allclose(m[:-1] / m[1:], o.p["broaden_ratio"], rtol=1e-14)
    assert m[-1] > 1e-4 >= m[-1] / o.p["broaden_ratio"]
    assert len(m) == math.ceil(math.log(2.5 / 1e-4) / math.log(1.05))


def _cdf_at(E, W, x):
    """Total weight at energies <= each x."""
    order = np.argsort(E)
    c = np.concatenate([[0.0], np.cumsum(W[order])])
    return c[np.searchsorted(E[order], x, side="right")]


@pytest.mark.parametrize("Nmax,T", [(3, 0.05), (4, 0.02)])
def test_dmnrg_last_shell_weights_are_exact_lehmann(tmp_path, Nmax, T):
    import test_oracle_ed as ted
    dg = ted._make(str(tmp_path), "QS", 1, Nmax, T=T, dmnrg=True)
    orc = load_oracle()
    o = orc.Oracle(str(tmp_path))
    rec = []
    orig = o.spec_dmnrg

    def spy(dIp, dI1, op1, op2, sf, cs, rhoIp, rhoI1, sc):
        real_add = o.cs_add
        o.cs_add = lambda cs_, E, W: (rec.append((sc, E.copy(), W.copy())), real_add(cs_, E, W))
        try:
            orig(dIp, dI1, op1, op2, sf, cs, rhoIp, rhoI1, sc)
        finally:
            del o.cs_add

    o.spec_dmnrg = spy
    o.run()
    sc = o.SCALE(Nmax)
    got = [(E, W) for s, E, W in rec if s == sc]
    assert got
    gE = np.concatenate([E for E, _ in got])
    gW = np.concatenate([W for _, W in got])
    C, H, cd, nd = ted._chain(dg, 1, Nmax, Nmax=Nmax)
    Eed, V = ed_chain.eig_sectors(C, H)
    om, w = ed_chain.lehmann(Eed, V, cd[("d", 1)], T)
    lo, hi = o.p["goodE"] * sc, o.p["goodE"] * o.p["Lambda"] ** 2 * sc
    ok = (np.abs(om) >= lo) & (np.abs(om) <= hi)
    assert not np.any(np.isclose(np.abs(om), lo, rtol=1e-8) | np.isclose(np.abs(om), hi, rtol=1e-8))
    gE, gW = gE / sc, gW
    om, w = om[ok] / sc, w[ok]
    # the weight distribution over energy: its cumulative sum at points away
    # from every transition (degenerate multiplets make single transitions
    # ambiguous), and its moments
    rng = np.random.default_rng(11)
    x = rng.uniform(-o.p["goodE"] * o.p["Lambda"] ** 2, o.p["goodE"] * o.p["Lambda"] ** 2, 400)
    x = x[np.min(np.abs(x[:, None] - om[None, :]), axis=1) > 1e-7]
    assert len(x) > 100
    np.testing.assert_allclose(_cdf_at(gE, gW, x), _cdf_at(om, w, x), rtol=1e-10, atol=1e-14)
    for k in range(5):
        assert abs((gW * gE ** k).sum() - (w * om ** k).sum()) <= 1e-11 * (np.abs(w * om ** k).sum())
    assert w.sum() > 0.05


def test_cluster_snapping_follows_find_clusters():
    """find_clusters / fix_splittings (src/splitting.cc:5-77, src/nrg.cc:561,
    2910-2918) by hand: closed clusters with distinct members snap to their
    first value, the last cluster stays open, equal members need no entry, and
    the mapping made in one iteration applies to the same value later."""
    orc = load_oracle()
    o = orc.Oracle.__new__(orc.Oracle)
    o.p = dict(orc.DEFAULTS, fixeps=1e-3)
    a, b = 0.5 + 1e-4, 1.0 + 2e-4
    diag = {(0, 1): dict(value=np.array([0.0, 0.5, 1.0, 2.0])),
            (1, 2): dict(value=np.array([a, b, 1.5, 1.5, 2.0 + 1e-4]))}
    o.fix_splittings(diag)
    np.testing.assert_array_equal(diag[(0, 1)]["value"], [0.0, 0.5, 1.0, 2.0])
    # 0.5+1e-4 -> 0.5 and 1.0+2e-4 -> 1.0; {1.5, 1.5} is not split; the top
    # cluster {2.0, 2.0+1e-4} is never closed
    np.testing.assert_array_equal(diag[(1, 2)]["value"], [0.5, 1.0, 1.5, 1.5, 2.0 + 1e-4])
    assert o.cluster_mapping == {a: 0.5, b: 1.0}
    # a later iteration: a lone value equal to an earlier key is still mapped
    later = {(0, 1): dict(value=np.array([0.0, a, 3.0]))}
    o.fix_splittings(later)
    np.testing.assert_array_equal(later[(0, 1)]["value"], [0.0, 0.5, 3.0])
