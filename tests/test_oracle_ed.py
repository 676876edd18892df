"""Pin the oracle (and the `data` generator) against brute-force ED of whole chains.

Without truncation NRG is exact: the multiplet-expanded eigenvalues at every
iteration, the thermodynamics computed from them, the singlet expectation
values and the FDM spectral weights (only the last shell carries weight when
nothing is discarded) all equal the exact diagonalisation of the impurity plus
the Wilson chain built up to that length.  A wrong coefficient table entry,
a transposed f-block, a wrong SCALE exponent or a dropped zeta term fails here.
"""
import math
import os

import numpy as np
import pytest

from conftest import load_nrgamd, load_oracle
import ed_chain


def _make(tmp, sym, nch, Nmax, U=1.0, eps_d=-0.4, Gamma=0.05, Lambda=2.0, T=0.3, specd=None, **kw):
    dg = load_nrgamd().datagen if hasattr(load_nrgamd(), "datagen") else __import__("nrgamd.datagen").datagen
    model = dg.siam(sym=sym, nch=nch, U=U, eps_d=eps_d, Gamma=Gamma, Lambda=Lambda)
    model.write(os.path.join(tmp, "data"), Lambda, 1.0, Nmax)
    dg.write_param(os.path.join(tmp, "param"), Lambda=Lambda, keep=100000, T=T, Nmax=Nmax,
                   specd=specd or ("A_d-A_d" if sym == "QS" else "A_d-A_d-u"), bins=50, **kw)
    return dg


def _chain(dg, nch, K, U=1.0, eps_d=-0.4, Gamma=0.05, Lambda=2.0, Nmax=None):
    xi, zeta = dg.chain_coefficients(Lambda, 1.0, Nmax)
    V = [math.sqrt(2 * Gamma * dg.gamma_A(Lambda) / math.pi)] * nch
    C, H, cd, nd = ed_chain.siam_chain(U, eps_d, V, [xi] * nch, [zeta] * nch, K, nch)
    return C, H, cd, nd


@pytest.mark.parametrize("sym", ["QS", "QSZ"])
def test_eigenvalues_td_expv_match_ed_1ch(tmp_path, sym):
    Nmax = 3
    dg = _make(str(tmp_path), sym, 1, Nmax)
    o = load_oracle().Oracle(str(tmp_path)).run()
    for N in range(Nmax):
        C, H, cd, nd = _chain(dg, 1, N + 1, Nmax=Nmax)
        E, V = ed_chain.eig_sectors(C, H)
        sc = o.SCALE(N + 1)
        ref = np.sort((E - E.min()) / sc)
        mine = np.sort(np.concatenate([np.repeat(v, o.tab.mult(I)) for I, (v, k) in o.iters[N].items()]))
        assert len(mine) == len(ref)
        np.testing.assert_allclose(mine, ref, atol=1e-10, rtol=0)
        # thermodynamics at T_N = SCALE(N+1)
        b = np.exp(-(E - E.min()) / sc)
        Z = b.sum()
        Eb = ((E - E.min()) / sc * b).sum() / Z
        row = o.td_rows[N + 1]
        assert abs(row["E"] - Eb) < 1e-9
        assert abs(row["F"] + math.log(Z)) < 1e-9
        # <n_d> in the custom file
        ndexp = np.einsum("ij,jk,ki->", V.T, nd, V * b[None, :]) / Z
        t, vals = o.custom_rows[N + 1]
        assert abs(vals["n_d"] - ndexp) < 1e-9


def test_eigenvalues_match_ed_2ch(tmp_path):
    Nmax = 1
    dg = _make(str(tmp_path), "QS", 2, Nmax)
    o = load_oracle().Oracle(str(tmp_path)).run()
    C, H, cd, nd = _chain(dg, 2, 1, Nmax=Nmax)
    E, V = ed_chain.eig_sectors(C, H)
    ref = np.sort((E - E.min()) / o.SCALE(1))
    mine = np.sort(np.concatenate([np.repeat(v, o.tab.mult(I)) for I, (v, k) in o.iters[0].items()]))
    np.testing.assert_allclose(mine, ref, atol=1e-10, rtol=0)


@pytest.mark.parametrize("sym", ["QS", "QSZ"])
def test_fdm_spectrum_is_exact_lehmann_without_truncation(tmp_path, sym):
    Nmax = 3
    T = 0.05
    dg = _make(str(tmp_path), sym, 1, Nmax, T=T, fdm=True)
    orc = load_oracle()
    o = orc.Oracle(str(tmp_path)).run()
    C, H, cd, nd = _chain(dg, 1, Nmax, Nmax=Nmax)
    E, V = ed_chain.eig_sectors(C, H)
    om, w = ed_chain.lehmann(E, V, cd[("d", 1)], T)
    p = o.p
    spos = orc.Bins(p, o.SCALE(0), o.SCALE(Nmax))
    sneg = orc.Bins(p, o.SCALE(0), o.SCALE(Nmax))
    pos = om >= 0
    spos.add_many(om[pos], w[pos])
    sneg.add_many(-om[~pos], w[~pos])
    bp, bn = o.spectra[("FDM", "A_d", "A_d", 0 if sym == "QS" else 1)]
    np.testing.assert_allclose(bp.w, spos.w, atol=1e-12)
    np.testing.assert_allclose(bn.w, sneg.w, atol=1e-12)
    assert abs(bp.w.sum() + bn.w.sum() - 1.0) < 1e-10


@pytest.mark.parametrize("sym", ["QS", "QSZ"])
def test_fdm_expectation_value_is_exact_without_truncation(tmp_path, sym):
    """customfdm (calc_trace_fdm_kept at fdmexpvn = 0): with nothing
    discarded the FDM density matrix of shell 0 is the whole chain's thermal
    density matrix traced down to it, so <n_d> equals ED's Tr(rho n_d) at T."""
    Nmax = 3
    T = 0.07
    dg = _make(str(tmp_path), sym, 1, Nmax, T=T, fdm=True, fdmexpv=True, fdmexpvn=0)
    o = load_oracle().Oracle(str(tmp_path)).run()
    C, H, cd, nd = _chain(dg, 1, Nmax, Nmax=Nmax)
    E, V = ed_chain.eig_sectors(C, H)
    b = np.exp(-(E - E.min()) / T)
    ref = np.einsum("ij,jk,ki->", V.T, nd, V * b[None, :]) / b.sum()
    assert abs(o.fdmexpv_values()["n_d"] - ref) < 1e-11


def _kondo_case(tmp, nch, Nmax, J=0.4, Lambda=4.0):
    dg = load_nrgamd().datagen if hasattr(load_nrgamd(), "datagen") else __import__("nrgamd.datagen").datagen
    model = dg.kondo(sym="QSZ", nch=nch, J=J, Lambda=Lambda)
    model.write(os.path.join(tmp, "data"), Lambda, 1.0, Nmax)
    dg.write_param(os.path.join(tmp, "param"), Lambda=Lambda, keep=100000, T=0.3, Nmax=Nmax)
    return dg


@pytest.mark.parametrize("nch,Nmax", [(1, 3), (3, 1)])
def test_kondo_eigenvalues_match_ed(tmp_path, nch, Nmax):
    """The Kondo generator (models.m:382-394: J S_d . s_0 per channel, the
    local moment carrying no charge) and the QSZ tables: NRG eigenvalues equal
    ED of the whole chain, one channel up to 5 sites, three channels at the
    first iteration (8192 states, the three-channel configuration's first
    shell)."""
    J, Lambda = 0.4, 4.0
    dg = _kondo_case(str(tmp_path), nch, Nmax, J=J, Lambda=Lambda)
    o = load_oracle().Oracle(str(tmp_path)).run()
    xi, zeta = dg.chain_coefficients(Lambda, 1.0, Nmax)
    for N in range(Nmax):
        C, H = ed_chain.kondo_chain([J] * nch, [xi] * nch, [zeta] * nch, N + 1, nch)
        E, V = ed_chain.eig_sectors(C, H)
        ref = np.sort((E - E.min()) / o.SCALE(N + 1))
        mine = np.sort(np.concatenate([np.repeat(v, o.tab.mult(I)) for I, (v, k) in o.iters[N].items()]))
        assert len(mine) == len(ref)
        np.testing.assert_allclose(mine, ref, atol=1e-10, rtol=0)


def test_chain_coefficients_closed_form():
    dg = __import__("nrgamd.datagen").datagen
    for L, N in ((2.0, 50), (3.0, 40), (4.0, 30)):
        xi, zeta = dg.chain_coefficients(L, 1.0, N)
        for n in range(N + 1):
            assert abs(xi[n] / dg.wilson_xi_closed(n, L) - 1) < 1e-14
        assert max(abs(z) for z in zeta) < 1e-15
