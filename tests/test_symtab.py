"""The engine's own coefficient tables against golden vectors of the reference's.

The engine derives its tables (nrgamd/symtab.py) from Clebsch-Gordan algebra;
the reference ships them precomputed (coefnew/qs/qs-1ch-offdiag.dat,
qs-1ch-spinupa.dat, qs-1ch-doubletp.dat, ...).  Both describe the same
operators in combination bases that may differ by one sign per combination,
so every entry must agree as ref[i,j] = g_i g_j * mine[i,j], with a single
gauge vector g for all tables, channels and spins.
"""
import json
import os
from fractions import Fraction as Fr

import pytest

from conftest import REPO, load_nrgamd

GOLD = os.path.join(REPO, "tests", "golden")


def _load(sym, nch):
    with open(os.path.join(GOLD, "tables_%s_%dch.json" % (sym.lower(), nch))) as f:
        return json.load(f)


def _ref_occ(g, nch):
    occ = [[None] * nch for _ in g["In"]]
    for r in g["diag"]:
        i, ch = r["idx"][0], r["idx"][1]
        occ[i - 1][ch] = int(round(r["values"][0]))
    return [tuple(o) for o in occ]


def _ref_occ_up(g, nch):
    if "diag_up" not in g:
        return [None] * len(g["In"])
    occ = [[None] * nch for _ in g["In"]]
    for r in g["diag_up"]:
        i, ch = r["idx"][0], r["idx"][2]
        occ[i - 1][ch] = int(round(r["values"][0]))
    return [tuple(o) for o in occ]


def _mapping(t, g, nch):
    """my comb index -> reference comb index (0-based)."""
    ref_occ = _ref_occ(g, nch)
    ref_up = _ref_occ_up(g, nch)
    ref_keys = [(tuple(g["In"][i]), tuple(g["QN"][i]), ref_occ[i], ref_up[i]) for i in range(len(g["In"]))]
    m = {}
    for a, c in enumerate(t.combs):
        key = (tuple(c["In"]), tuple(c["QN"]), tuple(c["occ"]), c.get("occ_up"))
        assert ref_keys.count(key) == 1, key
        m[a] = ref_keys.index(key)
    return m


def _collect(t, g, nch, sym):
    """List of (my_i, my_j, my_vals, ref_vals) to compare."""
    m = _mapping(t, g, nch)
    inv = {v: k for k, v in m.items()}
    pairs = []
    ref_off = {(r["idx"][0] - 1, r["idx"][1] - 1, r["idx"][2]): r["values"] for r in g["offdiag"]}
    mine_off = {(a, b, ch): vals for a, b, ch, vals in t.offdiag}
    for (a, b, ch), vals in mine_off.items():
        ra, rb = m[a], m[b]
        if (ra, rb, ch) in ref_off:
            pairs.append((a, b, vals, ref_off[(ra, rb, ch)], "offdiag"))
        else:
            # reference lists only i<j and takes the transpose for i>j
            pairs.append((a, b, vals, ref_off.get((rb, ra, ch)), "offdiag-T"))
    for ch in range(nch):
        for kind, key in (("up", "spinup"), ("down", "spindown")):
            ref = {(r["i1"] - 1, r["ip"] - 1): r["values"] for r in g[key][ch]}
            mine = {(a, b): vals for c, k, a, b, vals in t.recalcf if c == ch and k == kind}
            for (ra, rb), rv in ref.items():
                a, b = inv[ra], inv[rb]
                pairs.append((a, b, mine.get((a, b)), rv, key))
            for (a, b), vals in mine.items():
                assert (m[a], m[b]) in ref, ("extra recalc_f entry", ch, kind, a, b)
    for kind, key in (("p", "doubletp"), ("m", "doubletm")):
        ref = {(r["i1"] - 1, r["ip"] - 1): r["values"] for r in g[key]}
        mine = {(a, b): vals for k, a, b, vals in t.doublet if k == kind}
        for (ra, rb), rv in ref.items():
            a, b = inv[ra], inv[rb]
            pairs.append((a, b, mine.get((a, b)), rv, key))
        for (a, b), vals in mine.items():
            assert (m[a], m[b]) in ref, ("extra doublet entry", kind, a, b)
    return pairs


def _check_gauge(pairs, ncombs):
    """Find g with ref = g_a g_b mine on every defined entry."""
    rel = []
    for a, b, mine, ref, what in pairs:
        assert ref is not None, ("missing reference entry", a, b, what)
        for k in range(len(ref)):
            mv = None if mine is None else mine[k if len(mine) > 1 else 0]
            rv = ref[k]
            if mv is None or rv is None:
                continue
            if abs(mv) < 1e-12 and abs(rv) < 1e-12:
                continue
            assert abs(abs(mv) - abs(rv)) < 1e-12, (what, a, b, k + 1, mv, rv)
            rel.append((a, b, 1 if mv * rv > 0 else -1, what, k + 1))
    g = {0: 1}
    changed = True
    while changed:
        changed = False
        for a, b, s, what, ss in rel:
            if a in g and b not in g:
                g[b] = g[a] * s
                changed = True
            elif b in g and a not in g:
                g[a] = g[b] * s
                changed = True
    for a, b, s, what, ss in rel:
        assert g[a] * g[b] == s, ("gauge inconsistency", what, a, b, ss)
    return g


@pytest.mark.parametrize("sym,nch", [("QS", 1), ("QS", 2), ("QSZ", 1), ("QSZ", 3)])
def test_tables_match_reference_up_to_gauge(sym, nch):
    symtab = load_nrgamd().symtab
    t = symtab.generate(sym, nch, ssmax=24 if sym == "QS" else 64)
    g = _load(sym, nch)
    assert [tuple(x) for x in g["In"]] and len(g["In"]) == t.ncombs
    pairs = _collect(t, g, nch, sym)
    if sym == "QS":
        pairs = [(a, b, m, r[:24], w) for a, b, m, r, w in pairs]
    gauge = _check_gauge(pairs, t.ncombs)
    assert len(gauge) == t.ncombs


def test_clebsch_gordan_known_values():
    symtab = load_nrgamd().symtab
    h = Fr(1, 2)
    # <1/2 1/2, 1/2 -1/2 | 0 0> = 1/sqrt(2); <1 1, 1/2 -1/2 | 1/2 1/2> = sqrt(2/3)
    assert abs(symtab.cg(h, h, h, -h, Fr(0), Fr(0)) - 2 ** -0.5) < 1e-15
    assert abs(symtab.cg(Fr(1), Fr(1), h, -h, h, h) - (2 / 3) ** 0.5) < 1e-15
    assert abs(symtab.cg(Fr(1), Fr(0), h, h, h, h) + (1 / 3) ** 0.5) < 1e-15
    # orthogonality: sum_m1 CG^2 = 1 over coupled states
    j1, j2, J = Fr(3, 2), Fr(1), Fr(3, 2)
    tot = sum(symtab.cg(j1, m1, j2, J - m1, J, J) ** 2 for m1 in (Fr(3, 2), Fr(1, 2), Fr(-1, 2)))
    assert abs(tot - 1) < 1e-14
