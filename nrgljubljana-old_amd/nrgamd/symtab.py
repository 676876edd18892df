"""Derive the NRG coefficient tables for QS and QSZ symmetry from first principles.

The reference keeps these tables as precomputed expression fragments
(coefnew/qs/qs-<n>ch-{In2,QN,offdiag,diag,spinup*,spindown*,doublet*}.dat,
coefnew/qsz/qsz-<n>ch-*.dat) that are #included into
SymmetryQS::makematrix (src/sym-QS.cc:153-209), the RECALC_F_TAB /
RECALC_TAB macros (src/recalc-macros.m4:5-43) and the recalc drivers
(src/nrg-recalc-QS.cc.m4:17-103).  Here they are computed numerically
instead, for any number of channels, from:

* the Fock space of one new Wilson site (2*nch fermion modes, Jordan-Wigner
  ordered ch0-up, ch0-down, ch1-up, ...);
* its decomposition into multiplets of fixed charge and per-channel occupancy
  (QS: SU(2) multiplets built from highest-weight states with S-; QSZ: single
  Fock states);
* Clebsch-Gordan coupling |Q S M; a, r> = sum CG(S' M', s m | S M)
  c_site^dag(m) |Q' S' M'; r>, with site operators standing left of the
  block's operators, so block fermion operators pick up (-1)^{n_site};
* the reduced-matrix-element convention of the `data` file:
  <S1 M1|f^dag_sigma|S2 M2> = CG(S2 M2, 1/2 sigma | S1 M1) <S1||f^dag||S2>
  (initial.m ireducMatrixSpeedy for QS).

The combination basis can differ from the reference's by a sign (or, for
degenerate site multiplets, an orthogonal rotation) per combination; every
physical output (energies, thermodynamics, spectra) is invariant under that.

Output: a plain-text table file that the C++ host reads at start-up.
"""
from fractions import Fraction as Fr
from functools import lru_cache
from math import factorial, sqrt
import itertools
import numpy as np

HALF = Fr(1, 2)


@lru_cache(maxsize=None)
def cg(j1, m1, j2, m2, J, M):
    """Clebsch-Gordan <j1 m1, j2 m2 | J M> (Condon-Shortley), Racah's formula.
    Arguments are Fractions (integers or half-integers)."""
    if m1 + m2 != M:
        return 0.0
    if not (abs(j1 - j2) <= J <= j1 + j2):
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(M) > J:
        return 0.0
    for x in (j1 + m1, j2 + m2, J + M, j1 + j2 + J):
        if Fr(x).denominator != 1:
            return 0.0

    def I(x):
        x = Fr(x)
        assert x.denominator == 1
        return int(x)

    f = factorial
    pre = Fr((I(2 * J) + 1) * f(I(J + j1 - j2)) * f(I(J - j1 + j2)) * f(I(j1 + j2 - J)),
             f(I(j1 + j2 + J + 1)))
    pre *= f(I(J + M)) * f(I(J - M)) * f(I(j1 - m1)) * f(I(j1 + m1)) * f(I(j2 - m2)) * f(I(j2 + m2))
    s = Fr(0)
    for k in range(0, I(j1 + j2 - J) + 1):
        a = [k, I(j1 + j2 - J) - k, I(j1 - m1) - k, I(j2 + m2) - k,
             I(J - j2 + m1) + k, I(J - j1 - m2) + k]
        if min(a) < 0:
            continue
        d = 1
        for x in a:
            d *= f(x)
        s += Fr((-1) ** k, d)
    if s == 0:
        return 0.0
    val = sqrt(float(pre * s * s))
    return val if s > 0 else -val


class Site:
    """Fock space of one Wilson-chain site with nch spin-1/2 channels."""

    def __init__(self, nch):
        self.nch = nch
        self.nm = 2 * nch
        self.dim = 2 ** self.nm
        dim = self.dim
        self.cdag = []
        for k in range(self.nm):
            m = np.zeros((dim, dim))
            for n in range(dim):
                if not (n >> k) & 1:
                    sgn = -1.0 if bin(n & ((1 << k) - 1)).count("1") % 2 else 1.0
                    m[n | (1 << k), n] = sgn
            self.cdag.append(m)
        self.nocc = np.array([bin(n).count("1") for n in range(dim)])

    def mode(self, ch, sigma):
        """sigma = +1/2 (up) or -1/2 (down)."""
        return 2 * ch + (0 if sigma > 0 else 1)

    def occ(self, n):
        return tuple(((n >> (2 * c)) & 1) + ((n >> (2 * c + 1)) & 1) for c in range(self.nch))

    def sz(self, n):
        s = Fr(0)
        for c in range(self.nch):
            s += HALF * (((n >> (2 * c)) & 1) - ((n >> (2 * c + 1)) & 1))
        return s

    def splus(self):
        sp = np.zeros((self.dim, self.dim))
        for c in range(self.nch):
            sp += self.cdag[2 * c] @ self.cdag[2 * c + 1].T
        return sp


def _fix_sign(v):
    k = int(np.argmax(np.abs(v) > 1e-12))
    return v if v[k] > 0 else -v


def site_multiplets_su2(site):
    """List of (q, s, occ, {m: vector}) for SU(2) multiplets of fixed occupancy."""
    sp = site.splus()
    sm = sp.T
    occs = sorted(set(site.occ(n) for n in range(site.dim)))
    res = []
    for occ in occs:
        states = [n for n in range(site.dim) if site.occ(n) == occ]
        q = sum(occ) - site.nch
        smax = max(site.sz(n) for n in states)
        S = smax
        while S >= 0:
            sel = [n for n in states if site.sz(n) == S]
            if sel:
                A = sp[:, sel]
                u, sv, vt = np.linalg.svd(A)
                rank = int(np.sum(sv > 1e-10))
                null = vt[rank:].T  # columns: null vectors in 'sel' coordinates
                if null.shape[1] > 0:
                    # orthonormal, deterministic basis of the highest-weight space
                    q_, r_ = np.linalg.qr(null)
                    for k in range(q_.shape[1]):
                        v = np.zeros(site.dim)
                        v[sel] = q_[:, k]
                        v = _fix_sign(v)
                        mult = {S: v}
                        m = S
                        while m > -S:
                            w = sm @ mult[m]
                            nrm = sqrt(float(S * (S + 1) - m * (m - 1)))
                            mult[m - 1] = w / nrm
                            m -= 1
                        res.append((q, S, occ, mult))
            S -= 1
    return res


def site_states_u1(site):
    """QSZ: every Fock state is its own 'multiplet' (q, sz, occ, vector)."""
    res = []
    for occ in sorted(set(site.occ(n) for n in range(site.dim))):
        states = [n for n in range(site.dim) if site.occ(n) == occ]
        for n in sorted(states, key=lambda n: (-site.sz(n), n)):
            v = np.zeros(site.dim)
            v[n] = 1.0
            res.append((sum(occ) - site.nch, site.sz(n), occ, v))
    return res


class Tables:
    def __init__(self, sym, nch, ssmax=64):
        self.sym = sym
        self.nch = nch
        self.ssmax = ssmax
        self.site = Site(nch)
        self.combs = []  # dicts
        if sym == "QS":
            mults = site_multiplets_su2(self.site)
            mults.sort(key=lambda t: t[0])  # stable: charge ascending
            for mi, (q, s, occ, vecs) in enumerate(mults):
                d = -s
                while d <= s:
                    self.combs.append(dict(mu=mi, q=q, s=s, occ=occ, vecs=vecs, delta=d,
                                           In=(-q, int(2 * d)), QN=(q, int(2 * s) + 1)))
                    d += 1
        elif sym == "QSZ":
            sts = site_states_u1(self.site)
            sts.sort(key=lambda t: t[0])
            for mi, (q, sz, occ, v) in enumerate(sts):
                n = int(np.argmax(np.abs(v)))
                up = tuple((n >> (2 * c)) & 1 for c in range(nch))
                self.combs.append(dict(mu=mi, q=q, s=sz, occ=occ, occ_up=up, vecs={sz: v}, delta=None,
                                       In=(-q, int(-2 * sz)), QN=(q, int(2 * sz))))
        else:
            raise ValueError(sym)
        self.ncombs = len(self.combs)
        assert self.ncombs == 4 ** nch

    # ---- state construction ----------------------------------------------
    def state(self, a, S, M):
        """QS: |Q S M; a> as {(site_idx, M'): amp}; block spin S' = S + delta."""
        c = self.combs[a]
        Sp = S + c["delta"]
        out = {}
        for m, v in c["vecs"].items():
            Mp = M - m
            if abs(Mp) > Sp:
                continue
            w = cg(Sp, Mp, c["s"], m, S, M)
            if w == 0.0:
                continue
            for n in np.nonzero(np.abs(v) > 1e-14)[0]:
                out[(int(n), Mp)] = out.get((int(n), Mp), 0.0) + w * v[n]
        return out, Sp

    def valid(self, a, S):
        c = self.combs[a]
        Sp = S + c["delta"]
        return Sp >= 0 and abs(Sp - c["s"]) <= S <= Sp + c["s"]

    # ---- QS coefficients --------------------------------------------------
    def offdiag_qs(self, a, b, ch, S):
        """H block (a,b) gets c * xi * <anc_a||f^dag_ch||anc_b>."""
        ca, cb = self.combs[a], self.combs[b]
        if not (self.valid(a, S) and self.valid(b, S)):
            return None
        va, Sa = self.state(a, S, S)
        vb, Sb = self.state(b, S, S)
        tot = 0.0
        for sigma in (HALF, -HALF):
            cmat = self.site.cdag[self.site.mode(ch, sigma)].T  # annihilator
            for (n, Mb), amp in vb.items():
                col = cmat[:, n]
                for n2 in np.nonzero(np.abs(col) > 0)[0]:
                    Ma = Mb + sigma
                    if abs(Ma) > Sa:
                        continue
                    g = cg(Sb, Mb, HALF, sigma, Sa, Ma)
                    if g == 0.0:
                        continue
                    sgn = -1.0 if self.site.nocc[n2] % 2 else 1.0
                    tot += va.get((int(n2), Ma), 0.0) * amp * col[n2] * sgn * g
        return tot

    def recalcf_qs(self, a, b, ch, S1, Sp):
        """<I1 a||c^dag_ch||Ip b>, block parts identical (same ancestor)."""
        if not (self.valid(a, S1) and self.valid(b, Sp)):
            return None
        sigma = HALF
        Mp = S1 - sigma
        if abs(Mp) > Sp:
            return None
        norm = cg(Sp, Mp, HALF, sigma, S1, S1)
        if norm == 0.0:
            return None
        va, Sa = self.state(a, S1, S1)
        vb, Sb = self.state(b, Sp, Mp)
        if Sa != Sb:
            return 0.0
        cd = self.site.cdag[self.site.mode(ch, sigma)]
        tot = 0.0
        for (n, M), amp in vb.items():
            col = cd[:, n]
            for n2 in np.nonzero(np.abs(col) > 0)[0]:
                tot += va.get((int(n2), M), 0.0) * amp * col[n2]
        return tot / norm

    def doublet_qs(self, a, b, S1, Sp):
        """<I1 a||T||Ip b> for a fermionic doublet block operator T with unit
        reduced element between the two ancestors."""
        if not (self.valid(a, S1) and self.valid(b, Sp)):
            return None
        sigma = HALF
        Mp = S1 - sigma
        if abs(Mp) > Sp:
            return None
        norm = cg(Sp, Mp, HALF, sigma, S1, S1)
        if norm == 0.0:
            return None
        va, Sa = self.state(a, S1, S1)
        vb, Sb = self.state(b, Sp, Mp)
        tot = 0.0
        for (n, Mb), amp in vb.items():
            Ma = Mb + sigma
            if abs(Ma) > Sa:
                continue
            g = cg(Sb, Mb, HALF, sigma, Sa, Ma)
            if g == 0.0:
                continue
            sgn = -1.0 if self.site.nocc[n] % 2 else 1.0
            tot += va.get((n, Ma), 0.0) * amp * g * sgn
        return tot / norm

    # ---- QSZ coefficients (no spin coupling) ------------------------------
    def _vec(self, a):
        return next(iter(self.combs[a]["vecs"].values()))

    def offdiag_qsz(self, a, b, ch):
        ca, cb = self.combs[a], self.combs[b]
        va, vb = self._vec(a), self._vec(b)
        tot = 0.0
        for sigma in (HALF, -HALF):
            # block gains sigma: anc_a SSZ = anc_b SSZ + 2 sigma
            if ca["In"][1] + 0 != cb["In"][1] + 0:
                pass
            cmat = self.site.cdag[self.site.mode(ch, sigma)].T
            w = cmat @ vb
            sgn = np.where(self.site.nocc % 2 == 1, -1.0, 1.0)
            # spin bookkeeping: site loses sigma, block gains sigma
            if cb["s"] - sigma != ca["s"]:
                continue
            tot += float(va @ (w * sgn))
        return tot

    def recalcf_qsz(self, a, b, ch, sigma):
        va, vb = self._vec(a), self._vec(b)
        cd = self.site.cdag[self.site.mode(ch, sigma)]
        return float(va @ (cd @ vb))

    def doublet_qsz(self, a, b):
        va, vb = self._vec(a), self._vec(b)
        sgn = np.where(self.site.nocc % 2 == 1, -1.0, 1.0)
        return float(va @ (vb * sgn))

    # ---- table assembly ------------------------------------------------------
    def build(self):
        nc = self.ncombs
        ss_range = range(1, self.ssmax + 1)
        self.offdiag = []   # (a, b, ch, values)
        self.recalcf = []   # (ch, kind, a, b, values) kind 'up'/'down'
        self.doublet = []   # (kind, a, b, values) kind 'p'/'m'
        if self.sym == "QS":
            for a, b in itertools.product(range(nc), range(nc)):
                ca, cb = self.combs[a], self.combs[b]
                if ca["q"] != cb["q"] - 1:
                    continue  # ancestor a must carry one electron more than b
                for ch in range(self.nch):
                    vals = [self.offdiag_qs(a, b, ch, Fr(ss - 1, 2)) for ss in ss_range]
                    if any(v is not None and abs(v) > 1e-12 for v in vals):
                        self.offdiag.append((a, b, ch, vals))
            for ch in range(self.nch):
                for kind, dss in (("up", 1), ("down", -1)):
                    for a, b in itertools.product(range(nc), range(nc)):
                        ca, cb = self.combs[a], self.combs[b]
                        if ca["q"] != cb["q"] + 1:
                            continue
                        # same ancestor: I1 + In_a == Ip + In_b with I1 = Ip + (1, dss)
                        if (1 + ca["In"][0], dss + ca["In"][1]) != tuple(cb["In"]):
                            continue
                        vals = []
                        for ssp in ss_range:
                            Sp = Fr(ssp - 1, 2)
                            S1 = Sp + Fr(dss, 2)
                            vals.append(None if S1 < 0 else self.recalcf_qs(a, b, ch, S1, Sp))
                        if any(v is not None and abs(v) > 1e-12 for v in vals):
                            self.recalcf.append((ch, kind, a, b, vals))
            for kind, dss in (("p", 1), ("m", -1)):
                for a, b in itertools.product(range(nc), range(nc)):
                    ca, cb = self.combs[a], self.combs[b]
                    if ca["mu"] != cb["mu"]:
                        continue
                    # ancestors: IN1 = I1 + In_a, INp = Ip + In_b, Ip = I1 + (-1, dss)
                    dQ = ca["In"][0] - (-1 + cb["In"][0])
                    dS = ca["In"][1] - (dss + cb["In"][1])
                    if dQ != 1 or abs(dS) != 1:
                        continue
                    vals = []
                    for ss1 in ss_range:
                        S1 = Fr(ss1 - 1, 2)
                        Sp = S1 + Fr(dss, 2)
                        vals.append(None if Sp < 0 else self.doublet_qs(a, b, S1, Sp))
                    if any(v is not None and abs(v) > 1e-12 for v in vals):
                        self.doublet.append((kind, a, b, vals))
        else:
            for a, b in itertools.product(range(nc), range(nc)):
                ca, cb = self.combs[a], self.combs[b]
                if ca["q"] != cb["q"] - 1:
                    continue
                for ch in range(self.nch):
                    v = self.offdiag_qsz(a, b, ch)
                    if abs(v) > 1e-12:
                        self.offdiag.append((a, b, ch, [v]))
            for ch in range(self.nch):
                for kind, sigma in (("up", HALF), ("down", -HALF)):
                    dssz = int(2 * sigma)
                    for a, b in itertools.product(range(nc), range(nc)):
                        ca, cb = self.combs[a], self.combs[b]
                        if (1 + ca["In"][0], dssz + ca["In"][1]) != tuple(cb["In"]):
                            continue
                        v = self.recalcf_qsz(a, b, ch, sigma)
                        if abs(v) > 1e-12:
                            self.recalcf.append((ch, kind, a, b, [v]))
            for kind, dssz in (("p", 1), ("m", -1)):
                for a, b in itertools.product(range(nc), range(nc)):
                    ca, cb = self.combs[a], self.combs[b]
                    if ca["mu"] != cb["mu"]:
                        continue
                    dQ = ca["In"][0] - (-1 + cb["In"][0])
                    dS = ca["In"][1] - (dssz + cb["In"][1])
                    if dQ != 1 or dS != -dssz:
                        continue
                    v = self.doublet_qsz(a, b)
                    if abs(v) > 1e-12:
                        self.doublet.append((kind, a, b, [v]))
        return self

    def write(self, path):
        def fmt(vals):
            return " ".join("nan" if v is None else repr(float(v)) for v in vals)
        with open(path, "w") as f:
            f.write("NRGTAB 1\n")
            f.write("sym %s\nchannels %d\ncombs %d\nssmax %d\n"
                    % (self.sym, self.nch, self.ncombs, self.ssmax))
            for i, c in enumerate(self.combs):
                f.write("comb %d %d %d %d %d %s\n" % (i, c["In"][0], c["In"][1], c["QN"][0], c["QN"][1],
                                                   " ".join(str(x) for x in c["occ"])))
            for a, b, ch, vals in self.offdiag:
                f.write("offdiag %d %d %d %d %s\n" % (a, b, ch, len(vals), fmt(vals)))
            for ch, kind, a, b, vals in self.recalcf:
                f.write("recalcf %d %s %d %d %d %s\n" % (ch, kind, a, b, len(vals), fmt(vals)))
            for kind, a, b, vals in self.doublet:
                f.write("doublet %s %d %d %d %s\n" % (kind, a, b, len(vals), fmt(vals)))
            f.write("end\n")


def generate(sym, nch, ssmax=64):
    return Tables(sym, nch, ssmax).build()


if __name__ == "__main__":
    import sys
    out = sys.argv[1]
    for sym, nch in (("QS", 1), ("QS", 2), ("QSZ", 1), ("QSZ", 3)):
        t = generate(sym, nch)
        t.write("%s/%s-%dch.tab" % (out, sym.lower(), nch))
        print(sym, nch, len(t.offdiag), len(t.recalcf), len(t.doublet))
