"""Generator of NRG Ljubljana `data` files (format version 9).

The reference produces `data` with its Mathematica front end
(initial.m maketable 4309-4393, makeheader 4094-4101, makeenergies 4110-4122,
makeireducf 4234-4238, makedisctables 4241-4253); Mathematica is not
available, so this module builds the same file from an exact diagonalisation
of the initial cluster (impurity + site 0 of every channel) in Python:

* Hamiltonian: single-impurity Anderson impurity
  H_imp = delta n_d + U/2 (n_d - 1)^2, delta = eps_d + U/2 (initial.m:535),
  hybridisation V_ch sum_sigma (d^dag f_ch0 + h.c.) with
  V_ch = sqrt(theta0_ch * A_Lambda / pi), theta0_ch = 2 Gamma_ch for a flat
  band (initial.m:488, 2907, 3025, 1023), or the multi-channel Kondo coupling
  sum_ch J_ch S_d . s_ch0 with d singly occupied (models.m:382-394);
* energies (E - E_GS)/SCALE(0) per invariant subspace (initial.m:4117);
* reduced matrix elements <I1||f_ch0^dag||I2> and <I1||d^dag||I2> with
  <S1 S1|c^dag_up|S2 S1-1/2> = CG(S2 S1-1/2, 1/2 1/2|S1 S1) <I1||c^dag||I2>
  (initial.m ireducMatrixSpeedy, QS) or plain matrix elements of the creation
  operator whose spin is fixed by the SSZ difference (QSZ);
* Wilson-chain coefficients xi_n, zeta_n of a flat band discretised with
  Yoshida-Whitaker-Oliveira representative energies and twist z, obtained by a
  Lanczos tridiagonalisation in multiprecision (z=1 reproduces the closed
  form of scripts/wilsonxi:16-25).
"""
from fractions import Fraction as Fr
import math

import numpy as np

from .symtab import cg, HALF


# --------------------------------------------------------------------------
# Discretisation
# --------------------------------------------------------------------------

def scale(N, Lambda, z=1.0, bandrescale=1.0):
    """SCALE(N) for discretisation 'Y' (src/nrg.cc:481-497)."""
    return 0.5 * (1.0 + 1.0 / Lambda) * Lambda ** (-(N - 1) / 2.0 + 1.0 - z) * bandrescale


def wilson_xi_closed(n, Lambda):
    """Closed-form flat-band Wilson coefficient (scripts/wilsonxi:16-25)."""
    L = Lambda
    xin = (1.0 - L ** (-n - 1)) * (1.0 - L ** (-2 * n - 1)) ** -0.5 * (1.0 - L ** (-2 * n - 3)) ** -0.5
    return 0.5 * (1.0 + 1.0 / L) * L ** (-n / 2.0) * xin


def _lanczos_flat(Lambda, z, Nmax, M, dps):
    import mpmath as mp
    mp.mp.dps = dps
    L = mp.mpf(Lambda)
    zz = mp.mpf(z)
    eps = [mp.mpf(1)] + [L ** (-(zz + m - 1)) for m in range(1, M + 1)]
    E, w = [], []
    for m in range(M):
        e = (eps[m] + eps[m + 1]) / 2
        d = eps[m] - eps[m + 1]
        E += [e, -e]
        w += [d / 2, d / 2]
    v = [mp.sqrt(x) for x in w]
    nrm = mp.sqrt(sum(x * x for x in v))
    v = [x / nrm for x in v]
    vprev = [mp.mpf(0)] * len(v)
    xi, zeta = [], []
    beta = mp.mpf(0)
    for n in range(Nmax + 1):
        Hv = [E[i] * v[i] for i in range(len(v))]
        alpha = sum(v[i] * Hv[i] for i in range(len(v)))
        zeta.append(alpha)
        r = [Hv[i] - alpha * v[i] - beta * vprev[i] for i in range(len(v))]
        beta = mp.sqrt(sum(x * x for x in r))
        xi.append(beta)
        vprev, v = v, [x / beta for x in r]
    return [float(x) for x in xi], [float(x) for x in zeta]


def chain_coefficients(Lambda, z, Nmax):
    """xi_0..xi_Nmax and zeta_0..zeta_Nmax (flat band, Y discretisation).

    The star of 2M levels is truncated at Lambda^-M and the three-term
    recurrence cancels ~M*log10(Lambda) digits, so both M and the working
    precision are raised until two successive settings agree to 1e-15."""
    M = 3 * Nmax + 40
    dps = int(40 + 3 * M * math.log10(Lambda))
    prev = _lanczos_flat(Lambda, z, Nmax, M, dps)
    for _ in range(6):
        M, dps = int(M * 1.5), int(dps * 1.5)
        cur = _lanczos_flat(Lambda, z, Nmax, M, dps)
        if all(abs(a - b) <= 1e-15 * abs(b) for a, b in zip(prev[0], cur[0])) and \
           all(abs(a - b) <= 1e-15 * max(abs(b), 1e-300) + 1e-300 for a, b in zip(prev[1], cur[1])):
            return cur
        prev = cur
    raise RuntimeError("chain coefficients did not converge")


# --------------------------------------------------------------------------
# Fock space of the initial cluster
# --------------------------------------------------------------------------

class Fock:
    def __init__(self, modes):
        """modes: list of (orbital name, sigma) with sigma = +1/2 / -1/2."""
        self.modes = modes
        self.nm = len(modes)
        self.dim = 2 ** self.nm
        self.cdag = {}
        for k, md in enumerate(modes):
            m = np.zeros((self.dim, self.dim))
            for n in range(self.dim):
                if not (n >> k) & 1:
                    sgn = -1.0 if bin(n & ((1 << k) - 1)).count("1") % 2 else 1.0
                    m[n | (1 << k), n] = sgn
            self.cdag[md] = m

    def c(self, orb, s):
        return self.cdag[(orb, s)].T

    def cd(self, orb, s):
        return self.cdag[(orb, s)]

    def num(self, orb):
        return self.cd(orb, HALF) @ self.c(orb, HALF) + self.cd(orb, -HALF) @ self.c(orb, -HALF)

    def sz_of(self, n):
        s = Fr(0)
        for k, (orb, sg) in enumerate(self.modes):
            if (n >> k) & 1:
                s += sg
        return s

    def nel(self, n):
        return bin(n).count("1")

    def splus(self, orbs):
        sp = np.zeros((self.dim, self.dim))
        for o in orbs:
            sp += self.cd(o, HALF) @ self.c(o, -HALF)
        return sp


def spin_dot(F, a, b):
    """S_a . S_b for two spin-1/2 orbitals."""
    def S(o):
        sp = F.cd(o, HALF) @ F.c(o, -HALF)
        sm = sp.T
        sz = 0.5 * (F.cd(o, HALF) @ F.c(o, HALF) - F.cd(o, -HALF) @ F.c(o, -HALF))
        return sp, sm, sz
    pa, ma, za = S(a)
    pb, mb, zb = S(b)
    return 0.5 * (pa @ mb + ma @ pb) + za @ zb


class Model:
    """Initial cluster + operators; written as a `data` file."""

    def __init__(self, sym, nch, H, F, orbs, q_offset, allowed, ops_d, ops_s, GSenergy_shift=0.0):
        self.sym, self.nch, self.H, self.F = sym, nch, H, F
        self.orbs = orbs
        self.q_offset = q_offset
        self.allowed = allowed  # boolean mask of basis states kept (projection)
        self.ops_d = ops_d      # name -> orbital whose c^dag is the doublet operator
        self.ops_s = ops_s      # name -> matrix (singlet operators)

    # invariant subspaces at the highest-weight / fixed-Sz states
    def subspaces(self):
        F = self.F
        res = {}
        states = [n for n in range(F.dim) if self.allowed[n]]
        qs = sorted(set(F.nel(n) - self.q_offset for n in states))
        sp = F.splus(self.orbs)
        for q in qs:
            sec = [n for n in states if F.nel(n) - self.q_offset == q]
            szs = sorted(set(F.sz_of(n) for n in sec))
            for sz in szs:
                if self.sym == "QS" and sz < 0:
                    continue
                sel = [n for n in sec if F.sz_of(n) == sz]
                if self.sym == "QS":
                    A = sp[:, sel]
                    u, sv, vt = np.linalg.svd(A, full_matrices=True)
                    rank = int(np.sum(sv > 1e-10))
                    basis = vt[rank:].T
                    if basis.shape[1] == 0:
                        continue
                    inv = (q, int(2 * sz) + 1)
                else:
                    basis = np.eye(len(sel))
                    inv = (q, int(2 * sz))
                B = np.zeros((F.dim, basis.shape[1]))
                B[sel, :] = basis
                h = B.T @ self.H @ B
                h = 0.5 * (h + h.T)
                e, v = np.linalg.eigh(h)
                res[inv] = (e, B @ v, sz)
        return res

    def lower(self, vec):
        """Apply S- to a vector (for QS reduced matrix elements)."""
        return self.F.splus(self.orbs).T @ vec

    def reduced(self, subs, opmat_of_sigma, I1, I2):
        """<I1||c^dag||I2> block (size dim1 x dim2)."""
        e1, v1, sz1 = subs[I1]
        e2, v2, sz2 = subs[I2]
        if self.sym == "QS":
            S1 = Fr(I1[1] - 1, 2)
            S2 = Fr(I2[1] - 1, 2)
            sigma = HALF
            M2 = S1 - sigma
            vv2 = v2.copy()
            m = S2
            while m > M2:
                vv2 = self.lower(vv2) / math.sqrt(float(S2 * (S2 + 1) - m * (m - 1)))
                m -= 1
            norm = cg(S2, M2, HALF, sigma, S1, S1)
            return (v1.T @ opmat_of_sigma(sigma) @ vv2) / norm
        else:
            sigma = Fr(I1[1] - I2[1], 2)
            return v1.T @ opmat_of_sigma(sigma) @ v2

    def pairs_doublet(self, subs):
        out = []
        keys = sorted(subs)
        for I1 in keys:
            for I2 in keys:
                if I1[0] != I2[0] + 1:
                    continue
                if abs(I1[1] - I2[1]) != 1:
                    continue
                out.append((I1, I2))
        return out

    def write(self, path, Lambda, z, Nmax, xi=None, zeta=None):
        subs = self.subspaces()
        keys = sorted(subs)
        GS = min(subs[k][0][0] for k in keys)
        sc0 = scale(0, Lambda, z)
        if xi is None:
            xi, zeta = chain_coefficients(Lambda, z, Nmax)
        lines = []
        lines.append("# Input file for NRG Ljubljana")
        lines.append("# symtype %s" % self.sym)
        lines.append("# generated by nrgamd.datagen")
        lines.append("#!9")
        lines.append("# Number of channels, chain sites, subspaces: ")
        lines.append("%d %d %d" % (self.nch, Nmax, len(keys)))
        lines.append("# SCALE %.17g" % sc0)
        lines.append("# Energies (GS energy subtracted, multiplied by 1/SCALE):")
        for k in keys:
            e = subs[k][0]
            lines.append("%d %d" % k)
            lines.append("%d" % len(e))
            lines.append(" ".join("%.17g" % ((x - GS) / sc0) for x in e))
        lines.append("# Irreducible matrix elements for Wilson chains:")
        F = self.F
        pairs = self.pairs_doublet(subs)
        for ch in range(self.nch):
            lines.append("f %d 0" % ch)
            lines.append("%d" % len(pairs))
            for I1, I2 in pairs:
                m = self.reduced(subs, lambda s, ch=ch: F.cd(("f", ch), s), I1, I2)
                lines.append("%d %d %d %d" % (I1 + I2))
                for row in m:
                    lines.append(" ".join("%.17g" % x for x in row))
        lines.append("# GS energy in absolute units:")
        lines.append("e")
        lines.append("%.17g" % GS)
        lines.append("# Irreducible matrix elements for other operators:")
        for name, mat in self.ops_s.items():
            lines.append("s %s" % name)
            lines.append("%d" % len(keys))
            for I in keys:
                v = subs[I][1]
                m = v.T @ mat @ v
                lines.append("%d %d %d %d" % (I + I))
                for row in m:
                    lines.append(" ".join("%.17g" % x for x in row))
        for name, orb in self.ops_d.items():
            lines.append("d %s" % name)
            lines.append("%d" % len(pairs))
            for I1, I2 in pairs:
                m = self.reduced(subs, lambda s, orb=orb: F.cd(orb, s), I1, I2)
                lines.append("%d %d %d %d" % (I1 + I2))
                for row in m:
                    lines.append(" ".join("%.17g" % x for x in row))
        lines.append("# Discretization tables:")
        lines.append("z")
        for ch in range(self.nch):
            lines.append("%d" % Nmax)
            lines.extend("%.17g" % x for x in xi[:Nmax + 1])
        for ch in range(self.nch):
            lines.append("%d" % Nmax)
            lines.extend("%.17g" % x for x in zeta[:Nmax + 1])
        lines.append("")
        with open(path, "w") as f:
            f.write("\n".join(lines))
        return subs


def gamma_A(Lambda):
    """A_Lambda correction for discretisation Y (initial.m:1023)."""
    return 0.5 * (1 + 1 / Lambda) / (1 - 1 / Lambda) * math.log(Lambda)


def siam(sym="QS", nch=1, U=1.0, eps_d=-0.5, Gamma=0.04, Lambda=2.0, z=1.0, Nmax=30, lambda_correction=True):
    """Single Anderson impurity coupled to nch flat-band channels."""
    Gammas = list(Gamma) if isinstance(Gamma, (list, tuple)) else [Gamma] * nch
    modes = [("d", HALF), ("d", -HALF)]
    for ch in range(nch):
        modes += [(("f", ch), HALF), (("f", ch), -HALF)]
    F = Fock(modes)
    nd = F.num("d")
    delta = eps_d + U / 2
    I = np.eye(F.dim)
    H = delta * nd + U / 2 * (nd - I) @ (nd - I)
    A = gamma_A(Lambda) if lambda_correction else 1.0
    for ch in range(nch):
        V = math.sqrt(2 * Gammas[ch] * A / math.pi)
        for s in (HALF, -HALF):
            hop = F.cd("d", s) @ F.c(("f", ch), s)
            H += V * (hop + hop.T)
    orbs = ["d"] + [("f", ch) for ch in range(nch)]
    allowed = np.ones(F.dim, dtype=bool)
    model = Model(sym, nch, H, F, orbs, q_offset=1 + nch, allowed=allowed,
                  ops_d={"A_d": "d"}, ops_s={"n_d": nd})
    return model


def kondo(sym="QSZ", nch=3, J=0.4, Lambda=4.0, z=1.0, Nmax=30):
    """Spin-1/2 Kondo impurity coupled to nch channels (models.m:382-394)."""
    Js = list(J) if isinstance(J, (list, tuple)) else [J] * nch
    modes = [("d", HALF), ("d", -HALF)]
    for ch in range(nch):
        modes += [(("f", ch), HALF), (("f", ch), -HALF)]
    F = Fock(modes)
    H = np.zeros((F.dim, F.dim))
    for ch in range(nch):
        H += Js[ch] * spin_dot(F, ("f", ch), "d")
    allowed = np.array([((n & 1) + ((n >> 1) & 1)) == 1 for n in range(F.dim)])
    orbs = ["d"] + [("f", ch) for ch in range(nch)]
    # the local moment carries no charge: Q counts conduction electrons only
    model = Model(sym, nch, H, F, orbs, q_offset=1 + nch, allowed=allowed, ops_d={}, ops_s={})
    return model


def write_param(path, **kw):
    lines = ["[param]"]
    for k, v in kw.items():
        if isinstance(v, bool):
            v = "true" if v else "false"
        lines.append("%s=%s" % (k, v))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
