"""Brute-force exact diagonalisation of a whole Wilson chain (test helper).

Independent of both the oracle and the engine: it builds the full Fock-space
Hamiltonian of impurity + chain sites 0..K from the same physical parameters
and diagonalises it in (Q, Sz) sectors with numpy.  Without truncation NRG is
exact, so its eigenvalues, thermodynamics and (for FDM, where only the last
shell carries weight) spectral weights must reproduce these numbers.
"""
import math

import numpy as np


class Chain:
    def __init__(self, nch, K):
        # modes: d-up, d-down, then site n of channel ch: up, down
        self.modes = [("d", 1), ("d", -1)]
        for n in range(K + 1):
            for ch in range(nch):
                self.modes += [((n, ch), 1), ((n, ch), -1)]
        self.nm = len(self.modes)
        self.dim = 2 ** self.nm
        self.index = {m: k for k, m in enumerate(self.modes)}
        n = np.arange(self.dim, dtype=np.int64)
        self.N = np.array([bin(x).count("1") for x in range(self.dim)])
        self.Sz2 = np.zeros(self.dim, dtype=np.int64)
        for k, (o, s) in enumerate(self.modes):
            self.Sz2 += s * ((n >> k) & 1)

    def cdag(self, mode):
        """Sparse action of c^dag as (target, source, sign) arrays."""
        k = self.index[mode]
        n = np.arange(self.dim, dtype=np.int64)
        free = ((n >> k) & 1) == 0
        src = n[free]
        sign = np.where(np.array([bin(int(x) & ((1 << k) - 1)).count("1") for x in src]) % 2, -1.0, 1.0)
        return src | (1 << k), src, sign

    def dense(self, mode):
        t, s, g = self.cdag(mode)
        m = np.zeros((self.dim, self.dim))
        m[t, s] = g
        return m


def siam_chain(U, eps_d, V, xi, zeta, K, nch=1):
    """Dense H of the SIAM on impurity + K+1 chain sites per channel."""
    C = Chain(nch, K)
    H = np.zeros((C.dim, C.dim))
    cd = {m: C.dense(m) for m in C.modes}
    nd = cd[("d", 1)] @ cd[("d", 1)].T + cd[("d", -1)] @ cd[("d", -1)].T
    I = np.eye(C.dim)
    H += (eps_d + U / 2) * nd + U / 2 * (nd - I) @ (nd - I)
    for ch in range(nch):
        for s in (1, -1):
            h = V[ch] * cd[("d", s)] @ cd[((0, ch), s)].T
            H += h + h.T
            for n in range(K):
                h = xi[ch][n] * cd[((n, ch), s)] @ cd[((n + 1, ch), s)].T
                H += h + h.T
        for n in range(K + 1):
            nn = cd[((n, ch), 1)] @ cd[((n, ch), 1)].T + cd[((n, ch), -1)] @ cd[((n, ch), -1)].T
            H += zeta[ch][n] * (nn - I)
    return C, H, cd, nd


def eig_sectors(C, H):
    """All eigenpairs, block diagonal in (N, Sz)."""
    vals, vecs, keys = [], [], []
    E = np.empty(C.dim)
    V = np.zeros((C.dim, C.dim))
    pos = 0
    order = []
    for Nn in sorted(set(C.N)):
        for sz in sorted(set(C.Sz2[C.N == Nn])):
            sel = np.nonzero((C.N == Nn) & (C.Sz2 == sz))[0]
            h = H[np.ix_(sel, sel)]
            w, v = np.linalg.eigh(h)
            E[pos:pos + len(sel)] = w
            V[sel, pos:pos + len(sel)] = v
            pos += len(sel)
    return E, V


def lehmann(E, V, op, T):
    """Peaks (omega, weight) of A(w) for <<c; c^dag>> with op = c^dag:
    weight |<m|op|n>|^2 (e^{-E_n/T} + e^{-E_m/T}) / Z at omega = E_m - E_n."""
    E0 = E.min()
    b = np.exp(-(E - E0) / T)
    Z = b.sum()
    M = V.T @ op @ V  # <m|op|n>
    W = M * M * (b[None, :] + b[:, None]) / Z
    Om = E[:, None] - E[None, :]
    nz = W > 0
    return Om[nz], W[nz]


def kondo_chain(J, xi, zeta, K, nch):
    """H of the spin-1/2 Kondo impurity (J_ch S_d . s_ch(0), written
    independently of the generator) on K+1 chain sites per channel, restricted
    to the states with one electron on the impurity orbital (sparse operators:
    three channels and two sites are 2^14 Fock states); returns an object
    with N, Sz2 of the selected states and the dense restricted H."""
    import scipy.sparse as sp
    C = Chain(nch, K)

    def cdag(mode):
        t, s, g = C.cdag(mode)
        return sp.csr_matrix((g, (t, s)), shape=(C.dim, C.dim))
    cd = {m: cdag(m) for m in C.modes}
    n = np.arange(C.dim, dtype=np.int64)
    nd = ((n >> C.index[("d", 1)]) & 1) + ((n >> C.index[("d", -1)]) & 1)
    sel = np.nonzero(nd == 1)[0]

    def spin(o):
        up, dn = cd[(o, 1)], cd[(o, -1)]
        return up @ dn.T, 0.5 * (up @ up.T - dn @ dn.T)  # S+ = c_up^dag c_dn, Sz

    I = sp.identity(C.dim, format="csr")
    H = sp.csr_matrix((C.dim, C.dim))
    spd, szd = spin("d")
    for ch in range(nch):
        sp0, sz0 = spin((0, ch))
        H = H + J[ch] * (0.5 * (spd @ sp0.T + spd.T @ sp0) + szd @ sz0)
        for s in (1, -1):
            for k in range(K):
                h = xi[ch][k] * cd[((k, ch), s)] @ cd[((k + 1, ch), s)].T
                H = H + h + h.T
        for k in range(K + 1):
            nn = cd[((k, ch), 1)] @ cd[((k, ch), 1)].T + cd[((k, ch), -1)] @ cd[((k, ch), -1)].T
            H = H + zeta[ch][k] * (nn - I)
    Hs = H.tocsr()[sel][:, sel].toarray()

    class Sub:
        pass
    R = Sub()
    R.dim = len(sel)
    R.N = C.N[sel]
    R.Sz2 = C.Sz2[sel]
    return R, Hs
