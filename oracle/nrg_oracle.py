"""CPU oracle: a plain numpy restatement of the reference NRG hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the `nrg` binary, the HIP
library, the package) imports or runs this file; only tests/, smoke() and the
cpu_baseline leg of bench.py use it, as the checker.

It restates, function by function, the reference's two-pass NRG driver for
QS and QSZ symmetry (file:line citations are relative to the reference tree):

  read_data / parse_datafile_header   src/nrg.cc:3293-3401, src/read-input.cc:49-144
  SCALE                               src/nrg.cc:481-497
  nrg_make_subspaces_list             src/nrg.cc:2415-2430
  nrg_determine_tasks / determine_ranges  src/nrg.cc:2753-2781, 1479-1487
  nrg_prepare_task_for_diag           src/nrg.cc:2456-2478, src/matrix.cc:86-192
  diagonalise (dsyev)                 src/diag.h:54-89, 297-353  (numpy eigh = LAPACK)
  find_groundstate / subtract         src/nrg.cc:1420-1455
  find_clusters / fix_splittings      src/splitting.cc:5-77
  highest_retained_energy / truncate  src/nrg.cc:1919-1998
  nrg_calculate_TD + Sym TD extras    src/nrg.cc:2333-2369, src/sym-QS.cc:106-123, src/sym-QSZ.cc:72-91
  split_in_blocks                     src/recalc.cc:124-155
  recalc_f / recalc_general           src/recalc.cc:12-97, 164-281
  recalc_singlet                      src/nrg.cc:2112-2142
  calc_trace_singlet (custom)         src/nrg.cc:946-971
  init_rho / calc_ZnD / init_rho_FDM  src/nrg.cc:1877-1913, 2002-2041; src/dmnrg.h:487-525
  cdmI / DM and FDM sweeps            src/dmnrg.h:327-596
  SPEC_DMNRG::calc, SPEC_FDM::calc    src/spec_DMNRG.cc:13-52, src/spec_FDM.cc:41-119
  calc_generic                        src/spec.cc:26-62
  Bins / merge / trim / savebins      src/bins.h:50-190, src/spectrumrealfreq.cc:32-263
  continuous (broadening)             src/spectrumrealfreq.cc:271-306, src/spectral.h:56-102

Coefficient tables come from tests/golden/tables_*.json: the reference's own
table entries (coefnew/...) evaluated numerically.

Deliberate difference: at the last iteration of the DMNRG pass the reference
reuses whatever density matrix was loaded last (src/nrg.cc:2387-2393 skips the
load there); the oracle uses the density matrix built by init_rho for that
shell.  With NN2even and an even Nmax the last shell is not merged
(spectrumrealfreq.cc:164-185), so the outputs coincide.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(os.path.dirname(HERE), "tests", "golden")

DEFAULTS = dict(
    Lambda=1.0, z=1.0, discretization="Y", bandrescale=1.0, keep=100, keepenergy=-1.0, keepmin=0,
    T=-999.0, betabar=1.0, strategy="kept", dmnrg=False, fdm=False, dm=False, specd="", bins=100,
    savebins=False, broaden=True, goodE=2.0, NN2even=True, NN2avg=False, NN1=False, NNtanh=0.0,
    safeguard=1e-5, safeguardmax=200, fixeps=1e-15, lastall=False, lastalloverride=False,
    emin=-1.0, emax=-1.0, discard_trim=1e-16, discard_immediately=1e-16, alpha=0.3, omega0=-1.0,
    omega0_ratio=1.0, broaden_max=2.5, broaden_min=-99.0, broaden_min_ratio=3.0, broaden_ratio=1.05,
    width_td=16, prec_td=10, width_custom=16, prec_custom=10, prec_xy=10, Nmax=0, Ninit=0,
    calc0=True, tdht=0, dumpannotated=0, grouptol=1e-6, dumpsubspaces=False, fdmexpv=False, fdmexpvn=0,
)


def parse_param(path):
    """[param] block of keyword=value lines (src/misc.h:71-124)."""
    p = dict(DEFAULTS)
    inblock = False
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line or line[0] in "#;":
                continue
            if line.startswith("["):
                inblock = line == "[param]"
                continue
            if inblock and "=" in line:
                k, v = line.split("=", 1)
                k, v = k.strip(), v.strip()
                d = DEFAULTS.get(k)
                if isinstance(d, bool):
                    v = v.lower() in ("true", "1", "yes")
                elif isinstance(d, int):
                    v = int(v)
                elif isinstance(d, float):
                    v = float(v)
                p[k] = v
    # validateparameters (src/read-input.cc:6-36)
    cfs = p["fdm"]
    if p["dmnrg"] or cfs:
        p["dm"] = True
    if cfs and not p["lastalloverride"]:
        p["lastall"] = True
    return p


class Data:
    """`data` file, version 9 (src/nrg.cc:3293-3401, src/read-input.cc:98-144)."""

    def __init__(self, path):
        with open(path) as f:
            lines = f.read().split("\n")
        toks = []
        self.sym = None
        for l in lines:
            s = l.strip()
            if s.startswith("#"):
                if "symtype" in s:
                    self.sym = s.split()[-1]
                continue
            toks.append(s)
        it = iter(" ".join(toks).split())
        nxt = lambda: next(it)
        self.channels = int(nxt())
        self.Nmax = int(nxt())
        nsubs = int(nxt())
        self.energies = {}
        for _ in range(nsubs):
            I = (int(nxt()), int(nxt()))
            n = int(nxt())
            self.energies[I] = np.array([float(nxt()) for _ in range(n)])
        self.f = []
        for ch in range(self.channels):
            tag, i, j = nxt(), int(nxt()), int(nxt())
            assert tag == "f" and i == ch and j == 0
            self.f.append(self._matels(nxt))
        self.ops = {"s": {}, "d": {}}
        self.GS = 0.0
        self.xi = []
        self.zeta = []
        while True:
            try:
                tag = nxt()
            except StopIteration:
                break
            if tag == "e":
                self.GS = float(nxt())
            elif tag in ("s", "d", "p", "g", "t", "o", "q"):
                name = nxt()
                self.ops.setdefault(tag, {})[name] = self._matels(nxt)
            elif tag == "z":
                for tab in (self.xi, self.zeta):
                    for ch in range(self.channels):
                        n = int(nxt())
                        tab.append(np.array([float(nxt()) for _ in range(n + 1)]))
            else:
                raise ValueError("unknown block %s" % tag)

    def _matels(self, nxt):
        nf = int(nxt())
        m = {}
        for _ in range(nf):
            I1 = (int(nxt()), int(nxt()))
            I2 = (int(nxt()), int(nxt()))
            n1, n2 = len(self.energies[I1]), len(self.energies[I2])
            m[(I1, I2)] = np.array([float(nxt()) for _ in range(n1 * n2)]).reshape(n1, n2)
        return m


class Tables:
    """Reference coefficient tables (golden vectors of coefnew/<sym>/...)."""

    def __init__(self, sym, nch):
        with open(os.path.join(GOLD, "tables_%s_%dch.json" % (sym.lower(), nch))) as f:
            g = json.load(f)
        self.sym = sym
        self.In = [tuple(x) for x in g["In"]]
        self.QN = [tuple(x) for x in g["QN"]]
        self.combs = len(self.In)
        self.offdiag = [(r["idx"][0] - 1, r["idx"][1] - 1, r["idx"][2], r["values"]) for r in g["offdiag"]]
        self.diag = [(r["idx"][0] - 1, r["idx"][1], r["values"][0]) for r in g["diag"]]
        self.spinup = [[(r["i1"] - 1, r["ip"] - 1, r["values"]) for r in t] for t in g["spinup"]]
        self.spindown = [[(r["i1"] - 1, r["ip"] - 1, r["values"]) for r in t] for t in g["spindown"]]
        self.doubletp = [(r["i1"] - 1, r["ip"] - 1, tuple(r["IN1"]), tuple(r["INp"]), r["values"]) for r in g["doubletp"]]
        self.doubletm = [(r["i1"] - 1, r["ip"] - 1, tuple(r["IN1"]), tuple(r["INp"]), r["values"]) for r in g["doubletm"]]

    @staticmethod
    def at(values, ss):
        """Table values are stored for SS (or SSZ) = 1..; QSZ tables are constant."""
        if ss < 1:
            v = values[0]
        else:
            v = values[ss - 1]
        return v

    # symmetry plug-in (src/sym-QS.cc, src/sym-QSZ.cc)
    def mult(self, I):
        return I[1] if self.sym == "QS" else 1

    def allowed(self, I):
        return I[1] > 0 if self.sym == "QS" else True

    def triangle(self, I1, I2, I3):
        if I1[0] != I2[0] + I3[0]:
            return False
        if self.sym == "QS":
            a, b, c = I1[1], I2[1], I3[1]
            return abs(a - b) <= c - 1 and abs(b - c) <= a - 1 and abs(c - a) <= b - 1
        return I1[1] == I2[1] + I3[1]

    def specdens_factor(self, Ip, I1):
        if self.sym == "QS":
            ssp, ss1 = Ip[1], I1[1]
            Sp = (ssp - 1) / 2.0
            return Sp + 1.0 if ss1 == ssp + 1 else Sp
        return 1.0


def add(a, b):
    return (a[0] + b[0], a[1] + b[1])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1])


class Bins:
    """Logarithmic coarse-graining grid (src/bins.h:50-131)."""

    def __init__(self, p, scale0, scale_last):
        self.bpd = p["bins"]
        emax = p["emax"] if p["emax"] > 0 else scale0 * 10 ** 2.0
        emin = p["emin"] if p["emin"] > 0 else scale_last / 10 ** 2.0
        self.l10min = math.floor(math.log10(emin))
        self.l10max = math.ceil(math.log10(emax))
        self.nr = int((self.l10max - self.l10min) * self.bpd + 1.0)
        self.e = np.array([10.0 ** (self.l10min + i / self.bpd) for i in range(self.nr)])
        self.w = np.zeros(self.nr)
        self.discard = p["discard_immediately"]

    def add_many(self, E, W):
        """Bins::add + add_std for arrays of positive energies."""
        keep = ~(np.abs(W) < self.discard * E)
        E, W = E[keep], W[keep]
        if E.size == 0:
            return
        with np.errstate(divide="ignore"):
            pos = (np.log10(E) - self.l10min) * self.bpd
        # truncation toward zero, like long(pos); E == 0 gives -inf -> first bin
        idx = np.where(pos > -9e18, np.trunc(np.where(pos > -9e18, pos, 0.0)), -1).astype(np.int64)
        low = idx < 0
        high = idx >= self.nr - 1
        mid = ~(low | high)
        np.add.at(self.w, 0, W[low].sum()) if low.any() else None
        if high.any():
            self.w[self.nr - 1] += W[high].sum()
        rem = pos[mid] - idx[mid]
        np.add.at(self.w, idx[mid], (1.0 - rem) * W[mid])
        np.add.at(self.w, idx[mid] + 1, rem * W[mid])


class Oracle:
    def __init__(self, workdir, verbose=False):
        self.wd = workdir
        self.p = parse_param(os.path.join(workdir, "param"))
        self.verbose = verbose

    # ------------------------------------------------------------------ misc
    def SCALE(self, N):
        p = self.p
        s = 0.5 * (1.0 + 1.0 / p["Lambda"])
        s *= p["Lambda"] ** (-(N - 1) / 2.0 + 1.0 - p["z"])
        return s * p["bandrescale"]

    def load(self):
        d = Data(os.path.join(self.wd, "data"))
        self.data = d
        self.sym = d.sym
        self.nch = d.channels
        self.tab = Tables(d.sym, d.channels)
        if self.p["Nmax"] == 0:
            self.p["Nmax"] = d.Nmax
        assert len(d.xi[0]) - 1 == self.p["Nmax"]
        self.Nmax = self.p["Nmax"]
        self.combs = 4 ** self.nch
        self.Nlen = self.Nmax

    def fresh_state(self):
        d = self.data
        diagprev = {}
        for I, e in d.energies.items():
            diagprev[I] = dict(value=e.copy(), shift=0.0, absenergy=np.zeros(len(e)))
        f = [dict(m) for m in d.f]
        ops_s = {k: dict(v) for k, v in d.ops["s"].items()}
        ops_d = {k: dict(v) for k, v in d.ops["d"].items()}
        return diagprev, f, ops_s, ops_d

    # ------------------------------------------------------------- assembly
    def tasks(self, diagprev):
        T = self.tab
        subs = set()
        for Ip, e in diagprev.items():
            if len(e["value"]) == 0:
                continue
            for i in range(self.combs):
                I = sub(Ip, T.In[i])
                if T.allowed(I):
                    subs.add(I)
        res = []
        for I in sorted(subs):
            anc = [add(I, T.In[i]) for i in range(self.combs)]
            rmax = []
            for i in range(self.combs):
                ok = T.triangle(I, anc[i], T.QN[i])
                rmax.append(len(diagprev[anc[i]]["value"]) if (ok and anc[i] in diagprev) else 0)
            if sum(rmax):
                res.append((I, anc, rmax))
        return res

    def hamiltonian(self, N, I, anc, rmax, diagprev, f):
        T = self.tab
        p = self.p
        sc = self.SCALE(N + 1)
        off = np.concatenate([[0], np.cumsum(rmax)]).astype(int)
        d = off[-1]
        h = np.zeros((d, d))
        sf = math.sqrt(p["Lambda"])
        for i in range(self.combs):
            if rmax[i]:
                h[off[i]:off[i + 1], off[i]:off[i + 1]] += np.diag(sf * diagprev[anc[i]]["value"])
        ss = I[1]
        for i, j, ch, vals in T.offdiag:
            if not (rmax[i] and rmax[j]):
                continue
            fac = Tables.at(vals, ss)
            assert fac is not None and math.isfinite(fac)
            fs = fac * self.data.xi[ch][N] / sc
            if abs(fs) < 1e-8:
                continue
            m = f[ch][(anc[i], anc[j])]
            h[off[i]:off[i + 1], off[j]:off[j + 1]] += fs * m
            h[off[j]:off[j + 1], off[i]:off[i + 1]] += fs * m.T
        for i, ch, number in T.diag:
            if rmax[i]:
                shift = self.data.zeta[ch][N + 1] * (number - 1.0) / sc
                h[off[i]:off[i + 1], off[i]:off[i + 1]] += shift * np.eye(rmax[i])
        return h, off

    # ------------------------------------------------------------ truncation
    def find_clusters(self, energies):
        """find_clusters (src/splitting.cc:44-77) over ascending energies: a
        cluster runs while values stay less than fixeps above its first value;
        when a later value closes it and its members are not all equal
        (cluster_splitting), every member unlike the first maps to the first.
        The last cluster is never closed.  Entries go into the run-wide
        STAT::cluster_mapping (src/nrg.cc:561), an existing key is kept
        (std::map::insert)."""
        eps = self.p["fixeps"]
        mapping = self.__dict__.setdefault("cluster_mapping", {})
        e0 = energies[0]
        i0 = 0
        for i in range(len(energies)):
            if (energies[i] - e0) < eps:
                continue
            cl = energies[i0:i]
            for x in cl[1:]:
                if x != cl[0]:
                    mapping.setdefault(float(x), float(cl[0]))
            e0 = energies[i]
            i0 = i

    def fix_splittings(self, diag):
        """The sequence of src/nrg.cc:2910-2918: clusters of the sorted
        energies into the mapping, every value found in the mapping replaced
        (fix_splittings, src/splitting.cc:5-17), and after a replacement the
        clusters of the new values added too."""
        self.find_clusters(np.sort(np.concatenate([e["value"] for e in diag.values()])))
        mapping = self.cluster_mapping
        changed = False
        for e in diag.values():
            new = np.array([mapping.get(float(x), x) for x in e["value"]]) if mapping else e["value"]
            if mapping and np.any(new != e["value"]):
                changed = True
                e["value"] = new
        if changed:
            self.find_clusters(np.sort(np.concatenate([e["value"] for e in diag.values()])))

    def truncation(self, diag, last, dmnrgrun):
        p = self.p
        energies = np.sort(np.concatenate([e["value"] for e in diag.values()]))
        total = len(energies)
        if p["keepenergy"] <= 0.0:
            nrkeep = p["keep"]
        else:
            nrkeep = 1 + int(np.sum(energies <= p["keepenergy"]))
            nrkeep = min(max(nrkeep, p["keepmin"]), p["keep"])
        if p["safeguard"] > 0.0:
            extra = 0
            while nrkeep < total and energies[nrkeep] - energies[nrkeep - 1] <= p["safeguard"] \
                    and extra < p["safeguardmax"]:
                nrkeep += 1
                extra += 1
        nrkeep = min(max(nrkeep, 1), total)
        Emax = energies[nrkeep - 1]
        kept = {}
        for I, e in diag.items():
            c = int(np.sum(e["value"] <= Emax))
            if p["fdm"] and last and not p["lastalloverride"]:
                c = len(e["value"])
            if dmnrgrun and last and p["lastall"]:
                c = len(e["value"])
            kept[I] = c
        return kept

    # ------------------------------------------------------------------ TD
    def td(self, diag, scale_now, additional=1.0):
        T = self.tab
        rf = self.p["betabar"] * additional
        Z = E = E2 = 0.0
        trA = trB = trC = 0.0
        for I, e in diag.items():
            be = rf * e["value"]
            ex = np.exp(-be)
            m = T.mult(I)
            sZ = ex.sum()
            Z += m * sZ
            E += m * (be * ex).sum()
            E2 += m * (be * be * ex).sum()
            sumZ = m * sZ
            if self.sym == "QS":
                trA += sumZ * (I[1] * I[1] - 1) / 12.0
                trB += sumZ * I[0]
                trC += sumZ * I[0] * I[0]
            else:
                trA += sumZ * (I[1] / 2.0)      # <Sz>
                trB += sumZ * (I[1] / 2.0) ** 2  # <Sz^2>
                trC += sumZ * I[0]
        Em, E2m = E / Z, E2 / Z
        row = dict(T=scale_now / self.p["betabar"], E=Em, E2=E2m, C=E2m - Em * Em, F=-math.log(Z), S=Em + math.log(Z))
        if self.sym == "QS":
            row.update({"<Sz^2>": trA / Z, "<Q>": trB / Z, "<Q^2>": trC / Z})
        else:
            row.update({"<Sz>": trA / Z, "<Sz^2>": trB / Z, "<Q>": trC / Z})
        return row, Z

    # --------------------------------------------------------------- recalc
    def recalc_f(self, diag, rmaxs, ancs):
        T = self.tab
        newf = [dict() for _ in range(self.nch)]
        for Ip in diag:
            qp, ssp = Ip
            for I1, tabs in (((qp + 1, ssp + 1), T.spinup), ((qp + 1, ssp - 1), T.spindown)):
                if I1 not in diag:
                    continue
                for ch in range(self.nch):
                    dim1, dimp = diag[I1]["nr"], diag[Ip]["nr"]
                    if dim1 == 0 or dimp == 0:
                        continue
                    F = np.zeros((dim1, dimp))
                    for i1, ip, vals in tabs[ch]:
                        r1, rp = rmaxs[I1][i1], rmaxs[Ip][ip]
                        if not (r1 > 0 and rp > 0):
                            continue
                        assert ancs[I1][i1] == ancs[Ip][ip]
                        fac = Tables.at(vals, ssp)
                        F += fac * diag[I1]["blocks"][i1] @ diag[Ip]["blocks"][ip].T
                    newf[ch][(I1, Ip)] = F
        return newf

    def recalc_general(self, diag, rmaxs, ancs, cold, I1, Ip, table, ss_eval, Iop):
        T = self.tab
        if not T.triangle(I1, Ip, Iop):
            return None
        dim1, dimp = diag[I1]["nr"], diag[Ip]["nr"]
        cn = np.zeros((dim1, dimp))
        if dim1 == 0 or dimp == 0:
            return cn
        for i1, ip, dIN1, dINp, vals in table:
            if dIN1 is None:  # singlet table: same combination, same ancestor
                IN1, INp = ancs[I1][i1], ancs[Ip][ip]
            else:  # doublet tables give ancestors relative to (q1, ss1)
                IN1 = (I1[0] + dIN1[0], ss_eval[0] + dIN1[1])
                INp = (I1[0] + dINp[0], ss_eval[0] + dINp[1])
            if not (T.allowed(IN1) and T.allowed(INp)):
                continue
            r1, rp = rmaxs[I1][i1], rmaxs[Ip][ip]
            if r1 == 0 or rp == 0:
                continue
            assert ancs[I1][i1] == IN1 and ancs[Ip][ip] == INp, (I1, Ip, i1, ip, IN1, INp)
            cnt = 1 if (IN1, INp) in cold else 0
            tri = T.triangle(IN1, Iop, INp)
            if not tri and cnt == 0:
                continue
            assert cnt == 1, ("recalc bug trap", I1, Ip, IN1, INp)
            fac = 1.0 if vals is None else Tables.at(vals, ss_eval[1])
            assert math.isfinite(fac) and abs(fac) < 1000
            m = cold[(IN1, INp)]
            U1 = diag[I1]["blocks"][i1]
            Up = diag[Ip]["blocks"][ip]
            cn += fac * (U1 @ (m @ Up.T))
        return cn

    def recalc_doublet(self, diag, rmaxs, ancs, cold):
        T = self.tab
        new = {}
        for I1 in diag:
            q1, ss1 = I1
            for Ip, table in (((q1 - 1, ss1 + 1), T.doubletp), ((q1 - 1, ss1 - 1), T.doubletm)):
                if Ip not in diag:
                    continue
                r = self.recalc_general(diag, rmaxs, ancs, cold, I1, Ip, table, (ss1, ss1),
                                        (1, 2) if self.sym == "QS" else (1, ss1 - Ip[1]))
                if r is not None:
                    new[(I1, Ip)] = r
        return new

    def recalc_singlet(self, diag, rmaxs, ancs, cold):
        new = {}
        for I1 in diag:
            table = [(i, i, None, None, None) for i in range(self.combs)]
            r = self.recalc_general(diag, rmaxs, ancs, cold, I1, I1, table, (I1[1], I1[1]),
                                    (0, 1) if self.sym == "QS" else (0, 0))
            if r is not None:
                new[(I1, I1)] = r
        return new

    # ------------------------------------------------------------ NRG pass 1
    def run_pass1(self):
        p = self.p
        T = self.tab
        diagprev, f, ops_s, ops_d = self.fresh_state()
        self.td_rows = []
        self.custom_rows = []
        self.iters = []  # per N: dict(I -> (values_all, kept))
        self.store = {}  # per N: I -> full eigen data for DM and pass 2
        self.dm = {}     # per N: I -> (kept, total, rmax, absenergy)
        totalenergy = self.data.GS
        # calc0 (src/nrg.cc:3061-3076)
        if p["calc0"]:
            sc = self.SCALE(0)
            row, Z = self.td(diagprev, sc)
            self.td_rows.append(row)
            self.custom_rows.append((sc / p["betabar"], self.expv(diagprev, ops_s)))
        for N in range(p["Ninit"], self.Nmax):
            last = (N + 1 == self.Nmax)
            sc = self.SCALE(N + 1)
            tasks = self.tasks(diagprev)
            diag = {}
            rmaxs, ancs, offs = {}, {}, {}
            for I, anc, rmax in tasks:
                h, off = self.hamiltonian(N, I, anc, rmax, diagprev, f)
                w, v = np.linalg.eigh(h)
                diag[I] = dict(value=w, U=v.T.copy(), shift=0.0)
                rmaxs[I], ancs[I], offs[I] = rmax, anc, off
            Egs = min(e["value"][0] + e["shift"] for e in diag.values())
            for e in diag.values():
                e["value"] = e["value"] - Egs
                e["shift"] = Egs
            self.fix_splittings(diag)
            kept = self.truncation(diag, last, False)
            totalenergy += Egs * sc
            for I, e in diag.items():
                e["absenergy"] = totalenergy + e["value"] * sc
            # store (pre-truncation) for the DM sweeps and the DMNRG pass
            self.store[N] = {I: dict(value=e["value"].copy(), absenergy=e["absenergy"].copy(), U=e["U"],
                                     shift=Egs, rmax=rmaxs[I], anc=ancs[I], off=offs[I]) for I, e in diag.items()}
            self.iters.append({I: (e["value"].copy(), kept[I]) for I, e in diag.items()})
            row, Z = self.td(diag, sc)
            self.td_rows.append(row)
            # split + truncate
            for I, e in diag.items():
                k = kept[I]
                off = offs[I]
                e["blocks"] = [e["U"][:k, off[i]:off[i + 1]] for i in range(self.combs)]
                e["value"] = e["value"][:k]
                e["nr"] = k
            self.dm[N] = {I: dict(kept=kept[I], total=len(self.store[N][I]["value"]), rmax=rmaxs[I],
                                  absenergy=self.store[N][I]["absenergy"]) for I in diag}
            if not last:
                f = self.recalc_f(diag, rmaxs, ancs)
            ops_s = {k: self.recalc_singlet(diag, rmaxs, ancs, v) for k, v in ops_s.items()}
            if N == int(p["fdmexpvn"]):
                self.ops_s_fdmexpv = ops_s
            self.custom_rows.append((sc / p["betabar"], self.expv(diag, ops_s)))
            diagprev = {I: dict(value=e["value"], shift=e["shift"]) for I, e in diag.items()}
        self.GSenergy = totalenergy
        self.last_diag = diagprev
        self.scale_last = self.SCALE(self.Nmax)

    def expv(self, diag, ops_s):
        T = self.tab
        Z = 0.0
        for I, e in diag.items():
            Z += T.mult(I) * np.exp(-self.p["betabar"] * e["value"]).sum()
        out = {}
        for name, op in ops_s.items():
            tr = 0.0
            for I, e in diag.items():
                m = op[(I, I)]
                n = len(e["value"])
                tr += T.mult(I) * np.sum(np.exp(-self.p["betabar"] * e["value"]) * np.diag(m)[:n])
            out[name] = tr / Z
        return out

    # ----------------------------------------------------------- DM sweeps
    def cdmI(self, i, I1, rhoN, st, rhoNEW, coef, dmN):
        dim = rhoNEW.shape[0]
        nrom = rhoN.shape[0]
        if nrom == 0 or dim == 0:
            return
        rmax = dmN[I1]["rmax"][i]
        if rmax == 0:
            return
        assert rmax == dim
        off = st["off"]
        U = st["U"][:nrom, off[i]:off[i] + dim]
        Tm = U.T @ rhoN
        rhoNEW += coef * (Tm @ U)

    def init_rho(self):
        p = self.p
        diag = self.last_diag
        T = self.tab
        ZZ = sum(T.mult(I) * np.exp(-e["value"] * self.scale_last / p["T"]).sum() for I, e in diag.items())
        rho = {I: np.diag(np.exp(-e["value"] * self.scale_last / p["T"]) / ZZ) for I, e in diag.items()}
        return rho

    def dm_sweep(self, rho):
        T = self.tab
        rhos = {self.Nmax - 1: rho}
        for N in range(self.Nmax - 1, self.p["Ninit"], -1):
            st = self.store[N]
            prev = {}
            for I, info in self.dm[N - 1].items():
                dim = info["kept"]
                r = np.zeros((dim, dim))
                if dim:
                    for i in range(self.combs):
                        I1 = sub(I, T.In[i])
                        coef = T.mult(I1) / T.mult(I)
                        if I1 in rho and I1 in st:
                            self.cdmI(i, I1, rho[I1], st[I1], r, coef, self.dm[N])
                prev[I] = r
            tr = sum(T.mult(I) * np.trace(m) for I, m in prev.items())
            assert abs(tr - 1.0) < 1e-8, tr
            rhos[N - 1] = prev
            rho = prev
        return rhos

    def calc_ZnD(self):
        T = self.tab
        p = self.p
        self.ZnD = {}
        for N in range(p["Ninit"], self.Nlen):
            ZZ = np.longdouble(0)
            last = (N + 1 == self.Nmax)
            for I, info in self.dm[N].items():
                mn = 0 if last else info["kept"]
                Eabs = np.asarray(info["absenergy"][mn:info["total"]], dtype=np.longdouble) - np.longdouble(self.GSenergy)
                assert np.all(Eabs >= 0.0)
                ZZ += np.longdouble(T.mult(I)) * np.sum(np.exp(-Eabs / np.longdouble(p["T"])))
            self.ZnD[N] = float(ZZ)
        Z = 0.0
        for N in range(p["Ninit"], self.Nlen):
            Z += float(self.combs) ** (self.Nlen - N - 1) * self.ZnD[N]
        self.ZZ = Z
        self.wn = {N: float(self.combs) ** (self.Nlen - N - 1) * self.ZnD[N] / Z for N in range(p["Ninit"], self.Nlen)}
        assert abs(sum(self.wn.values()) - 1.0) < 1e-12
        self.wnfactor = {}
        for N in range(p["Ninit"], self.Nlen):
            w = self.wn[N] / self.ZnD[N] if self.ZnD[N] != 0 else float("nan")
            self.wnfactor[N] = 0.0 if math.isnan(w) else w

    def init_rho_FDM(self, N):
        T = self.tab
        p = self.p
        ZZ = self.ZnD[N]
        last = (N + 1 == self.Nmax)
        rho = {}
        for I, info in self.dm[N].items():
            mn = 0 if last else info["kept"]
            mx = info["total"]
            r = np.zeros((mx, mx))
            Eabs = np.asarray(info["absenergy"][mn:mx], dtype=np.longdouble) - np.longdouble(self.GSenergy)
            val2 = np.exp(-Eabs / np.longdouble(p["T"])) * (np.longdouble(self.wn[N]) / np.longdouble(ZZ))
            val2 = np.where(np.isfinite(val2), val2, 0.0)
            r[np.arange(mn, mx), np.arange(mn, mx)] = val2.astype(np.float64)
            rho[I] = r
        return rho

    def fdm_sweep(self):
        T = self.tab
        rho = self.init_rho_FDM(self.Nmax - 1)
        rhos = {}
        for N in range(self.Nmax - 1, self.p["Ninit"], -1):
            st = self.store[N]
            last = (N + 1 == self.Nmax)
            rhoDD = None if last else self.init_rho_FDM(N)
            prev = {}
            for I, info in self.dm[N - 1].items():
                dim = info["kept"]
                r = np.zeros((dim, dim))
                if dim:
                    for i in range(self.combs):
                        I1 = sub(I, T.In[i])
                        coef = T.mult(I1) / T.mult(I)
                        if I1 in st:
                            if I1 in rho:
                                self.cdmI(i, I1, rho[I1], st[I1], r, coef, self.dm[N])
                            if rhoDD is not None and I1 in rhoDD:
                                self.cdmI(i, I1, rhoDD[I1], st[I1], r, coef, self.dm[N])
                prev[I] = r
            tr = sum(T.mult(I) * np.trace(m) for I, m in prev.items())
            expected = sum(self.wn[n] for n in range(N, self.Nmax))
            assert abs((tr - expected) / expected) < 1e-8, (N, tr, expected)
            rhos[N - 1] = prev
            rho = prev
        return rhos

    # ---------------------------------------------------------- DMNRG pass
    def run_pass2(self):
        p = self.p
        T = self.tab
        diagprev, f, ops_s, ops_d = self.fresh_state()
        spec_ops = []
        if p["specd"]:
            for name in str(p["specd"]).replace(",", " ").split():
                parts = name.split("-")
                spin = {"u": 1, "d": -1}.get(parts[2], 0) if len(parts) == 3 else 0
                spec_ops.append((parts[0], parts[1], spin))
        kinds = []
        if p["dmnrg"]:
            kinds.append("DMNRG")
        if p["fdm"]:
            kinds.append("FDM")
        sc0, scl = self.SCALE(0), self.SCALE(self.Nmax)
        self.spectra = {(k, a, b, sp): (Bins(p, sc0, scl), Bins(p, sc0, scl)) for k in kinds for a, b, sp in spec_ops}
        recalc_all = p["fdm"]  # do_recalc_all(): strategy=kept overridden for CFS/FDM in this pass
        totalenergy = self.data.GS
        for N in range(p["Ninit"], self.Nmax):
            last = (N + 1 == self.Nmax)
            sc = self.SCALE(N + 1)
            st = self.store[N]
            diag = {}
            for I, s in st.items():
                diag[I] = dict(value=s["value"].copy(), shift=s["shift"], U=s["U"],
                               absenergy=s["absenergy"] - self.GSenergy)
            Egs = min(e["value"][0] + e["shift"] for e in diag.values())
            self.fix_splittings(diag)
            kept = self.truncation(diag, last, True)
            totalenergy += Egs * sc
            rmaxs = {I: st[I]["rmax"] for I in st}
            ancs = {I: st[I]["anc"] for I in st}
            for I, e in diag.items():
                off = st[I]["off"]
                e["nr"] = len(e["value"])
                e["blocks"] = [e["U"][:, off[i]:off[i + 1]] for i in range(self.combs)]

            def spectra_now(diag):
                for (kind, a, b, spin), (spos, sneg) in self.spectra.items():
                    op1, op2 = ops_d[a], ops_d[b]
                    cs = (Bins(p, sc0, scl), Bins(p, sc0, scl))
                    for Ii in diag:
                        for Ij in diag:
                            II = (Ij, Ii)
                            if II in op1 and II in op2 and (spin == 0 or Ij[1] - Ii[1] == spin):
                                sf = T.specdens_factor(Ii, Ij)
                                if kind == "DMNRG":
                                    self.spec_dmnrg(diag[Ii], diag[Ij], op1[II], op2[II], sf, cs, self.rho_now[Ii], self.rho_now[Ij], sc)
                                else:
                                    self.spec_fdm(N, last, diag[Ii], diag[Ij], op1[II], op2[II], sf, cs,
                                                  self.rhoFDM_now.get(Ii), self.rhoFDM_now.get(Ij), kept[Ii], kept[Ij])
                    if kind == "DMNRG":
                        self.merge_nn2(N, last, spos, sneg, cs, sc)
                    else:
                        spos.w += cs[0].w
                        sneg.w += cs[1].w

            if p["dmnrg"]:
                self.rho_now = self.rhos.get(N) if not last else self.rho_init
            if p["fdm"]:
                self.rhoFDM_now = self.rhoFDMs.get(N, {}) if not last else {}
            if recalc_all:
                ops_d = {k: self.recalc_doublet(diag, rmaxs, ancs, v) for k, v in ops_d.items()}
                spectra_now(diag)
            for I, e in diag.items():
                k = kept[I]
                e["blocks"] = [b[:k] for b in e["blocks"]]
                e["value"] = e["value"][:k]
                e["nr"] = k
            if not last:
                f = self.recalc_f(diag, rmaxs, ancs)
            if not recalc_all:
                ops_d = {k: self.recalc_doublet(diag, rmaxs, ancs, v) for k, v in ops_d.items()}
                spectra_now(diag)
            # trim to kept
            for k, op in ops_d.items():
                for II, m in op.items():
                    op[II] = m[:diag[II[0]]["nr"], :diag[II[1]]["nr"]]
            diagprev = diag
        self.totalenergy2 = totalenergy

    def spec_dmnrg(self, dIp, dI1, op1, op2, sf, cs, rhoIp, rhoI1, sc):
        p = self.p
        fac = self.getfactor()
        Emin, Emax = p["goodE"], p["goodE"] * fac * fac
        dimp, dim1 = rhoIp.shape[0], rhoI1.shape[0]
        Em = dIp["value"][:dimp]
        Ej = dI1["value"][:dim1]
        E = Ej[:, None] - Em[None, :]
        A = op2[:dim1, :dimp] @ rhoIp  # sumA(rj, rm) = sum_ri op2(rj,ri) rho(rm,ri)
        B = rhoI1 @ op1[:dim1, :dimp]  # sumB(rj, rm) = sum_ri op1(ri,rm) rho(rj,ri)
        W = sf * (A * op1[:dim1, :dimp] + B * op2[:dim1, :dimp])  # -sign = +1 (fermionic)
        absE = np.abs(E)
        ok = ~((absE < Emin) | (absE > Emax))
        self.cs_add(cs, sc * E[ok], W[ok])

    def spec_fdm(self, N, last, dI, dJ, op1, op2, sf, cs, rhoi, rhoj, kept_i, kept_j):
        p = self.p
        wnf = self.wnfactor[N]
        Tt = p["T"]
        reti = 0 if last else (rhoi.shape[0] if rhoi is not None else 0)
        retj = 0 if last else (rhoj.shape[0] if rhoj is not None else 0)
        alli, allj = dI["nr"], dJ["nr"]
        Ei = dI["absenergy"][:alli]
        Ej = dJ["absenergy"][:allj]
        E = Ej[:, None] - Ei[None, :]
        P = sf * op1[:allj, :alli] * op2[:allj, :alli]
        bi = np.exp(-Ei / Tt)
        bj = np.exp(-Ej / Tt)
        Dj = np.arange(allj) >= retj
        Di = np.arange(alli) >= reti
        # greater: A3 (D_i,D_j) + A2 (D_i,K_j): weight wnf exp(-Ei/T)
        m = Di[None, :] & np.ones((allj, 1), bool)
        self.cs_add(cs, E[m], (P * wnf * bi[None, :])[m])
        if allj > 0 and reti > 0:
            op2rho = op2[:allj, :reti] @ rhoi
            W = sf * op1[:allj, :reti] * op2rho
            mm = np.ones((allj, reti), bool) & Dj[:, None]
            self.cs_add(cs, E[:, :reti][mm], W[mm])
        # lesser: B3 (D_i,D_j) and B1 (K_i,D_j): wnf exp(-Ej/T); -sign = +1
        m = Dj[:, None] & np.ones((1, alli), bool)
        self.cs_add(cs, E[m], (P * wnf * bj[:, None])[m])
        if retj > 0 and alli > 0:
            rhoop2 = rhoj @ op2[:retj, :alli]
            W = sf * op1[:retj, :alli] * rhoop2
            mm = np.ones((retj, alli), bool) & Di[None, :]
            self.cs_add(cs, E[:retj, :][mm], W[mm])

    def cs_add(self, cs, E, W):
        pos = E >= 0.0
        cs[0].add_many(E[pos], W[pos])
        cs[1].add_many(-E[~pos], W[~pos])

    def getfactor(self):
        p = self.p
        return p["Lambda"] if (self.nch == 1 and not p["NN1"]) else math.sqrt(p["Lambda"])

    def merge_nn2(self, N, last, spos, sneg, cs, sc):
        p = self.p
        if not p["NN1"] and not p["NN2avg"]:
            even = (N % 2 == 0)
            if not (even if p["NN2even"] else not even):
                return
        fac = self.getfactor()
        Emin, Ex, Emax = sc * p["goodE"], sc * p["goodE"] * fac, sc * p["goodE"] * fac * fac
        first = (N == p["Ninit"])
        for full, c in ((spos, cs[0]), (sneg, cs[1])):
            E = c.e
            w = c.w
            win = np.zeros_like(E)
            for k in range(len(E)):
                e = E[k]
                if e <= Ex and last:
                    win[k] = 1.0
                elif e >= Ex and first:
                    win[k] = 1.0
                elif e <= Emin or e >= Emax:
                    win[k] = 0.0
                elif Emin < e <= Ex:
                    win[k] = (e - Emin) / (Ex - Emin)
                else:
                    win[k] = 1.0 - (e - Ex) / (Emax - Ex)
            ok = (Emin < E) & (E < Emax) & (w != 0.0)
            factor = 0.5 if p["NN2avg"] else 1.0
            full.w[ok] += factor * w[ok] * win[ok]

    # -------------------------------------------------- optional dump files
    @staticmethod
    def fcmp(x, y, eps):
        """my_fcmp (src/numerics.h:37-53)."""
        if x == 0.0 and y == 0.0:
            return 0
        if abs(x) < eps and abs(y) < eps:
            return 0
        if abs(x - y) / (abs(x) + abs(y)) < eps:
            return 0
        return 1 if x > y else -1

    def annotated_levels(self):
        """annotated.dat as data (dump_annotated, src/nrg.cc:2062-2107, called
        from nrg_perform_measurements 2403-2409 in the first run, calc0
        included): per iteration a list of (energy, sorted QN strings, total
        multiplicity) groups, energies as omega_N-scaled values (dumpscaled,
        no dumpabs)."""
        p = self.p
        out = []
        per_iter = []
        if p["calc0"]:
            per_iter.append({I: np.asarray(v) for I, v in self.data.energies.items()})
        per_iter += [{I: v for I, (v, k) in it.items()} for it in self.iters]
        for vals in per_iter:
            lst = sorted((float(e), I) for I, v in vals.items() for e in v)
            ln = min(len(lst), int(p["dumpannotated"]))
            i = ln
            while i < len(lst) - 1:  # the reference never adds the last level
                if self.fcmp(lst[i][0], lst[i - 1][0], p["grouptol"]) == 0:
                    ln += 1
                    i += 1
                else:
                    break
            groups = []
            i = 0
            while i < ln:
                i0 = i
                q, deg = [], 0
                while i < ln and self.fcmp(lst[i][0], lst[i0][0], p["grouptol"]) == 0:
                    q.append("%d %d" % lst[i][1])
                    deg += self.tab.mult(lst[i][1])
                    i += 1
                groups.append((lst[i0][0], sorted(q), deg))
            out.append(groups)
        return out

    def subspaces_lines(self):
        """subspaces.dat (dump_subspace_information, src/nrg.cc:3207-3218):
        per iteration the subspaces with len (= kept after truncate_perform,
        src/nrg.cc:359-367), kept and total."""
        lines = []
        for N in range(self.p["Ninit"], self.Nmax):
            dm = self.dm[N]
            lines.append("Iteration %d" % N)
            lines.append("len_dm=%d" % len(dm))
            for I in sorted(dm):
                lines.append("I=%d %d len=%d kept=%d total=%d" % (I[0], I[1], dm[I]["kept"], dm[I]["kept"],
                                                                   dm[I]["total"]))
            lines.append("")
        return lines

    def fdmexpv_values(self):
        """customfdm (calc_trace_fdm_kept, src/nrg.cc:973-991; evaluated at
        iteration fdmexpvn, src/nrg.cc:2239-2247): sum_I mult(I) sum_{i,j <
        kept} rho_FDM[I](i,j) O[I](j,i)."""
        N = int(self.p["fdmexpvn"])
        rho = self.rhoFDMs[N]
        out = {}
        for name, op in self.ops_s_fdmexpv.items():
            tr = 0.0
            for I, info in self.dm[N].items():
                k = info["kept"]
                if not k:
                    continue
                tr += self.tab.mult(I) * np.sum(rho[I][:k, :k] * op[(I, I)][:k, :k].T)
            out[name] = tr
        return out

    # --------------------------------------------------------------- output
    def trimmed(self, b):
        e, w = b.e, b.w
        keep = [i for i in range(len(e) - 1) if not (abs(w[i]) < self.p["discard_trim"] * (e[i + 1] - e[i]))]
        keep.append(len(e) - 1)
        return e[keep], w[keep]

    def run(self):
        self.load()
        self.run_pass1()
        p = self.p
        if p["fdm"]:
            self.calc_ZnD()
        if p["dm"]:
            self.rho_init = self.init_rho()
            if p["dmnrg"]:
                self.rhos = self.dm_sweep(self.rho_init)
            if p["fdm"]:
                self.rhoFDMs = self.fdm_sweep()
            self.run_pass2()
            assert abs(self.totalenergy2 - self.GSenergy) <= 1e-12 * max(1.0, abs(self.GSenergy))
        return self

    def spectrum_bins(self, kind, a, b, spin=0):
        """Trimmed (E, w) pairs as written to spec_<kind>_dens_<a>-<b>.bin."""
        spos, sneg = self.spectra[(kind, a, b, spin)]
        ep, wp = self.trimmed(spos)
        en, wn = self.trimmed(sneg)
        return np.concatenate([ep, -en]), np.concatenate([wp, wn])

    def broadening_mesh(self):
        """make_mesh + get_broaden_min (src/spectrumrealfreq.cc:212-229):
        broaden_max, divided by broaden_ratio while above the lowest energy
        (broaden_min, else broaden_min_ratio x SCALE(Nmax))."""
        p = self.p
        bmin = p["broaden_min"] if p["broaden_min"] > 0.0 else p["broaden_min_ratio"] * self.SCALE(self.Nmax)
        mesh = []
        E = p["broaden_max"]
        while E > bmin:
            mesh.append(E)
            E /= p["broaden_ratio"]
        return np.array(mesh)

    def br_new(self, e, ept):
        """BR_NEW (src/spectral.h:56-102) for arrays: the log-Gaussian
        BR_L(e, ept) = exp(-(ln(e/ept)/alpha - alpha/4)^2) / (alpha |e| sqrt(pi))
        (zero across the sign change and at ept = 0), blended below |e| = omega0
        into the Gaussian BR_G of width omega0 with h = exp(-(ln(|e|/omega0)/alpha)^2)."""
        p = self.p
        alpha = p["alpha"]
        gamma = alpha / 4.0
        omega0 = p["omega0_ratio"] * p["T"] if p["omega0"] < 0.0 else p["omega0"]
        e, ept = np.broadcast_arrays(np.asarray(e, float), np.asarray(ept, float))
        same = (e * ept) > 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            x = np.where(same, np.log(np.where(same, e / ept, 1.0)) / alpha - gamma, 0.0)
            bl = np.where(same, np.exp(-x * x) / (alpha * np.abs(e) * math.sqrt(math.pi)), 0.0)
            y = np.log(np.abs(e) / omega0) / alpha
            h = np.exp(-y * y)
            bg = np.exp(-((e - ept) / omega0) ** 2) / (omega0 * math.sqrt(math.pi))
        return np.where(np.abs(e) > omega0, bl, bl * h + bg * (1.0 - h))

    def broadened(self, kind, a, b, spin=0):
        """SpectrumRealFreq::continuous (src/spectrumrealfreq.cc:271-306): the
        trimmed bins of both signs broadened onto the mesh at +E and -E;
        returns (E, A) in the order spec_<kind>_dens_<a>-<b>.dat lists them
        (negative energies ascending, then positive ascending)."""
        spos, sneg = self.spectra[(kind, a, b, spin)]
        ep, wp = self.trimmed(spos)
        en, wn = self.trimmed(sneg)
        mesh = self.broadening_mesh()
        M = mesh[:, None]
        vpos = (wp[None, :] * self.br_new(M, ep[None, :])).sum(1) + (wn[None, :] * self.br_new(M, -en[None, :])).sum(1)
        vneg = (wp[None, :] * self.br_new(-M, ep[None, :])).sum(1) + (wn[None, :] * self.br_new(-M, -en[None, :])).sum(1)
        order = np.argsort(mesh)
        return (np.concatenate([-mesh[order][::-1], mesh[order]]),
                np.concatenate([vneg[order][::-1], vpos[order]]))


def run(workdir):
    return Oracle(workdir).run()
