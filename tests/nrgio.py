"""Readers for the engine's output files and helpers shared by the tests."""
import math
import os
import struct
import subprocess
import sys

import numpy as np

from conftest import REPO, PKG

TABLES = os.path.join(PKG, "tables")


def read_dump(path):
    """--dump-eigen file: per iteration N, per subspace (Q, SS, d, kept, values)."""
    out = []
    with open(path, "rb") as f:
        b = f.read()
    p = 0
    while p < len(b):
        N, ns = struct.unpack_from("ii", b, p)
        p += 8
        it = {}
        for _ in range(ns):
            q, ss, d, k = struct.unpack_from("iiii", b, p)
            p += 16
            v = np.frombuffer(b, dtype=np.float64, count=d, offset=p).copy()
            p += 8 * d
            it[(q, ss)] = (v, k)
        out.append((N, it))
    return out


def read_table(path):
    """td / custom: whitespace columns, '#' header lines."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(x) for x in line.split()])
    return np.array(rows)


def read_bins(path):
    a = np.fromfile(path, dtype=np.float64)
    return a[0::2], a[1::2]


def build_mock():
    """Host driver linked against the CPU stand-in of the C ABI (tests only)."""
    out = os.path.join(REPO, "tests", "mock", "_build", "nrg_mock")
    srcs = [os.path.join(PKG, "host", f) for f in ("main.cpp", "cli.cpp", "engine.cpp", "io.cpp")]
    srcs.append(os.path.join(REPO, "tests", "mock", "mock_gpu.cpp"))
    newest = max(os.path.getmtime(s) for s in srcs + [os.path.join(PKG, "host", "nrg_host.h"),
                                                      os.path.join(REPO, "include", "nrg_gpu.h")])
    if not os.path.exists(out) or os.path.getmtime(out) < newest:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-I" + os.path.join(REPO, "include")] + srcs + ["-o", out])
    return out


def build_mock_serve():
    """`nrg-serve` linked against the CPU stand-in of the C ABI (tests only)."""
    out = os.path.join(REPO, "tests", "mock", "_build", "nrg_serve_mock")
    srcs = [os.path.join(PKG, "host", f) for f in ("serve_main.cpp", "serve.cpp", "cli.cpp", "engine.cpp", "io.cpp")]
    srcs.append(os.path.join(REPO, "tests", "mock", "mock_gpu.cpp"))
    newest = max(os.path.getmtime(s) for s in srcs + [os.path.join(PKG, "host", "nrg_host.h"),
                                                      os.path.join(REPO, "include", "nrg_gpu.h")])
    if not os.path.exists(out) or os.path.getmtime(out) < newest:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-pthread", "-I" + os.path.join(REPO, "include")] + srcs +
                              ["-o", out])
    return out


def ensure_tables():
    if not os.path.exists(os.path.join(TABLES, "qs-1ch.tab")):
        subprocess.check_call([sys.executable, "-m", "nrgamd.symtab", TABLES], cwd=PKG)


def run_engine(binary, workdir, timeout=600, extra=()):
    ensure_tables()
    env = dict(os.environ)
    env["NRG_TABLES"] = TABLES
    r = subprocess.run([binary, "--dump-eigen", "eig.bin", "--stats", "stats.json"] + list(extra), cwd=workdir,
                       env=env,
                       capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("engine failed: %s\n%s" % (r.stdout[-2000:], r.stderr[-2000:]))
    return r.stdout


def make_case(workdir, model="siam", sym="QS", nch=1, Nmax=8, Lambda=2.0, z=1.0, model_kw=None, **param):
    sys.path.insert(0, PKG)
    from nrgamd import datagen
    model_kw = model_kw or {}
    if model == "siam":
        m = datagen.siam(sym=sym, nch=nch, Lambda=Lambda, z=z, **model_kw)
    else:
        m = datagen.kondo(sym=sym, nch=nch, Lambda=Lambda, z=z, **model_kw)
    m.write(os.path.join(workdir, "data"), Lambda, z, Nmax)
    datagen.write_param(os.path.join(workdir, "param"), Lambda=Lambda, z=z, **param)


def compare_with_oracle(workdir, o, spectra=(), eig_atol=1e-10, rel=1e-8, eigen=True):
    """Engine outputs in `workdir` vs an oracle run on the same inputs
    (eigen=False: a run without the --dump-eigen file, e.g. a served solve)."""
    if eigen:
        dump = read_dump(os.path.join(workdir, "eig.bin"))
        assert len(dump) == len(o.iters)
        for (N, it), ref in zip(dump, o.iters):
            assert set(it) == set(ref), ("subspace lists differ at N=%d" % N)
            for I, (v, k) in it.items():
                rv, rk = ref[I]
                assert len(v) == len(rv)
                assert k == rk, ("kept count differs", N, I, k, rk)
                np.testing.assert_allclose(v, rv, atol=eig_atol, rtol=0)
    td = read_table(os.path.join(workdir, "td"))
    assert len(td) == len(o.td_rows)
    keys = ["T", "<Sz^2>", "<Q>", "<Q^2>", "E", "E2", "C", "F", "S"] if o.sym == "QS" else \
        ["T", "<Sz>", "<Sz^2>", "<Q>", "E", "E2", "C", "F", "S"]
    for row, ref in zip(td, o.td_rows):
        for x, k in zip(row, keys):
            y = ref[k]
            assert abs(x - y) <= max(1e-9 * abs(y), 5e-10), (k, x, y)  # td is printed with 10 digits
    # custom: <O> of the singlet operators per iteration (printed with 10 digits)
    cu = read_table(os.path.join(workdir, "custom"))
    assert len(cu) == len(o.custom_rows)
    for row, (Tref, vals) in zip(cu, o.custom_rows):
        assert abs(row[0] - Tref) <= 1e-9 * abs(Tref)
        for x, name in zip(row[1:], sorted(vals)):
            y = vals[name]
            assert abs(x - y) <= max(1e-9 * abs(y), 1e-12), ("custom", name, x, y)
    for kind, a, b, spin in spectra:
        name = "%s-%s" % (a, b) + {1: "-u", -1: "-d", 0: ""}[spin]
        E, W = read_bins(os.path.join(workdir, "spec_%s_dens_%s.bin" % (kind, name)))
        rE, rW = o.spectrum_bins(kind, a, b, spin)
        assert len(E) == len(rE), (kind, len(E), len(rE))
        np.testing.assert_allclose(E, rE, rtol=1e-15)
        # every bin to `rel` of itself; the absolute floor only catches bins
        # that are zero up to rounding
        scale = max(np.abs(rW).max(), 1e-300)
        np.testing.assert_allclose(W, rW, rtol=rel, atol=1e-14 * scale)
        if o.p["broaden"]:
            # the broadened curve (.dat, printed with prec_xy = 10 digits)
            dat = np.loadtxt(os.path.join(workdir, "spec_%s_dens_%s.dat" % (kind, name)), ndmin=2)
            rE, rA = o.broadened(kind, a, b, spin)
            assert dat.shape == (len(rE), 2), (kind, dat.shape, len(rE))
            np.testing.assert_allclose(dat[:, 0], rE, rtol=1e-9)
            np.testing.assert_allclose(dat[:, 1], rA, rtol=rel + 1e-9, atol=1e-14 * max(np.abs(rA).max(), 1e-300))
    compare_dumps(workdir, o)


def read_annotated(path):
    """annotated.dat: per iteration (blocks separated by an empty line) the
    groups (energy, [QN strings], multiplicity)."""
    its, cur = [], []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line.strip():
                its.append(cur)
                cur = []
                continue
            e = float(line.split()[0])
            q = [x.strip() for x in line[line.index("(") + 1:line.rindex(")")].split(") (")]
            deg = int(line[line.rindex("[") + 1:line.rindex("]")])
            cur.append((e, q, deg))
    return its


def read_energies(path):
    """energies.nrg / energies.dmnrg: {N: {(Q, SS): values}}."""
    out, N, I = {}, None, None
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("====="):
                N = int(line.split()[-1])
                out[N] = {}
            elif line.startswith("Subspace:"):
                a, b = line.split()[1:3]
                I = (int(a), int(b))
            elif line:
                out[N][I] = np.array([float(x) for x in line.split()])
    return out


def compare_dumps(workdir, o):
    """annotated.dat, subspaces.dat, energies.*, customfdm when the run wrote them."""
    p = o.p
    path = os.path.join(workdir, "annotated.dat")
    if int(p.get("dumpannotated", 0)) > 0:
        got, ref = read_annotated(path), o.annotated_levels()
        assert len(got) == len(ref)
        for g, r in zip(got, ref):
            assert [(q, d) for e, q, d in g] == [(q, d) for e, q, d in r]
            for (e1, _, _), (e2, _, _) in zip(g, r):
                assert abs(e1 - e2) <= 1e-7 * max(abs(e2), 1e-6), (e1, e2)  # 8 digits printed
    if p.get("dumpsubspaces"):
        with open(os.path.join(workdir, "subspaces.dat")) as f:
            assert f.read().split("\n")[:-1] == o.subspaces_lines()
    if p.get("dumpenergies"):
        e1 = read_energies(os.path.join(workdir, "energies.nrg"))
        ref = {N: {I: v for I, (v, k) in it.items()} for N, it in zip(range(p["Ninit"], o.Nmax), o.iters)}
        for N in ref:
            assert set(e1[N]) == set(ref[N])
            for I in ref[N]:
                np.testing.assert_allclose(e1[N][I], ref[N][I], rtol=1e-5, atol=1e-5)  # 6 digits printed
        if p.get("dm"):
            with open(os.path.join(workdir, "energies.nrg")) as f1, open(os.path.join(workdir, "energies.dmnrg")) as f2:
                assert f1.read() == f2.read()  # the reference's check that both runs agree
    if p.get("fdmexpv"):
        cf = read_table(os.path.join(workdir, "customfdm"))
        ref = o.fdmexpv_values()
        assert len(cf) == 1 and abs(cf[0][0] - p["T"]) <= 1e-9 * p["T"]
        for x, name in zip(cf[0][1:], sorted(ref)):
            y = ref[name]
            assert abs(x - y) <= max(1e-9 * abs(y), 1e-12), ("customfdm", name, x, y)
