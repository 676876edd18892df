"""Serving workload driver: seeded request streams of whole NRG solves.

A request is one SIAM (QS) parameter set solved for several z-members
(discretisation offsets); each member is one solve.  `make_requests` writes
every solve's param and data into its own directory and a request list with
Poisson arrival times; `run` hands the list to `bin/nrg-serve` (continuous
batching of the solves' iterations on one GPU, admission by HBM footprint)
and then z-averages each request's broadened spectra (weight 1/Nz on the
common energy mesh, as the reference's averaging over z sub-runs).

    python3 -m nrgamd.serve --requests 64 --out /tmp/serve   # the full stream
"""
import argparse
import json
import math
import os
import subprocess
import sys

import numpy as np

from . import datagen

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
Z_MEMBERS = (0.25, 0.5, 0.75, 1.0)
SPEC = "spec_FDM_dens_A_d-A_d"


def draw_request(seed, keep_range=(1000, 8000)):
    """One request's physical parameters from its own seed (U ~ U[0.2, 2],
    eps_d = -U/2 + delta with delta ~ U[-0.2, 0.2], Gamma log-uniform in
    [0.01, 0.1], Lambda in {2, 3}, keep ~ U[keep_range])."""
    rng = np.random.default_rng(seed)
    U = float(rng.uniform(0.2, 2.0))
    delta = float(rng.uniform(-0.2, 0.2))
    Gamma = float(10.0 ** rng.uniform(-2.0, -1.0))
    Lambda = float(rng.choice([2.0, 3.0]))
    keep = int(rng.integers(keep_range[0], keep_range[1] + 1))
    return dict(U=U, eps_d=-U / 2 + delta, Gamma=Gamma, Lambda=Lambda, keep=keep)


def nmax_for(Lambda, T):
    """Chain length whose last energy scale reaches T."""
    return int(math.ceil(2.0 * math.log(1.0 / T) / math.log(Lambda)))


def make_requests(out, nreq=64, seed=1000, rate=None, keep_range=(1000, 8000), T=1e-6, zs=Z_MEMBERS,
                  keepenergy=10.0):
    """Writes out/req<i>/z<k>/{param,data} and out/requests.txt; returns the
    request descriptions.  Arrivals: Poisson with `rate` requests/s (all at 0
    when None); a request's z-members arrive together."""
    os.makedirs(out, exist_ok=True)
    arr_rng = np.random.default_rng(seed - 1)
    t = 0.0
    reqs = []
    lines = []
    for i in range(nreq):
        p = draw_request(seed + i, keep_range)
        if rate:
            t += float(arr_rng.exponential(1.0 / rate))
        Nmax = nmax_for(p["Lambda"], T)
        members = []
        for k, z in enumerate(zs):
            d = os.path.join(out, "req%03d" % i, "z%d" % k)
            os.makedirs(d, exist_ok=True)
            m = datagen.siam(sym="QS", nch=1, U=p["U"], eps_d=p["eps_d"], Gamma=p["Gamma"], Lambda=p["Lambda"])
            m.write(os.path.join(d, "data"), p["Lambda"], z, Nmax)
            datagen.write_param(os.path.join(d, "param"), Lambda=p["Lambda"], z=z, keep=p["keep"],
                                keepenergy=keepenergy, T=T, fdm=True, specd="A_d-A_d", savebins=True,
                                broaden=True)
            members.append(d)
            lines.append("req%03d/z%d %.6f %s" % (i, k, t, d))
        reqs.append(dict(id="req%03d" % i, arrival=t, Nmax=Nmax, members=members, **p))
    with open(os.path.join(out, "requests.txt"), "w") as f:
        f.write("# id arrival_seconds directory\n" + "\n".join(lines) + "\n")
    return reqs


def read_dat(path):
    a = np.loadtxt(path)
    return a[:, 0], a[:, 1]


def z_average(req):
    """Mean of the members' broadened spectra on the energies they share
    (the mesh is the same geometric sequence for every z; its low end depends
    on the last shell's scale)."""
    curves = [read_dat(os.path.join(d, SPEC + ".dat")) for d in req["members"]]
    common = set(np.round(curves[0][0], 12))
    for e, _ in curves[1:]:
        common &= set(np.round(e, 12))
    keys = np.array(sorted(common))
    acc = np.zeros(len(keys))
    for e, v in curves:
        lut = dict(zip(np.round(e, 12), v))
        acc += np.array([lut[k] for k in keys])
    acc /= len(curves)
    out = os.path.join(os.path.dirname(req["members"][0]), SPEC + ".zavg.dat")
    np.savetxt(out, np.column_stack([keys, acc]), fmt="%.12e")
    return out


def run(out, reqs, slots=8, window_ms=5.0, mem_fraction=0.85, timeout=3600, exe=None, reqfile=None):
    """Runs bin/nrg-serve on out/requests.txt, z-averages, returns the report."""
    exe = exe or os.path.join(PKG, "bin", "nrg-serve")
    if not os.path.exists(exe):
        raise RuntimeError("bin/nrg-serve is not built; run __graft_entry__.build()")
    report = os.path.join(out, "serve_report.json")
    env = dict(os.environ, NRG_TABLES=os.path.join(PKG, "tables"))
    # status 1 when a request failed: the report lists it with its error and
    # the other requests are complete
    subprocess.run([exe, reqfile or os.path.join(out, "requests.txt"), "--report", report, "--slots", str(slots),
                    "--window-ms", str(window_ms), "--mem-fraction", str(mem_fraction)], env=env, check=False,
                   timeout=timeout)
    if not os.path.exists(report):
        raise RuntimeError("nrg-serve wrote no report")
    with open(report) as f:
        rep = json.load(f)
    bad = {s["dir"] for s in rep["requests"] if s["rc"] != 0}
    for r in reqs:
        r["zavg"] = None if any(m in bad for m in r["members"]) else z_average(r)
    # request latency: last member finished minus arrival
    fin = {}
    for s in rep["requests"]:
        rid = s["id"].split("/")[0]
        fin[rid] = max(fin.get(rid, 0.0), s["finished"])
    lat = [fin[r["id"]] - r["arrival"] for r in reqs]
    rep["request_latency_p50_s"] = float(np.percentile(lat, 50))
    rep["request_latency_p99_s"] = float(np.percentile(lat, 99))
    rep["requests_per_s"] = len(reqs) / rep["makespan_s"]
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=None, help="requests per second (Poisson); default: all at once")
    ap.add_argument("--keep-min", type=int, default=1000)
    ap.add_argument("--keep-max", type=int, default=8000)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    reqs = make_requests(a.out, a.requests, a.seed, a.rate, (a.keep_min, a.keep_max))
    rep = run(a.out, reqs, slots=a.slots)
    summary = {k: v for k, v in rep.items() if k != "requests"}
    print(json.dumps(summary))


if __name__ == "__main__":
    sys.exit(main())
