"""Serving loop (continuous batching of whole solves) on the CPU stand-in of
the C ABI: every solve of a mixed stream gives the same outputs as when it
runs alone, the shared eigensolver really merges iterations of different
solves, latencies are reported and z-averaged spectra are written."""
import json
import os
import shutil

import numpy as np
import pytest

import nrgio
from conftest import load_nrgamd

serve = load_nrgamd().serve if hasattr(load_nrgamd(), "serve") else __import__("nrgamd.serve").serve


def _outputs(d):
    td = open(os.path.join(d, "td")).read()
    bins = np.fromfile(os.path.join(d, serve.SPEC + ".bin"), dtype=np.float64)
    return td, bins


@pytest.fixture(scope="module")
def stream(tmp_path_factory):
    exe = nrgio.build_mock_serve()
    nrgio.ensure_tables()
    root = str(tmp_path_factory.mktemp("serve"))
    mix = os.path.join(root, "mix")
    reqs = serve.make_requests(mix, nreq=3, seed=77, rate=None, keep_range=(40, 90), T=1e-3, zs=(0.5, 1.0))
    rep = serve.run(mix, reqs, slots=4, window_ms=50.0, exe=exe)
    # the first request once more, alone, from fresh copies of its inputs
    alone = os.path.join(root, "alone")
    lines = []
    for k, d in enumerate(reqs[0]["members"]):
        nd = os.path.join(alone, "z%d" % k)
        os.makedirs(nd)
        for f in ("param", "data"):
            shutil.copy(os.path.join(d, f), nd)
        lines.append("%s/z%d 0.0 %s" % (reqs[0]["id"], k, nd))
    with open(os.path.join(alone, "requests.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    one = dict(reqs[0], members=[os.path.join(alone, "z%d" % k) for k in range(len(lines))])
    rep1 = serve.run(alone, [one], slots=4, exe=exe)
    return reqs, rep, one, rep1


def test_every_solve_completes_and_latency_is_reported(stream):
    reqs, rep, _, _ = stream
    assert rep["solves"] == 6 and rep["failed"] == 0
    assert all(s["rc"] == 0 and s["latency"] >= 0 for s in rep["requests"])
    assert 0 <= rep["latency_p50_s"] <= rep["latency_p99_s"]
    assert rep["request_latency_p50_s"] <= rep["request_latency_p99_s"]
    assert rep["eigenpairs"] > 0 and rep["solves_per_s"] > 0


def test_iterations_of_different_solves_share_launches(stream):
    _, rep, _, _ = stream
    assert rep["merged_iterations"] > rep["shared_launches"]


def test_result_alone_equals_result_in_the_mix(stream):
    reqs, _, one, _ = stream
    for mixed, alone in zip(reqs[0]["members"], one["members"]):
        td_m, b_m = _outputs(mixed)
        td_a, b_a = _outputs(alone)
        assert td_m == td_a
        np.testing.assert_array_equal(b_m, b_a)


def test_z_average_is_the_mean_of_the_members(stream):
    reqs, _, _, _ = stream
    r = reqs[1]
    e, v = serve.read_dat(r["zavg"])
    curves = [serve.read_dat(os.path.join(d, serve.SPEC + ".dat")) for d in r["members"]]
    lut = [dict(zip(np.round(ce, 12), cv)) for ce, cv in curves]
    for k in (0, len(e) // 2, len(e) - 1):
        key = np.round(e[k], 12)
        assert v[k] == pytest.approx(np.mean([l[key] for l in lut]), rel=1e-10)


def test_failed_request_is_reported_and_the_rest_complete(tmp_path):
    """A solve whose input is broken fails alone: the server reports it with
    its error and finishes the other requests (no process exit)."""
    exe = nrgio.build_mock_serve()
    nrgio.ensure_tables()
    out = str(tmp_path)
    reqs = serve.make_requests(out, nreq=2, seed=5, rate=None, keep_range=(40, 60), T=1e-3, zs=(1.0,))
    bad = reqs[1]["members"][0]
    with open(os.path.join(bad, "data"), "w") as f:
        f.write("# broken\n")
    rep = serve.run(out, reqs, slots=2, exe=exe)
    assert rep["solves"] == 2 and rep["failed"] == 1
    by_dir = {s["dir"]: s for s in rep["requests"]}
    assert by_dir[bad]["rc"] == 1 and "version" in by_dir[bad]["error"]
    good = reqs[0]["members"][0]
    assert by_dir[good]["rc"] == 0 and os.path.exists(os.path.join(good, "DONE"))
    assert reqs[0]["zavg"] is not None and reqs[1]["zavg"] is None
