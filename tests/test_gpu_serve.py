"""Serving loop on the GPU: solves of a mixed stream share the batched
eigensolver launches and each gives the same result as when it runs alone
(td identical, spectral bins within 1e-12 relative), and the z-averaged
spectrum is written."""
import os
import shutil

import numpy as np
import pytest

from conftest import load_nrgamd

pytestmark = pytest.mark.gpu
serve = load_nrgamd().serve if hasattr(load_nrgamd(), "serve") else __import__("nrgamd.serve").serve


def _outputs(d):
    td = np.loadtxt(os.path.join(d, "td"))
    bins = np.fromfile(os.path.join(d, serve.SPEC + ".bin"), dtype=np.float64)
    return td, bins


def test_gpu_serve_alone_equals_mixed(tmp_path):
    print("case: 4 requests x 2 z-members, keep 100-300", flush=True)
    mix = str(tmp_path / "mix")
    reqs = serve.make_requests(mix, nreq=4, seed=123, rate=None, keep_range=(100, 300), T=1e-4, zs=(0.5, 1.0))
    rep = serve.run(mix, reqs, slots=8, window_ms=20.0, timeout=300)
    assert rep["failed"] == 0 and rep["solves"] == 8
    assert rep["merged_iterations"] > rep["shared_launches"]
    alone = str(tmp_path / "alone")
    lines = []
    for k, d in enumerate(reqs[2]["members"]):
        nd = os.path.join(alone, "z%d" % k)
        os.makedirs(nd)
        for f in ("param", "data"):
            shutil.copy(os.path.join(d, f), nd)
        lines.append("%s/z%d 0.0 %s" % (reqs[2]["id"], k, nd))
    with open(os.path.join(alone, "requests.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    one = dict(reqs[2], members=[os.path.join(alone, "z%d" % k) for k in range(len(lines))])
    rep1 = serve.run(alone, [one], slots=8, timeout=300)
    assert rep1["failed"] == 0
    for m, a in zip(reqs[2]["members"], one["members"]):
        td_m, b_m = _outputs(m)
        td_a, b_a = _outputs(a)
        np.testing.assert_allclose(td_m, td_a, rtol=1e-12, atol=0)
        scale = np.abs(b_a).max()
        np.testing.assert_allclose(b_m, b_a, rtol=1e-12, atol=1e-12 * scale)
    assert os.path.exists(reqs[2]["zavg"])
