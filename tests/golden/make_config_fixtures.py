"""Writes the oracle's results for the keep = 10000 two-channel Anderson
configuration cut to Nmax = 3 (tests/golden/anderson2ch_keep10000_nmax3.npz).

The oracle needs minutes for it (its last iteration has 92 378 eigenpairs in
blocks of up to 8820 rows), so the GPU parity test compares against these
stored numbers instead of running it: per iteration and subspace the
dimension, the kept count, the lowest eigenvalues (all of them up to 400 rows,
the lowest 400 above) and the sum and sum of squares of all of them; the td
and custom rows; the FDM spectral bins.

    python3 tests/golden/make_config_fixtures.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "nrgljubljana-old_amd"))
sys.path.insert(0, REPO)

NLOW = 400
C3_CUT = dict(sym="QS", nch=2, Lambda=3.0, Nmax=3, model_kw=dict(U=0.5, eps_d=-0.25, Gamma=0.02),
              param=dict(keep=10000, keepenergy=100.0, T=1e-10, fdm=True, specd="A_d-A_d", bins=100,
                         savebins=True))


def write_inputs(wd, c):
    from nrgamd import datagen
    m = datagen.siam(sym=c["sym"], nch=c["nch"], Lambda=c["Lambda"], **c["model_kw"])
    m.write(os.path.join(wd, "data"), c["Lambda"], 1.0, c["Nmax"])
    datagen.write_param(os.path.join(wd, "param"), Lambda=c["Lambda"], **c["param"])


def main():
    from oracle import nrg_oracle
    out = {}
    with tempfile.TemporaryDirectory() as wd:
        write_inputs(wd, C3_CUT)
        o = nrg_oracle.Oracle(wd).run()
    keys, dims, kept, s1, s2, low = [], [], [], [], [], []
    for N, it in enumerate(o.iters):
        for I in sorted(it):
            v, k = it[I]
            keys.append((N, I[0], I[1]))
            dims.append(len(v))
            kept.append(k)
            s1.append(float(np.sum(v)))
            s2.append(float(np.sum(v * v)))
            lo = np.full(NLOW, np.nan)
            lo[:min(NLOW, len(v))] = v[:NLOW]
            low.append(lo)
    out["keys"] = np.array(keys, dtype=np.int64)
    out["dims"] = np.array(dims, dtype=np.int64)
    out["kept"] = np.array(kept, dtype=np.int64)
    out["sum"] = np.array(s1)
    out["sumsq"] = np.array(s2)
    out["low"] = np.array(low)
    td_keys = ["T", "<Sz^2>", "<Q>", "<Q^2>", "E", "E2", "C", "F", "S"]
    out["td"] = np.array([[r[k] for k in td_keys] for r in o.td_rows])
    names = sorted(o.custom_rows[0][1])
    out["custom"] = np.array([[t] + [vals[n] for n in names] for t, vals in o.custom_rows])
    E, W = o.spectrum_bins("FDM", "A_d", "A_d", 0)
    out["fdm_E"], out["fdm_W"] = E, W
    path = os.path.join(HERE, "anderson2ch_keep10000_nmax3.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
