"""The drop-in on the GPU on the configurations of the measured workloads,
against the oracle (eigenvalues within 1e-10 omega_N, identical kept counts,
td / custom / spectral bins within 1e-8 relative):

- SIAM QS keep <= 4000 (Ecut 10), Nmax 50, FDM at T = 1e-10: full size, the
  oracle run beside it;
- two-channel Anderson QS keep 10000 (Lambda 3, FDM at T = 1e-10) cut to
  Nmax 3 (92 378 eigenpairs in blocks up to 8820 rows at N = 2), against the
  oracle's numbers stored by tests/golden/make_config_fixtures.py;
- three-channel Kondo QSZ (J rho = 0.2 per channel, Lambda 4, keep 12000) at
  Nmax 1: its first shell of 8192 states (the next shell holds 524 288 states
  in blocks of up to 31 752 rows, beyond a CPU oracle);
- served SIAM requests drawn as the serving stream draws them (seeded U,
  eps_d, Gamma, Lambda, keep in 1000..8000), each z-member against the oracle.
"""
import os

import numpy as np
import pytest

from conftest import PKG, REPO, load_oracle
import nrgio

pytestmark = pytest.mark.gpu

NRG = os.path.join(PKG, "bin", "nrg")


def test_siam_keep4000_fdm_full_size(tmp_path):
    wd = str(tmp_path)
    nrgio.make_case(wd, sym="QS", nch=1, Nmax=50, Lambda=2.0, model_kw=dict(U=1.0, eps_d=-0.5, Gamma=0.04),
                    keep=4000, keepenergy=10.0, T=1e-10, fdm=True, specd="A_d-A_d", bins=100, savebins=True)
    print("case siam keep4000 Nmax50 FDM", flush=True)
    nrgio.run_engine(NRG, wd)
    o = load_oracle().Oracle(wd).run()
    nrgio.compare_with_oracle(wd, o, spectra=[("FDM", "A_d", "A_d", 0)])


def test_anderson2ch_keep10000_cut_matches_stored_oracle(tmp_path):
    import sys
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    import make_config_fixtures as mk
    fx = np.load(os.path.join(REPO, "tests", "golden", "anderson2ch_keep10000_nmax3.npz"))
    wd = str(tmp_path)
    mk.write_inputs(wd, mk.C3_CUT)
    print("case anderson2ch keep10000 Nmax3", flush=True)
    nrgio.run_engine(NRG, wd)
    dump = nrgio.read_dump(os.path.join(wd, "eig.bin"))
    got = {}
    for N, it in dump:
        for I, (v, k) in it.items():
            got[(N, I[0], I[1])] = (v, k)
    keys = [tuple(k) for k in fx["keys"]]
    assert set(got) == set(keys)
    for r, key in enumerate(keys):
        v, k = got[key]
        assert len(v) == fx["dims"][r] and k == fx["kept"][r], key
        n = min(len(v), mk.NLOW)
        np.testing.assert_allclose(v[:n], fx["low"][r][:n], atol=1e-10, rtol=0)
        # all eigenvalues through two invariants (sum: trace; sum of squares:
        # Frobenius norm), each to 1e-10 per eigenvalue
        assert abs(v.sum() - fx["sum"][r]) <= 1e-10 * len(v) * max(1.0, np.abs(v).max())
        assert abs((v * v).sum() - fx["sumsq"][r]) <= 1e-10 * len(v) * max(1.0, np.abs(v).max()) ** 2
    td = nrgio.read_table(os.path.join(wd, "td"))
    assert td.shape == fx["td"].shape
    np.testing.assert_allclose(td, fx["td"], rtol=1e-9, atol=5e-10)
    cu = nrgio.read_table(os.path.join(wd, "custom"))
    np.testing.assert_allclose(cu, fx["custom"], rtol=1e-9, atol=1e-12)
    E, W = nrgio.read_bins(os.path.join(wd, "spec_FDM_dens_A_d-A_d.bin"))
    np.testing.assert_allclose(E, fx["fdm_E"], rtol=1e-15)
    np.testing.assert_allclose(W, fx["fdm_W"], rtol=1e-8, atol=1e-14 * np.abs(fx["fdm_W"]).max())


def test_kondo3ch_qsz_keep12000_first_shell(tmp_path):
    wd = str(tmp_path)
    nrgio.make_case(wd, model="kondo", sym="QSZ", nch=3, Nmax=1, Lambda=4.0, model_kw=dict(J=0.4), keep=12000,
                    T=1e-6)
    print("case kondo3ch QSZ keep12000 Nmax1", flush=True)
    nrgio.run_engine(NRG, wd)
    o = load_oracle().Oracle(wd).run()
    nrgio.compare_with_oracle(wd, o)


def test_served_requests_match_oracle(tmp_path):
    from nrgamd import serve
    import bench_params
    out = str(tmp_path)
    reqs = serve.make_requests(out, nreq=1, seed=bench_params.SERVE_SEED, rate=None, keep_range=(1000, 8000),
                               zs=(0.25, 1.0))
    print("case served request", reqs[0], flush=True)
    rep = serve.run(out, reqs, slots=4)
    assert rep["failed"] == 0
    for d in reqs[0]["members"]:
        o = load_oracle().Oracle(d).run()
        nrgio.compare_with_oracle(d, o, spectra=[("FDM", "A_d", "A_d", 0)], eigen=False)
