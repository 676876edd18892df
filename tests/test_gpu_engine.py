"""The drop-in `nrg` (host + libnrg_gpu.so on the GPU) against the oracle.

Bar (from the task statement): eigenvalues within 1e-10 in units of omega_N,
identical kept-multiplet counts at every iteration, thermodynamics and
spectral bins within 1e-8 relative.  Cases cover QS one- and two-channel
SIAM with DM-NRG and FDM spectra, QSZ with spin-resolved spectra, and the
single-channel configuration of the BASELINE (Lambda=2, keep<=500,
Ecut=8, Nmax=30, 50 bins/decade, T=1e-8, DM-NRG) at full size.
"""
import os

import pytest

from conftest import PKG, load_oracle
import nrgio
from test_host_mock import CASES, _spectra

pytestmark = pytest.mark.gpu

NRG = os.path.join(PKG, "bin", "nrg")


def _run(wd, c):
    nrgio.make_case(wd, sym=c["sym"], nch=c["nch"], Nmax=c["Nmax"], Lambda=c.get("Lambda", 2.0),
                    model_kw=c.get("model_kw"), **c["param"])
    print("case", wd, c, flush=True)
    nrgio.run_engine(NRG, wd)
    o = load_oracle().Oracle(wd).run()
    nrgio.compare_with_oracle(wd, o, spectra=_spectra(c["param"]))


@pytest.mark.parametrize("case", sorted(CASES))
def test_engine_small_cases(tmp_path, case):
    _run(str(tmp_path), CASES[case])


def test_engine_siam_keep500_dmnrg(tmp_path):
    c = dict(sym="QS", nch=1, Nmax=30, model_kw=dict(U=1.0, eps_d=-0.5, Gamma=0.04),
             param=dict(keep=500, keepenergy=8.0, T=1e-8, dmnrg=True, specd="A_d-A_d", bins=50, savebins=True,
                        broaden=True))
    _run(str(tmp_path), c)


def test_engine_siam_fdm_keep300(tmp_path):
    c = dict(sym="QS", nch=1, Nmax=24, model_kw=dict(U=1.0, eps_d=-0.5, Gamma=0.04),
             param=dict(keep=300, keepenergy=10.0, T=1e-6, fdm=True, specd="A_d-A_d", bins=50, savebins=True))
    _run(str(tmp_path), c)


@pytest.mark.parametrize("pageable", [False, True])
def test_engine_spill_to_host_matches_oracle(tmp_path, pageable):
    """Eigenvectors spilled to host memory after every iteration (pinned, or
    pageable as once pinned memory runs out) and brought back for the DM/FDM
    sweeps and pass 2 give the oracle's results."""
    import json
    c = dict(sym="QS", nch=1, Nmax=20, model_kw=dict(U=1.0, eps_d=-0.5, Gamma=0.04),
             param=dict(keep=300, keepenergy=8.0, T=1e-6, fdm=True, dmnrg=True, specd="A_d-A_d", bins=50,
                        savebins=True))
    wd = str(tmp_path)
    nrgio.make_case(wd, sym=c["sym"], nch=c["nch"], Nmax=c["Nmax"], model_kw=c["model_kw"], **c["param"])
    print("case spill-all pageable=%s" % pageable, wd, flush=True)
    nrgio.run_engine(NRG, wd, extra=("--spill-all",) + (("--spill-pageable",) if pageable else ()))
    st = json.load(open(os.path.join(wd, "stats.json")))
    assert st["spills"] > 0 and st["restores"] > 0
    o = load_oracle().Oracle(wd).run()
    nrgio.compare_with_oracle(wd, o, spectra=_spectra(c["param"]))
