"""Host driver (C++) against the oracle, with the C ABI served by a CPU stand-in.

These tests exercise everything the host decides (subspace lists, truncation,
thermodynamics, recalculation and density-matrix task lists, spectrum
merging, output formats) without a GPU.  The GPU-side arithmetic is tested
separately in test_gpu_*.py through the real library.
"""
import os

import pytest

from conftest import load_oracle
import nrgio

CASES = {
    "qs_fdm": dict(sym="QS", nch=1, Nmax=10, param=dict(keep=80, keepenergy=8.0, T=0.01, fdm=True,
                                                       specd="A_d-A_d", bins=50, savebins=True)),
    "qs_dmnrg": dict(sym="QS", nch=1, Nmax=10, param=dict(keep=80, keepenergy=8.0, T=1e-3, dmnrg=True,
                                                         specd="A_d-A_d", bins=50, savebins=True)),
    "qsz_fdm": dict(sym="QSZ", nch=1, Nmax=8, param=dict(keep=80, keepenergy=8.0, T=0.01, fdm=True,
                                                        specd="A_d-A_d-u A_d-A_d-d", bins=50, savebins=True)),
    # the optional dump files: annotated.dat, subspaces.dat, energies.nrg /
    # energies.dmnrg and the FDM expectation values (customfdm)
    "qs_dumps": dict(sym="QS", nch=1, Nmax=9, param=dict(keep=60, keepenergy=7.0, T=0.02, fdm=True, dmnrg=True,
                                                        specd="A_d-A_d", bins=50, savebins=True,
                                                        dumpannotated=25, dumpsubspaces=True, dumpenergies=True,
                                                        fdmexpv=True, fdmexpvn=2)),
    # cluster snapping with a coarse fixeps: many clusters snap, the mapping
    # carries over iterations and passes, the top cluster stays open
    "qs_fixeps": dict(sym="QS", nch=1, Nmax=10, param=dict(keep=80, keepenergy=8.0, T=0.01, fdm=True, dmnrg=True,
                                                          specd="A_d-A_d", bins=50, savebins=True, fixeps=2e-3)),
    "qs2ch_fdm": dict(sym="QS", nch=2, Nmax=5, Lambda=3.0,
                      model_kw=dict(U=0.5, eps_d=-0.25, Gamma=0.02),
                      param=dict(keep=120, keepenergy=6.0, T=0.01, fdm=True, specd="A_d-A_d", bins=50,
                                 savebins=True)),
}


def _spectra(p):
    out = []
    for name in p["specd"].split():
        parts = name.split("-")
        spin = {"u": 1, "d": -1}.get(parts[2], 0) if len(parts) == 3 else 0
        for kind in (["DMNRG"] if p.get("dmnrg") else []) + (["FDM"] if p.get("fdm") else []):
            out.append((kind, parts[0], parts[1], spin))
    return out


@pytest.mark.parametrize("case", sorted(CASES))
def test_host_matches_oracle(tmp_path, case):
    c = CASES[case]
    wd = str(tmp_path)
    nrgio.make_case(wd, sym=c["sym"], nch=c["nch"], Nmax=c["Nmax"], Lambda=c.get("Lambda", 2.0),
                    model_kw=c.get("model_kw"), **c["param"])
    nrgio.run_engine(nrgio.build_mock(), wd)
    o = load_oracle().Oracle(wd).run()
    nrgio.compare_with_oracle(wd, o, spectra=_spectra(c["param"]))
    assert os.path.exists(os.path.join(wd, "DONE"))


def test_param_duplicate_is_fatal(tmp_path):
    wd = str(tmp_path)
    nrgio.make_case(wd, Nmax=4, keep=50, T=0.1)
    with open(os.path.join(wd, "param"), "a") as f:
        f.write("keep=60\n")
    with pytest.raises(RuntimeError):
        nrgio.run_engine(nrgio.build_mock(), wd)


def test_missing_table_entry_is_an_error(tmp_path):
    wd = str(tmp_path)
    nrgio.make_case(wd, Nmax=4, keep=50, T=0.1)
    with open(os.path.join(wd, "data")) as f:
        txt = f.read()
    # a data file whose Nmax disagrees with its coefficient tables
    with open(os.path.join(wd, "param"), "a") as f:
        f.write("Nmax=6\n")
    with pytest.raises(RuntimeError):
        nrgio.run_engine(nrgio.build_mock(), wd)


@pytest.mark.parametrize("case", ["qs_dmnrg", "qs_fdm"])
def test_spill_to_host_changes_nothing(tmp_path, case):
    """With a zero device budget every iteration's eigenvectors go to host
    memory after use (pinned, or pageable as after pinned memory runs out)
    and come back for the sweeps and pass 2: all outputs are identical to the
    resident run."""
    import json
    c = CASES[case]
    outs = []
    for tag, extra in (("resident", ()), ("spilled", ("--spill-all",)),
                       ("pageable", ("--spill-all", "--spill-pageable"))):
        wd = str(tmp_path / tag)
        os.makedirs(wd)
        nrgio.make_case(wd, sym=c["sym"], nch=c["nch"], Nmax=c["Nmax"], **c["param"])
        nrgio.run_engine(nrgio.build_mock(), wd, extra=extra)
        st = json.load(open(os.path.join(wd, "stats.json")))
        files = sorted(f for f in os.listdir(wd) if f.startswith(("td", "custom", "spec_")))
        outs.append((st, {f: open(os.path.join(wd, f), "rb").read() for f in files}))
    (st0, f0), (st1, f1), (st2, f2) = outs
    assert st0["spills"] == 0 and st1["spills"] > 0 and st1["restores"] > 0 and st2["spills"] == st1["spills"]
    assert f0.keys() == f1.keys() == f2.keys() and len(f0) >= 3
    for f in f0:
        assert f0[f] == f1[f] == f2[f], f


def _run_rc(wd, extra=()):
    import subprocess
    nrgio.ensure_tables()
    env = dict(os.environ, NRG_TABLES=nrgio.TABLES)
    return subprocess.run([nrgio.build_mock()] + list(extra), cwd=wd, env=env, capture_output=True, text=True,
                          timeout=300)


def test_forcestop_stops_at_the_iteration(tmp_path):
    """forcestop=N (src/nrg.cc:3113-3114): status 1 with the reference's
    message at the start of iteration N; td holds the rows before it."""
    wd = str(tmp_path)
    nrgio.make_case(wd, Nmax=8, keep=60, keepenergy=7.0, T=0.05, fdm=True, specd="A_d-A_d", forcestop=4)
    r = _run_rc(wd)
    assert r.returncode == 1 and "*** Stop request at iteration 4" in r.stdout
    td = nrgio.read_table(os.path.join(wd, "td"))
    assert len(td) == 1 + 4  # calc0 and iterations 0..3
    assert not os.path.exists(os.path.join(wd, "DONE"))
    assert not any(f.startswith("spec_") for f in os.listdir(wd))


@pytest.mark.parametrize("what,msg", [("nrg", "*** Stopped after the first sweep."),
                                      ("rho", "*** Stopped after the DM calculation.")])
def test_stopafter(tmp_path, what, msg):
    """stopafter=nrg|rho (src/nrg.cc:3453, 3462): the first pass completes
    (td, custom, subspaces.dat), no spectra, status 1."""
    wd = str(tmp_path)
    nrgio.make_case(wd, Nmax=7, keep=60, keepenergy=7.0, T=0.05, fdm=True, specd="A_d-A_d", stopafter=what,
                    dumpsubspaces=True)
    r = _run_rc(wd)
    assert r.returncode == 1 and msg in r.stdout
    assert len(nrgio.read_table(os.path.join(wd, "td"))) == 1 + 7
    assert os.path.exists(os.path.join(wd, "subspaces.dat"))
    assert not os.path.exists(os.path.join(wd, "DONE"))
    assert not any(f.startswith("spec_") for f in os.listdir(wd))


def test_error_returns_status_without_exit(tmp_path):
    """An input error is reported with status 1 and the message (the
    engine library throws; nrg_main catches)."""
    wd = str(tmp_path)
    nrgio.make_case(wd, Nmax=4, keep=50, T=0.1)
    with open(os.path.join(wd, "data"), "w") as f:
        f.write("# not a data file\n")
    r = _run_rc(wd)
    assert r.returncode == 1 and "Error:" in r.stderr and "version" in r.stderr


def test_tuning_table_read_and_checked(tmp_path):
    # tables/tune.txt (written by nrgamd.autotune on the device) is read at
    # start: known settings go to the environment, anything else is an error
    import shutil
    import subprocess
    wd = str(tmp_path / "run")
    os.makedirs(wd)
    nrgio.make_case(wd, sym="QS", nch=1, Nmax=4, keep=40, keepenergy=6.0, T=0.05)
    nrgio.ensure_tables()
    tables = str(tmp_path / "tables")
    shutil.copytree(nrgio.TABLES, tables)
    binary = nrgio.build_mock()
    env = dict(os.environ, NRG_TABLES=tables)
    with open(os.path.join(tables, "tune.txt"), "w") as f:
        f.write("# measured\nNRG_BLK_MIN_N 1280  # two-stage from here\nNRG_GEMM_MFMA_MIN_TILES 256\n")
    r = subprocess.run([binary], cwd=wd, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with open(os.path.join(tables, "tune.txt"), "w") as f:
        f.write("NRG_NO_SUCH_KNOB 3\n")
    r = subprocess.run([binary], cwd=wd, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "unknown setting NRG_NO_SUCH_KNOB" in r.stderr
