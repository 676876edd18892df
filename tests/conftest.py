import importlib
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "nrgljubljana-old_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)
if os.path.join(REPO, "tests") not in sys.path:
    sys.path.insert(0, os.path.join(REPO, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def load_nrgamd():
    return importlib.import_module("nrgamd")


def load_oracle():
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    return importlib.import_module("oracle.nrg_oracle")
