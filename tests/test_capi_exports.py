"""The C-ABI library loads and exports every entry point include/nrg_gpu.h declares.

No compute calls: this runs without a GPU.
"""
import ctypes
import os
import re

from conftest import REPO, PKG


def _declared():
    with open(os.path.join(REPO, "include", "nrg_gpu.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^(?:nrg_status|void|int|const char \*)\s*\**(nrg_\w+)\s*\(", src, re.M)))


def test_header_declares_the_path():
    names = _declared()
    for n in ("nrg_assemble_batch", "nrg_diag_batch", "nrg_gemm_grouped", "nrg_spec_bin", "nrg_ctx_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = os.path.join(PKG, "lib", "libnrg_gpu.so")
    assert os.path.exists(lib), "build the library first (__graft_entry__.build())"
    h = ctypes.CDLL(lib)
    missing = [n for n in _declared() if not hasattr(h, n)]
    assert not missing, missing


def test_binding_lists_match_header():
    import nrgamd.capi as capi
    assert sorted(capi.EXPORTS) == _declared()
