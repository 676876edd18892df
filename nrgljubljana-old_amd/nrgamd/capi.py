"""ctypes binding of the C ABI in include/nrg_gpu.h (libnrg_gpu.so).

Used by the GPU tests and bench.py to drive single kernels; the `nrg` binary
links the same library directly.  Loading fails loudly when the library has
not been built: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(PKG, "lib", "libnrg_gpu.so")

EXPORTS = ["nrg_ctx_create", "nrg_ctx_destroy", "nrg_last_error", "nrg_sync", "nrg_ctx_release_workspace", "nrg_dev_alloc", "nrg_dev_free",
           "nrg_h2d", "nrg_d2h", "nrg_d2h_strided", "nrg_zero", "nrg_gather_d2h", "nrg_gather_dev", "nrg_scatter_h2d", "nrg_mem_info", "nrg_host_alloc", "nrg_host_free", "nrg_assemble_batch", "nrg_diag_batch", "nrg_diag_values",
           "nrg_gemm_grouped", "nrg_gemm_grouped_noalloc", "nrg_spec_bin", "nrg_timing_enable", "nrg_timing_filter", "nrg_timing_reset",
           "nrg_timing_get"]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int64), ("ms", C.c_double), ("flops", C.c_double),
                ("bytes", C.c_double)]


def kernel_stats(lib):
    arr = (KernelStat * 64)()
    n = lib.nrg_timing_get(arr, 64)
    return {arr[i].name.decode(): dict(launches=arr[i].launches, ms=arr[i].ms, flops=arr[i].flops,
                                       bytes=arr[i].bytes) for i in range(min(n, 64))}


class AsmBlock(C.Structure):
    _fields_ = [("H", C.c_void_p), ("d", C.c_int32), ("term_begin", C.c_int32), ("term_count", C.c_int32),
                ("pad", C.c_int32)]


class AsmTerm(C.Structure):
    _fields_ = [("src", C.c_void_p), ("scale", C.c_double), ("kind", C.c_int32), ("r0", C.c_int32),
                ("c0", C.c_int32), ("rows", C.c_int32), ("cols", C.c_int32), ("ld", C.c_int32)]


class EigTask(C.Structure):
    _fields_ = [("A", C.c_void_p), ("W", C.c_void_p), ("n", C.c_int32), ("pad", C.c_int32)]


class GemmOut(C.Structure):
    _fields_ = [("C", C.c_void_p), ("m", C.c_int32), ("n", C.c_int32), ("ldc", C.c_int32),
                ("term_begin", C.c_int32), ("term_count", C.c_int32), ("pad", C.c_int32), ("beta", C.c_double)]


class GemmTerm(C.Structure):
    _fields_ = [("A", C.c_void_p), ("B", C.c_void_p), ("s", C.c_void_p), ("alpha", C.c_double), ("k", C.c_int32),
                ("lda", C.c_int32), ("ldb", C.c_int32), ("transA", C.c_int32), ("transB", C.c_int32),
                ("pad", C.c_int32)]


class SpecPair(C.Structure):
    _fields_ = [("op1", C.c_void_p), ("op2", C.c_void_p), ("M1", C.c_void_p), ("M2", C.c_void_p),
                ("Ei", C.c_void_p), ("Ej", C.c_void_p), ("alli", C.c_int32), ("allj", C.c_int32),
                ("reti", C.c_int32), ("retj", C.c_int32), ("ld1", C.c_int32), ("ld2", C.c_int32),
                ("ldm1", C.c_int32), ("ldm2", C.c_int32), ("sf", C.c_double)]


class SpecParams(C.Structure):
    _fields_ = [("mode", C.c_int32), ("nbins", C.c_int32), ("log10emin", C.c_double),
                ("bins_per_decade", C.c_double), ("discard", C.c_double), ("T", C.c_double), ("wnf", C.c_double),
                ("escale", C.c_double), ("emin_w", C.c_double), ("emax_w", C.c_double)]


def load(path=LIB):
    if not os.path.exists(path):
        raise RuntimeError("libnrg_gpu.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    for name in EXPORTS:
        getattr(lib, name)  # raises AttributeError when an export is missing
    lib.nrg_last_error.restype = C.c_char_p
    return lib


class Ctx:
    """One context on one device; device buffers as numpy-shaped handles."""

    def __init__(self, device=0, lib=None):
        self.lib = lib or load()
        self.ctx = C.c_void_p()
        self.check(self.lib.nrg_ctx_create(C.byref(self.ctx), device), "nrg_ctx_create")
        self.bufs = []

    def check(self, st, what):
        if st != 0:
            msg = self.lib.nrg_last_error(self.ctx).decode() if self.ctx else ""
            raise RuntimeError("%s failed (%d): %s" % (what, st, msg))

    def alloc(self, n):
        p = C.c_void_p()
        self.check(self.lib.nrg_dev_alloc(self.ctx, C.c_size_t(8 * max(n, 1)), C.byref(p)), "alloc")
        self.bufs.append(p.value)
        return p.value

    def put(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(a.size)
        self.check(self.lib.nrg_h2d(self.ctx, C.c_void_p(p), a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * a.size)), "h2d")
        return p

    def get(self, p, shape):
        a = np.empty(shape, dtype=np.float64)
        self.check(self.lib.nrg_d2h(self.ctx, a.ctypes.data_as(C.c_void_p), C.c_void_p(p), C.c_size_t(8 * a.size)), "d2h")
        return a

    def sync(self):
        self.check(self.lib.nrg_sync(self.ctx), "sync")

    def release_workspace(self):
        self.check(self.lib.nrg_ctx_release_workspace(self.ctx), "release_workspace")

    def close(self):
        for p in self.bufs:
            self.lib.nrg_dev_free(self.ctx, C.c_void_p(p))
        self.bufs = []
        self.lib.nrg_ctx_destroy(self.ctx)
        self.ctx = None

    # -- batched entry points --------------------------------------------
    def diag_batch(self, mats):
        """Eigen-decompose a list of symmetric matrices; returns (W, U rows)."""
        tasks = (EigTask * len(mats))()
        ptrs = []
        for t, m in zip(tasks, mats):
            n = m.shape[0]
            A = self.put(m)
            W = self.alloc(n)
            t.A, t.W, t.n = A, W, n
            ptrs.append((A, W, n))
        self.check(self.lib.nrg_diag_batch(self.ctx, tasks, len(mats)), "diag_batch")
        return [(self.get(W, (n,)), self.get(A, (n, n))) for A, W, n in ptrs]

    def diag_values(self, count):
        """Eigenvalues of the last diag batch, concatenated in task order, as
        soon as they are final (nrg_diag_values)."""
        out = np.empty(count, dtype=np.float64)
        self.check(self.lib.nrg_diag_values(self.ctx, out.ctypes.data_as(C.c_void_p), C.c_size_t(count)),
                   "diag_values")
        return out

    def gemm_grouped(self, outs, terms):
        o = (GemmOut * len(outs))(*outs)
        t = (GemmTerm * max(1, len(terms)))(*terms)
        self.check(self.lib.nrg_gemm_grouped(self.ctx, o, len(outs), t, len(terms)), "gemm_grouped")

    def gemm_grouped_noalloc(self, outs, terms):
        """The no-allocation variant (fixed device scratch array); returns the
        status instead of raising, so that callers can test its limit."""
        o = (GemmOut * len(outs))(*outs)
        t = (GemmTerm * max(1, len(terms)))(*terms)
        return self.lib.nrg_gemm_grouped_noalloc(self.ctx, o, len(outs), t, len(terms))

    def assemble(self, blocks, terms):
        b = (AsmBlock * len(blocks))(*blocks)
        t = (AsmTerm * max(1, len(terms)))(*terms)
        self.check(self.lib.nrg_assemble_batch(self.ctx, b, len(blocks), t, len(terms)), "assemble")

    def spec_bin(self, prm, pairs, dbins):
        p = (SpecPair * len(pairs))(*pairs)
        self.check(self.lib.nrg_spec_bin(self.ctx, C.byref(prm), p, len(pairs), C.c_void_p(dbins)), "spec_bin")
