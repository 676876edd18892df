/* nrg_gpu.h - C ABI of the MI355X NRG hot path (libnrg_gpu.so).
 *
 * The reference NRG Ljubljana back end is one C++ program; its hot path is
 * reached through four internal seams, each of which has a batched,
 * device-resident replacement here:
 *
 *   Matrix nrg_prepare_task_for_diag(const Invar&)   src/nrg.cc:2456-2478
 *     + Symmetry::makematrix / offdiag_build          src/sym-QS.cc:153-209,
 *                                                     src/matrix.cc:86-192
 *       -> nrg_assemble_batch
 *   Eigen diagonalise(Matrix&)   (LAPACK dsyev)       src/diag.h:297-353
 *       -> nrg_diag_batch
 *   recalc_f / recalc_general / cdmI  (BLAS dgemm)    src/recalc.cc:12-97,
 *                                                     164-281, src/dmnrg.h:327-388
 *       -> nrg_gemm_grouped
 *   SPEC_DMNRG::calc, SPEC_FDMgt/ls::calc +
 *     ChainSpectrumBinning::add / Bins::add_std       src/spec_DMNRG.cc:13-52,
 *                                                     src/spec_FDM.cc:41-119,
 *                                                     src/nrg.cc:1017-1022,
 *                                                     src/bins.h:102-131
 *       -> nrg_spec_bin
 *
 * Conventions: all matrices are row-major FP64 in device memory (the
 * reference's uBLAS matrices are row-major); device pointers are plain
 * addresses returned by nrg_dev_alloc.  Every batched call takes a host array
 * of descriptors, copies it to the device and launches asynchronously on the
 * context's stream; nrg_sync waits.  Errors are returned as negative status
 * codes with a message from nrg_last_error (the reference calls exit(1)).
 * Calls on one context must be serialised by the caller.
 */
#ifndef NRG_GPU_H
#define NRG_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nrg_ctx nrg_ctx;
typedef int nrg_status;

#define NRG_OK 0
#define NRG_ERR_ARG (-1)
#define NRG_ERR_HIP (-2)
#define NRG_ERR_NOMEM (-3)
#define NRG_ERR_NOCONV (-4)

/* ---- context, memory ---------------------------------------------------- */

/* Creates a context on HIP device `device` with its own stream. */
nrg_status nrg_ctx_create(nrg_ctx **ctx, int device);
void nrg_ctx_destroy(nrg_ctx *ctx);
const char *nrg_last_error(const nrg_ctx *ctx);
nrg_status nrg_sync(nrg_ctx *ctx);
/* No counterpart in the reference (its LAPACK calls allocate per call):
 * waits for the stream, then returns the context's eigensolver workspace and
 * grouped-GEMM partial buffer to the device.  The next call that needs them
 * allocates them again.  For callers done with eigensolves for a while (the
 * density-matrix and spectral passes), whose HBM goes to other data. */
nrg_status nrg_ctx_release_workspace(nrg_ctx *ctx);

nrg_status nrg_dev_alloc(nrg_ctx *ctx, size_t bytes, void **dptr);
nrg_status nrg_dev_free(nrg_ctx *ctx, void *dptr);
/* Synchronous copies (the host side of the path is small vectors only). */
nrg_status nrg_h2d(nrg_ctx *ctx, void *dst, const void *src, size_t bytes);
nrg_status nrg_d2h(nrg_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Strided device->host gather: `count` doubles at src[k*stride]. */
nrg_status nrg_d2h_strided(nrg_ctx *ctx, double *dst, const double *src, size_t count, size_t stride);
/* Pinned host memory (for spilling device data that does not fit in HBM, as
 * the reference writes workdir/unitaryN to disk, src/dmnrg.h:198-304);
 * nrg_h2d / nrg_d2h copy to and from it at full link speed. */
nrg_status nrg_host_alloc(nrg_ctx *ctx, size_t bytes, void **hptr);
nrg_status nrg_host_free(nrg_ctx *ctx, void *hptr);
/* Free and total device memory of the context's device (bytes); used to admit
 * solves by HBM footprint.  No counterpart in the reference. */
nrg_status nrg_mem_info(nrg_ctx *ctx, size_t *free_bytes, size_t *total_bytes);
/* Fills `count` doubles with zero on the stream (a kernel, not a memset node). */
nrg_status nrg_zero(nrg_ctx *ctx, double *dst, size_t count);

/* Batched transfers of many small strided device vectors (eigenvalues of all
 * subspaces, density-matrix and operator diagonals): item k is `count`
 * doubles at ptr[0], ptr[stride], ...; the host side is all items back to
 * back.  One gather/scatter kernel per call.  The gather returns once the
 * values are on the host (it waits for the stream); the scatter is queued
 * like a launch (the values travel with its staged descriptors, so `host`
 * may be reused when it returns).  No counterpart in the reference, which
 * keeps these vectors in host memory. */
typedef struct {
    double *ptr;
    int64_t count;
    int64_t stride;
} nrg_strided;
nrg_status nrg_gather_d2h(nrg_ctx *ctx, const nrg_strided *items, int nitems, double *host);
/* The same gather into device memory `dst`, queued like a launch: values
 * that the host needs only later (a trace per iteration) are collected on
 * the device and read back together. */
nrg_status nrg_gather_dev(nrg_ctx *ctx, const nrg_strided *items, int nitems, double *dst);
nrg_status nrg_scatter_h2d(nrg_ctx *ctx, const nrg_strided *items, int nitems, const double *host);

/* ---- assembly ----------------------------------------------------------- */
/* Replaces nrg_prepare_task_for_diag + makematrix (src/nrg.cc:2456-2478,
 * src/matrix.cc:86-192).  Each block is a dense symmetric d x d matrix that
 * is zeroed and then receives its terms in order:
 *   NRG_TERM_DIAGVEC : H[r0+r][r0+r] += scale * src[r],        r < rows
 *   NRG_TERM_BLOCK   : H[r0+r][c0+c] += scale * src[r*ld+c] and the mirrored
 *                      H[c0+c][r0+r] (r0 != c0), r < rows, c < cols
 *   NRG_TERM_SHIFT   : H[r0+r][r0+r] += scale,                 r < rows   */
#define NRG_TERM_DIAGVEC 0
#define NRG_TERM_BLOCK 1
#define NRG_TERM_SHIFT 2

typedef struct {
    double *H;
    int32_t d;
    int32_t term_begin;
    int32_t term_count;
    int32_t pad;
} nrg_asm_block;

typedef struct {
    const double *src;
    double scale;
    int32_t kind;
    int32_t r0, c0, rows, cols, ld;
} nrg_asm_term;

nrg_status nrg_assemble_batch(nrg_ctx *ctx, const nrg_asm_block *blocks, int nblocks,
                              const nrg_asm_term *terms, int nterms);

/* ---- eigensolver -------------------------------------------------------- */
/* Replaces diagonalise() with dsyev (src/diag.h:54-89, 297-353) for a batch
 * of independent symmetric matrices.  On entry A (n x n, row-major, leading
 * dimension n) holds the full symmetric matrix; on exit row r of A is the
 * normalised eigenvector for W[r], with W ascending (the layout of
 * Eigen::matrix0 / Eigen::value, src/nrg.cc:285-376).  Workspace is internal
 * (3 n^2 + O(n) doubles per matrix).  n is limited to 32768; larger matrices
 * return NRG_ERR_ARG before any device access.  NRG_ERR_NOCONV is reserved (the divide-and-conquer
 * solver has no failing iteration).  Queued like the other batched calls: a
 * tridiagonalisation team that does not assemble within its bounded wait is
 * reported (NRG_ERR_HIP) by the next call that waits for the stream
 * (nrg_sync, the transfers to the host). */
typedef struct {
    double *A;
    double *W;
    int32_t n;
    int32_t pad;
} nrg_eig_task;

nrg_status nrg_diag_batch(nrg_ctx *ctx, const nrg_eig_task *tasks, int ntasks);

/* No counterpart in the reference (diagonalise() returns values and vectors
 * together, src/diag.h:297-353): after nrg_diag_batch, waits until that
 * batch's eigenvalues are final — before its eigenvectors, whose
 * back-transformation is still running — and copies them to host, the W of
 * each task concatenated in the order the tasks were passed (count = sum of
 * the n).  The caller can truncate (nrg_truncate_prepare, src/nrg.cc:1919-1998)
 * while the GPU finishes the eigenvectors.  Reports the batch's device-side
 * status like a wait; NRG_ERR_ARG without a queued batch or for another count. */
nrg_status nrg_diag_values(nrg_ctx *ctx, double *host, size_t count);

/* ---- grouped GEMM ------------------------------------------------------- */
/* Replaces the dgemm calls of recalc_f (src/recalc.cc:90-91), recalc_general
 * (src/recalc.cc:268-272) and cdmI (src/dmnrg.h:382-387) for a whole
 * iteration at once.  Output o is
 *   C_o = beta_o * C_o + sum_{t in terms of o} alpha_t * op(A_t) diag(s_t) op(B_t)
 * with op(X) = X or X^T (row-major), s_t an optional length-k scale vector
 * (NULL = identity; used for diagonal density matrices).  Outputs must not
 * alias each other; each output is accumulated by exactly one workgroup per
 * tile, so the result is deterministic. */
typedef struct {
    double *C;
    int32_t m, n, ldc;
    int32_t term_begin, term_count;
    int32_t pad;
    double beta;
} nrg_gemm_out;

typedef struct {
    const double *A;
    const double *B;
    const double *s;
    double alpha;
    int32_t k, lda, ldb;
    int32_t transA, transB;
    int32_t pad;
} nrg_gemm_term;

nrg_status nrg_gemm_grouped(nrg_ctx *ctx, const nrg_gemm_out *outs, int nouts,
                            const nrg_gemm_term *terms, int nterms);

/* No counterpart in the reference.  nrg_gemm_grouped for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and results, with the
 * launch's descriptors and the partial tiles of split outputs in a fixed
 * device array of NRG_NOALLOC_SCRATCH_BYTES per device instead of the
 * context's staging ring and partial-tile buffer.  The caller passes inputs
 * and outputs as for nrg_gemm_grouped; no hipMalloc, hipFree or other
 * allocation is made.  A launch needing more scratch returns NRG_ERR_ARG
 * with the sizes in nrg_last_error.  Every call on one device shares that
 * device's array: calls on one context are ordered by its stream, calls that
 * overlap on one device from different contexts race. */
#define NRG_NOALLOC_SCRATCH_BYTES (16u << 20)
nrg_status nrg_gemm_grouped_noalloc(nrg_ctx *ctx, const nrg_gemm_out *outs, int nouts,
                                    const nrg_gemm_term *terms, int nterms);

/* ---- spectral weights + binning ----------------------------------------- */
/* Replaces the delta-peak loops of SPEC_FDMgt::calc / SPEC_FDMls::calc
 * (src/spec_FDM.cc:41-119) and SPEC_DMNRG::calc (src/spec_DMNRG.cc:13-52)
 * together with ChainSpectrumBinning::add -> Bins::add (src/nrg.cc:1017-1022,
 * src/bins.h:102-131).  One pair = one coupled subspace pair (Ij, Ii) of
 * calc_generic (src/spec.cc:26-62); element (j, i) of the pair contributes
 * peaks at energy Ej[j] - Ei[i] (times escale) to the positive or negative
 * histogram of `bins` (pos[0..nbins), neg[nbins..2 nbins)), accumulated.
 *
 * mode NRG_SPEC_FDM (energies absolute, T = temperature, wnf = w_N / Z_N^D):
 *   i >= reti            : + sf*op1*op2*wnf*exp(-Ei/T)      (A3, A2)
 *   i <  reti, j >= retj : + sf*op1*M1                       (A1, M1 = op2.rho_i)
 *   j >= retj            : + sf*op1*op2*wnf*exp(-Ej/T)       (B3, B1)
 *   j <  retj, i >= reti : + sf*op1*M2                       (B2, M2 = rho_j.op2)
 * mode NRG_SPEC_DMNRG (energies in units of the shell scale, escale = scale):
 *   |Ej-Ei| in [emin_w, emax_w]: + sf*(op1*M1 + op2*M2), M1 = op2.rho_p,
 *   M2 = rho_1.op1
 * All operands are row-major with the given leading dimensions; M1/M2 may be
 * NULL when unused.  Peaks with |w| < discard*|E| are dropped
 * (P::DISCARD_IMMEDIATELY).  The grid is x_k = 10^(log10emin + k/bpd). */
#define NRG_SPEC_FDM 0
#define NRG_SPEC_DMNRG 1

typedef struct {
    const double *op1, *op2, *M1, *M2;
    const double *Ei, *Ej;
    int32_t alli, allj, reti, retj;
    int32_t ld1, ld2, ldm1, ldm2;
    double sf;
} nrg_spec_pair;

typedef struct {
    int32_t mode;
    int32_t nbins;
    double log10emin;
    double bins_per_decade;
    double discard;
    double T;
    double wnf;
    double escale;
    double emin_w, emax_w;
} nrg_spec_params;

nrg_status nrg_spec_bin(nrg_ctx *ctx, const nrg_spec_params *prm, const nrg_spec_pair *pairs,
                        int npairs, double *bins);

/* ---- measurement (no counterpart in the reference) --------------------- */
/* Process-wide per-kernel accounting for benchmarks and profiling: when
 * enabled, every launch is bracketed by HIP events on the stream it runs on,
 * and the algorithmic work of the launch (flops of the FP64 arithmetic the
 * reference performs, bytes of operands it must read/write) is recorded.
 * Event times are folded in at nrg_sync / nrg_timing_get. */
typedef struct {
    char name[32];
    int64_t launches;
    double ms;     /* summed event time */
    double flops;  /* summed algorithmic flops */
    double bytes;  /* summed algorithmic bytes */
} nrg_kernel_stat;

void nrg_timing_enable(int on);
/* Records events only for launches of the named kernel (NULL or "" = all),
 * so that a timed run carries the events of the kernel under study alone. */
void nrg_timing_filter(const char *kernel);
void nrg_timing_reset(void);
/* Fills up to `max` entries, returns the number of kernels with launches. */
int nrg_timing_get(nrg_kernel_stat *out, int max);

#ifdef __cplusplus
}
#endif
#endif
