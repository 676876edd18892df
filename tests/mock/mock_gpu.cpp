// mock_gpu.cpp - CPU stand-in for the nrg_gpu.h C ABI.
//
// TEST INFRASTRUCTURE ONLY: it lets the CPU test suite exercise the host
// driver (task lists, truncation, recalculation descriptors, density-matrix
// sweeps, output files) without a GPU.  It is never linked into the product;
// the product `nrg` links libnrg_gpu.so and fails if that cannot be loaded.
// Arithmetic is plain loops; the eigensolver is cyclic Jacobi.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "nrg_gpu.h"

struct nrg_ctx {
    std::string err;
};

extern "C" {

nrg_status nrg_ctx_create(nrg_ctx **c, int) {
    *c = new nrg_ctx();
    return NRG_OK;
}
void nrg_ctx_destroy(nrg_ctx *c) { delete c; }
const char *nrg_last_error(const nrg_ctx *c) { return c->err.c_str(); }
nrg_status nrg_sync(nrg_ctx *) { return NRG_OK; }
nrg_status nrg_ctx_release_workspace(nrg_ctx *) { return NRG_OK; }
nrg_status nrg_dev_alloc(nrg_ctx *, size_t bytes, void **p) {
    *p = std::calloc(std::max<size_t>(bytes, 8), 1);
    return *p ? NRG_OK : NRG_ERR_NOMEM;
}
nrg_status nrg_dev_free(nrg_ctx *, void *p) {
    std::free(p);
    return NRG_OK;
}
nrg_status nrg_h2d(nrg_ctx *, void *d, const void *s, size_t b) {
    if (b) std::memcpy(d, s, b);
    return NRG_OK;
}
nrg_status nrg_d2h(nrg_ctx *, void *d, const void *s, size_t b) {
    if (b) std::memcpy(d, s, b);
    return NRG_OK;
}
nrg_status nrg_d2h_strided(nrg_ctx *, double *d, const double *s, size_t n, size_t st) {
    for (size_t i = 0; i < n; ++i) d[i] = s[i * st];
    return NRG_OK;
}
nrg_status nrg_gather_d2h(nrg_ctx *, const nrg_strided *it, int n, double *host) {
    for (int k = 0; k < n; ++k)
        for (int64_t e = 0; e < it[k].count; ++e) *host++ = it[k].ptr[e * it[k].stride];
    return NRG_OK;
}
nrg_status nrg_gather_dev(nrg_ctx *c, const nrg_strided *it, int n, double *dst) {
    return nrg_gather_d2h(c, it, n, dst);  // device memory is host memory here
}
nrg_status nrg_scatter_h2d(nrg_ctx *, const nrg_strided *it, int n, const double *host) {
    for (int k = 0; k < n; ++k)
        for (int64_t e = 0; e < it[k].count; ++e) it[k].ptr[e * it[k].stride] = *host++;
    return NRG_OK;
}
nrg_status nrg_host_alloc(nrg_ctx *, size_t bytes, void **p) {
    *p = std::malloc(std::max<size_t>(bytes, 8));
    return *p ? NRG_OK : NRG_ERR_NOMEM;
}
nrg_status nrg_host_free(nrg_ctx *, void *p) {
    std::free(p);
    return NRG_OK;
}
nrg_status nrg_mem_info(nrg_ctx *, size_t *f, size_t *t) {
    *f = size_t(1) << 34;
    *t = size_t(1) << 35;
    return NRG_OK;
}
nrg_status nrg_zero(nrg_ctx *, double *d, size_t n) {
    std::fill(d, d + n, 0.0);
    return NRG_OK;
}

nrg_status nrg_assemble_batch(nrg_ctx *, const nrg_asm_block *blocks, int nb, const nrg_asm_term *terms, int) {
    for (int b = 0; b < nb; ++b) {
        const nrg_asm_block &B = blocks[b];
        const int d = B.d;
        std::fill(B.H, B.H + (size_t)d * d, 0.0);
        for (int t = B.term_begin; t < B.term_begin + B.term_count; ++t) {
            const nrg_asm_term &T = terms[t];
            if (T.kind == NRG_TERM_DIAGVEC)
                for (int r = 0; r < T.rows; ++r) B.H[(size_t)(T.r0 + r) * d + T.r0 + r] += T.scale * T.src[r];
            else if (T.kind == NRG_TERM_SHIFT)
                for (int r = 0; r < T.rows; ++r) B.H[(size_t)(T.r0 + r) * d + T.r0 + r] += T.scale;
            else
                for (int r = 0; r < T.rows; ++r)
                    for (int c = 0; c < T.cols; ++c) {
                        const double v = T.scale * T.src[(size_t)r * T.ld + c];
                        B.H[(size_t)(T.r0 + r) * d + T.c0 + c] += v;
                        B.H[(size_t)(T.c0 + c) * d + T.r0 + r] += v;
                    }
        }
    }
    return NRG_OK;
}

static void jacobi(int n, double *A, double *W) {
    std::vector<double> a(A, A + (size_t)n * n), v((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - s * akq;
                    a[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - s * aqk;
                    a[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
                    v[(size_t)k * n + p] = c * vkp - s * vkq;
                    v[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
    for (int r = 0; r < n; ++r) {
        W[r] = a[(size_t)idx[r] * n + idx[r]];
        for (int k = 0; k < n; ++k) A[(size_t)r * n + k] = v[(size_t)k * n + idx[r]];
    }
}

static std::vector<nrg_eig_task> g_last_diag;  // for nrg_diag_values
nrg_status nrg_diag_batch(nrg_ctx *, const nrg_eig_task *t, int nt) {
    for (int i = 0; i < nt; ++i) jacobi(t[i].n, t[i].A, t[i].W);
    g_last_diag.assign(t, t + nt);
    return NRG_OK;
}

nrg_status nrg_diag_values(nrg_ctx *, double *host, size_t count) {
    size_t pos = 0;
    for (const nrg_eig_task &t : g_last_diag) {
        if (pos + t.n > count) return NRG_ERR_ARG;
        for (int r = 0; r < t.n; ++r) host[pos + r] = t.W[r];
        pos += t.n;
    }
    g_last_diag.clear();
    return pos == count ? NRG_OK : NRG_ERR_ARG;
}

nrg_status nrg_gemm_grouped(nrg_ctx *, const nrg_gemm_out *outs, int no, const nrg_gemm_term *terms, int) {
    for (int o = 0; o < no; ++o) {
        const nrg_gemm_out &O = outs[o];
        std::vector<double> acc((size_t)O.m * O.n, 0.0);
        for (int t = O.term_begin; t < O.term_begin + O.term_count; ++t) {
            const nrg_gemm_term &T = terms[t];
            for (int r = 0; r < O.m; ++r)
                for (int kk = 0; kk < T.k; ++kk) {
                    double a = T.transA ? T.A[(size_t)kk * T.lda + r] : T.A[(size_t)r * T.lda + kk];
                    if (T.s) a *= T.s[kk];
                    a *= T.alpha;
                    if (a == 0.0) continue;
                    for (int c = 0; c < O.n; ++c)
                        acc[(size_t)r * O.n + c] += a * (T.transB ? T.B[(size_t)c * T.ldb + kk] : T.B[(size_t)kk * T.ldb + c]);
                }
        }
        for (int r = 0; r < O.m; ++r)
            for (int c = 0; c < O.n; ++c) {
                double &x = O.C[(size_t)r * O.ldc + c];
                x = (O.beta == 0.0 ? 0.0 : O.beta * x) + acc[(size_t)r * O.n + c];
            }
    }
    return NRG_OK;
}

static void bin_add(double *h, const nrg_spec_params &P, double E, double w) {
    const bool pos = E >= 0;
    const double x = pos ? E : -E;
    if (std::fabs(w) < P.discard * x) return;
    double *b = pos ? h : h + P.nbins;
    const double p = (std::log10(x) - P.log10emin) * P.bins_per_decade;
    long long idx = (p > -9e18) ? (long long)p : -1;
    if (idx < 0) {
        b[0] += w;
        return;
    }
    if (idx >= P.nbins - 1) {
        b[P.nbins - 1] += w;
        return;
    }
    const double rem = p - (double)idx;
    b[idx] += (1 - rem) * w;
    b[idx + 1] += rem * w;
}

nrg_status nrg_spec_bin(nrg_ctx *, const nrg_spec_params *P, const nrg_spec_pair *pairs, int np, double *bins) {
    for (int k = 0; k < np; ++k) {
        const nrg_spec_pair &q = pairs[k];
        for (int j = 0; j < q.allj; ++j)
            for (int i = 0; i < q.alli; ++i) {
                const double o1 = q.op1[(size_t)j * q.ld1 + i], o2 = q.op2[(size_t)j * q.ld2 + i];
                const double E = (q.Ej[j] - q.Ei[i]) * P->escale;
                if (P->mode == NRG_SPEC_FDM) {
                    const bool Di = i >= q.reti, Dj = j >= q.retj;
                    if (Di) bin_add(bins, *P, E, q.sf * o1 * o2 * P->wnf * std::exp(-q.Ei[i] / P->T));
                    else if (Dj) bin_add(bins, *P, E, q.sf * o1 * q.M1[(size_t)j * q.ldm1 + i]);
                    if (Dj) bin_add(bins, *P, E, q.sf * o1 * o2 * P->wnf * std::exp(-q.Ej[j] / P->T));
                    else if (Di) bin_add(bins, *P, E, q.sf * o1 * q.M2[(size_t)j * q.ldm2 + i]);
                } else {
                    const double aE = std::fabs(q.Ej[j] - q.Ei[i]);
                    if (aE < P->emin_w || aE > P->emax_w) continue;
                    bin_add(bins, *P, E, q.sf * (q.M1[(size_t)j * q.ldm1 + i] * o1 + q.M2[(size_t)j * q.ldm2 + i] * o2));
                }
            }
    }
    return NRG_OK;
}

}  // extern "C"

extern "C" {
void nrg_timing_enable(int) {}
void nrg_timing_filter(const char *) {}
void nrg_timing_reset(void) {}
int nrg_timing_get(nrg_kernel_stat *, int) { return 0; }
}
