// nrg_gpu.hip - CDNA4 (gfx950) kernels and C ABI of the NRG hot path.
//
// Kernels (one file, one code object):
//   k_zero          fill kernel (used instead of memset nodes)
//   k_assemble      per-subspace Hamiltonian assembly
//                   (src/nrg.cc:2456-2478, src/matrix.cc:86-192)
//   k_eig_*         batched FP64 symmetric eigensolver replacing dsyev
//                   (src/diag.h:54-89): Householder tridiagonalisation,
//                   explicit accumulation of Q, implicit QL with Wilkinson
//                   shift whose rotations are applied to Q^T rows, rank sort.
//   k_gemm          grouped FP64 GEMM with per-output accumulation lists
//                   (recalc_f, recalc_general, cdmI)
//   k_spec_bin      spectral weights of SPEC_FDM / SPEC_DMNRG fused with the
//                   logarithmic binning of Bins::add_std into an LDS histogram
//
// Every launch is on the context's stream; descriptor arrays are staged in
// a context-owned device buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nrg_gpu.h"

struct nrg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // staging for descriptor arrays
    void *desc = nullptr;
    size_t desc_bytes = 0;
    // eigensolver workspace
    double *ews = nullptr;
    size_t ews_doubles = 0;
    int *dstatus = nullptr;
};

static nrg_status fail(nrg_ctx *c, nrg_status code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(c, NRG_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------------------
// per-kernel accounting (HIP events on the launch stream)
// ---------------------------------------------------------------------------
enum KernelId { K_ZERO, K_ASSEMBLE, K_EIG, K_GEMM, K_SPEC, K_COUNT };
static const char *kKernelNames[K_COUNT] = {"k_zero", "k_assemble", "k_eig", "k_gemm", "k_spec_bin"};
struct PendingLaunch {
    int kid;
    hipEvent_t a, b;
    double flops, bytes;
};
static bool g_timing = false;
static nrg_kernel_stat g_stats[K_COUNT];
static std::vector<PendingLaunch> g_pending;

static void timing_fold() {
    for (auto &p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess)
            g_stats[p.kid].ms += ms;
        g_stats[p.kid].launches += 1;
        g_stats[p.kid].flops += p.flops;
        g_stats[p.kid].bytes += p.bytes;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    g_pending.clear();
}

struct LaunchTimer {
    bool on;
    PendingLaunch p;
    hipStream_t s;
    LaunchTimer(int kid, hipStream_t st, double flops, double bytes) : on(g_timing), s(st) {
        if (!on) return;
        p.kid = kid;
        p.flops = flops;
        p.bytes = bytes;
        (void)hipEventCreate(&p.a);
        (void)hipEventCreate(&p.b);
        (void)hipEventRecord(p.a, s);
    }
    ~LaunchTimer() {
        if (!on) return;
        (void)hipEventRecord(p.b, s);
        g_pending.push_back(p);
    }
};

static nrg_status stage(nrg_ctx *c, const void *host, size_t bytes, void **dev) {
    if (bytes > c->desc_bytes) {
        if (c->desc) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipFree(c->desc));
        }
        size_t nb = std::max<size_t>(bytes, 1 << 20);
        HIPCHK(c, hipMalloc(&c->desc, nb));
        c->desc_bytes = nb;
    }
    // The staging buffer is reused by the next call: wait for previous users.
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpyAsync(c->desc, host, bytes, hipMemcpyHostToDevice, c->stream));
    *dev = c->desc;
    return NRG_OK;
}

// ---------------------------------------------------------------------------
// block-level reduction helpers (wave64, 256 threads)
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int NT>
__device__ inline double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < NT / 64; ++k) s += red[k];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// fill
// ---------------------------------------------------------------------------
__global__ void k_zero(double *p, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t st = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_assemble(const nrg_asm_block *blocks, const nrg_asm_term *terms) {
    const nrg_asm_block b = blocks[blockIdx.x];
    const int d = b.d;
    double *H = b.H;
    const size_t dd = (size_t)d * d;
    for (size_t i = threadIdx.x; i < dd; i += blockDim.x) H[i] = 0.0;
    __syncthreads();
    for (int t = 0; t < b.term_count; ++t) {
        const nrg_asm_term tm = terms[b.term_begin + t];
        if (tm.kind == NRG_TERM_DIAGVEC) {
            for (int r = threadIdx.x; r < tm.rows; r += blockDim.x)
                H[(size_t)(tm.r0 + r) * d + tm.r0 + r] += tm.scale * tm.src[r];
        } else if (tm.kind == NRG_TERM_SHIFT) {
            for (int r = threadIdx.x; r < tm.rows; r += blockDim.x) H[(size_t)(tm.r0 + r) * d + tm.r0 + r] += tm.scale;
        } else {
            const int n = tm.rows * tm.cols;
            for (int e = threadIdx.x; e < n; e += blockDim.x) {
                const int r = e / tm.cols, cc = e - r * tm.cols;
                const double v = tm.scale * tm.src[(size_t)r * tm.ld + cc];
                H[(size_t)(tm.r0 + r) * d + tm.c0 + cc] += v;
                H[(size_t)(tm.c0 + cc) * d + tm.r0 + r] += v;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// eigensolver: one 1024-thread workgroup per matrix.
//
//  1. Householder tridiagonalisation A <- H_j A H_j (j = 0..n-3).  Step j's
//     rank-2 update of the trailing matrix is fused with step j+1's
//     matrix-vector product, so each step streams the trailing matrix once
//     (one read + one write).  Threads own columns (coalesced rows).
//  2. Q = H_0 ... H_{n-3} accumulated backwards into Z, again one fused pass
//     (update of step j + reduction of step j-1) per reflector.
//  3. A <- Q^T; implicit QL with Wilkinson shifts on (d, e).  Wave 0 computes
//     the scalar recurrence of sweep t+1 while waves 1..15 apply the rotations
//     of sweep t to the rows of A (double-buffered cosines/sines).
//  4. Stable rank sort; rows gathered through Z.
// Vectors live in LDS for n <= EIG_LDS_N, else in the global workspace.
// ---------------------------------------------------------------------------
constexpr int EIG_NT = 1024;
constexpr int EIG_NW = EIG_NT / 64;
constexpr int EIG_LDS_N = 1344;  // 6 vectors x 1344 x 8 B = 63 KB dynamic LDS
constexpr int EIG_NVEC = 12 + EIG_NW;  // vectors per matrix in the global workspace (+ row-group partials)

__host__ __device__ inline size_t eig_ws_doubles(int n) { return (size_t)n * n + EIG_NVEC * (size_t)n; }

__device__ inline double eig_block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < EIG_NW; ++k) s += red[k];
    return s;
}

// Decomposition of a pass over an (mr x mc) block whose threads own columns:
// waves take 64-column chunks, and when there are fewer chunks than waves the
// rows are split into R contiguous groups so every wave keeps loads in flight.
struct PassMap {
    int ncc, R, rpg;
};
__device__ inline PassMap pass_map(int mr, int mc) {
    PassMap p;
    p.ncc = (mc + 63) / 64;
    p.R = (p.ncc >= EIG_NW) ? 1 : EIG_NW / p.ncc;
    if (p.R > mr) p.R = mr > 0 ? mr : 1;
    p.rpg = (mr + p.R - 1) / p.R;
    return p;
}

// IN_LDS selects at compile time where the work vectors live, so that the
// LDS variant addresses them with ds_* instructions (a run-time choice would
// turn every access into a flat access queued behind the global traffic).
template <bool IN_LDS>
__device__ __forceinline__ void eig_core(const nrg_eig_task &tk, double *ws_base, const size_t ws_off, int *status,
                                         double *red, int *s_ctl, double *s_scal) {
    extern __shared__ double lds[];
    const int n = tk.n;
    double *__restrict__ A = tk.A;
    double *__restrict__ W = tk.W;
    const int tid = threadIdx.x;
    if (n <= 0) return;
    if (n == 1) {
        if (tid == 0) {
            W[0] = A[0];
            A[0] = 1.0;
        }
        return;
    }
    double *base = ws_base + ws_off;
    double *__restrict__ Z = base;
    // global vectors: tau [0,n), pv [n,2n), spare [2n,3n), then 9 more if not in LDS
    double *gv = base + (size_t)n * n;
    constexpr bool in_lds = IN_LDS;
    double *vecs = in_lds ? lds : gv + 3 * n;
    double *va = vecs;          // reflector j (length m)
    double *vb = vecs + n;      // reflector j+1 / j-1
    double *wv = vecs + 2 * n;  // w = p - K v  /  accumulation row vector
    double *dg = vecs + 3 * n;
    double *od = vecs + 4 * n;
    // QL rotation buffers reuse the reflector vectors (phases are disjoint)
    double *rc0 = va, *rs0 = vb, *rc1 = wv, *rs1 = vecs + 5 * n;
    double *tau = gv;           // n, global
    double *pv = gv + n;        // p vector (global); rank array at the end
    double *part = gv + 12 * (size_t)n;  // EIG_NW x n row-group partial sums
    const int wave = tid >> 6, lane = tid & 63;

    // ---- 1. tridiagonalisation
    bool p_ready = false;
    for (int j = 0; j < n - 2; ++j) {
        const int m = n - j - 1;
        double *xr = A + (size_t)j * n + j + 1;  // row j right of the diagonal (final)
        double *A22 = A + (size_t)(j + 1) * n + (j + 1);
        // reflector of step j (may have been prepared by the previous fused pass)
        if (!p_ready) {
            double psum = 0.0;
            for (int k = 1 + tid; k < m; k += EIG_NT) psum += xr[k] * xr[k];
            const double sigma = eig_block_sum(psum, red);
            const double x0 = xr[0];
            double t = 0.0;
            if (sigma != 0.0) {
                const double alpha = -copysign(sqrt(x0 * x0 + sigma), x0);
                const double v0 = x0 - alpha;
                t = 2.0 / (v0 * v0 + sigma);
                for (int k = tid; k < m; k += EIG_NT) va[k] = (k == 0) ? v0 : xr[k];
                if (tid == 0) od[j] = alpha;
            } else if (tid == 0) {
                od[j] = x0;
            }
            if (tid == 0) tau[j] = t;
            __syncthreads();
            if (t == 0.0) continue;  // nothing to do; next step reads row j+1 as is
            if (tid == 0) xr[0] = va[0];
            // p = tau * A22 * v (row groups -> partials -> sum)
            {
                const PassMap pm = pass_map(m, m);
                for (int ch = (pm.ncc >= EIG_NW ? wave : wave % pm.ncc); ch < pm.ncc; ch += (pm.ncc >= EIG_NW ? EIG_NW : pm.ncc * EIG_NW)) {
                    const int g = (pm.ncc >= EIG_NW) ? 0 : wave / pm.ncc;
                    if (g >= pm.R) break;
                    const int c = ch * 64 + lane;
                    if (c >= m) continue;
                    const int r0 = g * pm.rpg, r1 = min(m, r0 + pm.rpg);
                    double acc = 0.0;
                    int r = r0;
                    for (; r + 8 <= r1; r += 8) {
                        const double *q = A22 + (size_t)r * n + c;
                        double a[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) a[u] = q[(size_t)u * n];
#pragma unroll
                        for (int u = 0; u < 8; ++u) acc += a[u] * va[r + u];
                    }
                    for (; r < r1; ++r) acc += A22[(size_t)r * n + c] * va[r];
                    part[(size_t)g * n + c] = acc;
                }
                __syncthreads();
                for (int c = tid; c < m; c += EIG_NT) {
                    double acc = 0.0;
                    for (int g = 0; g < pm.R; ++g) acc += part[(size_t)g * n + c];
                    pv[c] = t * acc;
                }
                __syncthreads();
            }
        }
        const double t = tau[j];
        // K = tau/2 v.p ; w = p - K v
        double pp = 0.0;
        for (int k = tid; k < m; k += EIG_NT) pp += va[k] * pv[k];
        const double K = 0.5 * t * eig_block_sum(pp, red);
        for (int k = tid; k < m; k += EIG_NT) wv[k] = pv[k] - K * va[k];
        __syncthreads();
        // row 0 of A22 (row j+1 of A) first: it carries the next reflector
        for (int c = tid; c < m; c += EIG_NT) A22[c] -= va[0] * wv[c] + wv[0] * va[c];
        __syncthreads();
        const bool has_next = (j + 1) < n - 2;
        double tn = 0.0;
        if (has_next) {
            const int mn = m - 1;
            double *xn = A22 + 1;  // row j+1, columns j+2..
            double psum = 0.0;
            for (int k = 1 + tid; k < mn; k += EIG_NT) psum += xn[k] * xn[k];
            const double sigma = eig_block_sum(psum, red);
            const double x0 = xn[0];
            if (sigma != 0.0) {
                const double alpha = -copysign(sqrt(x0 * x0 + sigma), x0);
                const double v0 = x0 - alpha;
                tn = 2.0 / (v0 * v0 + sigma);
                for (int k = tid; k < mn; k += EIG_NT) vb[k] = (k == 0) ? v0 : xn[k];
                if (tid == 0) od[j + 1] = alpha;
            } else if (tid == 0) {
                od[j + 1] = x0;
            }
            if (tid == 0) tau[j + 1] = tn;
            __syncthreads();
            if (tid == 0 && tn != 0.0) xn[0] = vb[0];
        }
        // fused pass over rows/cols 1..m-1 of A22: update, and p' = tn * A33 v'
        const bool fuse = has_next && tn != 0.0;
        {
            const int mm = m - 1;
            const PassMap pm = pass_map(mm, mm);
            for (int ch = (pm.ncc >= EIG_NW ? wave : wave % pm.ncc); ch < pm.ncc; ch += (pm.ncc >= EIG_NW ? EIG_NW : pm.ncc * EIG_NW)) {
                const int g = (pm.ncc >= EIG_NW) ? 0 : wave / pm.ncc;
                if (g >= pm.R) break;
                const int c = 1 + ch * 64 + lane;
                if (c >= m) continue;
                const double wc = wv[c], vc = va[c];
                const int r0 = 1 + g * pm.rpg, r1 = min(m, r0 + pm.rpg);
                double acc = 0.0;
                int r = r0;
                for (; r + 8 <= r1; r += 8) {
                    double *q = A22 + (size_t)r * n + c;
                    double a[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) a[u] = q[(size_t)u * n];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        a[u] -= va[r + u] * wc + wv[r + u] * vc;
                        q[(size_t)u * n] = a[u];
                        acc += a[u] * vb[r + u - 1];
                    }
                }
                for (; r < r1; ++r) {
                    double *q = A22 + (size_t)r * n + c;
                    const double a0 = q[0] - (va[r] * wc + wv[r] * vc);
                    q[0] = a0;
                    acc += a0 * vb[r - 1];
                }
                if (fuse) part[(size_t)g * n + (c - 1)] = acc;
            }
            __syncthreads();
            if (fuse) {
                for (int c = tid; c < mm; c += EIG_NT) {
                    double acc = 0.0;
                    for (int g = 0; g < pm.R; ++g) acc += part[(size_t)g * n + c];
                    pv[c] = tn * acc;
                }
            }
        }
        __syncthreads();
        if (fuse) {
            // the next reflector becomes current
            for (int k = tid; k < m - 1; k += EIG_NT) va[k] = vb[k];
            __syncthreads();
        }
        p_ready = fuse;
    }
    if (tid == 0) {
        od[n - 2] = A[(size_t)(n - 2) * n + n - 1];
        od[n - 1] = 0.0;
    }
    for (int i = tid; i < n; i += EIG_NT) dg[i] = A[(size_t)i * n + i];
    // ---- 2. Q = H_0 ... H_{n-3}: Z = I, then backwards, one fused pass per step
    for (size_t e = tid; e < (size_t)n * n; e += EIG_NT) {
        const size_t r = e / n, cc = e - r * n;
        Z[e] = (r == cc) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (n >= 3) {
        // wt_j for the first processed reflector j = n-3 on an identity block: wt = v
        int j = n - 3;
        {
            const int m = n - j - 1;
            const double *v = A + (size_t)j * n + j + 1;
            for (int k = tid; k < m; k += EIG_NT) {
                va[k] = v[k];
                wv[k] = v[k];
            }
        }
        __syncthreads();
        for (; j >= 0; --j) {
            const int m = n - j - 1;
            const double t = tau[j];
            const bool has_prev = j >= 1;
            const double tp = has_prev ? tau[j - 1] : 0.0;
            if (has_prev) {
                const double *v = A + (size_t)(j - 1) * n + j;  // reflector j-1, length m+1
                for (int k = tid; k <= m; k += EIG_NT) vb[k] = v[k];
            }
            __syncthreads();
            double *Zb = Z + (size_t)(j + 1) * n + (j + 1);
            // thread column c: Z[j+1+r][j+1+c] -= t*va[r]*wv[c]; acc += vb[1+r]*Znew
            const PassMap pm = pass_map(m, m);
            for (int ch = (pm.ncc >= EIG_NW ? wave : wave % pm.ncc); ch < pm.ncc; ch += (pm.ncc >= EIG_NW ? EIG_NW : pm.ncc * EIG_NW)) {
                const int g = (pm.ncc >= EIG_NW) ? 0 : wave / pm.ncc;
                if (g >= pm.R) break;
                const int c = ch * 64 + lane;
                if (c >= m) continue;
                const double tw = t * wv[c];
                const int r0 = g * pm.rpg, r1 = min(m, r0 + pm.rpg);
                double acc = 0.0;
                int r = r0;
                for (; r + 8 <= r1; r += 8) {
                    double *q = Zb + (size_t)r * n + c;
                    double z[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) z[u] = q[(size_t)u * n];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        z[u] -= va[r + u] * tw;
                        if (t != 0.0) q[(size_t)u * n] = z[u];
                        acc += vb[1 + r + u] * z[u];
                    }
                }
                for (; r < r1; ++r) {
                    double *q = Zb + (size_t)r * n + c;
                    const double z0 = q[0] - va[r] * tw;
                    if (t != 0.0) q[0] = z0;
                    acc += vb[1 + r] * z0;
                }
                if (has_prev) part[(size_t)g * n + c] = acc;
            }
            __syncthreads();
            if (has_prev) {
                for (int c = tid; c < m; c += EIG_NT) {
                    double acc = 0.0;
                    for (int g = 0; g < pm.R; ++g) acc += part[(size_t)g * n + c];
                    pv[1 + c] = acc;
                }
            }
            __syncthreads();
            if (has_prev) {
                // next step: reflector j-1 becomes current, wt from this pass
                for (int k = tid; k <= m; k += EIG_NT) {
                    va[k] = vb[k];
                    wv[k] = (k == 0) ? vb[0] : pv[k];
                }
                __syncthreads();
            }
            (void)tp;
        }
    }
    // ---- 3. A <- Z^T (row i of A = column i of Q)
    for (size_t e = tid; e < (size_t)n * n; e += EIG_NT) {
        const size_t r = e / n, cc = e - r * n;
        A[cc * n + r] = Z[e];
    }
    __syncthreads();
    // ---- 4. implicit QL; wave 0 computes sweep t+1 while waves 1.. apply sweep t
    // s_ctl: [0] l, [1] iteration count, [2+b] lo of buffer b, [4+b] hi of buffer b, [6] done
    if (tid == 0) {
        s_ctl[0] = 0;
        s_ctl[1] = 0;
        s_ctl[2] = s_ctl[3] = -1;
        s_ctl[4] = s_ctl[5] = -2;
        s_ctl[6] = 0;
    }
    __syncthreads();
    const double eps = 2.220446049250313e-16;
    for (int round = 0;; ++round) {
        const int b = round & 1;
        const int done_before = s_ctl[6];
        __syncthreads();  // everyone has read the flag before thread 0 may set it
        if (wave == 0 && !done_before) {
            // the whole wave looks for the deflation point; lane 0 runs the sweep
            double *rc = b ? rc1 : rc0;
            double *rs = b ? rs1 : rs0;
            int l = s_ctl[0];
            int iter = s_ctl[1];
            int m = l, done = 0;
            for (;;) {
                if (l >= n) {
                    done = 1;
                    break;
                }
                m = n - 1;
                for (int base = l; base < n - 1; base += 64) {
                    const int idx = base + lane;
                    bool small = false;
                    if (idx < n - 1) {
                        const double dd = fabs(dg[idx]) + fabs(dg[idx + 1]);
                        small = fabs(od[idx]) <= eps * dd;
                    }
                    const unsigned long long bal = __ballot(small);
                    if (bal) {
                        m = base + __ffsll((long long)bal) - 1;
                        break;
                    }
                }
                if (m == l) {
                    ++l;
                    iter = 0;
                    continue;
                }
                if (++iter > 60) {
                    if (lane == 0) atomicExch(status, NRG_ERR_NOCONV);
                    done = 1;
                }
                break;
            }
            if (lane == 0) {
                int lo = -1, hi = -2;
                if (!done) {
                    double g = (dg[l + 1] - dg[l]) / (2.0 * od[l]);
                    double r = sqrt(g * g + 1.0);
                    g = dg[m] - dg[l] + od[l] / (g + copysign(r, g));
                    double s = 1.0, c = 1.0, p = 0.0;
                    int i;
                    bool early = false;
                    double odi = od[m - 1], dgi = dg[m - 1], dgi1 = dg[m];
                    for (i = m - 1; i >= l; --i) {
                        const double f = s * odi;
                        const double bb = c * odi;
                        // prefetch the next iteration's operands off the critical path
                        const double odn = (i > l) ? od[i - 1] : 0.0;
                        const double dgn = (i > l) ? dg[i - 1] : 0.0;
                        const double x = f * f + g * g;
                        if (x == 0.0) {
                            od[i + 1] = 0.0;
                            dg[i + 1] = dgi1 - p;
                            od[m] = 0.0;
                            early = true;
                            break;
                        }
                        // 1/sqrt with one Newton step: |c^2 + s^2 - 1| ~ 1 ulp
                        double y = __builtin_amdgcn_rsq(x);
                        y = y * (1.5 - 0.5 * x * y * y);
                        r = x * y;
                        od[i + 1] = r;
                        s = f * y;
                        c = g * y;
                        g = dgi1 - p;
                        r = (dgi - g) * s + 2.0 * c * bb;
                        p = s * r;
                        dg[i + 1] = g + p;
                        g = c * r - bb;
                        rc[i] = c;
                        rs[i] = s;
                        dgi1 = dgi;
                        dgi = dgn;
                        odi = odn;
                    }
                    hi = m - 1;
                    lo = early ? i + 1 : l;
                    if (!early) {
                        dg[l] -= p;
                        od[l] = g;
                        od[m] = 0.0;
                    }
                }
                s_ctl[0] = l;
                s_ctl[1] = iter;
                s_ctl[2 + b] = lo;
                s_ctl[4 + b] = hi;
                if (done) {
                    s_ctl[6] = 1;
                    s_ctl[2 + b] = -1;
                    s_ctl[4 + b] = -2;
                }
            }
        }
        if (tid >= 64) {
            // apply the sweep computed in the previous round
            const int pb = b ^ 1;
            const int lo = s_ctl[2 + pb], hi = s_ctl[4 + pb];
            if (round > 0 && hi >= lo && lo >= 0) {
                const double *rc = pb ? rc1 : rc0;
                const double *rs = pb ? rs1 : rs0;
                for (int k = tid - 64; k < n; k += EIG_NT - 64) {
                    double *col = A + k;
                    double up = col[(size_t)(hi + 1) * n];
                    int i = hi;
                    // 8 rows at a time: loads of the block issued before its stores
                    for (; i - 7 >= lo; i -= 8) {
                        double cur[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) cur[u] = col[(size_t)(i - u) * n];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const double c = rc[i - u], s = rs[i - u];
                            col[(size_t)(i - u + 1) * n] = s * cur[u] + c * up;
                            up = c * cur[u] - s * up;
                        }
                    }
                    for (; i >= lo; --i) {
                        const double cur = col[(size_t)i * n];
                        const double c = rc[i], s = rs[i];
                        col[(size_t)(i + 1) * n] = s * cur + c * up;
                        up = c * cur - s * up;
                    }
                    col[(size_t)lo * n] = up;
                }
            }
        }
        __syncthreads();
        if (done_before) break;  // the last computed sweep has now been applied
    }
    // ---- 5. sort ascending (stable rank), rows gathered through Z
    int *rank = reinterpret_cast<int *>(pv);
    for (int i = tid; i < n; i += EIG_NT) {
        const double di = dg[i];
        int rk = 0;
        for (int jj = 0; jj < n; ++jj) {
            const double dj = dg[jj];
            rk += (dj < di) || (dj == di && jj < i);
        }
        rank[i] = rk;
        W[rk] = di;
    }
    __syncthreads();
    for (size_t e = tid; e < (size_t)n * n; e += EIG_NT) {
        const size_t i = e / n, k = e - i * n;
        Z[(size_t)rank[i] * n + k] = A[e];
    }
    __syncthreads();
    for (size_t e = tid; e < (size_t)n * n; e += EIG_NT) A[e] = Z[e];
    (void)s_scal;
}

__global__ __launch_bounds__(EIG_NT) void k_eig_lds(const nrg_eig_task *tasks, double *ws_base, const size_t *ws_off,
                                                     int *status) {
    __shared__ double red[EIG_NW];
    __shared__ double s_scal[4];
    __shared__ int s_ctl[8];
    const nrg_eig_task tk = tasks[blockIdx.x];
    eig_core<true>(tk, ws_base, ws_off[blockIdx.x], status, red, s_ctl, s_scal);
}

__global__ __launch_bounds__(EIG_NT) void k_eig_glb(const nrg_eig_task *tasks, double *ws_base, const size_t *ws_off,
                                                     int *status) {
    __shared__ double red[EIG_NW];
    __shared__ double s_scal[4];
    __shared__ int s_ctl[8];
    const nrg_eig_task tk = tasks[blockIdx.x];
    eig_core<false>(tk, ws_base, ws_off[blockIdx.x], status, red, s_ctl, s_scal);
}

// ---------------------------------------------------------------------------
// grouped GEMM (FP64, LDS tiled, 64x64 tile, 4x4 per thread)
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 16;

struct GemmItem {
    int out, tm, tn, pad;
};

__global__ __launch_bounds__(256) void k_gemm(const nrg_gemm_out *outs, const nrg_gemm_term *terms,
                                              const GemmItem *items) {
    __shared__ double As[GK][GT + 1];
    __shared__ double Bs[GK][GT + 1];
    const GemmItem it = items[blockIdx.x];
    const nrg_gemm_out o = outs[it.out];
    const int tid = threadIdx.x;
    const int tr = tid / 16, tc = tid % 16;  // 16x16 threads, 4x4 each
    const int row0 = it.tm * GT, col0 = it.tn * GT;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int t = 0; t < o.term_count; ++t) {
        const nrg_gemm_term tm = terms[o.term_begin + t];
        for (int k0 = 0; k0 < tm.k; k0 += GK) {
            // load A tile: rows row0..+64, k k0..+16 (1024 elements, 4 per thread)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = tid + q * 256;
                int r, kk;
                if (tm.transA) {  // contiguous along rows of op(A)
                    kk = e / GT;
                    r = e % GT;
                } else {
                    r = e / GK;
                    kk = e % GK;
                }
                const int gr = row0 + r, gk = k0 + kk;
                double v = 0.0;
                if (gr < o.m && gk < tm.k) {
                    v = tm.transA ? tm.A[(size_t)gk * tm.lda + gr] : tm.A[(size_t)gr * tm.lda + gk];
                    if (tm.s) v *= tm.s[gk];
                }
                As[kk][r] = v;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = tid + q * 256;
                int cidx, kk;
                if (tm.transB) {
                    cidx = e / GK;
                    kk = e % GK;
                } else {
                    kk = e / GT;
                    cidx = e % GT;
                }
                const int gc = col0 + cidx, gk = k0 + kk;
                double v = 0.0;
                if (gc < o.n && gk < tm.k) v = tm.transB ? tm.B[(size_t)gc * tm.ldb + gk] : tm.B[(size_t)gk * tm.ldb + gc];
                Bs[kk][cidx] = v * tm.alpha;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < GK; ++kk) {
                double a[4], b[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a[q] = As[kk][tr + 16 * q];
                    b[q] = Bs[kk][tc + 16 * q];
                }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[x][y] = fma(a[x], b[y], acc[x][y]);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
        const int gr = row0 + tr + 16 * x;
        if (gr >= o.m) continue;
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int gc = col0 + tc + 16 * y;
            if (gc >= o.n) continue;
            double *cp = o.C + (size_t)gr * o.ldc + gc;
            *cp = (o.beta == 0.0 ? 0.0 : o.beta * *cp) + acc[x][y];
        }
    }
}

// ---------------------------------------------------------------------------
// spectral weights + LDS histogram
// ---------------------------------------------------------------------------
struct SpecItem {
    int pair;
    int pad;
    long long e0, e1;
};

__device__ inline void bin_add(double *hist, const nrg_spec_params &P, double E, double wgt) {
    const bool posside = E >= 0.0;
    const double x = posside ? E : -E;
    if (fabs(wgt) < P.discard * x) return;
    double *h = posside ? hist : hist + P.nbins;
    const double pos = (log10(x) - P.log10emin) * P.bins_per_decade;
    long long index;
    if (!(pos > -9.0e18)) index = -1;  // x == 0 -> -inf: first bin
    else index = (long long)pos;       // truncation toward zero, as long(pos)
    if (index < 0) {
        atomicAdd(&h[0], wgt);
        return;
    }
    if (index >= P.nbins - 1) {
        atomicAdd(&h[P.nbins - 1], wgt);
        return;
    }
    const double rem = pos - (double)index;
    atomicAdd(&h[index], (1.0 - rem) * wgt);
    atomicAdd(&h[index + 1], rem * wgt);
}

__global__ __launch_bounds__(256) void k_spec_bin(const nrg_spec_params P, const nrg_spec_pair *pairs,
                                                  const SpecItem *items, double *bins) {
    extern __shared__ double hist[];
    const int nb2 = 2 * P.nbins;
    for (int b = threadIdx.x; b < nb2; b += blockDim.x) hist[b] = 0.0;
    __syncthreads();
    const SpecItem it = items[blockIdx.x];
    const nrg_spec_pair q = pairs[it.pair];
    for (long long e = it.e0 + threadIdx.x; e < it.e1; e += blockDim.x) {
        const int j = (int)(e / q.alli);
        const int i = (int)(e - (long long)j * q.alli);
        const double o1 = q.op1[(size_t)j * q.ld1 + i];
        const double o2 = q.op2[(size_t)j * q.ld2 + i];
        const double Ei = q.Ei[i], Ej = q.Ej[j];
        const double E = (Ej - Ei) * P.escale;
        if (P.mode == NRG_SPEC_FDM) {
            const bool Di = i >= q.reti, Dj = j >= q.retj;
            if (Di) bin_add(hist, P, E, q.sf * o1 * o2 * P.wnf * exp(-Ei / P.T));
            else if (Dj) bin_add(hist, P, E, q.sf * o1 * q.M1[(size_t)j * q.ldm1 + i]);
            if (Dj) bin_add(hist, P, E, q.sf * o1 * o2 * P.wnf * exp(-Ej / P.T));
            else if (Di) bin_add(hist, P, E, q.sf * o1 * q.M2[(size_t)j * q.ldm2 + i]);
        } else {
            const double aE = fabs(Ej - Ei);
            if (aE < P.emin_w || aE > P.emax_w) continue;
            const double wA = q.M1[(size_t)j * q.ldm1 + i] * o1;
            const double wB = q.M2[(size_t)j * q.ldm2 + i] * o2;
            bin_add(hist, P, E, q.sf * (wA + wB));
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb2; b += blockDim.x) {
        const double v = hist[b];
        if (v != 0.0) atomicAdd(&bins[b], v);
    }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

nrg_status nrg_ctx_create(nrg_ctx **out, int device) {
    if (!out) return NRG_ERR_ARG;
    nrg_ctx *c = new nrg_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->dstatus, sizeof(int));
    if (e != hipSuccess) {
        fprintf(stderr, "nrg_ctx_create: %s\n", hipGetErrorString(e));
        delete c;
        *out = nullptr;
        return NRG_ERR_HIP;
    }
    *out = c;
    return NRG_OK;
}

void nrg_ctx_destroy(nrg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->desc) (void)hipFree(c->desc);
    if (c->ews) (void)hipFree(c->ews);
    if (c->dstatus) (void)hipFree(c->dstatus);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *nrg_last_error(const nrg_ctx *c) { return c ? c->err.c_str() : "null context"; }

nrg_status nrg_sync(nrg_ctx *c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!g_pending.empty()) timing_fold();
    return NRG_OK;
}

void nrg_timing_enable(int on) { g_timing = on != 0; }

void nrg_timing_reset(void) {
    timing_fold();
    memset(g_stats, 0, sizeof g_stats);
}

int nrg_timing_get(nrg_kernel_stat *out, int max) {
    timing_fold();
    int n = 0;
    for (int k = 0; k < K_COUNT; ++k) {
        if (!g_stats[k].launches) continue;
        if (n < max && out) {
            out[n] = g_stats[k];
            snprintf(out[n].name, sizeof out[n].name, "%s", kKernelNames[k]);
        }
        ++n;
    }
    return n;
}

nrg_status nrg_dev_alloc(nrg_ctx *c, size_t bytes, void **dptr) {
    if (!c || !dptr) return NRG_ERR_ARG;
    hipError_t e = hipMalloc(dptr, std::max<size_t>(bytes, 8));
    if (e != hipSuccess) return fail(c, NRG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return NRG_OK;
}

nrg_status nrg_dev_free(nrg_ctx *c, void *dptr) {
    if (!dptr) return NRG_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(dptr));
    return NRG_OK;
}

nrg_status nrg_h2d(nrg_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!bytes) return NRG_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return NRG_OK;
}

nrg_status nrg_d2h(nrg_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!bytes) return NRG_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return NRG_OK;
}

nrg_status nrg_d2h_strided(nrg_ctx *c, double *dst, const double *src, size_t count, size_t stride) {
    if (!count) return NRG_OK;
    HIPCHK(c, hipMemcpy2DAsync(dst, sizeof(double), src, stride * sizeof(double), sizeof(double), count,
                               hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return NRG_OK;
}

nrg_status nrg_zero(nrg_ctx *c, double *dst, size_t count) {
    if (!count) return NRG_OK;
    const int nb = (int)std::min<size_t>((count + 255) / 256, 4096);
    LaunchTimer lt(K_ZERO, c->stream, 0.0, 8.0 * count);
    hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, c->stream, dst, count);
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

nrg_status nrg_assemble_batch(nrg_ctx *c, const nrg_asm_block *blocks, int nblocks, const nrg_asm_term *terms,
                              int nterms) {
    if (nblocks <= 0) return NRG_OK;
    for (int b = 0; b < nblocks; ++b) {
        const nrg_asm_block &B = blocks[b];
        if (B.d <= 0 || B.term_begin < 0 || B.term_begin + B.term_count > nterms) return fail(c, NRG_ERR_ARG, "assemble: bad block");
        for (int t = B.term_begin; t < B.term_begin + B.term_count; ++t) {
            const nrg_asm_term &T = terms[t];
            const int ce = (T.kind == NRG_TERM_BLOCK) ? T.c0 + T.cols : 0;
            if (T.r0 < 0 || T.r0 + T.rows > B.d || ce > B.d || (T.kind == NRG_TERM_BLOCK && T.ld < T.cols))
                return fail(c, NRG_ERR_ARG, "assemble: term out of range");
        }
    }
    const size_t bb = sizeof(nrg_asm_block) * nblocks, tb = sizeof(nrg_asm_term) * nterms;
    std::vector<char> buf(bb + tb);
    memcpy(buf.data(), blocks, bb);
    if (tb) memcpy(buf.data() + bb, terms, tb);
    void *d;
    nrg_status s = stage(c, buf.data(), buf.size(), &d);
    if (s) return s;
    double abytes = 0;
    for (int b = 0; b < nblocks; ++b) abytes += 8.0 * blocks[b].d * blocks[b].d;
    LaunchTimer lt(K_ASSEMBLE, c->stream, 0.0, abytes);
    hipLaunchKernelGGL(k_assemble, dim3(nblocks), dim3(256), 0, c->stream, (const nrg_asm_block *)d,
                       (const nrg_asm_term *)((char *)d + bb));
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

nrg_status nrg_diag_batch(nrg_ctx *c, const nrg_eig_task *tasks_in, int ntasks) {
    if (ntasks <= 0) return NRG_OK;
    // largest first: better load balance, and the LDS-sized tail is contiguous
    std::vector<nrg_eig_task> sorted(tasks_in, tasks_in + ntasks);
    std::stable_sort(sorted.begin(), sorted.end(),
                     [](const nrg_eig_task &a, const nrg_eig_task &b) { return a.n > b.n; });
    const nrg_eig_task *tasks = sorted.data();
    std::vector<size_t> off(ntasks);
    size_t tot = 0;
    for (int i = 0; i < ntasks; ++i) {
        if (tasks[i].n <= 0 || !tasks[i].A || !tasks[i].W) return fail(c, NRG_ERR_ARG, "diag: bad task");
        off[i] = tot;
        tot += eig_ws_doubles(tasks[i].n);
    }
    if (tot > c->ews_doubles) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->ews) HIPCHK(c, hipFree(c->ews));
        c->ews = nullptr;
        HIPCHK(c, hipMalloc(&c->ews, tot * sizeof(double)));
        c->ews_doubles = tot;
    }
    const size_t tb = sizeof(nrg_eig_task) * ntasks, ob = sizeof(size_t) * ntasks;
    std::vector<char> buf(tb + ob);
    memcpy(buf.data(), tasks, tb);
    memcpy(buf.data() + tb, off.data(), ob);
    void *d;
    nrg_status s = stage(c, buf.data(), buf.size(), &d);
    if (s) return s;
    int zero = 0;
    HIPCHK(c, hipMemcpyAsync(c->dstatus, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
    double eflops = 0, ebytes = 0;
    for (int i = 0; i < ntasks; ++i) {
        const double n = tasks[i].n;
        eflops += 10.0 / 3.0 * n * n * n;  // tridiagonalisation + back-transform
        ebytes += 8.0 / 3.0 * n * n * n;   // trailing matrix read once per column
    }
    // tasks arrive largest first; the ones that fit the LDS variant form the tail
    int split = ntasks;
    while (split > 0 && tasks[split - 1].n <= EIG_LDS_N) --split;
    for (int i = 0; i < split; ++i)
        if (tasks[i].n <= EIG_LDS_N) return fail(c, NRG_ERR_ARG, "diag: tasks must be sorted by decreasing size");
    int nl = 0;
    for (int i = split; i < ntasks; ++i) nl = std::max(nl, tasks[i].n);
    const size_t lds = 6 * (size_t)nl * sizeof(double);
    {
    LaunchTimer lt(K_EIG, c->stream, eflops, ebytes);
    if (split > 0)
        hipLaunchKernelGGL(k_eig_glb, dim3(split), dim3(EIG_NT), 0, c->stream, (const nrg_eig_task *)d, c->ews,
                           (const size_t *)((char *)d + tb), c->dstatus);
    if (split < ntasks)
        hipLaunchKernelGGL(k_eig_lds, dim3(ntasks - split), dim3(EIG_NT), lds, c->stream,
                           (const nrg_eig_task *)d + split, c->ews, (const size_t *)((char *)d + tb) + split,
                           c->dstatus);
    }
    HIPCHK(c, hipGetLastError());
    int st = 0;
    HIPCHK(c, hipMemcpyAsync(&st, c->dstatus, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st) return fail(c, NRG_ERR_NOCONV, "diag: QL iteration did not converge");
    return NRG_OK;
}

nrg_status nrg_gemm_grouped(nrg_ctx *c, const nrg_gemm_out *outs, int nouts, const nrg_gemm_term *terms,
                            int nterms) {
    if (nouts <= 0) return NRG_OK;
    std::vector<GemmItem> items;
    for (int o = 0; o < nouts; ++o) {
        const nrg_gemm_out &O = outs[o];
        if (O.m < 0 || O.n < 0 || O.term_begin < 0 || O.term_begin + O.term_count > nterms || (O.m > 0 && O.ldc < O.n))
            return fail(c, NRG_ERR_ARG, "gemm: bad output");
        for (int t = O.term_begin; t < O.term_begin + O.term_count; ++t) {
            const nrg_gemm_term &T = terms[t];
            const int needA = T.transA ? O.m : T.k, needB = T.transB ? T.k : O.n;
            if (T.k < 0 || T.lda < needA || T.ldb < needB) return fail(c, NRG_ERR_ARG, "gemm: bad term");
        }
        for (int tm = 0; tm * GT < O.m; ++tm)
            for (int tn = 0; tn * GT < O.n; ++tn) items.push_back({o, tm, tn, 0});
    }
    if (items.empty()) return NRG_OK;
    const size_t ob = sizeof(nrg_gemm_out) * nouts, tb = sizeof(nrg_gemm_term) * nterms,
                 ib = sizeof(GemmItem) * items.size();
    std::vector<char> buf(ob + tb + ib);
    memcpy(buf.data(), outs, ob);
    if (tb) memcpy(buf.data() + ob, terms, tb);
    memcpy(buf.data() + ob + tb, items.data(), ib);
    void *d;
    nrg_status s = stage(c, buf.data(), buf.size(), &d);
    if (s) return s;
    double gflops = 0, gbytes = 0;
    for (int o = 0; o < nouts; ++o)
        for (int t = outs[o].term_begin; t < outs[o].term_begin + outs[o].term_count; ++t) {
            const double m = outs[o].m, n = outs[o].n, k = terms[t].k;
            gflops += 2.0 * m * n * k;
            gbytes += 8.0 * (m * k + k * n);
        }
    for (int o = 0; o < nouts; ++o) gbytes += 8.0 * outs[o].m * outs[o].n * (outs[o].beta != 0.0 ? 2 : 1);
    LaunchTimer lt(K_GEMM, c->stream, gflops, gbytes);
    hipLaunchKernelGGL(k_gemm, dim3(items.size()), dim3(256), 0, c->stream, (const nrg_gemm_out *)d,
                       (const nrg_gemm_term *)((char *)d + ob), (const GemmItem *)((char *)d + ob + tb));
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

nrg_status nrg_spec_bin(nrg_ctx *c, const nrg_spec_params *prm, const nrg_spec_pair *pairs, int npairs,
                        double *bins) {
    if (npairs <= 0) return NRG_OK;
    if (!prm || prm->nbins < 2 || 2 * (size_t)prm->nbins * sizeof(double) > 64 * 1024)
        return fail(c, NRG_ERR_ARG, "spec: bad bin grid");
    std::vector<SpecItem> items;
    const long long chunk = 256 * 32;
    for (int p = 0; p < npairs; ++p) {
        const nrg_spec_pair &q = pairs[p];
        if (q.alli < 0 || q.allj < 0 || q.reti > q.alli || q.retj > q.allj) return fail(c, NRG_ERR_ARG, "spec: bad pair");
        const long long ne = (long long)q.alli * q.allj;
        for (long long e = 0; e < ne; e += chunk) items.push_back({p, 0, e, std::min(ne, e + chunk)});
    }
    if (items.empty()) return NRG_OK;
    const size_t pb = sizeof(nrg_spec_pair) * npairs, ib = sizeof(SpecItem) * items.size();
    std::vector<char> buf(pb + ib);
    memcpy(buf.data(), pairs, pb);
    memcpy(buf.data() + pb, items.data(), ib);
    void *d;
    nrg_status s = stage(c, buf.data(), buf.size(), &d);
    if (s) return s;
    const size_t lds = 2 * (size_t)prm->nbins * sizeof(double);
    double sbytes = 0;
    for (int p = 0; p < npairs; ++p) {
        const nrg_spec_pair &q = pairs[p];
        const double el = (double)q.alli * q.allj;
        if (prm->mode == NRG_SPEC_FDM)
            sbytes += 16.0 * el + 8.0 * ((double)q.allj * q.reti + (double)q.retj * q.alli);
        else
            sbytes += 32.0 * el;
        sbytes += 8.0 * (q.alli + q.allj);
    }
    LaunchTimer lt(K_SPEC, c->stream, 0.0, sbytes);
    hipLaunchKernelGGL(k_spec_bin, dim3(items.size()), dim3(256), lds, c->stream, *prm, (const nrg_spec_pair *)d,
                       (const SpecItem *)((char *)d + pb), bins);
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

}  // extern "C"
