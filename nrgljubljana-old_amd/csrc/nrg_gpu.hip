// nrg_gpu.hip - CDNA4 (gfx950) kernels and C ABI of the NRG hot path.
//
// Kernels (one file, one code object):
//   k_zero          fill kernel (used instead of memset nodes)
//   k_assemble      per-subspace Hamiltonian assembly
//                   (src/nrg.cc:2456-2478, src/matrix.cc:86-192)
//   (the batched eigensolver replacing dsyev lives in nrg_eig.hip)
//   k_gemm          grouped FP64 GEMM with per-output accumulation lists
//                   (recalc_f, recalc_general, cdmI)
//   k_spec_bin      spectral weights of SPEC_FDM / SPEC_DMNRG fused with the
//                   logarithmic binning of Bins::add_std into an LDS histogram
//
// Every launch is on the context's stream; descriptor arrays are staged in
// a context-owned device buffer.
#include <chrono>
#include <thread>

#include "nrg_internal.h"

const char *kKernelNames[K_COUNT] = {"k_zero",       "k_assemble", "k_eig_tri_coop", "k_eig_tri_one", "k_eig_tri_reg", "k_eig_tri_blk", "k_dc_leaf", "k_dc_small", "k_dc_deflate",
                                     "k_dc_secular", "k_dc_zhat",  "k_dc_unorm", "k_gemm_dc", "k_gemm_bt", "k_dc_rank",
                                     "k_eig_bt",     "k_eig_copy", "k_gemm",     "k_spec_bin", "k_gather_scatter",
                                     "k_sbr_chase", "k_sbr_bt"};
bool g_timing = false;
int g_timing_only = -1;
nrg_kernel_stat g_stats[K_COUNT];
std::vector<PendingLaunch> g_pending;

void timing_fold() {
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

nrg_status wait_event(nrg_ctx *c, hipEvent_t e) {
    // spin for a short while, then yield between polls (solves running in
    // other threads of a serving process keep their cores)
    const auto t0 = std::chrono::steady_clock::now();
    for (bool spin = true;;) {
        const hipError_t r = hipEventQuery(e);
        if (r == hipSuccess) return NRG_OK;
        if (r != hipErrorNotReady) return fail(c, NRG_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(r));
        if (spin) spin = std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(200);
        else std::this_thread::yield();
    }
}

nrg_status wait_stream(nrg_ctx *c) {
    HIPCHK(c, hipEventRecord(c->ev_wait, c->stream));
    nrg_status s = wait_event(c, c->ev_wait);
    if (s) return s;
    if (c->status_pending) {
        c->status_pending = false;
        if (*c->hstat) return fail(c, NRG_ERR_HIP, "diag: a tridiagonalisation team did not assemble in time");
    }
    return NRG_OK;
}

nrg_status stage(nrg_ctx *c, const void *host, size_t bytes, void **dev) {
    // the current slot is released by everything queued so far
    if (c->desc_cur >= 0) {
        HIPCHK(c, hipEventRecord(c->desc_ev[c->desc_cur], c->stream));
        c->desc_busy[c->desc_cur] = true;
    }
    const int k = (c->desc_cur + 1) % nrg_ctx::NSTAGE;
    c->desc_cur = k;
    if (c->desc_busy[k]) {
        nrg_status s = wait_event(c, c->desc_ev[k]);
        if (s) return s;
        c->desc_busy[k] = false;
    }
    if (bytes > c->desc_bytes[k]) {
        if (c->desc[k]) {
            HIPCHK(c, hipFree(c->desc[k]));
            HIPCHK(c, hipHostFree(c->hdesc[k]));
            c->desc[k] = c->hdesc[k] = nullptr;
        }
        const size_t nb = std::max<size_t>(bytes + bytes / 2, 1 << 20);
        HIPCHK(c, hipMalloc(&c->desc[k], nb));
        HIPCHK(c, hipHostMalloc(&c->hdesc[k], nb, hipHostMallocDefault));
        c->desc_bytes[k] = nb;
    }
    memcpy(c->hdesc[k], host, bytes);
    HIPCHK(c, hipMemcpyAsync(c->desc[k], c->hdesc[k], bytes, hipMemcpyHostToDevice, c->stream));
    *dev = c->desc[k];
    return NRG_OK;
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
// batched strided transfers (through one contiguous staging buffer)
// ---------------------------------------------------------------------------
struct XItem {
    double *ptr;
    long long count, stride, off;
};
struct XChunk {
    int item, pad;
    long long e0, e1;
};
constexpr long long kXChunk = 4096;

__global__ __launch_bounds__(256) void k_gather(const XItem *items, const XChunk *chunks, double *buf) {
    const XChunk ck = chunks[blockIdx.x];
    const XItem it = items[ck.item];
    for (long long e = ck.e0 + threadIdx.x; e < ck.e1; e += 256) buf[it.off + e] = it.ptr[e * it.stride];
}

__global__ __launch_bounds__(256) void k_scatter(const XItem *items, const XChunk *chunks, const double *buf) {
    const XChunk ck = chunks[blockIdx.x];
    const XItem it = items[ck.item];
    for (long long e = ck.e0 + threadIdx.x; e < ck.e1; e += 256) it.ptr[e * it.stride] = buf[it.off + e];
}

// Stages the item list (and, for a scatter, the host values after it);
// returns the total element count.  A gather to the host needs the pinned
// buffer sized (to_host).
static nrg_status xfer_prepare(nrg_ctx *c, const nrg_strided *items, int nitems, const double *payload, bool to_host,
                               size_t *total, XItem **d_items, XChunk **d_chunks, int *nchunks,
                               const double **d_payload) {
    std::vector<XItem> xi(nitems);
    std::vector<XChunk> xc;
    long long off = 0;
    for (int k = 0; k < nitems; ++k) {
        if (items[k].count < 0 || (items[k].count > 0 && (!items[k].ptr || items[k].stride < 1)))
            return fail(c, NRG_ERR_ARG, "transfer: bad item");
        xi[k] = {items[k].ptr, items[k].count, items[k].stride, off};
        for (long long e = 0; e < items[k].count; e += kXChunk)
            xc.push_back({k, 0, e, std::min<long long>(items[k].count, e + kXChunk)});
        off += items[k].count;
    }
    *total = (size_t)off;
    *nchunks = (int)xc.size();
    if (!off) return NRG_OK;
    if (to_host && (size_t)off > c->hpin_doubles) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->hpin) HIPCHK(c, hipHostFree(c->hpin));
        c->hpin = nullptr;
        const size_t nd = std::max<size_t>((size_t)off * 2, 1 << 16);
        HIPCHK(c, hipHostMalloc((void **)&c->hpin, nd * sizeof(double), hipHostMallocDefault));
        c->hpin_doubles = nd;
    }
    const size_t ib = sizeof(XItem) * xi.size(), cb = sizeof(XChunk) * xc.size();
    const size_t pb = (ib + cb + 7) & ~(size_t)7, vb = payload ? sizeof(double) * (size_t)off : 0;
    std::vector<char> buf(pb + vb);
    memcpy(buf.data(), xi.data(), ib);
    memcpy(buf.data() + ib, xc.data(), cb);
    if (vb) memcpy(buf.data() + pb, payload, vb);
    void *d;
    nrg_status s = stage(c, buf.data(), buf.size(), &d);
    if (s) return s;
    *d_items = (XItem *)d;
    *d_chunks = (XChunk *)((char *)d + ib);
    if (d_payload) *d_payload = (const double *)((char *)d + pb);
    return NRG_OK;
}

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
// One workgroup per band of ASM_BAND rows of a block: the band is zeroed,
// then every term adds what falls into its rows, in term order (for a block
// term its direct part, then its transposed part), so each element sums the
// same contributions in the same order as a whole-block pass would.
constexpr int ASM_BAND = 16;
struct AsmItem {
    int block, r0, r1, pad;  // rows [r0, r1) of block `block`
};

__global__ __launch_bounds__(256) void k_assemble(const nrg_asm_block *blocks, const nrg_asm_term *terms,
                                                  const AsmItem *items) {
    const AsmItem it = items[blockIdx.x];
    const nrg_asm_block b = blocks[it.block];
    const int d = b.d, R0 = it.r0, R1 = it.r1;
    double *H = b.H;
    const int tid = threadIdx.x;
    for (int i = tid; i < (R1 - R0) * d; i += blockDim.x) H[(size_t)R0 * d + i] = 0.0;
    __syncthreads();
    for (int t = 0; t < b.term_count; ++t) {
        const nrg_asm_term tm = terms[b.term_begin + t];
        if (tm.kind == NRG_TERM_DIAGVEC || tm.kind == NRG_TERM_SHIFT) {
            const int lo = max(R0, tm.r0), hi = min(R1, tm.r0 + tm.rows);
            for (int r = lo + tid; r < hi; r += blockDim.x)
                H[(size_t)r * d + r] += tm.kind == NRG_TERM_DIAGVEC ? tm.scale * tm.src[r - tm.r0] : tm.scale;
        } else {
            // direct part: rows r0 + r of the band, columns c0 + cc
            const int lo = max(R0, tm.r0), hi = min(R1, tm.r0 + tm.rows);
            if (lo < hi) {
                const int n = (hi - lo) * tm.cols;
                for (int e = tid; e < n; e += blockDim.x) {
                    const int r = lo - tm.r0 + e / tm.cols, cc = e % tm.cols;
                    H[(size_t)(tm.r0 + r) * d + tm.c0 + cc] += tm.scale * tm.src[(size_t)r * tm.ld + cc];
                }
            }
            __syncthreads();
            // transposed part: rows c0 + cc of the band, columns r0 + r
            const int clo = max(R0, tm.c0), chi = min(R1, tm.c0 + tm.cols);
            if (clo < chi) {
                const int n = (chi - clo) * tm.rows;
                for (int e = tid; e < n; e += blockDim.x) {
                    const int cc = clo - tm.c0 + e / tm.rows, r = e % tm.rows;
                    H[(size_t)(tm.c0 + cc) * d + tm.r0 + r] += tm.scale * tm.src[(size_t)r * tm.ld + cc];
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// grouped GEMM (FP64, LDS tiled, GT x GT tile, GMT x GMT per thread; 32 x 32:
// the stage launches have few tiles, so small tiles buy parallelism)
// A workgroup walks the (term, k) sequence of its tile, or of its slice of
// it, in steps of GK.  Software-pipelined: the global loads of step i+1 are
// issued right after step i's tile is in LDS and complete while step i is
// computed; two LDS buffers, one barrier per step.
// ---------------------------------------------------------------------------
constexpr int GLQ = GT * GK / 256;  // tile elements per thread and operand
constexpr int GMT = GT / 16;        // outputs per thread along each tile dimension

// op(A) rows row0.., k k0.. and op(B) k k0.., columns col0.. of one step of a
// TILE x TILE output tile (TILE * GK / 256 elements per thread and operand)
template <int TILE, int NQ = TILE * GK / 256>
__device__ __forceinline__ void gemm_load_step(const nrg_gemm_term &tm, int m, int n, int row0, int col0, int k0,
                                               int tid, double (&ra)[NQ], double (&rb)[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = tid + q * 256;
        int r, kk;
        if (tm.transA) {  // contiguous along rows of op(A)
            kk = e / TILE;
            r = e % TILE;
        } else {
            r = e / GK;
            kk = e % GK;
        }
        const int gr = row0 + r, gk = k0 + kk;
        double v = 0.0;
        if (gr < m && gk < tm.k) {
            v = tm.transA ? tm.A[(size_t)gk * tm.lda + gr] : tm.A[(size_t)gr * tm.lda + gk];
            if (tm.s) v *= tm.s[gk];
        }
        ra[q] = v;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = tid + q * 256;
        int cidx, kk;
        if (tm.transB) {
            cidx = e / GK;
            kk = e % GK;
        } else {
            kk = e / TILE;
            cidx = e % TILE;
        }
        const int gc = col0 + cidx, gk = k0 + kk;
        double v = 0.0;
        if (gc < n && gk < tm.k) v = tm.transB ? tm.B[(size_t)gc * tm.ldb + gk] : tm.B[(size_t)gk * tm.ldb + gc];
        rb[q] = v * tm.alpha;
    }
}

template <int TILE, int NQ = TILE * GK / 256>
__device__ __forceinline__ void gemm_store_step(const nrg_gemm_term &tm, int tid, const double (&ra)[NQ],
                                                const double (&rb)[NQ], double (*As)[TILE + 1], double (*Bs)[TILE + 1]) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = tid + q * 256;
        if (tm.transA) As[e / TILE][e % TILE] = ra[q];
        else As[e % GK][e / GK] = ra[q];
        if (tm.transB) Bs[e % GK][e / GK] = rb[q];
        else Bs[e / TILE][e % TILE] = rb[q];
    }
}

__global__ __launch_bounds__(256) void k_gemm(const nrg_gemm_out *outs, const nrg_gemm_term *terms,
                                              const GemmItem *items, double *part) {
    __shared__ double As[2][GK][GT + 1];
    __shared__ double Bs[2][GK][GT + 1];
    const GemmItem it = items[blockIdx.x];
    const nrg_gemm_out o = outs[it.out];
    const int tid = threadIdx.x;
    const int tr = tid / 16, tc = tid % 16;  // 16x16 threads, GMT x GMT each
    const int row0 = it.tm * GT, col0 = it.tn * GT;
    if (row0 >= o.m || col0 >= o.n) return;  // m may have been lowered on the device
    // k-steps [g0, g1) of the output's (term, k) sequence belong to this item
    int g0 = 0, g1 = INT_MAX;
    if (it.ns > 1) {
        int S = 0;  // k may have been lowered on the device
        for (int t = 0; t < o.term_count; ++t) S += (terms[o.term_begin + t].k + GK - 1) / GK;
        g0 = (int)((long long)S * it.s / it.ns);
        g1 = (int)((long long)S * (it.s + 1) / it.ns);
    }
    double acc[GMT][GMT];
#pragma unroll
    for (int a = 0; a < GMT; ++a)
#pragma unroll
        for (int b = 0; b < GMT; ++b) acc[a][b] = 0.0;
    // first step of the slice: term t, offset k0
    int t = 0, k0 = 0, g = 0;
    nrg_gemm_term tm{};
    bool have = false;
    for (; t < o.term_count && g < g1; ++t) {
        tm = terms[o.term_begin + t];
        const int st = (tm.k + GK - 1) / GK;
        if (g + st <= g0) {  // the whole term lies before this slice
            g += st;
            continue;
        }
        k0 = (g0 - g) * GK;  // (g <= g0 here)
        g = g0;
        have = g0 < g1;  // an empty slice still writes its (zero) partial tile
        break;
    }
    double ra[GLQ], rb[GLQ];
    if (have) gemm_load_step<GT>(tm, o.m, o.n, row0, col0, k0, tid, ra, rb);
    int buf = 0;
    while (have) {
        gemm_store_step<GT>(tm, tid, ra, rb, As[buf], Bs[buf]);
        __syncthreads();
        // the next step (possibly of a later term), loaded while this one is computed
        ++g;
        k0 += GK;
        bool next = g < g1;
        if (next && k0 >= tm.k) {
            next = false;
            for (++t; t < o.term_count; ++t) {
                tm = terms[o.term_begin + t];
                if (tm.k > 0) {
                    next = true;
                    k0 = 0;
                    break;
                }
            }
        }
        if (next) gemm_load_step<GT>(tm, o.m, o.n, row0, col0, k0, tid, ra, rb);
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            double a[GMT], b[GMT];
#pragma unroll
            for (int q = 0; q < GMT; ++q) {
                a[q] = As[buf][kk][tr + 16 * q];
                b[q] = Bs[buf][kk][tc + 16 * q];
            }
#pragma unroll
            for (int x = 0; x < GMT; ++x)
#pragma unroll
                for (int y = 0; y < GMT; ++y) acc[x][y] = fma(a[x], b[y], acc[x][y]);
        }
        buf ^= 1;
        have = next;
    }
    if (it.ns > 1) {
        double *pt = part + (size_t)(it.part + it.s) * GT * GT;
#pragma unroll
        for (int x = 0; x < GMT; ++x)
#pragma unroll
            for (int y = 0; y < GMT; ++y) pt[(tr + 16 * x) * GT + tc + 16 * y] = acc[x][y];
        return;
    }
#pragma unroll
    for (int x = 0; x < GMT; ++x) {
        const int gr = row0 + tr + 16 * x;
        if (gr >= o.m) continue;
#pragma unroll
        for (int y = 0; y < GMT; ++y) {
            const int gc = col0 + tc + 16 * y;
            if (gc >= o.n) continue;
            double *cp = o.C + (size_t)gr * o.ldc + gc;
            *cp = (o.beta == 0.0 ? 0.0 : o.beta * *cp) + acc[x][y];
        }
    }
}

// C tile = beta C + sum of the slice partials, in slice order
__global__ __launch_bounds__(256) void k_gemm_reduce(const nrg_gemm_out *outs, const GemmItem *items,
                                                     const double *part) {
    const GemmItem it = items[blockIdx.x];
    const nrg_gemm_out o = outs[it.out];
    const int row0 = it.tm * GT, col0 = it.tn * GT;
    if (row0 >= o.m || col0 >= o.n) return;
    const double *pt = part + (size_t)it.part * GT * GT;
    for (int e = threadIdx.x; e < GT * GT; e += 256) {
        const int gr = row0 + e / GT, gc = col0 + e % GT;
        if (gr >= o.m || gc >= o.n) continue;
        double v = 0.0;
        for (int s = 0; s < it.ns; ++s) v += pt[(size_t)s * GT * GT + e];
        double *cp = o.C + (size_t)gr * o.ldc + gc;
        *cp = (o.beta == 0.0 ? 0.0 : o.beta * *cp) + v;
    }
}

// ---------------------------------------------------------------------------
// grouped GEMM on FP64 MFMA for launches with many output tiles: one 64 x 64
// tile per workgroup of four wavefronts, each wavefront a 32 x 32 quarter as
// 2 x 2 blocks of v_mfma_f64_16x16x4f64 (A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k, D[i][j] in lane j + 16 (i % 4), element i / 4).  The same
// per-output term lists and operand staging (two LDS buffers, one barrier per
// k-step of GK) as k_gemm; tiles are never split, so the accumulation order
// of an output is fixed.
// ---------------------------------------------------------------------------
typedef double gemm_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_gemm_mfma(const nrg_gemm_out *outs, const nrg_gemm_term *terms,
                                                   const GemmItem *items) {
    __shared__ double As[2][GK][GTM + 1];
    __shared__ double Bs[2][GK][GTM + 1];
    const GemmItem it = items[blockIdx.x];
    const nrg_gemm_out o = outs[it.out];
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int wr = w >> 1, wc = w & 1;
    const int row0 = it.tm * GTM, col0 = it.tn * GTM;
    if (row0 >= o.m || col0 >= o.n) return;  // m may have been lowered on the device
    gemm_d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = gemm_d4{0.0, 0.0, 0.0, 0.0};
    int t = 0, k0 = 0;
    nrg_gemm_term tm{};
    bool have = false;
    for (; t < o.term_count; ++t) {
        tm = terms[o.term_begin + t];
        if (tm.k > 0) {
            have = true;
            break;
        }
    }
    constexpr int NQ = GTM * GK / 256;
    double ra[NQ], rb[NQ];
    if (have) gemm_load_step<GTM>(tm, o.m, o.n, row0, col0, 0, tid, ra, rb);
    int buf = 0;
    const int li = lane & 15, lk = lane >> 4;
    while (have) {
        gemm_store_step<GTM>(tm, tid, ra, rb, As[buf], Bs[buf]);
        __syncthreads();
        k0 += GK;
        bool next = true;
        if (k0 >= tm.k) {
            next = false;
            for (++t; t < o.term_count; ++t) {
                tm = terms[o.term_begin + t];
                if (tm.k > 0) {
                    next = true;
                    k0 = 0;
                    break;
                }
            }
        }
        if (next) gemm_load_step<GTM>(tm, o.m, o.n, row0, col0, k0, tid, ra, rb);
#pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            const double a0 = As[buf][kk + lk][wr * 32 + li], a1 = As[buf][kk + lk][wr * 32 + 16 + li];
            const double b0 = Bs[buf][kk + lk][wc * 32 + li], b1 = Bs[buf][kk + lk][wc * 32 + 16 + li];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        buf ^= 1;
        have = next;
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gr = row0 + wr * 32 + bi * 16 + lk + 4 * v, gc = col0 + wc * 32 + bj * 16 + li;
                if (gr >= o.m || gc >= o.n) continue;
                double *cp = o.C + (size_t)gr * o.ldc + gc;
                *cp = (o.beta == 0.0 ? 0.0 : o.beta * *cp) + acc[bi][bj][v];
            }
}

bool gemm_plan(const std::vector<nrg_gemm_out> &outs, size_t o0, const std::vector<int> &ksteps,
               std::vector<GemmItem> &items, std::vector<GemmItem> &reduce, int &slots, int min_tiles) {
    size_t big = 0;
    for (size_t o = o0; o < outs.size(); ++o)
        big += (size_t)((outs[o].m + GTM - 1) / GTM) * ((outs[o].n + GTM - 1) / GTM);
    if (big >= (size_t)std::max(1, min_tiles)) {
        for (size_t o = o0; o < outs.size(); ++o)
            for (int tm = 0; tm * GTM < outs[o].m; ++tm)
                for (int tn = 0; tn * GTM < outs[o].n; ++tn) items.push_back({(int)o, tm, tn, 0, 0, 0});
        return true;
    }
    const size_t first = items.size();
    for (size_t o = o0; o < outs.size(); ++o)
        for (int tm = 0; tm * GT < outs[o].m; ++tm)
            for (int tn = 0; tn * GT < outs[o].n; ++tn) items.push_back({(int)o, tm, tn, 0, 0, 0});
    slots = std::max(slots, gemm_split(items, first, ksteps, reduce));
    return false;
}

std::vector<int> gemm_ksteps(const nrg_gemm_out *outs, int nouts, const nrg_gemm_term *terms) {
    std::vector<int> ks(nouts, 0);
    for (int o = 0; o < nouts; ++o)
        for (int t = outs[o].term_begin; t < outs[o].term_begin + outs[o].term_count; ++t)
            ks[o] += (terms[t].k + GK - 1) / GK;
    return ks;
}

int gemm_split(std::vector<GemmItem> &items, size_t first, const std::vector<int> &ksteps,
               std::vector<GemmItem> &reduce) {
    const size_t ntiles = items.size() - first;
    if (ntiles == 0 || ntiles >= (size_t)GEMM_SPLIT_MIN) return 0;
    const int want = (int)((GEMM_SPLIT_TARGET + ntiles - 1) / ntiles);
    const std::vector<GemmItem> tiles(items.begin() + first, items.end());
    items.resize(first);
    int slots = 0;
    for (const GemmItem &t : tiles) {
        const int ns = std::max(1, std::min(want, ksteps[t.out] / GEMM_SLICE_MIN));
        if (ns <= 1) {
            items.push_back(t);
            continue;
        }
        for (int s = 0; s < ns; ++s) items.push_back({t.out, t.tm, t.tn, s, ns, slots});
        reduce.push_back({t.out, t.tm, t.tn, 0, ns, slots});
        slots += ns;
    }
    return slots;
}

nrg_status gemm_reserve(nrg_ctx *c, int slots) {
    if ((size_t)slots <= c->gpart_tiles) return NRG_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->gpart) HIPCHK(c, hipFree(c->gpart));
    c->gpart = nullptr;
    c->gpart_tiles = std::max<size_t>((size_t)slots * 3 / 2, 1024);
    HIPCHK(c, hipMalloc(&c->gpart, c->gpart_tiles * GT * GT * sizeof(double)));
    return NRG_OK;
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

nrg_status gemm_launch_dev(nrg_ctx *c, int kid, const nrg_gemm_out *d_outs, const nrg_gemm_term *d_terms,
                           const GemmItem *d_items, int nitems, const GemmItem *d_reduce, int nreduce,
                           double flops, double bytes, bool mfma, double *part) {
    if (!part) part = c->gpart;
    if (nitems <= 0) return NRG_OK;
    LaunchTimer lt(kid, c->stream, flops, bytes);
    if (mfma) {
        hipLaunchKernelGGL(k_gemm_mfma, dim3(nitems), dim3(256), 0, c->stream, d_outs, d_terms, d_items);
        HIPCHK(c, hipGetLastError());
        return NRG_OK;
    }
    hipLaunchKernelGGL(k_gemm, dim3(nitems), dim3(256), 0, c->stream, d_outs, d_terms, d_items, part);
    if (nreduce > 0)
        hipLaunchKernelGGL(k_gemm_reduce, dim3(nreduce), dim3(256), 0, c->stream, d_outs, d_reduce, part);
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

// Scratch of nrg_gemm_grouped_noalloc: one array per device, held until the
// process ends; host code uses the address hipGetSymbolAddress returns for
// the context's device (c->scratch), never the array's name.
static __device__ __align__(256) double scratch[NRG_NOALLOC_SCRATCH_BYTES / sizeof(double)];

nrg_status nrg_ctx_create(nrg_ctx **out, int device) {
    if (!out) return NRG_ERR_ARG;
    nrg_ctx *c = new nrg_ctx();
    c->device = device;
    if (const char *t = getenv("NRG_GEMM_MFMA_MIN_TILES")) c->gemm_mfma_min_tiles = std::max(1, atoi(t));
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_wait, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&c->dstatus, sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->hstat, sizeof(int), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_values, hipEventDisableTiming);
    if (e == hipSuccess) e = hipGetSymbolAddress((void **)&c->scratch, HIP_SYMBOL(scratch));
    for (int k = 0; k < nrg_ctx::NSTAGE && e == hipSuccess; ++k)
        e = hipEventCreateWithFlags(&c->desc_ev[k], hipEventDisableTiming);
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
    for (int k = 0; k < nrg_ctx::NSTAGE; ++k) {
        if (c->desc[k]) (void)hipFree(c->desc[k]);
        if (c->hdesc[k]) (void)hipHostFree(c->hdesc[k]);
        if (c->desc_ev[k]) (void)hipEventDestroy(c->desc_ev[k]);
    }
    if (c->ews) (void)hipFree(c->ews);
    if (c->cnt) (void)hipFree(c->cnt);
    if (c->gpart) (void)hipFree(c->gpart);
    if (c->dstatus) (void)hipFree(c->dstatus);
    if (c->hstat) (void)hipHostFree(c->hstat);
    if (c->hval) (void)hipHostFree(c->hval);
    if (c->ev_values) (void)hipEventDestroy(c->ev_values);
    if (c->hpin) (void)hipHostFree(c->hpin);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_wait) (void)hipEventDestroy(c->ev_wait);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *nrg_last_error(const nrg_ctx *c) { return c ? c->err.c_str() : "null context"; }

nrg_status nrg_diag_values(nrg_ctx *c, double *host, size_t count) {
    if (!c->values_pending) return fail(c, NRG_ERR_ARG, "diag_values: no eigensolve is queued");
    if (count != c->values_count) return fail(c, NRG_ERR_ARG, "diag_values: count differs from the batch's total dimension");
    nrg_status s = wait_event(c, c->ev_values);
    if (s) return s;
    c->values_pending = false;
    if (*c->hstat) return fail(c, NRG_ERR_HIP, "diag: a tridiagonalisation team did not assemble in time");
    if (count) memcpy(host, c->hval, count * sizeof(double));
    return NRG_OK;
}

nrg_status nrg_sync(nrg_ctx *c) {
    nrg_status s = wait_stream(c);
    if (s) return s;
    if (!g_pending.empty()) timing_fold();
    return NRG_OK;
}

nrg_status nrg_ctx_release_workspace(nrg_ctx *c) {
    if (!c) return NRG_ERR_ARG;
    nrg_status s = nrg_sync(c);
    if (s) return s;
    if (c->ews) HIPCHK(c, hipFree(c->ews));
    c->ews = nullptr;
    c->ews_doubles = 0;
    if (c->gpart) HIPCHK(c, hipFree(c->gpart));
    c->gpart = nullptr;
    c->gpart_tiles = 0;
    return NRG_OK;
}

void nrg_timing_enable(int on) { g_timing = on != 0; }

void nrg_timing_filter(const char *kernel) {
    g_timing_only = -1;
    if (!kernel || !*kernel) return;
    for (int k = 0; k < K_COUNT; ++k)
        if (!strcmp(kernel, kKernelNames[k])) g_timing_only = k;
}

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
    if (e != hipSuccess) {
        // the refusal is reported here; clear it, or the next launch check
        // (hipGetLastError) would report it again after the caller recovered
        (void)hipGetLastError();
        return fail(c, NRG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
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
    return wait_stream(c);
}

nrg_status nrg_d2h(nrg_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!bytes) return NRG_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return wait_stream(c);
}

nrg_status nrg_d2h_strided(nrg_ctx *c, double *dst, const double *src, size_t count, size_t stride) {
    if (!count) return NRG_OK;
    HIPCHK(c, hipMemcpy2DAsync(dst, sizeof(double), src, stride * sizeof(double), sizeof(double), count,
                               hipMemcpyDeviceToHost, c->stream));
    return wait_stream(c);
}

nrg_status nrg_gather_d2h(nrg_ctx *c, const nrg_strided *items, int nitems, double *host) {
    if (nitems <= 0) return NRG_OK;
    size_t tot = 0;
    XItem *di = nullptr;
    XChunk *dc = nullptr;
    int nch = 0;
    nrg_status s = xfer_prepare(c, items, nitems, nullptr, true, &tot, &di, &dc, &nch, nullptr);
    if (s || !tot) return s;
    {
        // the kernel writes straight into the pinned host buffer (no copy)
        LaunchTimer lt(K_XFER, c->stream, 0.0, 16.0 * tot);
        hipLaunchKernelGGL(k_gather, dim3(nch), dim3(256), 0, c->stream, di, dc, c->hpin);
    }
    HIPCHK(c, hipGetLastError());
    s = wait_stream(c);
    if (s) return s;
    memcpy(host, c->hpin, tot * sizeof(double));
    return NRG_OK;
}

nrg_status nrg_gather_dev(nrg_ctx *c, const nrg_strided *items, int nitems, double *dst) {
    if (nitems <= 0) return NRG_OK;
    if (!dst) return fail(c, NRG_ERR_ARG, "gather_dev: no destination");
    size_t tot = 0;
    XItem *di = nullptr;
    XChunk *dc = nullptr;
    int nch = 0;
    nrg_status s = xfer_prepare(c, items, nitems, nullptr, false, &tot, &di, &dc, &nch, nullptr);
    if (s || !tot) return s;
    {
        LaunchTimer lt(K_XFER, c->stream, 0.0, 16.0 * tot);
        hipLaunchKernelGGL(k_gather, dim3(nch), dim3(256), 0, c->stream, di, dc, dst);
    }
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

nrg_status nrg_scatter_h2d(nrg_ctx *c, const nrg_strided *items, int nitems, const double *host) {
    if (nitems <= 0) return NRG_OK;
    size_t tot = 0;
    XItem *di = nullptr;
    XChunk *dc = nullptr;
    int nch = 0;
    // the values travel with the staged items (asynchronous: the host
    // buffer may be reused as soon as the call returns)
    const double *dv = nullptr;
    nrg_status s = xfer_prepare(c, items, nitems, host, false, &tot, &di, &dc, &nch, &dv);
    if (s || !tot) return s;
    {
        LaunchTimer lt(K_XFER, c->stream, 0.0, 16.0 * tot);
        hipLaunchKernelGGL(k_scatter, dim3(nch), dim3(256), 0, c->stream, di, dc, dv);
    }
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

nrg_status nrg_host_alloc(nrg_ctx *c, size_t bytes, void **hptr) {
    if (!c || !hptr) return NRG_ERR_ARG;
    hipError_t e = hipHostMalloc(hptr, std::max<size_t>(bytes, 8), hipHostMallocDefault);
    if (e != hipSuccess) {
        // a refusal the caller may recover from (the engine then spills to
        // pageable memory): cleared, as in nrg_dev_alloc
        (void)hipGetLastError();
        return fail(c, NRG_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    return NRG_OK;
}

nrg_status nrg_host_free(nrg_ctx *c, void *hptr) {
    if (!hptr) return NRG_OK;
    HIPCHK(c, hipHostFree(hptr));
    return NRG_OK;
}

nrg_status nrg_mem_info(nrg_ctx *c, size_t *free_bytes, size_t *total_bytes) {
    if (!c || !free_bytes || !total_bytes) return NRG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemGetInfo(free_bytes, total_bytes));
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
    std::vector<AsmItem> items;
    for (int b = 0; b < nblocks; ++b)
        for (int r0 = 0; r0 < blocks[b].d; r0 += ASM_BAND) items.push_back({b, r0, std::min(r0 + ASM_BAND, blocks[b].d), 0});
    const size_t bb = sizeof(nrg_asm_block) * nblocks, tb = sizeof(nrg_asm_term) * nterms,
                 ib = sizeof(AsmItem) * items.size();
    const size_t to = (bb + 15) & ~(size_t)15, io = (to + tb + 15) & ~(size_t)15;
    std::vector<char> buf(io + ib);
    memcpy(buf.data(), blocks, bb);
    if (tb) memcpy(buf.data() + to, terms, tb);
    memcpy(buf.data() + io, items.data(), ib);
    void *d;
    nrg_status s = stage(c, buf.data(), buf.size(), &d);
    if (s) return s;
    double abytes = 0;
    for (int b = 0; b < nblocks; ++b) abytes += 8.0 * blocks[b].d * blocks[b].d;
    LaunchTimer lt(K_ASSEMBLE, c->stream, 0.0, abytes);
    hipLaunchKernelGGL(k_assemble, dim3(items.size()), dim3(256), 0, c->stream, (const nrg_asm_block *)d,
                       (const nrg_asm_term *)((char *)d + to), (const AsmItem *)((char *)d + io));
    HIPCHK(c, hipGetLastError());
    return NRG_OK;
}

// Checks the descriptors and builds one grouped launch: the packed
// descriptors (outputs, terms, work items, reduce items) in `buf`, their
// offsets, the kernel choice and the partial slots split tiles need.
struct GemmLaunch {
    std::vector<char> buf;
    size_t ob = 0, tb = 0, ib = 0, rb = 0;
    int nitems = 0, nreduce = 0, slots = 0;
    bool mfma = false;
    double flops = 0, bytes = 0;
};
static nrg_status gemm_prepare(nrg_ctx *c, const nrg_gemm_out *outs, int nouts, const nrg_gemm_term *terms,
                               int nterms, GemmLaunch &L) {
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
    }
    std::vector<GemmItem> reduce;
    L.mfma = gemm_plan(std::vector<nrg_gemm_out>(outs, outs + nouts), 0, gemm_ksteps(outs, nouts, terms), items,
                       reduce, L.slots, c->gemm_mfma_min_tiles);
    L.ob = sizeof(nrg_gemm_out) * nouts;
    L.tb = sizeof(nrg_gemm_term) * nterms;
    L.ib = sizeof(GemmItem) * items.size();
    L.rb = sizeof(GemmItem) * reduce.size();
    L.nitems = (int)items.size();
    L.nreduce = (int)reduce.size();
    L.buf.resize(L.ob + L.tb + L.ib + L.rb);
    memcpy(L.buf.data(), outs, L.ob);
    if (L.tb) memcpy(L.buf.data() + L.ob, terms, L.tb);
    if (L.ib) memcpy(L.buf.data() + L.ob + L.tb, items.data(), L.ib);
    if (L.rb) memcpy(L.buf.data() + L.ob + L.tb + L.ib, reduce.data(), L.rb);
    for (int o = 0; o < nouts; ++o)
        for (int t = outs[o].term_begin; t < outs[o].term_begin + outs[o].term_count; ++t) {
            const double m = outs[o].m, n = outs[o].n, k = terms[t].k;
            L.flops += 2.0 * m * n * k;
            L.bytes += 8.0 * (m * k + k * n);
        }
    for (int o = 0; o < nouts; ++o) L.bytes += 8.0 * outs[o].m * outs[o].n * (outs[o].beta != 0.0 ? 2 : 1);
    return NRG_OK;
}

static nrg_status gemm_launch_staged(nrg_ctx *c, const GemmLaunch &L, const char *d, double *part) {
    return gemm_launch_dev(c, K_GEMM, (const nrg_gemm_out *)d, (const nrg_gemm_term *)(d + L.ob),
                           (const GemmItem *)(d + L.ob + L.tb), L.nitems, (const GemmItem *)(d + L.ob + L.tb + L.ib),
                           L.nreduce, L.flops, L.bytes, L.mfma, part);
}

nrg_status nrg_gemm_grouped(nrg_ctx *c, const nrg_gemm_out *outs, int nouts, const nrg_gemm_term *terms,
                            int nterms) {
    if (nouts <= 0) return NRG_OK;
    GemmLaunch L;
    nrg_status s = gemm_prepare(c, outs, nouts, terms, nterms, L);
    if (s) return s;
    if (!L.nitems) return NRG_OK;
    if (L.slots) {
        s = gemm_reserve(c, L.slots);
        if (s) return s;
    }
    void *d;
    s = stage(c, L.buf.data(), L.buf.size(), &d);
    if (s) return s;
    return gemm_launch_staged(c, L, (const char *)d, c->gpart);
}

nrg_status nrg_gemm_grouped_noalloc(nrg_ctx *c, const nrg_gemm_out *outs, int nouts, const nrg_gemm_term *terms,
                                    int nterms) {
    if (!c || !c->scratch) return fail(c, NRG_ERR_ARG, "gemm_noalloc: no device scratch array on this context");
    if (nouts <= 0) return NRG_OK;
    GemmLaunch L;
    nrg_status s = gemm_prepare(c, outs, nouts, terms, nterms, L);
    if (s) return s;
    if (!L.nitems) return NRG_OK;
    // [descriptors | partial tiles of split outputs] in the fixed array
    const size_t desc = (L.buf.size() + 255) & ~(size_t)255;
    const size_t need = desc + (size_t)L.slots * GT * GT * sizeof(double);
    if (need > NRG_NOALLOC_SCRATCH_BYTES)
        return fail(c, NRG_ERR_ARG, "gemm_noalloc: the launch needs " + std::to_string(need) +
                                        " bytes of scratch, the fixed device array holds " +
                                        std::to_string((size_t)NRG_NOALLOC_SCRATCH_BYTES));
    char *base = (char *)c->scratch;
    // a copy on the context's stream: ordered after the work of the previous
    // call that still reads the array
    HIPCHK(c, hipMemcpyAsync(base, L.buf.data(), L.buf.size(), hipMemcpyHostToDevice, c->stream));
    return gemm_launch_staged(c, L, base, (double *)(base + desc));
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
