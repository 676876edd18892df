// nrg_eig.hip - batched FP64 symmetric eigensolver (replaces the dsyev call
// of diagonalise(), src/diag.h:54-89 and 297-353).
//
// All matrices of one NRG iteration are solved together, stage by stage:
//   k_eig_tri_*  Householder tridiagonalisation, step j's rank-2 update fused
//                with step j+1's matrix-vector product.  Up to 128 rows: one
//                workgroup per matrix, rows in registers (k_eig_tri_one); up
//                to 576: a team of workgroups with the rows in registers
//                (k_eig_tri_reg); larger: a team streaming its rows through
//                L2/HBM (k_eig_tri_coop).  One exchange + barrier per step.
//   k_dc_*       Cuppen divide and conquer on the tridiagonal matrix: the
//                tridiagonal is cut into leaves of 1 or 2 rows, then merged
//                level by level for all matrices at once.  A merge sorts the
//                poles, deflates (small z components, close poles by a Givens
//                rotation), solves every secular equation in parallel
//                (bracketed two-pole rational iteration in the coordinates of
//                the nearer pole), forms Gu-Eisenstat z-hat from the computed
//                roots (Loewner), normalises the eigenvectors of D + rho z z^T
//                and multiplies them into the merged block with the grouped
//                GEMM.  Eigenvectors of the tridiagonal are kept column-major
//                (one contiguous row per eigenvector).
//   k_eig_bt     back-transformation (n < 1025): one wavefront per eigenvector
//                applies the reflectors H_{n-3} .. H_0 (staged eight at a time
//                in LDS, shared by the eight eigenvectors of a workgroup) with
//                the vector held in registers, and writes it as row rank(c)
//                of the result; larger n: compact WY blocks and GEMMs.
// Output convention as diagonalise(): eigenvalues ascending in W, row r of A
// = eigenvector r.
#include <type_traits>

#include "nrg_internal.h"

constexpr int TRI_NT = 1024;
constexpr int TRI_NW = TRI_NT / 64;
constexpr int DEFL_NT = 256;
constexpr int DEFL_LDS_K = 5000;  // 2 x K doubles + 4 x K ints = 32 B x K of dynamic LDS
constexpr int SEG_L = 32;  // gather segments: about one per SEG_L sorted poles
constexpr int COPY_ROWS = 16;  // deflated rows copied back per workgroup
// Kind codes of a sorted pole of a merge: bits 0-1 how its row enters the
// gather (0 deflated by a small z, 1 new pole, 2 rotated into the previous
// pole); bit 2 set when a later pole rotates it away (deflated); bit 3 marks
// a rotation candidate of the parallel pass.
constexpr int KIND_MASK = 3, KIND_SRC = 4, KIND_CAND = 8;
constexpr int SEC_NT = 64;
constexpr double kEps = 2.220446049250313e-16;

// Per-matrix workspace: Vc, S1, S2 (n x n each), then vectors of n doubles.
//  Vc  eigenvectors of the tridiagonal, column-major (row c = eigenvector c)
//  S1  gathered / rotated eigenvector block of a merge; result of the
//      back-transformation
//  S2  Delta_ji = d_i - lambda_j of a merge, overwritten by the normalised
//      eigenvectors of D + rho z z^T
enum VecSlot {
    V_TAU = 0,  // Householder scalars
    V_PV,       // team exchange: p = tau A v (even steps)
    V_VB,       //   next row (odd steps)
    V_DGT,      //   p (odd steps)
    V_ODT,      //   next row (even steps)
    V_DG,    // diagonal of T (after tridiagonalisation)
    V_OD,    // off-diagonal of T
    V_D,     // current eigenvalues (per position)
    V_BETA,  // coupling cut at a merge boundary, stored at index mid
    V_DD,    // sorted non-deflated poles of a merge (at a..a+K')
    V_ZZ,    // their z components
    V_ZH,    // z-hat
    V_TS,    // secular root offsets
    V_ORG,   // int: origin pole of each root
    V_META,  // int: [a] = K', [a+1] = number of rotations (K >= 2, so [a+1] is the merge's own)
    V_RHO,   // rho of the merge at [a]
    V_RANK,  // int: final rank of each eigenvalue
    V_KIND,  // int per sorted pole: 0 deflated (small z), 1 pole, 2 rotated into the previous pole
    V_RC,    // Givens c, s of a kind-2 pole (at its sorted index)
    V_RS,
    V_SDS,   // deflation arrays (global variant)
    V_SZS,
    V_PERM,
    V_POS,
    V_SEG,   // int: first sorted pole of each gather segment
    V_NSEG,  // int: [a] = number of gather segments of the merge at a
    V_COUNT
};

__host__ __device__ inline size_t eig_ws_doubles(int n) { return 3 * (size_t)n * n + (size_t)V_COUNT * n; }
__device__ inline double *vslot(double *ws, int n, int s) { return ws + 3 * (size_t)n * n + (size_t)s * n; }
__device__ inline int *islot(double *ws, int n, int s) { return reinterpret_cast<int *>(vslot(ws, n, s)); }

struct EigMat {
    double *A;
    double *W;
    double *ws;
    int n;
    int pad;  // 1: reduced by a team (reflectors in S2); 2: two-stage reduction
};

struct DcMerge {
    double *ws;
    int n, a, mid, b;
    int gemm;  // index of this merge's output descriptor
    int pad;
};

struct DcChunk {
    int merge;
    int j0;
};

struct BtItem {
    int mat;
    int c0;
};

__device__ inline double tri_block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < TRI_NW; ++k) s += red[k];
    return s;
}

// ---------------------------------------------------------------------------
// 1. tridiagonalisation: A <- H_j A H_j, j = 0..n-3, H_j = I - tau_j v_j v_j^T
//    acting on rows/columns j+1..n-1.  v_j ends in row j of A (columns
//    j+1..n-1, first element included; the streaming teams leave it in S2
//    for k_dc_leaf to copy), tau_j in V_TAU, T in V_DG / V_OD.
// ---------------------------------------------------------------------------

__global__ void k_fill_int(int *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// a failure flag joins the host-visible status (cleared by the host when no
// eigensolve is queued)
__global__ void k_status_or(const int *src, int *dst) {
    if (threadIdx.x == 0 && *src) *dst = *src;
}

// eigenvalues of matrix k (final once ranked) to dst + voff[k] (pinned host
// memory, in the caller's task order)
__global__ __launch_bounds__(256) void k_values_out(const EigMat *mats, const long long *voff, double *dst) {
    const EigMat mt = mats[blockIdx.x];
    for (int r = threadIdx.x; r < mt.n; r += blockDim.x) dst[voff[blockIdx.x] + r] = mt.W[r];
}

// ---------------------------------------------------------------------------
// 1b. team tridiagonalisation: T workgroups share one matrix; row r of A
//     belongs to rank r mod T for the whole reduction (cyclic, so the active
//     rows stay spread over the team).  A is stored in full, so the
//     matrix-vector product is a dot product per owned row.
//     Per step j (j = -1 .. n-3) every workgroup holds v_j (redundant
//     copies), reads p_j = tau_j A22 v_j and row j+1 of A (as left by step
//     j-1), computes w = p - K v, the updated row j+1 and from it v_{j+1}
//     (all redundant, O(n)), and finally updates its own rows (columns j+2..)
//     while computing p_{j+1} for them.  One barrier per step.
//     Only p and the next row cross workgroups; they are written and read
//     with device-scope atomic stores / loads, which are coherent across
//     XCDs without cache write-back.  Everything else a workgroup touches is
//     its own (its rows of A, private vectors), so the barrier needs no
//     fence: each wave waits for its stores, then the workgroups meet on a
//     relaxed atomic counter.  Reflectors go to S2 (rank 0) and are copied
//     into the rows of A by the next kernel.
// ---------------------------------------------------------------------------
constexpr int TEAM_W = 16;                   // rows per workgroup when sizing a team
constexpr int TEAM_RW = 2;                   // rows a wavefront updates together
constexpr int TEAM_CU = 4;                   // column chunks of 64 per round
constexpr int TEAM_LDS_N = 6800;             // v, v', w of the team in LDS (160 KB)
constexpr size_t TRI_COOP_LDS = 3 * (size_t)TEAM_LDS_N * sizeof(double);
constexpr int TEAM_MAX = 32;                 // workgroups of one XCD
constexpr int TEAM_GLOBAL_N = 1024;          // larger matrices: teams across XCDs
constexpr int TEAM_GLOBAL_MAX = 128;
constexpr int N_XCD = 8;                     // workgroup b runs on XCD b mod 8
// Bounded team waits (polls of ~1 us each, about two minutes): a member may
// first finish a long job of its own (a large blocked reduction in the same
// launch), so the bound only ends a team that could never assemble.
constexpr long long TEAM_SPIN_LIMIT = 1LL << 27;
constexpr int BAR_STRIDE = 32;                    // ints between team barrier counters (128 B)

struct TriJob {
    int mat, T;
    int blocked;  // 2: two-stage (sbr_blk), 1: blocked reduction (tri_blk), 0: streaming rows (tri_team)
    int next;     // job the same team runs next (a chain of jobs on one team), -1: none
};

// (global address space, so that they are global_* rather than flat_*
// instructions: a scalar base plus a 32-bit lane offset instead of a 64-bit
// address per element)
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) int gint;
__device__ inline void st_dev(double *p, double v) {
    __hip_atomic_store((gdouble *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_dev(const double *p) {
    return __hip_atomic_load((gdouble *)const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Memory-model assumption of the team kernels.  Everything one workgroup
// hands to another goes through agent-scope atomic stores and loads
// (st_dev / ld_dev: sc1 accesses, which bypass the non-coherent vector and
// L2 caches on gfx950), and the team barrier orders them with s_waitcnt
// vmcnt(0) before a relaxed counter increment and __syncthreads after the
// poll: the hardware completes an acknowledged agent-scope store before any
// later agent-scope load of another CU can miss it.  The C++/HIP model
// would ask for release on the increment and acquire on the poll; on
// gfx950 those add an L2 writeback (buffer_wbl2 sc1) per arrival and an L2
// invalidate (buffer_inv sc1) per exit, measured at +2 % (one 13136-row
// matrix) to +4 % (the keep-10000 blocks) of the blocked reduction.
// Checked with ROCm 7.2 (AMD clang 22); the build refuses other versions
// until the assumption has been checked there.
static_assert(HIP_VERSION_MAJOR == 7 && HIP_VERSION_MINOR >= 2,
              "team barriers rely on gfx950 agent-scope atomic coherence as checked with ROCm 7.2");

// Team barrier with a bounded wait: returns false (and flags the status) if
// the team did not arrive, so that every workgroup still terminates.
__device__ inline bool team_barrier(int *cnt, int target, int *status, int *s_flag,
                                    long long limit = TEAM_SPIN_LIMIT) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's device-scope stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long long spins = 0;
        int ok = 1;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > limit || __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                atomicCAS(status, 0, NRG_ERR_HIP);
                break;
            }
        }
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag != 0;
}

template <bool IN_LDS>
__device__ __forceinline__ void tri_team(const EigMat &mt, int rank, int T, int *bar, int *status, double *red,
                                         int *s_flag) {
    extern __shared__ double lds[];
    const int n = mt.n;
    double *__restrict__ A = mt.A;
    double *ws = mt.ws;
    const int tid = threadIdx.x;
    double *priv = ws + (size_t)n * n + (size_t)rank * 4 * n;  // private area in S1
    double *vecs = IN_LDS ? lds : priv;
    double *va = vecs, *vb = vecs + n, *wv = vecs + 2 * n, *xr = priv + 3 * n;
    double *const pv0 = vslot(ws, n, V_PV), *const pv1 = vslot(ws, n, V_DGT);  // p, by step parity
    double *const rb0 = vslot(ws, n, V_ODT), *const rb1 = vslot(ws, n, V_VB);  // next row, by parity
    double *Y = ws + 2 * (size_t)n * n;  // reflectors (row j, columns j+1..)
    double *tau = vslot(ws, n, V_TAU);
    double *gdg = vslot(ws, n, V_DG), *god = vslot(ws, n, V_OD);
    double tj = 0.0;  // tau_j
    int phase = 0;
    for (int j = -1; j <= n - 3; ++j) {
        const int m = n - j - 1;  // length of v_j / p_j; row j+1 has m entries from column j+1
        if (rank == 0 && j >= 0 && tj != 0.0)
            for (int k = tid; k < m; k += TRI_NT) Y[(size_t)j * n + j + 1 + k] = va[k];
        // row j+1 as left by step j-1 (row 0: the assembled matrix)
        if (j < 0) {
            for (int k = tid; k < m; k += TRI_NT) xr[k] = A[k];
        } else {
            const double *rg = ((j + 2) & 1) ? rb1 : rb0;
            for (int k = tid; k < m; k += TRI_NT) xr[k] = ld_dev(rg + k);
        }
        if (tj != 0.0) {
            const double *pg = ((j + 2) & 1) ? pv1 : pv0;
            double pp = 0.0;
            for (int k = tid; k < m; k += TRI_NT) {
                const double x = ld_dev(pg + k);
                wv[k] = x;
                pp += va[k] * x;
            }
            const double Kc = 0.5 * tj * tri_block_sum(pp, red);
            for (int k = tid; k < m; k += TRI_NT) wv[k] -= Kc * va[k];
            __syncthreads();
            // x = row j+1 after step j's update
            const double v0 = va[0], w0 = wv[0];
            for (int k = tid; k < m; k += TRI_NT) xr[k] -= v0 * wv[k] + w0 * va[k];
        }
        __syncthreads();
        double psum = 0.0;
        for (int k = 2 + tid; k < m; k += TRI_NT) psum += xr[k] * xr[k];
        const double sigma = tri_block_sum(psum, red);
        const double x0 = xr[0], x1 = (m >= 2) ? xr[1] : 0.0;
        const bool has_next = (j + 1) <= n - 3;
        double tn = 0.0, alpha = x1;
        if (has_next && sigma != 0.0) {
            alpha = -copysign(sqrt(x1 * x1 + sigma), x1);
            const double v0 = x1 - alpha;
            tn = 2.0 / (v0 * v0 + sigma);
            for (int k = tid; k < m - 1; k += TRI_NT) vb[k] = (k == 0) ? v0 : xr[k + 1];
        }
        if (rank == 0 && tid == 0) {
            gdg[j + 1] = x0;
            god[j + 1] = alpha;
            if (has_next) tau[j + 1] = tn;
        }
        __syncthreads();
        double *rnext = ((j + 3) & 1) ? rb1 : rb0;
        double *pnext = ((j + 3) & 1) ? pv1 : pv0;
        // own rows (cyclic, r mod T = rank): update columns j+2..,
        // p_{j+1}[r] = tn * (row r . v_{j+1}) (A is symmetric), and the owner
        // of row j+2 publishes it for the next step.  A wavefront takes up to
        // TEAM_RW of its rows at once, so each column's vector entries are
        // read once for all of them and their loads are in flight together.
        {
            const int first = j + 2;
            const int wave = tid >> 6, lane = tid & 63;
            const int r0 = first + ((rank - first % T) + T) % T;
            const int rstep = TRI_NW * T;
            for (int rg = r0 + wave * T; rg < n; rg += TEAM_RW * rstep) {
                double vr[TEAM_RW], wr[TEAM_RW], acc[TEAM_RW];
#pragma unroll
                for (int k = 0; k < TEAM_RW; ++k) {
                    const int r = rg + k * rstep;
                    const bool ok = r < n && tj != 0.0;
                    vr[k] = ok ? va[r - j - 1] : 0.0;
                    wr[k] = ok ? wv[r - j - 1] : 0.0;
                    acc[k] = 0.0;
                }
                int cidx = first + lane;
                // TEAM_CU column chunks x TEAM_RW rows of independent loads per round
                for (; cidx + 64 * (TEAM_CU - 1) < n; cidx += 64 * TEAM_CU) {
                    double a[TEAM_CU][TEAM_RW];
#pragma unroll
                    for (int u = 0; u < TEAM_CU; ++u)
#pragma unroll
                        for (int k = 0; k < TEAM_RW; ++k) {
                            const int r = rg + k * rstep;
                            a[u][k] = (r < n) ? A[(size_t)r * n + cidx + 64 * u] : 0.0;
                        }
#pragma unroll
                    for (int u = 0; u < TEAM_CU; ++u) {
                        const int cu = cidx + 64 * u;
                        const double vc = (tj != 0.0) ? va[cu - j - 1] : 0.0;
                        const double wc = (tj != 0.0) ? wv[cu - j - 1] : 0.0;
                        const double bc = (tn != 0.0) ? vb[cu - first] : 0.0;
#pragma unroll
                        for (int k = 0; k < TEAM_RW; ++k) {
                            const int r = rg + k * rstep;
                            if (r < n) {
                                if (tj != 0.0) {
                                    a[u][k] -= vr[k] * wc + wr[k] * vc;
                                    A[(size_t)r * n + cu] = a[u][k];
                                }
                                if (r == first) st_dev(rnext + (cu - first), a[u][k]);
                                acc[k] += a[u][k] * bc;
                            }
                        }
                    }
                }
                for (; cidx < n; cidx += 64) {
                    const double vc = (tj != 0.0) ? va[cidx - j - 1] : 0.0;
                    const double wc = (tj != 0.0) ? wv[cidx - j - 1] : 0.0;
                    const double bc = (tn != 0.0) ? vb[cidx - first] : 0.0;
#pragma unroll
                    for (int k = 0; k < TEAM_RW; ++k) {
                        const int r = rg + k * rstep;
                        if (r < n) {
                            double a = A[(size_t)r * n + cidx];
                            if (tj != 0.0) {
                                a -= vr[k] * wc + wr[k] * vc;
                                A[(size_t)r * n + cidx] = a;
                            }
                            if (r == first) st_dev(rnext + (cidx - first), a);
                            acc[k] += a * bc;
                        }
                    }
                }
                if (tn != 0.0) {
#pragma unroll
                    for (int k = 0; k < TEAM_RW; ++k) {
                        const int r = rg + k * rstep;
                        const double sacc = wave_sum(acc[k]);
                        if (r < n && lane == 0) st_dev(pnext + (r - first), tn * sacc);
                    }
                }
            }
        }
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag)) return;
        // v_{j+1} becomes current
        if (has_next) {
            for (int k = tid; k < m - 1; k += TRI_NT) va[k] = vb[k];
            __syncthreads();
        }
        tj = tn;
    }
    if (rank == 0 && tid == 0) {
        gdg[n - 1] = ld_dev((n & 1) ? rb1 : rb0);  // A[n-1][n-1], published by the last pass
        god[n - 1] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// 1c. register-resident reduction for n <= REG_MAX_N.  A team of T
//     workgroups of NW wavefronts (T = 1: one workgroup, exchange through
//     LDS) owns the rows of A in registers: row r = rank + T (wave + NW s)
//     sits in slot s of that wavefront, lane l holding columns l + 64 i.
//     Teams run in their own persistent kernel of REG_NT threads, whose
//     register budget (256 per lane) holds RPW x NC doubles of rows plus the
//     step's vectors without spilling.  Per step j every
//     wavefront redundantly rebuilds w_j, the updated row j+1 and v_{j+1}
//     for its own columns from p_j and row j+1 as published by step j-1
//     (read once per workgroup into LDS, beside v_j, which wavefront 0 leaves
//     there), so a step synchronises only at the exchange (one team barrier)
//     and at that staging.  The own-row update is issued before the
//     reflector's reductions and square root, whose latency it fills.
//     Reductions are wave butterflies, which leave the same bits in every
//     lane, so all copies of v_{j+1} and tau_{j+1} agree.
//     Output as tri_core: reflector j in row j of A (columns j+1..), tau,
//     diagonal and off-diagonal.
// ---------------------------------------------------------------------------
constexpr int REG_ONE_N = 128;  // one workgroup: 16 waves x 8 rows, 2 column registers
constexpr int REG_MAX_N = 576;  // largest team variant: 9 column registers
constexpr int REG_NT = 512;     // team workgroups: 8 waves, two per SIMD (one per SIMD exposes the
                                // latency of each step's reduction chain: 12% slower)
// Variants by column registers NC (d <= 64 NC): two rows per wavefront up to
// d = 512 (teams of up to 32 workgroups, one XCD), three at d <= 576.  Few
// rows keep the per-step row reductions short and a tight NC keeps the
// per-step vector work small (d = 416 with NC = 9 and three rows: 3.8 us
// per step; NC = 7 and two rows on 26 workgroups is the same work per lane
// as d <= 448's share).  One row at d <= 256 leaves too few CUs to the
// one-workgroup matrices on the side stream.
constexpr int REG_TWO_ROWS_N = 512;
__host__ __device__ constexpr int reg_rpw(int n) { return n <= REG_TWO_ROWS_N ? 2 : 3; }

// team size of the register variant for dimension n (REG_ONE_N < n <=
// REG_MAX_N): workgroups of 8 waves x reg_rpw(n) rows
__host__ __device__ inline int reg_team(int n) {
    const int rows = (REG_NT / 64) * reg_rpw(n);
    return (n + rows - 1) / rows;
}

// Wave sums of R values at once (as wave_sum: the same bits in every lane);
// the R chains are independent and overlap.
template <int R>
__device__ __forceinline__ void wave_sum_n(double (&v)[R]) {
#pragma unroll
    for (int s = 0; s < R; ++s) v[s] = wave_sum(v[s]);
}

// Calls f(integral_constant<P>, j) with P the largest multiple of PS not
// above the run-time p (0 <= p < NC): at most ceil(NC / PS) copies of the
// step, few enough to keep the register allocation of one.
template <int P, int PS, int NC, class F>
__device__ __forceinline__ bool tri_reg_dispatch(int p, F &f, int j) {
    if constexpr (P + PS >= NC) {
        return f(std::integral_constant<int, P>{}, j);
    } else {
        if (p < P + PS) return f(std::integral_constant<int, P>{}, j);
        return tri_reg_dispatch<P + PS, PS, NC>(p, f, j);
    }
}

template <int NC, int RPW, int NW, bool TEAM>
__device__ __forceinline__ void tri_reg(const EigMat &mt, int rank, int T, int *bar, int *status, int *s_flag,
                                        double *lbuf, double *vsh) {
    const int n = mt.n;
    double *__restrict__ A = mt.A;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (n == 1) {
        if (rank == 0 && tid == 0) {
            mt.W[0] = A[0];
            A[0] = 1.0;
        }
        return;
    }
    if (n > 64 * NC || T * NW * RPW < n) {  // the same for every workgroup of the team
        if (tid == 0) atomicCAS(status, 0, NRG_ERR_ARG);
        return;
    }
    double *ws = mt.ws;
    double *tau = vslot(ws, n, V_TAU), *gdg = vslot(ws, n, V_DG), *god = vslot(ws, n, V_OD);
    double *const pb0 = TEAM ? vslot(ws, n, V_PV) : lbuf, *const pb1 = TEAM ? vslot(ws, n, V_DGT) : lbuf + n;
    double *const rb0 = TEAM ? vslot(ws, n, V_ODT) : lbuf + 2 * n, *const rb1 = TEAM ? vslot(ws, n, V_VB) : lbuf + 3 * n;
    const bool lead = rank == 0 && wave == 0;
    if constexpr (TEAM) {
        // staging buffer: columns outside the active range read as zero
        for (int k = tid; k < 128 * NC; k += 64 * NW) lbuf[k] = 0.0;  // first read after the step's barrier
    }
    int row[RPW];
#pragma unroll
    for (int s = 0; s < RPW; ++s) row[s] = rank + T * (wave + NW * s);
    // loads use clamped indices and are masked afterwards, so that they
    // are unconditional (atomic loads are never speculated)
    double a[RPW][NC];
#pragma unroll
    for (int s = 0; s < RPW; ++s) {
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const int c = lane + 64 * i;
            const double x = A[(size_t)min(row[s], n - 1) * n + min(c, n - 1)];
            a[s][i] = (row[s] < n && c < n) ? x : 0.0;
        }
    }
    // v_j (columns j+1..), its entries at column j+1 and at the own rows
    double vv[NC], v1 = 0.0, vrow[RPW];
#pragma unroll
    for (int i = 0; i < NC; ++i) vv[i] = 0.0;
#pragma unroll
    for (int s = 0; s < RPW; ++s) vrow[s] = 0.0;
    double tj = 0.0;
    int phase = 0;
    // step j; registers below (j + 1) / 64 hold only finished columns
    // (<= j), whose operands are all zero: the work of those below P is
    // skipped
    auto step = [&](auto Pc, const int j) -> bool {
        constexpr int P = decltype(Pc)::value;
        double w[NC], xr[NC], w1 = 0.0, wrow[RPW], x0, x1;
        if (j < 0) {
#pragma unroll
            for (int i = P; i < NC; ++i) {
                const int c = lane + 64 * i;
                const double x = A[min(c, n - 1)];
                xr[i] = c < n ? x : 0.0;
                w[i] = 0.0;
            }
#pragma unroll
            for (int s = 0; s < RPW; ++s) wrow[s] = 0.0;
            x0 = A[0];
            x1 = A[1];
        } else {
            const double *pg = ((j + 2) & 1) ? pb1 : pb0, *rg = ((j + 2) & 1) ? rb1 : rb0;
            if constexpr (TEAM) {
                // the workgroup reads the published p_j and row j+1 (columns
                // j+1..n-1) once into LDS, rather than every wavefront from L2
                constexpr int NL = (2 * NC + NW - 1) / NW;
                const int m = n - j - 1;
                double t[NL];
#pragma unroll
                for (int u = 0; u < NL; ++u) {
                    const int k = min(tid + 64 * NW * u, 2 * m - 1);
                    t[u] = k < m ? ld_dev(pg + j + 1 + k) : ld_dev(rg + j + 1 + k - m);
                }
                // (p in lbuf[0, 64 NC), the row in lbuf[64 NC, 128 NC): zero
                // outside columns j+1..n-1, column j cleared here)
#pragma unroll
                for (int u = 0; u < NL; ++u) {
                    const int k = tid + 64 * NW * u;
                    if (k < 2 * m) lbuf[k < m ? j + 1 + k : 64 * NC + j + 1 + k - m] = t[u];
                }
                if (tid == 0) {
                    lbuf[j] = 0.0;
                    lbuf[64 * NC + j] = 0.0;
                }
                __syncthreads();
                pg = lbuf;
                rg = lbuf + 64 * NC;
            }
            double pp = 0.0;
#pragma unroll
            for (int i = P; i < NC; ++i) {
                const int c = lane + 64 * i;
                if constexpr (TEAM) {
                    w[i] = pg[c];
                    xr[i] = rg[c];
                } else {
                    const bool ok = c > j && c < n;
                    const unsigned cc = (unsigned)min(c, n - 1);
                    const double pl = pg[cc];
                    const double rl = rg[cc];
                    w[i] = ok ? pl : 0.0;
                    xr[i] = ok ? rl : 0.0;
                }
                pp += vv[i] * w[i];
            }
            // p at column j+1 and at the own rows (read in LDS), and v_j at
            // the own rows (written to LDS by wavefront 0 at the end of the
            // previous step)
            const double p1 = pg[j + 1];
            double prow[RPW];
#pragma unroll
            for (int s = 0; s < RPW; ++s) {
                const int r = row[s];
                const double x = pg[min(r, n - 1)];
                prow[s] = (r > j && r < n) ? x : 0.0;
                vrow[s] = vsh[(j & 1) * 64 * NC + min(r, 64 * NC - 1)];
            }
            const double K = 0.5 * tj * wave_sum(pp);
#pragma unroll
            for (int i = P; i < NC; ++i) w[i] -= K * vv[i];
            w1 = p1 - K * v1;
#pragma unroll
            for (int s = 0; s < RPW; ++s) wrow[s] = prow[s] - K * vrow[s];
#pragma unroll
            for (int i = P; i < NC; ++i) xr[i] -= v1 * w[i] + w1 * vv[i];
            // row j+1 at columns j+1 and j+2 from the LDS copies (the same
            // inputs in every wavefront; these two entries are used only for
            // the diagonal and the next reflector's pivot)
            const double vv2 = vsh[(j & 1) * 64 * NC + j + 2];
            const double w2 = pg[j + 2] - K * vv2;
            x0 = rg[j + 1] - (v1 * w1 + w1 * v1);
            x1 = rg[j + 2] - (v1 * w2 + w1 * vv2);
        }
        // own rows: update with (v_j, w_j), independent of the reflector
        // below, so that it fills the latency of that chain (finished or
        // padding rows are updated too, but never published)
#pragma unroll
        for (int s = 0; s < RPW; ++s) {
#pragma unroll
            for (int i = P; i < NC; ++i) a[s][i] -= vrow[s] * w[i] + wrow[s] * vv[i];
        }
        // row j+1 is final: diagonal x0, and the reflector that zeroes it
        // beyond column j+2 (branch-free: a uniform choice of results)
        double ps = 0.0;
#pragma unroll
        for (int i = P; i < NC; ++i) {
            const int c = lane + 64 * i;
            if (c > j + 2 && (TEAM || c < n)) ps += xr[i] * xr[i];  // (team rows are zero past n)
        }
        const double sigma = wave_sum(ps);
        const bool has_next = j + 1 <= n - 3;
        const bool refl = has_next && sigma != 0.0;
        const double sg = refl ? sigma : 1.0;
        const double at = -copysign(sqrt(x1 * x1 + sg), x1);
        const double alpha = refl ? at : x1;
        const double v0 = refl ? x1 - at : 0.0;
        const double tn = refl ? 2.0 / (v0 * v0 + sg) : 0.0;
        const bool live = tn != 0.0;
        double nv[NC];
#pragma unroll
        for (int i = 0; i < P; ++i) nv[i] = 0.0;
#pragma unroll
        for (int i = P; i < NC; ++i) {
            const int c = lane + 64 * i;
            nv[i] = !live ? 0.0 : (c == j + 2) ? v0 : ((c > j + 2 && (TEAM || c < n)) ? xr[i] : 0.0);
        }
        // p_{j+1}[r] = tn row_r . v_{j+1}; the owner of row j+2 publishes
        // that row
        double acc[RPW];
#pragma unroll
        for (int s = 0; s < RPW; ++s) {
            acc[s] = 0.0;
#pragma unroll
            for (int i = P; i < NC; ++i) acc[s] += a[s][i] * nv[i];
        }
        wave_sum_n<RPW>(acc);
        double *pn = ((j + 3) & 1) ? pb1 : pb0, *rn = ((j + 3) & 1) ? rb1 : rb0;
#pragma unroll
        for (int s = 0; s < RPW; ++s) {
            const int r = row[s];
            if (r >= j + 2 && r < n) {
                if (lane == 0) {
                    if (TEAM) st_dev(pn + r, tn * acc[s]);
                    else pn[r] = tn * acc[s];
                }
                if (r == j + 2) {
#pragma unroll
                    for (int i = P; i < NC; ++i) {
                        const int c = lane + 64 * i;
                        if (c >= j + 2 && c < n) {
                            if (TEAM) st_dev(rn + c, a[s][i]);
                            else rn[c] = a[s][i];
                        }
                    }
                }
            }
        }
        // the lead wavefront's outputs, off the critical path: reflector j,
        // diagonal, off-diagonal and tau
        if (lead) {
            if (j >= 0) {
#pragma unroll
                for (int i = P; i < NC; ++i) {
                    const int c = lane + 64 * i;
                    if (c > j && c < n) A[(size_t)j * n + c] = vv[i];
                }
            }
            if (lane == 0) {
                gdg[j + 1] = x0;
                god[j + 1] = alpha;
                if (has_next) tau[j + 1] = tn;
            }
        }
        // v_{j+1} becomes current; wavefront 0 leaves it in LDS for the
        // entries at the own rows (read after the step's barrier)
#pragma unroll
        for (int i = P; i < NC; ++i) vv[i] = nv[i];
        v1 = live ? v0 : 0.0;
        if (wave == 0) {
#pragma unroll
            for (int i = 0; i < NC; ++i) vsh[((j + 1) & 1) * 64 * NC + lane + 64 * i] = nv[i];
        }
        tj = tn;
        if (TEAM) {
            ++phase;
            if (!team_barrier(bar, phase * T, status, s_flag)) return false;
        } else {
            __syncthreads();
        }
            return true;
    };
    for (int j = -1; j <= n - 3; ++j)
        if (!tri_reg_dispatch<0, TEAM ? (NC + 2) / 3 : NC, NC>((j + 1) >> 6, step, j)) return;

    if (rank == 0 && tid == 0) {
        const double *rl = (n & 1) ? rb1 : rb0;  // row n-1, published by the last step
        gdg[n - 1] = TEAM ? ld_dev(rl + n - 1) : rl[n - 1];
        god[n - 1] = 0.0;
    }
}

// One persistent launch for the team matrices of a batch.  Every workgroup
// first draws tickets from the global queue (teams of the largest matrices,
// spread over XCDs), then workgroups of XCD x (blockIdx mod 8) draw from queue
// x; a job of T tickets is a team of T workgroups.  Tickets are handed out in
// job order and a workgroup finishes its job before it draws again, so every
// team assembles.
// Matrices up to REG_ONE_N: one workgroup each.
__global__ __launch_bounds__(TRI_NT) void k_eig_tri_one(const EigMat *mats, const int *idx, int *status) {
    __shared__ double lbuf[4 * REG_ONE_N];
    __shared__ double vsh[2 * 64 * 2];
    const EigMat mt = mats[idx[blockIdx.x]];
    tri_reg<2, 8, TRI_NW, false>(mt, 0, 1, nullptr, status, nullptr, lbuf, vsh);
}

// One workgroup on a matrix of the queue that follows the teams; not inlined,
// so that its registers are allocated apart from the team variants'.
template <int NC, int RPW, int NW>
__device__ __noinline__ void tri_one_wg(const EigMat &mt, int *status, double *lbuf, double *vsh) {
    tri_reg<NC, RPW, NW, false>(mt, 0, 1, nullptr, status, nullptr, lbuf, vsh);
}

// One persistent launch for the register-resident teams: workgroups of XCD
// x draw tickets from queue x (a team stays on one XCD), as k_eig_tri_coop.
__global__ __launch_bounds__(REG_NT) void k_eig_tri_reg(const EigMat *mats, const TriJob *jobs, const int *tbase,
                                                        const int *qjob, int *counters, int *status, const int *ones,
                                                        int nones) {
    __shared__ int s_ticket, s_flag;
    __shared__ double lbuf[2 * 64 * (REG_MAX_N / 64)];
    __shared__ double vsh[2 * 64 * (REG_MAX_N / 64)];
    const int x = blockIdx.x % N_XCD;
    int *ticket = counters + x;
    int *bars = counters + N_XCD + 1;
    const int j0 = qjob[1 + x], j1 = qjob[2 + x];
    const int t0 = tbase[j0], total = tbase[j1] - t0;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(ticket, 1);
        __syncthreads();
        const int t = s_ticket;
        __syncthreads();
        if (t >= total) break;
        int lo = j0, hi = j1;  // last job of the queue with tbase <= t0 + t
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (tbase[mid] <= t0 + t) lo = mid;
            else hi = mid;
        }
        const TriJob jb = jobs[lo];
        const EigMat mt = mats[jb.mat];
        const int rank = t0 + t - tbase[lo];
        int *const bar = bars + BAR_STRIDE * lo;
        constexpr int NW = REG_NT / 64;
        if (mt.n <= 256) tri_reg<4, reg_rpw(256), NW, true>(mt, rank, jb.T, bar, status, &s_flag, lbuf, vsh);
        else if (mt.n <= 320) tri_reg<5, reg_rpw(320), NW, true>(mt, rank, jb.T, bar, status, &s_flag, lbuf, vsh);
        else if (mt.n <= 384) tri_reg<6, reg_rpw(384), NW, true>(mt, rank, jb.T, bar, status, &s_flag, lbuf, vsh);
        else if (mt.n <= 448) tri_reg<7, reg_rpw(448), NW, true>(mt, rank, jb.T, bar, status, &s_flag, lbuf, vsh);
        else if (mt.n <= 512) tri_reg<8, reg_rpw(512), NW, true>(mt, rank, jb.T, bar, status, &s_flag, lbuf, vsh);
        else tri_reg<9, reg_rpw(576), NW, true>(mt, rank, jb.T, bar, status, &s_flag, lbuf, vsh);
        __syncthreads();
    }
    // then the one-workgroup matrices (n <= REG_ONE_N, largest first) from a
    // queue shared by all XCDs: every CU already holds a workgroup of this
    // launch, so they need no CU of their own
    int *oticket = counters + N_XCD;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(oticket, 1);
        __syncthreads();
        const int t = s_ticket;
        __syncthreads();
        if (t >= nones) break;
        const EigMat mt = mats[ones[t]];
        constexpr int NW = REG_NT / 64;
        if (mt.n <= 64) tri_one_wg<1, 64 / NW, NW>(mt, status, lbuf, vsh);
        else tri_one_wg<2, REG_ONE_N / NW, NW>(mt, status, lbuf, vsh);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 1d. blocked tridiagonalisation for large matrices (the dsytrd / dlatrd
//     scheme).  Panels of BLK_NB columns: inside a panel the trailing matrix
//     keeps its value at the panel start, and column j and the product A v_j
//     are corrected by the panel's earlier reflectors (A - V W^T - W V^T, with
//     W kept as p - K v: p = tau A v, K = tau (p.v) / 2 per column); the panel
//     ends with one rank-2 BLK_NB update of the trailing matrix (SYR2K) on
//     FP64 MFMA.
//     The product reads only the upper triangle of the row-major storage
//     (c >= r: the lower triangle of the symmetric matrix read column-major).
//     An off-diagonal 64 x 64 tile (R < C) gives both A x (row part, rows of
//     R) and A^T x (column part, columns of C); diagonal tiles are kept whole
//     (the SYR2K updates both of their halves) and give the row part alone.
//     This halves the HBM traffic of the product, which bounds the whole
//     reduction (sum_j 4 (n-j)^2 bytes, (4/3) n^3 per matrix).
//     Ownership: 64-column block B (columns 64B .. 64B+63) belongs to team
//     rank blk_owner(B) for the whole reduction; only that rank touches A's
//     columns of B and rows B of the panel buffers with plain loads and
//     stores (tile (R, C) is processed by owner(C)).  What crosses workgroups
//     (column j, partial sums, the panel's V and p, the tiles' partial
//     products) is written with device-scope stores and read by other ranks
//     with device-scope loads.  Per column:
//       P1  owners: column j corrected by the panel, its tail stored as the
//           reflector, per-block partial sums |a|^2, W^T a, V^T a  -> barrier
//       P2  all ranks: block sums in block order and the reflector (the same
//           bits everywhere); owners: partial products of the tiles -> barrier
//       P3  owners: p = tau (A v - V W^T v - W V^T v) from the partial
//           products in a fixed order, per-block p.v                -> barrier
//     Every sum runs in an order fixed by (n, j) and the team size, so a
//     matrix reduced by a team of the same size gives the same bits whatever
//     else runs beside it; another team size changes the column parts' order
//     of summation (per wavefront of the owner) and so the last bits.
//     Output as tri_team: reflector j in row j of A (columns j+1.., first
//     element included), tau, diagonal and off-diagonal of T.
// ---------------------------------------------------------------------------
constexpr int BLK_NB = 32;   // panel width
constexpr int BLK_TWS = 65;  // row stride of a wavefront's transpose buffer (16 rows of 64)
constexpr int BLK_TW = 16 * BLK_TWS;
constexpr int BLK_P1W = 1 + 2 * BLK_NB;          // partial sums per block: |a|^2, W^T a, V^T a
constexpr long long BLK_SPIN_LIMIT = TEAM_SPIN_LIMIT;
// dynamic LDS (doubles): 16 transpose buffers, 16 vector buffers, coefficient arrays
constexpr int BLK_LDS_TW = 0, BLK_LDS_XS = BLK_LDS_TW + TRI_NW * BLK_TW, BLK_LDS_CF = BLK_LDS_XS + TRI_NW * 64;
enum BlkCoef { CF_GAM = 0, CF_BET, CF_K, CF_UW, CF_UV, CF_WJ, CF_VJ1, CF_WJ1, CF_SC, CF_COUNT };
constexpr size_t BLK_LDS = sizeof(double) * (BLK_LDS_CF + 64 * CF_COUNT);
static_assert(BLK_LDS <= TRI_COOP_LDS, "the blocked reduction fits the team launch's LDS");
static_assert(2 * 2 * BLK_NB * 64 <= TRI_NW * BLK_TW, "SYR2K operands fit in the transpose buffers");
static_assert(16 * 16 * 64 <= TRI_NW * BLK_TW, "P3 partial sums fit in the transpose buffers");
static_assert(BLK_NB <= 64 && 1 + 2 * BLK_NB <= 16 * 64, "coefficient arrays");

// S1 layout of a blocked reduction (doubles from a 512-byte aligned start):
// V and p of the panel (BLK_NB x npad each), the column, partial sums (P1,
// P3), row partial products of the tiles (nt x nt x 64), and per block the
// column partial products summed by each wavefront of its owner
// (nt x TRI_NW x 64).
__host__ __device__ inline size_t blk_ws_doubles(int n) {
    const size_t npad = ((size_t)n + 63) & ~(size_t)63, nt = npad / 64;
    return 64 + 2 * BLK_NB * npad + npad + nt * (BLK_P1W + 1) + nt * nt * 64 + nt * TRI_NW * 64;
}
__host__ __device__ inline int blk_owner(int B, int T) {
    const int p = B % (2 * T);
    return p < T ? p : 2 * T - 1 - p;  // snake: pairs of blocks balance the shrinking tiles
}
typedef double blk_d4 __attribute__((ext_vector_type(4)));

// LDS written by some lanes of a wavefront and read by others: the wave
// executes its LDS operations in order; this keeps the compiler from moving
// them across each other.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v_j at index i (zero outside [jj, n); v[jj] = v0, the tail from the column)
__device__ __forceinline__ double blk_x(const double *av, int i, int jj, int n, double v0) {
    return i == jj ? v0 : (i > jj && i < n) ? ld_dev(av + i) : 0.0;
}

// Partial products of tile (R, C) with x = v_j: row part sum_c A[r][c] x[c]
// -> Prow[R][C]; the column part sum_r A[r][c] x[r] (lane = column) is
// returned (zero for a diagonal tile).  One wavefront, lane = column; rows in chunks of 16, which
// pass through the wave's LDS buffer for the row part (lane = row x quarter).
// Edge tiles (rows or columns beyond n) load at clamped addresses and zero
// the values outside the matrix.
template <bool EDGE>
__device__ __forceinline__ void blk_chunk_load(const gdouble *A, int cc, int rb, int n, bool cok, double (&a)[16]) {
    // a uniform row base and 32-bit lane offsets (scalar-base addressing).
    // Edge chunks clamp both the base row and every row to n - 1, so that
    // the offset from the base is never negative (a chunk may start below
    // the matrix's last row).
    const int rbase = EDGE ? min(rb, n - 1) : rb;
    const gdouble *rowb = A + (size_t)rbase * n;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) a[rr] = rowb[(uint32_t)((EDGE ? min(rb + rr, n - 1) - rbase : rr) * n + cc)];
    if (EDGE) {
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) a[rr] = (cok && rb + rr < n) ? a[rr] : 0.0;
    }
}

// Row part of a chunk (rows 16q..16q+15 of the tile) through the wave's
// transpose buffer: the sum of row 16q + r ends up in lanes r, r+16, r+32,
// r+48, and lane 16q + r keeps it in `rowp`, so after four chunks lane l
// holds row l of the tile's row partial product (no store per chunk: a
// store's completion would be waited for with the next chunk's loads).
__device__ __forceinline__ void blk_chunk_use(const double (&a)[16], int q, double xr, double &colacc, double *Tw,
                                              const double *xs, int lane, double &rowp) {
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
        colacc = fma(a[rr], readlane_d(xr, 16 * q + rr), colacc);
        Tw[rr * BLK_TWS + lane] = a[rr];
    }
    wave_lds_sync();
    const int r2 = lane & 15, g = lane >> 4;
    double s = 0.0;
#pragma unroll
    for (int cc = 0; cc < 16; ++cc) s = fma(Tw[r2 * BLK_TWS + 16 * g + cc], xs[16 * g + cc], s);
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    rowp = g == q ? s : rowp;
    wave_lds_sync();
}

// One tile by one wavefront.  The previous tile's row partial (pend) is
// stored after this tile's first loads are issued, so no load waits for it.
template <bool EDGE>
__device__ __forceinline__ double blk_tile_t(const double *__restrict__ A, int n, int nt, int R, int C, int jj,
                                             double v0, const double *av, double *Prow, double *Tw, double *xs,
                                             int lane, double *&pend_p, double &pend_v) {
    const gdouble *p = (const gdouble *)A;
    const int c = 64 * C + lane, r0 = 64 * R;
    const bool cok = c < n;
    const int cc = cok ? c : n - 1;
    const double xc = blk_x(av, c, jj, n, v0);
    const double xr = blk_x(av, r0 + lane, jj, n, v0);
    double a[16];
    blk_chunk_load<EDGE>(p, cc, r0, n, cok, a);
    if (pend_p) st_dev(pend_p + lane, pend_v);
    xs[lane] = xc;
    double colacc = 0.0, rowp = 0.0;
    blk_chunk_use(a, 0, xr, colacc, Tw, xs, lane, rowp);
#pragma unroll 1
    for (int q = 1; q < 4; ++q) {
        blk_chunk_load<EDGE>(p, cc, r0 + 16 * q, n, cok, a);
        blk_chunk_use(a, q, xr, colacc, Tw, xs, lane, rowp);
    }
    pend_p = Prow + ((size_t)R * nt + C) * 64;
    pend_v = rowp;
    return R != C ? colacc : 0.0;
}

__device__ __forceinline__ double blk_tile(const double *__restrict__ A, int n, int nt, int R, int C, int jj,
                                           double v0, const double *av, double *Prow, double *Tw, double *xs, int lane,
                                           double *&pend_p, double &pend_v) {
    if (64 * C + 63 < n && 64 * R + 63 < n)
        return blk_tile_t<false>(A, n, nt, R, C, jj, v0, av, Prow, Tw, xs, lane, pend_p, pend_v);
    return blk_tile_t<true>(A, n, nt, R, C, jj, v0, av, Prow, Tw, xs, lane, pend_p, pend_v);
}

// The workgroup's tiles of the product (tile columns it owns, rows R0..C),
// flattened over its wavefronts.  The column parts of block C stay with its
// owner: each wavefront sums those of its tiles (in row order) and stores the
// sum to Pcw[C][wavefront] with a plain store (read back by the same
// wavefront).  Not inlined: the tile loop gets registers of its own.
template <class P>
__device__ __forceinline__ P *uni_ptr(P *p) {
    const uint64_t u = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u), hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return (P *)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __noinline__ void blk_symv(const double *A_, int n_, int nt_, int R0_, int rank_, int T_, int jj_, double v0,
                                      const double *av_, double *Prow_, double *Pcw_) {
    extern __shared__ double lds[];
    // arguments arrive in vector registers; they are uniform, and saying so
    // keeps tile indices and addresses scalar.  So is the wavefront index.
    // The LDS buffers are taken from the LDS array here (a pointer argument
    // would lose its address space and every access would be a flat one).
    const double *A = uni_ptr(A_), *av = uni_ptr(av_);
    double *Prow = uni_ptr(Prow_);
    gdouble *Pcw = (gdouble *)uni_ptr(Pcw_);
    const int n = uni(n_), nt = uni(nt_), R0 = uni(R0_), rank = uni(rank_), T = uni(T_), jj = uni(jj_);
    const int w = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double *const Tw = lds + BLK_LDS_TW + w * BLK_TW;
    double *const xs = lds + BLK_LDS_XS + w * 64;
    int f = 0;
    double *pend_p = nullptr, pend_v = 0.0;
    for (int C = R0; C < nt; ++C) {
        if (blk_owner(C, T) != rank) continue;
        const int cnt = C - R0 + 1;
        double col = 0.0;
        for (int i = ((w - f) % TRI_NW + TRI_NW) % TRI_NW; i < cnt; i += TRI_NW)
            col += blk_tile(A, n, nt, R0 + i, C, jj, v0, av, Prow, Tw, xs, lane, pend_p, pend_v);
        Pcw[((size_t)C * TRI_NW + w) * 64 + lane] = col;
        f += cnt;
    }
    if (pend_p) st_dev(pend_p + lane, pend_v);
}

// SYR2K of the finished panel on tile (R, C): A[r][c] -= sum_k V[r][k] W[c][k]
// + W[r][k] V[c][k] for r, c >= j1 (both halves of a diagonal tile).  The
// whole workgroup: the operands go to LDS (Ls = -[V W] at rows R, k-major;
// Rs = [W V] at columns C, staged when C changes), then wavefront w computes
// the 16 x 16 block (w / 4, w % 4) with v_mfma_f64_16x16x4f64 (A[i][k] in
// lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16 (i % 4),
// element i / 4: measured on gfx950), the old tile as the accumulator.
__device__ __forceinline__ void blk_syr2k_tile(double *__restrict__ A, int n, int npad, int R, int C, int j1,
                                               const double *Vb, const double *Wb, bool loadR) {
    extern __shared__ double lds[];
    double *const Ls = lds + BLK_LDS_TW, *const Rs = Ls + 2 * BLK_NB * 64;
    const double *const Kc = lds + BLK_LDS_CF + 64 * CF_K;
    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int bi = w >> 2, bj = w & 3;
    const int col = 64 * C + 16 * bj + (lane & 15);
    const int row0 = 64 * R + 16 * bi + (lane >> 4);  // element v: row row0 + 4 v
    // the old tile first: its loads are in flight while the operands are staged
    blk_d4 acc;
    gdouble *Ag = (gdouble *)A;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double x = Ag[(size_t)min(row0 + 4 * v, n - 1) * n + min(col, n - 1)];
        acc[v] = (row0 + 4 * v < n && col < n) ? x : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 2 * BLK_NB * 64 / TRI_NT; ++q) {
        const int e = tid + q * TRI_NT, k = e >> 6, i = e & 63;
        const int kk = k < BLK_NB ? k : k - BLK_NB;
        const int r = 64 * R + i, cidx = 64 * C + i;
        const double vr = ld_dev(Vb + (size_t)kk * npad + r), pr = ld_dev(Wb + (size_t)kk * npad + r);
        double vc = 0.0, pc = 0.0;
        if (loadR) {
            vc = ld_dev(Vb + (size_t)kk * npad + cidx);
            pc = ld_dev(Wb + (size_t)kk * npad + cidx);
        }
        Ls[e] = -(k < BLK_NB ? vr : pr - Kc[kk] * vr);
        if (loadR) Rs[e] = k < BLK_NB ? pc - Kc[kk] * vc : vc;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2 * BLK_NB / 4; ++s) {
        const int k = 4 * s + (lane >> 4);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[k * 64 + 16 * bi + (lane & 15)], Rs[k * 64 + 16 * bj + (lane & 15)],
                                                   acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int r = row0 + 4 * v;
        if (r >= j1 && r < n && col >= j1 && col < n) Ag[(size_t)r * n + col] = acc[v];
    }
    __syncthreads();
}

#ifdef NRG_BLK_PROF
#define BLK_PROF(i) do { __syncthreads(); if (tid == 0) { const long long now_ = wall_clock64(); pacc[i] += now_ - plast; plast = now_; } } while (0)
#else
#define BLK_PROF(i) do {} while (0)
#endif
// a - sum_{k < kc} (V[c][k] gam_k + p[c][k] bet_k) (coefficient k in lane k),
// in k order; the loads
// of 16 panel columns are issued together (clamped indices, products of
// columns beyond kc discarded by a select, so stale values never enter)
__device__ __forceinline__ double blk_corr(double a, const gdouble *Vb, const gdouble *Wb, int npad, int c, int kc,
                                           double gam_l, double bet_l) {
    for (int k0 = 0; k0 < kc; k0 += 16) {
        double v[16], p[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = min(k0 + u, BLK_NB - 1);
            v[u] = Vb[(size_t)k * npad + c];
            p[u] = Wb[(size_t)k * npad + c];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = min(k0 + u, BLK_NB - 1);
            const double t = v[u] * readlane_d(gam_l, k) + p[u] * readlane_d(bet_l, k);
            a -= k0 + u < kc ? t : 0.0;
        }
    }
    return a;
}

__device__ __noinline__ void tri_blk(const EigMat &mt, int rank, int T, int *bar, int *status, int *s_flag) {
    extern __shared__ double lds[];
    const int n = mt.n, npad = (n + 63) & ~63, nt = npad >> 6;
    double *__restrict__ A = mt.A;
    double *ws = mt.ws;
    double *const base = (double *)(((uintptr_t)(ws + (size_t)n * n) + 511) & ~(uintptr_t)511);
    double *const Vb = base;
    double *const Wb = Vb + (size_t)BLK_NB * npad;
    double *const av = Wb + (size_t)BLK_NB * npad;
    double *const part1 = av + npad;
    double *const part2 = part1 + (size_t)nt * BLK_P1W;
    double *const Prow = part2 + nt;
    double *const Pcw = Prow + (size_t)nt * nt * 64;
    double *const tau_g = vslot(ws, n, V_TAU), *const gdg = vslot(ws, n, V_DG), *const god = vslot(ws, n, V_OD);
    // plain accesses to global memory through global-address-space pointers
    // (global_* instructions rather than flat_*)
    gdouble *const gA = (gdouble *)A;
    const gdouble *const gV = (const gdouble *)Vb, *const gW = (const gdouble *)Wb, *const gav = (const gdouble *)av;
    const gdouble *const gPcw = (const gdouble *)Pcw;
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    double *const cf = lds + BLK_LDS_CF;
    double *const Kc = cf + 64 * CF_K;
    double *const uW = cf + 64 * CF_UW, *const uV = cf + 64 * CF_UV;
    double *const vj1 = cf + 64 * CF_VJ1, *const wj1 = cf + 64 * CF_WJ1, *const sc = cf + 64 * CF_SC;
    int phase = 0, j0 = 0;
    double tau_prev = 0.0;
#ifdef NRG_BLK_PROF
    long long pacc[12] = {0}, plast = wall_clock64();
#endif
    for (int j = 0; j < n; ++j) {
        int kc = j - j0;
        const int R0 = (j + 1) >> 6;  // first block of the trailing range [j+1, n)
        // K of column j-1 = tau (p . v) / 2, p.v summed over its blocks in
        // order: every wavefront computes it (the same bits), wavefront 0
        // keeps it in LDS for the later phases
        // wavefront 0 reads the partial sums and the panel's pivot-row values
        // for the workgroup (one reader per workgroup: every rank reads the
        // same few lines) and leaves them in LDS
        if (w == 0) {
            double Knew = 0.0;
            if (kc > 0) {
                double sp = 0.0;
                for (int B = (j >> 6) + lane; B < nt; B += 64) sp += ld_dev(part2 + B);
                Knew = 0.5 * tau_prev * wave_sum(sp);
                if (lane == 0) Kc[kc - 1] = Knew;
            }
            const int kn = kc == BLK_NB && j < n - 1 ? 0 : kc;  // (a panel ends here: none)
            if (lane < kn) {
                const int k = lane;
                const double b = ld_dev(Vb + (size_t)k * npad + j), wp = ld_dev(Wb + (size_t)k * npad + j);
                double v1 = 0.0, p1 = 0.0;
                if (j + 1 < n) {
                    v1 = ld_dev(Vb + (size_t)k * npad + j + 1);
                    p1 = ld_dev(Wb + (size_t)k * npad + j + 1);
                }
                const double Kk = k == kc - 1 ? Knew : Kc[k];
                cf[64 * CF_BET + k] = b;
                cf[64 * CF_GAM + k] = wp - 2.0 * Kk * b;
                cf[64 * CF_WJ + k] = wp - Kk * b;
                vj1[k] = v1;
                wj1[k] = p1 - Kk * v1;
            }
        }
        __syncthreads();
        BLK_PROF(0);
        // panel end: rank-2 BLK_NB update of the trailing matrix [j, n)
        if (kc == BLK_NB && j < n - 1) {
            const int R0s = j >> 6;
            for (int C = R0s; C < nt; ++C) {
                if (blk_owner(C, T) != rank) continue;
                for (int R = R0s; R <= C; ++R) blk_syr2k_tile(A, n, npad, R, C, j, Vb, Wb, R == R0s);
            }
            j0 = j;
            kc = 0;
        }
        BLK_PROF(1);
        // ---- P1: the panel's earlier columns at rows j (column correction,
        // diagonal) and j+1 (v0's share of W^T v and V^T v): lane k of every
        // wavefront holds column k's coefficients
        const double gam_l = lane < kc ? cf[64 * CF_GAM + lane] : 0.0, bet_l = lane < kc ? cf[64 * CF_BET + lane] : 0.0,
                     wj_l = lane < kc ? cf[64 * CF_WJ + lane] : 0.0;
        if (w == 0 && blk_owner(j >> 6, T) == rank) {
            const double d = gA[(size_t)j * n + j] - 2.0 * wave_sum(bet_l * wj_l);
            if (lane == 0) gdg[j] = d;
        }
        if (j == n - 1) {
            if (rank == 0 && tid == 0) god[j] = 0.0;
#ifdef NRG_BLK_PROF
            if (rank == 0 && tid == 0)
                printf("blkprof n=%d T=%d K %lld syr2k %lld p1coef %lld p1 %lld bar1 %lld p2sum %lld symv %lld bar2 %lld p3 %lld bar3 %lld\n",
                       n, T, pacc[0], pacc[1], pacc[2], pacc[3], pacc[4], pacc[5], pacc[6], pacc[7], pacc[8], pacc[9]);
#endif
            break;
        }
        BLK_PROF(2);
        {
            // groups of up to 16 owned blocks: the column by one wavefront per
            // block, then the block's partial sums (|a|^2, W^T a, V^T a) as
            // (block, quantity) items over all wavefronts, loads issued in
            // batches (no branch between them)
            double *abuf = lds + BLK_LDS_XS;  // [block of the group][64]: the column's tail
            __shared__ int gblk[16];  // the group's blocks
            const int nq = 1 + 2 * kc;
            for (int B0 = R0; B0 < nt;) {
                int nb = 0, B1 = B0;
                for (; B1 < nt && nb < 16; ++B1)
                    if (blk_owner(B1, T) == rank) {
                        if (tid == 0) gblk[nb] = B1;
                        ++nb;
                    }
                if (nb == 0) break;
                __syncthreads();
                if (w < nb) {
                    const int c = 64 * gblk[w] + lane;
                    const bool act = c > j && c < n, tail = c > j + 1 && c < n;
                    double a = act ? gA[(size_t)j * n + c] : 0.0;
                    a = blk_corr(a, gV, gW, npad, c, kc, gam_l, bet_l);
                    if (act) st_dev(av + c, a);
                    if (tail) gA[(size_t)j * n + c] = a;
                    abuf[w * 64 + lane] = tail ? a : 0.0;
                }
                __syncthreads();
                const int ni = nb * nq;
                for (int p0 = w; p0 < ni; p0 += 16 * 8) {
                    double val[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int pi = min(p0 + 16 * u, ni - 1), b = pi / nq, q = pi - b * nq;
                        const int k = max(q <= kc ? q - 1 : q - 1 - kc, 0);
                        const int c = 64 * gblk[b] + lane;
                        const double at = abuf[b * 64 + lane];
                        const double v = gV[(size_t)k * npad + c], wp = gW[(size_t)k * npad + c];
                        val[u] = q == 0 ? at * at : q <= kc ? (wp - Kc[k] * v) * at : v * at;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int pi = p0 + 16 * u;
                        if (pi >= ni) break;
                        const int b = pi / nq, q = pi - b * nq;
                        const double sm = wave_sum(val[u]);
                        const int slot = q == 0 ? 0 : q <= kc ? q : 1 + BLK_NB + (q - 1 - kc);
                        if (lane == 0) st_dev(part1 + (size_t)gblk[b] * BLK_P1W + slot, sm);
                    }
                }
                __syncthreads();
                B0 = B1;
            }
        }
        BLK_PROF(3);
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return;
        BLK_PROF(4);
        // ---- P2: block sums in block order (every workgroup the same bits)
        {
            // quantity t (0: |a|^2, then W^T a, V^T a) by wavefront t % 16,
            // lanes over the blocks; a wavefront's loads are issued together
            const int nq = 1 + 2 * kc;
            constexpr int QW = (1 + 2 * BLK_NB + TRI_NW - 1) / TRI_NW;
            double acc[QW];
#pragma unroll
            for (int u = 0; u < QW; ++u) acc[u] = 0.0;
            if (tid == 0) sc[1] = ld_dev(av + j + 1);
            for (int B = R0 + lane; B < nt; B += 64) {
#pragma unroll
                for (int u = 0; u < QW; ++u) {
                    const int t = min(w + TRI_NW * u, nq - 1);
                    const int slot = t == 0 ? 0 : t <= kc ? t : 1 + BLK_NB + (t - 1 - kc);
                    acc[u] += ld_dev(part1 + (size_t)B * BLK_P1W + slot);
                }
            }
#pragma unroll
            for (int u = 0; u < QW; ++u) {
                const int t = w + TRI_NW * u;
                if (t >= nq) break;
                const double sm = wave_sum(acc[u]);
                if (lane == 0) {
                    if (t == 0) sc[0] = sm;
                    else if (t <= kc) uW[t - 1] = sm;
                    else uV[t - 1 - kc] = sm;
                }
            }
        }
        __syncthreads();
        const double sigma = sc[0], x1 = sc[1];
        double tau = 0.0, v0 = 0.0, alpha = x1;
        if (j <= n - 3 && sigma != 0.0) {
            alpha = -copysign(sqrt(x1 * x1 + sigma), x1);
            v0 = x1 - alpha;
            tau = 2.0 / (v0 * v0 + sigma);
        }
        if (rank == 0 && tid == 0) {
            god[j] = alpha;
            if (j <= n - 3) tau_g[j] = tau;
        }
        if (tid == 0 && blk_owner((j + 1) >> 6, T) == rank) gA[(size_t)j * n + j + 1] = tau != 0.0 ? v0 : 0.0;
        if (tid < kc) {   // W^T v and V^T v include v0 at row j+1
            uW[tid] += wj1[tid] * v0;
            uV[tid] += vj1[tid] * v0;
        }
        {
            int idx = 0;  // column kc of the panel's V
            for (int B = R0; B < nt; ++B) {
                if (blk_owner(B, T) != rank) continue;
                if ((idx++ & (TRI_NW - 1)) != w) continue;
                const int c = 64 * B + lane;
                double v = 0.0;
                if (tau != 0.0 && c > j && c < n) v = c == j + 1 ? v0 : gav[c];
                st_dev(Vb + (size_t)kc * npad + c, v);
            }
        }
        BLK_PROF(5);
        if (tau != 0.0) blk_symv(A, n, nt, R0, rank, T, j + 1, v0, av, Prow, Pcw);
        BLK_PROF(6);
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return;
        BLK_PROF(7);
        // ---- P3: p on the owned blocks.  Row part i of block B (tile
        // (B, B+i)) is summed by wavefront i % 16, which then adds the column
        // parts its namesake summed in the product; the 16 partial sums are
        // added in wavefront order.
        {
            double *red = lds + BLK_LDS_TW;  // [block of the group][wavefront][64]
            for (int B0 = R0; B0 < nt;) {
                // a group of up to 16 owned blocks [B0, B1)
                int nb = 0, B1 = B0, mine = -1;
                for (; B1 < nt && nb < 16; ++B1)
                    if (blk_owner(B1, T) == rank) {
                        if (nb == w) mine = B1;
                        ++nb;
                    }
                if (nb == 0) break;
                if (tau != 0.0) {
                    int b = 0;
                    for (int Bb = B0; Bb < B1; ++Bb) {
                        if (blk_owner(Bb, T) != rank) continue;
                        const int nrow = nt - Bb;
                        double s = gPcw[((size_t)Bb * TRI_NW + w) * 64 + lane];
                        for (int i0 = w; i0 < nrow; i0 += 16 * 16) {
                            double t[16];
#pragma unroll
                            for (int u = 0; u < 16; ++u) {
                                // clamped addresses, then a select (no branch, so the
                                // 16 items' loads are in flight together)
                                const int i = i0 + 16 * u;
                                t[u] = ld_dev(Prow + ((size_t)Bb * nt + Bb + min(i, nrow - 1)) * 64 + lane);
                            }
#pragma unroll
                            for (int u = 0; u < 16; ++u) s += i0 + 16 * u < nrow ? t[u] : 0.0;
                        }
                        red[(b * 16 + w) * 64 + lane] = s;
                        ++b;
                    }
                }
                B0 = B1;
                __syncthreads();
                if (mine >= 0) {
                    const int Bb = mine;
                    const int c = 64 * Bb + lane;
                    const bool act = c > j && c < n;
                    double p = 0.0, v = 0.0;
                    if (tau != 0.0) {
                        double y = 0.0;
                        for (int u = 0; u < 16; ++u) y += red[(w * 16 + u) * 64 + lane];
                        for (int k0 = 0; k0 < kc; k0 += 16) {
                            double vv[16], pp[16];
#pragma unroll
                            for (int u = 0; u < 16; ++u) {
                                const int k = min(k0 + u, BLK_NB - 1);
                                vv[u] = gV[(size_t)k * npad + c];
                                pp[u] = gW[(size_t)k * npad + c];
                            }
#pragma unroll
                            for (int u = 0; u < 16; ++u) {
                                const int k = min(k0 + u, BLK_NB - 1);
                                const double t = vv[u] * uW[k] + (pp[u] - Kc[k] * vv[u]) * uV[k];
                                y -= k0 + u < kc ? t : 0.0;
                            }
                        }
                        if (act) {
                            p = tau * y;
                            v = c == j + 1 ? v0 : gav[c];
                        }
                    }
                    st_dev(Wb + (size_t)kc * npad + c, p);
                    const double pv = wave_sum(p * v);
                    if (lane == 0) st_dev(part2 + Bb, pv);
                }
                __syncthreads();
            }
        }
        BLK_PROF(8);
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return;
        BLK_PROF(9);
        tau_prev = tau;
    }
}

// ---------------------------------------------------------------------------
// 1e. two-stage reduction, stage 1 (sbr_blk): dense -> symmetric band of
//     half-width SBR_B (successive band reduction; the band then goes to
//     tridiagonal form by bulge chasing, section 1f).  The lower triangle is
//     kept up to date (the upper one is mirrored into it at the start), in
//     64 x 64 tiles with the diagonal tiles whole; team rank blk_owner(B)
//     alone writes the rows of block B, and tiles above the diagonal are read
//     as the transposes of their mirror tiles.  Panel p (columns
//     j0 = SBR_B p .. j0 + SBR_B - 1, trailing range [c0 = j0 + SBR_B, n)):
//       QR of the panel's lower copy A[c0:n, j0:j0+SBR_B] with its rows
//         spread over their owners (registers, lane = row): per reflector one
//         barrier for the partial sums |x|^2 and x . a_i of every block,
//         summed in block order by every rank (the same bits everywhere);
//       Y = A22 V on the owned tile rows (FP64 MFMA), per-block Gram
//         partials V^T V and V^T Y -> barrier -> every rank forms T (dlarft)
//         and S = T^T (V^T Y) T and writes W = Y T - V S / 2 for its rows
//         -> barrier;
//       A22 -= V W^T + W V^T on the owned tile rows (lower triangle, MFMA).
//     Output: row j0 + k of S2 holds R at columns c0 .. c0 + k (beta at
//     c0 + k) and reflector k at columns c0 + k + 1 .. n - 1 (implicit 1 at
//     c0 + k); tau in V_TAU; A's upper triangle the rest of the band
//     (columns j .. c0 - 1 of a panel's row j, all of the rows below the
//     last panel).
// ---------------------------------------------------------------------------
constexpr int SBR_B = 32;
constexpr int SBR_MAXQ = 8;  // owned blocks of a trailing range (wavefronts holding panel rows)
// S1 layout (doubles from a 512-byte aligned start): V, W (SBR_B x npad,
// reflector-major), Y (npad x SBR_B), QR partial sums (2 x nt x SBR_B),
// pivot rows (2 x SBR_B), Gram partials (nt x 2 x SBR_B^2) and their totals
__host__ __device__ inline size_t sbr_ws_doubles(int n) {
    const size_t npad = ((size_t)n + 63) & ~(size_t)63, nt = npad / 64;
    return 64 + 3 * (size_t)SBR_B * npad + 2 * nt * SBR_B + 2 * SBR_B + (nt + 1) * 2 * SBR_B * SBR_B;
}
// reflectors of stage 1: SBR_B per panel, panels while the trailing range
// [SBR_B (p + 1), n) has at least two rows
__host__ __device__ inline int sbr_nrefl(int n) { return n >= SBR_B + 2 ? SBR_B * ((n - SBR_B - 2) / SBR_B + 1) : 0; }
// ranks of a stage-1 team: at most SBR_MAXQ owned blocks per rank
__host__ __device__ inline int sbr_owned_max(int nt, int T) {
    int mx = 0;
    for (int r = 0; r < T; ++r) {
        int c = 0;
        for (int B = 0; B < nt; ++B) c += blk_owner(B, T) == r;
        mx = c > mx ? c : mx;
    }
    return mx;
}
// LDS map (doubles)
constexpr int SB_LA = 0;                      // product tiles (2 x 64 x 65)
constexpr int SB_LV = SB_LA + 2 * 64 * 65;    // V of the tiles' columns (2 x 32 x 65)
constexpr int SB_RED = SB_LV + 2 * 32 * 65;   // halves of the product (8 x 256)
constexpr int SB_T = SB_RED + 8 * 256;        // T, S, G, M (32 x 33 each)
constexpr int SB_S = SB_T + 32 * 33;
constexpr int SB_G = SB_S + 32 * 33;
constexpr int SB_M = SB_G + 32 * 33;
constexpr int SB_CO = SB_M + 32 * 33;         // per-reflector coefficients, taus
constexpr int SB_PS = SB_CO + 128;            // partial-sum chunks (32 x 33)
constexpr int SB_TOT = SB_PS + 32 * 33;          // their totals
constexpr int SB_END = SB_TOT + 64;  // (totals, then the pivot row)
constexpr int SB_LS = 0, SB_RS = 64 * 64;     // rank-2k operands: rows (64 x 64), columns (2 x 64 x 64), k-major
constexpr int SB_P = 0;  // panel QR: the rows of its wavefronts, [wavefront][column][lane]
static_assert(SB_P + SBR_MAXQ * SBR_B * 64 <= SB_CO, "the QR's panel rows leave its coefficients and sums alone");
static_assert(SB_RS + 2 * 64 * 64 <= SB_T, "rank-2k operands overlap only the product's buffers");
static_assert(sizeof(double) * SB_END <= TRI_COOP_LDS, "stage 1 fits the team launch's LDS");

// The tile loops of the product and the rank-2k update walk the trailing
// tile row of an owned block C: tile R's operands are loaded three tiles
// ahead into a four-deep register ring, staged in one of two LDS buffers, and
// one barrier per tile separates the staging from the MFMAs; so about three
// tiles of loads are in flight while a tile is computed.  The loads and the
// staging of every step are unconditional (a load past the row reads one
// fixed address, a staging past the row fills the idle buffer), and the
// steps run in groups of four without guards, the last group with them: a
// skipped load or staging in the loop body would make the compiler assume
// the ring's loads may still be pending at every later step and wait for
// them all before each new load.
#define SB_RING4(STEP)                                                          \
    {                                                                           \
        int i0_ = 0;                                                            \
        for (; i0_ + 4 <= cnt; i0_ += 4) {                                      \
            STEP(0, true) STEP(1, true) STEP(2, true) STEP(3, true)             \
        }                                                                       \
        if (i0_ < cnt) {                                                        \
            STEP(0, true) STEP(1, i0_ + 1 < cnt) STEP(2, i0_ + 2 < cnt)         \
        }                                                                       \
    }

// Y[C] = sum_R A[C rows][R columns] V[R] over R in [Bt, nt) (64 x 32): wavefront
// w computes the 16 x 16 block (w & 3, (w >> 2) & 1) over half w >> 3 of each
// tile's 64 columns (v_mfma_f64_16x16x4f64: A[i][k] in lane i + 16 k, B[k][j]
// in lane j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4); the
// halves are added in order.  Only the lower triangle (with whole diagonal
// tiles) is kept up to date: tiles R <= C are the rank's own rows, tiles
// R > C are read as the transpose of tile (R, C), another rank's rows (all
// through agent-scope loads, which see the other ranks' write-through
// stores; no tile is read twice by one rank in a panel).
__device__ __noinline__ void sbr_symm_row(const double *A_, int n_, int npad_, int nt_, int C_, int Bt_,
                                          const double *Vb_, double *Yb_) {
    extern __shared__ double lds[];
    const gdouble *const gA = (const gdouble *)uni_ptr(A_);
    const double *const Vb = uni_ptr(Vb_);
    double *const Yb = uni_ptr(Yb_);
    const int n = uni(n_), npad = uni(npad_), nt = uni(nt_), C = uni(C_), Bt = uni(Bt_);
    const int tid = threadIdx.x, w = uni(tid >> 6), lane = tid & 63;
    const int bi = w & 3, bj = (w >> 2) & 1, h = w >> 3;
    const int cnt = nt - Bt;
    double ra[4][4], rv[4][2];
    blk_d4 acc = {0.0, 0.0, 0.0, 0.0};
    // Loads outside the matrix read a clamped address; the staging multiplies
    // them by zero.  The ring registers hold the raw loaded values, so the
    // loads stay in flight until their tile is staged (a factor applied at the
    // load, or a guarded load, makes the compiler wait for each load at once).
    // Tiles R > C are the transposes of tile (R, C): the same agent-scope load
    // from the mirrored address, transposed when staged.
#define SY_LOAD(S, R)                                                                           \
    {                                                                                           \
        const bool ok_ = (R) < nt;                                                              \
        const int rb_ = (R) <= C ? C : (R), cb_ = (R) <= C ? (R) : C;                           \
        _Pragma("unroll") for (int u = 0; u < 4; ++u) {                                         \
            const int e = tid + TRI_NT * u, rw = 64 * rb_ + (e >> 6), cl = 64 * cb_ + (e & 63); \
            const size_t o_ = ok_ ? (size_t)min(rw, n - 1) * n + min(cl, n - 1) : 0;            \
            ra[S][u] = ld_dev((const double *)(gA + o_));                                       \
        }                                                                                       \
        _Pragma("unroll") for (int u = 0; u < 2; ++u) {                                         \
            const int e = tid + TRI_NT * u;                                                     \
            rv[S][u] = ld_dev(Vb + (ok_ ? (size_t)(e >> 6) * npad + 64 * (R) + (e & 63) : 0));   \
        }                                                                                       \
    }
#define SY_PUT(S, BUF, R)                                                                       \
    {                                                                                           \
        const bool tr_ = (R) > C;                                                               \
        const int rb_ = tr_ ? (R) : C, cb_ = tr_ ? C : (R);                                     \
        _Pragma("unroll") for (int u = 0; u < 4; ++u) {                                         \
            const int e = tid + TRI_NT * u, rw = 64 * rb_ + (e >> 6), cl = 64 * cb_ + (e & 63); \
            const double f = (rw < n && cl < n) ? 1.0 : 0.0;                                    \
            lds[SB_LA + (BUF) * 64 * 65 + (tr_ ? (e & 63) * 65 + (e >> 6) : (e >> 6) * 65 + (e & 63))] = f * ra[S][u]; \
        }                                                                                       \
        _Pragma("unroll") for (int u = 0; u < 2; ++u) {                                         \
            const int e = tid + TRI_NT * u;                                                     \
            lds[SB_LV + (BUF) * 32 * 65 + (e >> 6) * 65 + (e & 63)] = rv[S][u];                 \
        }                                                                                       \
    }
#define SY_STEP(J, G)                                                                           \
    if (G) {                                                                                    \
        const int R = Bt + i0_ + (J), buf = (i0_ + (J)) & 1;                                    \
        SY_PUT(((J) + 1) & 3, buf ^ 1, R + 1)                                                   \
        SY_LOAD(((J) + 3) & 3, R + 3)                                                           \
        _Pragma("unroll") for (int s = 8 * h; s < 8 * h + 8; ++s) {                             \
            const double a = lds[SB_LA + buf * 64 * 65 + (16 * bi + (lane & 15)) * 65 + 4 * s + (lane >> 4)]; \
            const double b = lds[SB_LV + buf * 32 * 65 + (16 * bj + (lane & 15)) * 65 + 4 * s + (lane >> 4)]; \
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);                    \
        }                                                                                       \
        __syncthreads();                                                                        \
    }
    SY_LOAD(0, Bt)
    SY_LOAD(1, Bt + 1)
    SY_LOAD(2, Bt + 2)
    SY_PUT(0, 0, Bt)
    __syncthreads();
    SB_RING4(SY_STEP)
#undef SY_LOAD
#undef SY_PUT
#undef SY_STEP
    if (h == 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) lds[SB_RED + (w & 7) * 256 + v * 64 + lane] = acc[v];
    }
    __syncthreads();
    if (h == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double y = acc[v] + lds[SB_RED + (w & 7) * 256 + v * 64 + lane];
            const int rr = 64 * C + 16 * bi + 4 * v + (lane >> 4), kk = 16 * bj + (lane & 15);
            Yb[(size_t)rr * SBR_B + kk] = y;
        }
    }
    __syncthreads();
}

// A[r][c] -= sum_k V[r][k] W[c][k] + W[r][k] V[c][k] on the trailing tile
// row of block C, rows and columns >= c0, tiles R <= C (the lower triangle
// and the whole diagonal tile; stores write through, since other ranks read
// these tiles transposed).  The row
// operands -[V W] of block C are staged once (k-major); per tile R the
// column operands [W V] go through a double LDS buffer and the old tile is
// the MFMA accumulator of wavefront w's 16 x 16 block (w / 4, w % 4).  Old
// entries outside the matrix are loaded from a clamped address and their
// outputs are not stored; the ring holds the loaded values unchanged, so the
// loads stay in flight until the accumulator takes them.
__device__ __noinline__ void sbr_syr2k_row(double *A_, int n_, int npad_, int nt_, int C_, int Bt_, int c0_,
                                           const double *Vb_, const double *Wb_) {
    extern __shared__ double lds[];
    gdouble *const gA = (gdouble *)uni_ptr(A_);
    const double *const Vb = uni_ptr(Vb_), *const Wb = uni_ptr(Wb_);
    const int n = uni(n_), npad = uni(npad_), nt = uni(nt_), C = uni(C_), Bt = uni(Bt_), c0 = uni(c0_);
    const int tid = threadIdx.x, w = uni(tid >> 6), lane = tid & 63;
    const int bi = w >> 2, bj = w & 3;
    const int row0 = 64 * C + 16 * bi + (lane >> 4);  // element v: row row0 + 4 v
    const int cnt = C - Bt + 1;  // tiles R = Bt .. C (the lower triangle)
    double *const Ls = lds + SB_LS;
    __syncthreads();
    {  // all loads first, then the stores (one wait instead of one per load)
        double ls[2 * SBR_B * 64 / TRI_NT];
#pragma unroll
        for (int u = 0; u < 2 * SBR_B * 64 / TRI_NT; ++u) {
            const int e = tid + TRI_NT * u, k = e >> 6, i = e & 63;
            const int kk = k < SBR_B ? k : k - SBR_B;
            ls[u] = ld_dev((k < SBR_B ? Vb : Wb) + (size_t)kk * npad + 64 * C + i);
        }
#pragma unroll
        for (int u = 0; u < 2 * SBR_B * 64 / TRI_NT; ++u) Ls[tid + TRI_NT * u] = -ls[u];
    }
    double ro[4][4], rr[4][4];
#define SR_LOAD(S, R)                                                                           \
    {                                                                                           \
        const bool ok_ = (R) <= C;                                                              \
        const int col = 64 * (R) + 16 * bj + (lane & 15);                                       \
        _Pragma("unroll") for (int v = 0; v < 4; ++v)                                          \
            ro[S][v] = gA[ok_ ? (size_t)min(row0 + 4 * v, n - 1) * n + min(col, n - 1) : 0];    \
        _Pragma("unroll") for (int u = 0; u < 4; ++u) {                                         \
            const int e = tid + TRI_NT * u, k = e >> 6, i = e & 63;                             \
            const int kk = k < SBR_B ? k : k - SBR_B;                                           \
            const int c = 64 * (R) + i;                                                         \
            rr[S][u] = ld_dev((k < SBR_B ? Wb : Vb) + (ok_ ? (size_t)kk * npad + c : 0));        \
        }                                                                                       \
    }
#define SR_PUT(S, BUF)                                                                          \
    {                                                                                           \
        _Pragma("unroll") for (int u = 0; u < 4; ++u) lds[SB_RS + (BUF) * 64 * 64 + tid + TRI_NT * u] = rr[S][u]; \
    }
#define SR_STEP(J, G)                                                                           \
    if (G) {                                                                                    \
        const int R = Bt + i0_ + (J), buf = (i0_ + (J)) & 1;                                    \
        SR_PUT(((J) + 1) & 3, buf ^ 1)                                                          \
        SR_LOAD(((J) + 3) & 3, R + 3)                                                           \
        blk_d4 acc;                                                                             \
        _Pragma("unroll") for (int v = 0; v < 4; ++v) acc[v] = ro[(J) & 3][v];                  \
        const double *Rs = lds + SB_RS + buf * 64 * 64;                                         \
        _Pragma("unroll") for (int s = 0; s < 2 * SBR_B / 4; ++s) {                             \
            const int k = 4 * s + (lane >> 4);                                                  \
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ls[k * 64 + 16 * bi + (lane & 15)],     \
                                                       Rs[k * 64 + 16 * bj + (lane & 15)], acc, 0, 0, 0); \
        }                                                                                       \
        const int col = 64 * R + 16 * bj + (lane & 15);                                         \
        _Pragma("unroll") for (int v = 0; v < 4; ++v) {                                         \
            const int r = row0 + 4 * v;                                                         \
            if (r >= c0 && r < n && col >= c0 && col < n) st_dev((double *)(gA + (size_t)r * n + col), acc[v]); \
        }                                                                                       \
        __syncthreads();                                                                        \
    }
    SR_LOAD(0, Bt)
    SR_LOAD(1, Bt + 1)
    SR_LOAD(2, Bt + 2)
    SR_PUT(0, 0)
    __syncthreads();
    SB_RING4(SR_STEP)
#undef SR_LOAD
#undef SR_PUT
#undef SR_STEP
}
#undef SB_RING4

// the number of owned blocks in [B0, nt), and the q-th of them (-1: none)
__device__ __forceinline__ int sbr_nowned(int B0, int nt, int rank, int T) {
    int nq = 0;
    for (int B = B0; B < nt; ++B) nq += blk_owner(B, T) == rank;
    return nq;
}
__device__ __forceinline__ int sbr_qblock(int B0, int nt, int rank, int T, int q) {
    for (int B = B0; B < nt; ++B)
        if (blk_owner(B, T) == rank && q-- == 0) return B;
    return -1;
}

// the sum over blocks B = Bs + g + 32 u < nt of p[B * SBR_B] (NU loads issued together)
template <int NU>
__device__ __forceinline__ double sbr_bsum(const double *p, int Bs, int g, int nt) {
    double t[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) t[u] = ld_dev(p + (size_t)min(Bs + g + 32 * u, nt - 1) * SBR_B);
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < NU; ++u) s += Bs + g + 32 * u < nt ? t[u] : 0.0;
    return s;
}

// Per-job constants of a stage-1 team member, kept in LDS (the phases are
// separate functions, each with registers of its own: as one function the
// phases spilled heavily).
struct SbrCtx {
    double *A, *ws, *Vb, *Wb, *Yb, *part, *piv, *gram, *gtot, *tau_g;
    int *bar, *status, *s_flag;
    int n, npad, nt, rank, T;
#ifdef NRG_SBR_PROF
    long long pacc[12], plast;
#endif
};
#ifdef NRG_SBR_PROF
#define SB_PROF(i) do { __syncthreads(); if (tid == 0) { const long long now_ = wall_clock64(); sc->pacc[i] += now_ - sc->plast; sc->plast = now_; } } while (0)
#else
#define SB_PROF(i) do {} while (0)
#endif

__device__ __noinline__ void sbr_mirror(const SbrCtx *sc) {
    extern __shared__ double lds[];
    const int n = sc->n, npad = sc->npad, nt = sc->nt, rank = sc->rank, T = sc->T;
    double *__restrict__ A = sc->A;
    double *ws = sc->ws;
    double *const Vb = sc->Vb, *const Wb = sc->Wb, *const Yb = sc->Yb, *const part = sc->part, *const piv = sc->piv;
    double *const gram = sc->gram, *const gtot = sc->gtot, *const tau_g = sc->tau_g;
    int *const bar = sc->bar, *const status = sc->status, *const s_flag = sc->s_flag;
    gdouble *const gA = (gdouble *)A;
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    double *const co = lds + SB_CO;  // [0..31] s_i, [32] tau, [33] beta, [34] scale, [64..95] taus
    (void)ws; (void)Vb; (void)Wb; (void)Yb; (void)part; (void)piv; (void)gram; (void)gtot; (void)tau_g;
    (void)bar; (void)status; (void)s_flag; (void)gA; (void)w; (void)lane; (void)co; (void)npad; (void)rank;
    // 0. mirror the upper triangle into the lower one on the owned rows:
    //    tile (C', B) (another rank's rows, written before this launch) is
    //    read whole and written transposed into tile (B, C')
    for (int B = 0; B < nt; ++B) {
        if (blk_owner(B, T) != rank) continue;
        for (int Cp = 0; Cp <= B; ++Cp) {
            double *const La = lds + SB_LA;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = tid + TRI_NT * u, rr = e >> 6, cc = e & 63;  // row 64 Cp + rr, column 64 B + cc
                const int r = 64 * Cp + rr, c = 64 * B + cc;
                La[rr * 65 + cc] = (r < n && c < n) ? gA[(size_t)r * n + c] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = tid + TRI_NT * u, ii = e >> 6, cc = e & 63;  // A[64 B + ii][64 Cp + cc]
                const int r = 64 * B + ii, c = 64 * Cp + cc;
                if (r < n && c < n && c < r) st_dev((double *)(gA + (size_t)r * n + c), La[cc * 65 + ii]);
            }
            __syncthreads();
        }
    }

}

// panel j0: QR of its lower copy (one barrier per reflector), the write-back
// of R, the reflectors and V; returns the team phase, -1 if a barrier failed
__device__ __noinline__ int sbr_qr(SbrCtx *sc, int j0, int phase) {
    extern __shared__ double lds[];
    const int n = sc->n, npad = sc->npad, nt = sc->nt, rank = sc->rank, T = sc->T;
    double *__restrict__ A = sc->A;
    double *ws = sc->ws;
    double *const Vb = sc->Vb, *const Wb = sc->Wb, *const Yb = sc->Yb, *const part = sc->part, *const piv = sc->piv;
    double *const gram = sc->gram, *const gtot = sc->gtot, *const tau_g = sc->tau_g;
    int *const bar = sc->bar, *const status = sc->status, *const s_flag = sc->s_flag;
    gdouble *const gA = (gdouble *)A;
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    double *const co = lds + SB_CO;  // [0..31] s_i, [32] tau, [33] beta, [34] scale, [64..95] taus
    (void)ws; (void)Vb; (void)Wb; (void)Yb; (void)part; (void)piv; (void)gram; (void)gtot; (void)tau_g;
    (void)bar; (void)status; (void)s_flag; (void)gA; (void)w; (void)lane; (void)co; (void)npad; (void)rank;
        const int c0 = j0 + SBR_B, m = n - c0;
        const int Bt = c0 >> 6;
        const int kmax = min(SBR_B, m - 1);
        const int nq = min(sbr_nowned(Bt, nt, rank, T), SBR_MAXQ);
        // ---- QR of the panel: wavefront q < nq holds rows of the q-th owned
        // block (lane = row), the panel's 32 columns in LDS (P[i * 64] =
        // column i of the lane's row; in registers, beside the block sums,
        // they spilled inside the reflector loop)
        const bool qw = w < nq;
        const int myB = qw ? sbr_qblock(Bt, nt, rank, T, w) : 0;
        const int r = 64 * myB + lane;
        const bool rok = qw && r >= c0 && r < n;
        double *const P = lds + SB_P + (size_t)(qw ? w : 0) * SBR_B * 64 + lane;
        if (qw) {
#pragma unroll
            for (int i0 = 0; i0 < SBR_B; i0 += 16) {
                double t[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) t[i] = gA[(size_t)min(r, n - 1) * n + j0 + i0 + i];
#pragma unroll
                for (int i = 0; i < 16; ++i) P[(i0 + i) * 64] = rok ? t[i] : 0.0;
            }
        }
        for (int k = 0; k < kmax; ++k) {
            const int par = k & 1, rp = c0 + k;
            const double xk = qw ? P[k * 64] : 0.0;
            if (qw) {
                const bool below = rok && r > rp;
                const double sg = wave_sum(below ? xk * xk : 0.0);
                if (lane == 0) st_dev(part + ((size_t)par * nt + myB) * SBR_B, sg);
#pragma unroll 2
                for (int i = k + 1; i < SBR_B; ++i) {
                    const double d = wave_sum(below ? xk * P[i * 64] : 0.0);
                    if (lane == 0) st_dev(part + ((size_t)par * nt + myB) * SBR_B + i, d);
                }
                if (r == rp)
                    for (int i = k; i < SBR_B; ++i) st_dev(piv + par * SBR_B + i, P[i * 64]);
            }
            SB_PROF(1);
            ++phase;
            if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return -1;
            SB_PROF(2);
            // block sums in block order: value i (0: |x|^2) by lanes i, blocks
            // B = Bs + g + 32 u by thread group g, the 32 group sums added in
            // order; the pivot row's values come in the same round trip
            {
                const int Bs = (rp + 1) >> 6, nbk = nt - Bs;
                const int i = tid & 31, g = tid >> 5;
                const double pv = g == 0 ? ld_dev(piv + par * SBR_B + i) : 0.0;
                double s = 0.0;
                if (i == 0 || i > k) {
                    const double *pp = part + (size_t)par * nt * SBR_B + i;
                    if (nbk <= 4 * 32) s = sbr_bsum<4>(pp, Bs, g, nt);
                    else if (nbk <= 8 * 32) s = sbr_bsum<8>(pp, Bs, g, nt);
                    else s = sbr_bsum<16>(pp, Bs, g, nt);
                }
                lds[SB_PS + g * 33 + i] = s;
                if (g == 0) lds[SB_TOT + 32 + i] = pv;
                __syncthreads();
            }
            SB_PROF(11);
            if (tid < 32) {
                double tot = 0.0;
                for (int gg = 0; gg < 32; ++gg) tot += lds[SB_PS + gg * 33 + tid];
                const double sigma = readlane_d(tot, 0);
                const double alpha = lds[SB_TOT + 32 + k];
                double tau = 0.0, beta = alpha, scale = 0.0;
                if (sigma != 0.0) {
                    beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
                    tau = (beta - alpha) / beta;
                    scale = 1.0 / (alpha - beta);
                }
                const int i = tid;
                double si = 0.0;
                if (i > k) si = tau * (lds[SB_TOT + 32 + i] + tot * scale);
                co[i] = si;
                if (i == 0) {
                    co[32] = tau;
                    co[33] = beta;
                    co[34] = scale;
                    co[64 + k] = tau;
                    if (rank == 0) tau_g[j0 + k] = tau;
                }
            }
            __syncthreads();
            if (qw) {
                const double scale = co[34], beta = co[33];
                const double v = r > rp ? xk * scale : 1.0;
                P[k * 64] = r > rp ? v : r == rp ? beta : xk;
                if (r >= rp && rok)
                    for (int i = k + 1; i < SBR_B; ++i) P[i * 64] -= co[i] * v;
            }
            // (co and the partial-sum table are next written after the next
            // team barrier, which every wavefront reaches after this update)
            SB_PROF(3);
        }
        for (int k = kmax + tid; k < SBR_B; k += TRI_NT) {
            co[64 + k] = 0.0;
            if (rank == 0) tau_g[j0 + k] = 0.0;
        }
        // write-back: the panel rows' R and reflectors into S2 (the rows of
        // A belong to another rank, whose XCD's L2 may hold dirty lines of
        // them; k_sbr_band copies them into A after this launch), and V
        if (qw) {
            double *const P2 = ws + 2 * (size_t)n * n;
#pragma unroll
            for (int i = 0; i < SBR_B; ++i) {
                const double pi = P[i * 64];
                if (rok) P2[(size_t)(j0 + i) * n + r] = pi;
                const double v = (i < kmax && rok) ? (r > c0 + i ? pi : r == c0 + i ? 1.0 : 0.0) : 0.0;
                st_dev(Vb + (size_t)i * npad + r, v);
            }
        }
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return -1;
        SB_PROF(4);
    return phase;
}

// panel j0: Y = A22 V, the Gram matrices, T, S, W, and the rank-64 update of
// the owned tile rows; returns the team phase, -1 if a barrier failed
__device__ __noinline__ int sbr_upd(SbrCtx *sc, int j0, int phase) {
    extern __shared__ double lds[];
    const int n = sc->n, npad = sc->npad, nt = sc->nt, rank = sc->rank, T = sc->T;
    double *__restrict__ A = sc->A;
    double *ws = sc->ws;
    double *const Vb = sc->Vb, *const Wb = sc->Wb, *const Yb = sc->Yb, *const part = sc->part, *const piv = sc->piv;
    double *const gram = sc->gram, *const gtot = sc->gtot, *const tau_g = sc->tau_g;
    int *const bar = sc->bar, *const status = sc->status, *const s_flag = sc->s_flag;
    gdouble *const gA = (gdouble *)A;
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    double *const co = lds + SB_CO;  // [0..31] s_i, [32] tau, [33] beta, [34] scale, [64..95] taus
    (void)ws; (void)Vb; (void)Wb; (void)Yb; (void)part; (void)piv; (void)gram; (void)gtot; (void)tau_g;
    (void)bar; (void)status; (void)s_flag; (void)gA; (void)w; (void)lane; (void)co; (void)npad; (void)rank;
        const int c0 = j0 + SBR_B;
        const int Bt = c0 >> 6;
        const int nq = min(sbr_nowned(Bt, nt, rank, T), SBR_MAXQ);
        // ---- Y = A22 V on the owned tile rows, then per-block Gram partials
        for (int q = 0; q < nq; ++q) {
            const int C = sbr_qblock(Bt, nt, rank, T, q);
            sbr_symm_row(A, n, npad, nt, C, Bt, Vb, Yb);
            // Gram partials of block C: G = V^T V, M = V^T Y over its rows
            {
                double *const Lv = lds + SB_LA;         // V[C] as [row][k] (64 x 33)
                double *const Ly = lds + SB_LA + 64 * 33;  // Y[C] (64 x 33)
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int e = tid + TRI_NT * u, kk = e >> 6, rr = e & 63;
                    Lv[rr * 33 + kk] = ld_dev(Vb + (size_t)kk * npad + 64 * C + rr);
                    const int rr2 = e >> 5, kk2 = e & 31;
                    Ly[rr2 * 33 + kk2] = Yb[(size_t)(64 * C + rr2) * SBR_B + kk2];
                }
                __syncthreads();
                const int i = tid >> 5, jj = tid & 31;
                double g = 0.0, mm = 0.0;
                for (int rr = 0; rr < 64; ++rr) {
                    const double vi = Lv[rr * 33 + i];
                    g = fma(vi, Lv[rr * 33 + jj], g);
                    mm = fma(vi, Ly[rr * 33 + jj], mm);
                }
                st_dev(gram + (size_t)C * 2 * SBR_B * SBR_B + tid, g);
                st_dev(gram + (size_t)C * 2 * SBR_B * SBR_B + SBR_B * SBR_B + tid, mm);
                __syncthreads();
            }
        }
        SB_PROF(5);
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return -1;
        SB_PROF(6);
        // ---- G, M summed over the blocks: element e of the 2 x 32 x 32 by
        // rank e % T (blocks B = Bt + g + 32 u by thread group g, the 32
        // group sums added in order, so the bits do not depend on the rank),
        // published -> barrier -> read by every rank
        {
            const int nme = (2 * SBR_B * SBR_B + T - 1 - rank) / T;  // elements of this rank
            const int g = tid & 31;
            for (int e0 = 0; e0 < nme; e0 += TRI_NT / 32) {
                const int le = e0 + (tid >> 5), e = rank + T * min(le, nme - 1);
                double t[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int B = min(Bt + g + 32 * u, nt - 1);
                    t[u] = ld_dev(gram + (size_t)B * 2 * SBR_B * SBR_B + e);
                }
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < 16; ++u) s += Bt + g + 32 * u < nt ? t[u] : 0.0;
                lds[SB_PS + (tid >> 5) * 33 + g] = s;
                __syncthreads();
                if (g == 0 && le < nme) {
                    double tt = 0.0;
                    for (int gg = 0; gg < 32; ++gg) tt += lds[SB_PS + (tid >> 5) * 33 + gg];
                    st_dev(gtot + e, tt);
                }
                __syncthreads();
            }
        }
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return -1;
        SB_PROF(7);
        // T (dlarft, forward columnwise); S = T^T M T; W = Y T - V S / 2
        {
            const int i = tid >> 5, jj = tid & 31;
            lds[SB_G + i * 33 + jj] = ld_dev(gtot + tid);
            lds[SB_M + i * 33 + jj] = ld_dev(gtot + SBR_B * SBR_B + tid);
        }
        __syncthreads();
        __syncthreads();
        // T in one wavefront: lane i < 32 builds row i in LDS, column by
        // column (T is upper triangular, T[i][k] = -tau_k sum_{l<k} T[i][l]
        // G[l][k]); a row held in registers through the unrolled triangle
        // spilled
        if (w == 0 && lane < SBR_B) {
            double *const Trow = lds + SB_T + lane * 33;
#pragma unroll 1
            for (int k = 0; k < SBR_B; ++k) {
                const double tk = co[64 + k];
                double t = 0.0;
#pragma unroll 4
                for (int l = 0; l < k; ++l) t = fma(Trow[l], lds[SB_G + l * 33 + k], t);
                Trow[k] = lane < k ? -tk * t : lane == k ? tk : 0.0;
            }
        }
        __syncthreads();
        {
            const int i = tid >> 5, jj = tid & 31;
            double t = 0.0;  // (M T)[i][jj]
            for (int l = 0; l < SBR_B; ++l) t = fma(lds[SB_M + i * 33 + l], lds[SB_T + l * 33 + jj], t);
            lds[SB_S + i * 33 + jj] = t;  // temporarily M T
        }
        __syncthreads();
        {
            const int i = tid >> 5, jj = tid & 31;
            double t = 0.0;  // (T^T (M T))[i][jj]
            for (int l = 0; l < SBR_B; ++l) t = fma(lds[SB_T + l * 33 + i], lds[SB_S + l * 33 + jj], t);
            __syncthreads();
            lds[SB_S + i * 33 + jj] = t;
        }
        __syncthreads();
        for (int q = 0; q < nq; ++q) {
            const int C = sbr_qblock(Bt, nt, rank, T, q);
            double *const Lv = lds + SB_LA;            // V[C] as [row][k] (64 x 33)
            double *const Ly = lds + SB_LA + 64 * 33;  // Y[C] (64 x 33)
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = tid + TRI_NT * u, kk = e >> 6, rr = e & 63;
                Lv[rr * 33 + kk] = ld_dev(Vb + (size_t)kk * npad + 64 * C + rr);
                const int rr2 = e >> 5, kk2 = e & 31;
                Ly[rr2 * 33 + kk2] = Yb[(size_t)(64 * C + rr2) * SBR_B + kk2];
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int e = tid + TRI_NT * u, jj = e >> 6, rr = e & 63;  // W[64 C + rr][jj]
                const int rw = 64 * C + rr;
                double t1 = 0.0, t2 = 0.0;
                // (rolled: fully unrolled, the T and S reads were hoisted
                // out of the block loop into registers, which spilled)
#pragma unroll 4
                for (int l = 0; l < SBR_B; ++l) {
                    t1 = fma(Ly[rr * 33 + l], lds[SB_T + l * 33 + jj], t1);
                    t2 = fma(Lv[rr * 33 + l], lds[SB_S + l * 33 + jj], t2);
                }
                const double x = (rw >= c0 && rw < n) ? t1 - 0.5 * t2 : 0.0;
                st_dev(Wb + (size_t)jj * npad + rw, x);
            }
        }
        SB_PROF(8);
        ++phase;
        if (!team_barrier(bar, phase * T, status, s_flag, BLK_SPIN_LIMIT)) return -1;
        SB_PROF(9);
        // ---- A22 -= V W^T + W V^T on the owned tile rows
        for (int q = 0; q < nq; ++q) sbr_syr2k_row(A, n, npad, nt, sbr_qblock(Bt, nt, rank, T, q), Bt, c0, Vb, Wb);
        SB_PROF(10);
    return phase;
}

__device__ __noinline__ void sbr_blk(const EigMat &mt, int rank, int T, int *bar, int *status, int *s_flag) {
    __shared__ SbrCtx sc;
    const int n = mt.n, npad = (n + 63) & ~63, nt = npad >> 6;
    if (threadIdx.x == 0) {
        double *const base = (double *)(((uintptr_t)(mt.ws + (size_t)n * n) + 511) & ~(uintptr_t)511);
        sc.A = mt.A;
        sc.ws = mt.ws;
        sc.Vb = base;
        sc.Wb = sc.Vb + (size_t)SBR_B * npad;
        sc.Yb = sc.Wb + (size_t)SBR_B * npad;
        sc.part = sc.Yb + (size_t)SBR_B * npad;          // [parity][block][SBR_B]
        sc.piv = sc.part + 2 * (size_t)nt * SBR_B;       // [parity][SBR_B]
        sc.gram = sc.piv + 2 * SBR_B;                    // [block][2][SBR_B][SBR_B]
        sc.gtot = sc.gram + (size_t)nt * 2 * SBR_B * SBR_B;  // [2][SBR_B][SBR_B]
        sc.tau_g = vslot(mt.ws, n, V_TAU);
        sc.bar = bar;
        sc.status = status;
        sc.s_flag = s_flag;
        sc.n = n;
        sc.npad = npad;
        sc.nt = nt;
        sc.rank = rank;
        sc.T = T;
#ifdef NRG_SBR_PROF
        for (int i = 0; i < 12; ++i) sc.pacc[i] = 0;
        sc.plast = wall_clock64();
#endif
    }
    __syncthreads();
    const int tid = threadIdx.x;
    (void)tid;
    SbrCtx *const scp = &sc;
    {
        SbrCtx *sc = scp;
        (void)sc;
        sbr_mirror(scp);
        SB_PROF(0);
    }
    int phase = 0;
    for (int j0 = 0;; j0 += SBR_B) {
        if (n - (j0 + SBR_B) < 2) break;
        phase = sbr_qr(scp, j0, phase);
        if (phase < 0) return;
        phase = sbr_upd(scp, j0, phase);
        if (phase < 0) return;
    }
#ifdef NRG_SBR_PROF
    if (rank == 0 && tid == 0)
        printf("sbrprof n=%d T=%d mirror %lld qr %lld qrbar %lld qrupd %lld wb+bar %lld symm %lld bar %lld gram+bar %lld tsw %lld bar %lld syr2k %lld (x10ns)\n",
               n, T, sc.pacc[0], sc.pacc[1], sc.pacc[2], sc.pacc[3], sc.pacc[4], sc.pacc[5], sc.pacc[6], sc.pacc[7], sc.pacc[8], sc.pacc[9], sc.pacc[10]);
    if (rank == 0 && tid == 0) printf("sbrprof2 n=%d T=%d bsum %lld\n", n, T, sc.pacc[11]);
#endif
}
#undef SB_PROF

__device__ __forceinline__ void tri_teams(const EigMat *mats, const TriJob *jobs, const int *tbase, const int *qjob,
                                          int *counters, int *status) {
    __shared__ double red[TRI_NW];
    __shared__ int s_ticket, s_flag;
    const int x = blockIdx.x % N_XCD;
    int *bars = counters + N_XCD + 1;     // counters[9 + BAR_STRIDE job]: team barrier
    for (int q = 0; q < 2; ++q) {
    // queue 0: global (jobs qjob[0]..qjob[1]); queue 1: this XCD's
    int *ticket = q == 0 ? counters + N_XCD : counters + x;
    const int j0 = q == 0 ? qjob[0] : qjob[1 + x], j1 = q == 0 ? qjob[1] : qjob[2 + x];
    const int t0 = tbase[j0], total = tbase[j1] - t0;
    for (;;) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(ticket, 1);
        __syncthreads();
        const int t = s_ticket;
        __syncthreads();
        if (t >= total) break;
        int lo = j0, hi = j1;  // last job of the queue with tbase <= t0 + t
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (tbase[mid] <= t0 + t) lo = mid;
            else hi = mid;
        }
        const int rank = t0 + t - tbase[lo];
        // the job, then the jobs chained after it on the same team (same
        // size, same ranks: no wait for members to arrive)
        for (int cur = lo; cur >= 0;) {
            const TriJob jb = jobs[cur];
            const EigMat mt = mats[jb.mat];
            int *const bar = bars + BAR_STRIDE * cur;
            if (jb.blocked == 2) sbr_blk(mt, rank, jb.T, bar, status, &s_flag);
            else if (jb.blocked) tri_blk(mt, rank, jb.T, bar, status, &s_flag);
            else if (mt.n <= TEAM_LDS_N) tri_team<true>(mt, rank, jb.T, bar, status, red, &s_flag);
            else tri_team<false>(mt, rank, jb.T, bar, status, red, &s_flag);
            __syncthreads();
            cur = jb.next;
        }
    }
    }
}

// Persistent team launches (the same body): k_eig_tri_blk when some job uses
// the blocked reduction, so that its time is reported under its own name.
__global__ __launch_bounds__(TRI_NT) void k_eig_tri_coop(const EigMat *mats, const TriJob *jobs, const int *tbase,
                                                         const int *qjob, int *counters, int *status) {
    tri_teams(mats, jobs, tbase, qjob, counters, status);
}
__global__ __launch_bounds__(TRI_NT) void k_eig_tri_blk(const EigMat *mats, const TriJob *jobs, const int *tbase,
                                                        const int *qjob, int *counters, int *status) {
    tri_teams(mats, jobs, tbase, qjob, counters, status);
}

// ---------------------------------------------------------------------------
// 1f. two-stage reduction, stage 2: band -> tridiagonal by bulge chasing.
//     Sweep s (s = 0 .. n - 2) removes column s below the subdiagonal with a
//     reflector on rows s + 1 .. s + b (task 0) and then chases the bulge:
//     task k >= 1 removes the fill of column c = s + 1 + (k - 1) b with a
//     reflector on rows r0 = s + 1 + k b .. r1 - 1 (r1 = min(r0 + b, n)),
//     applied two-sided.  Task (s, k) reads and writes only columns [c, r1)
//     of the lower band (offsets d < 2b below the diagonal), so sweep s may
//     run task k once sweep s - 1 has finished its tasks 0 .. k + 2: sweeps
//     run concurrently, one per chaser workgroup, each waiting on its
//     predecessor's progress counter.  A chaser keeps the 2b columns of its
//     current task in LDS (ring of 2b slots): task k loads the b new columns
//     [r0, r1) and writes back the b columns [c, r0) it finishes with.
//     Output: d, e (V_DG, V_OD); reflector (s, k) (tau in place of its unit
//     first element) in row n - 1 - s of A's strict lower triangle, columns
//     k b .. k b + len - 1 (the upper triangle holds stage 1's reflectors).
// ---------------------------------------------------------------------------
constexpr int CH_NT = 256;
constexpr int CH_W = 2 * SBR_B;     // columns in a chaser's ring, entries per band column
constexpr int CH_DONE = 1 << 30;    // progress of a finished sweep
// LDS row stride of a chaser's ring (doubles): even, so that the accesses
// along anti-diagonals (next column, one entry up: the dot products and the
// two-sided update) step by an odd number of doubles and spread over the
// banks (an odd stride put them all in one bank)
constexpr int CH_RS = CH_W + 2;
constexpr long long CH_SPIN_LIMIT = 1LL << 23;  // polls of ~1 us: seconds, far beyond any wait of a running team

struct SbrJob {
    int mat, P;  // matrix, chasers of its team
};

// band buffer in S1: column j holds A[j + d][j], d < CH_W (zero beyond n), then n progress ints
__host__ __device__ inline double *sbr_band(double *ws, int n) {
    return (double *)(((uintptr_t)(ws + (size_t)n * n) + 511) & ~(uintptr_t)511);
}
__host__ __device__ inline int *sbr_prog(double *ws, int n) { return (int *)(sbr_band(ws, n) + (size_t)n * CH_W); }

// After stage 1: the band (row j, columns j .. j + SBR_B) into the band
// buffer, taking a panel row's columns from the panel's trailing start on
// from S2 and the rest from A's lower triangle (A[j + d][j]); the panel
// rows' R and reflectors copied from S2 into A's upper triangle (S2 is
// the divide-and-conquer workspace later); progress counters zeroed.
// Item: (matrix, 16 rows).
__global__ __launch_bounds__(256) void k_sbr_band(const EigMat *mats, const BtItem *items) {
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n, nr = sbr_nrefl(n);
    double *band = sbr_band(mt.ws, n);
    int *prog = sbr_prog(mt.ws, n);
    const double *P2 = mt.ws + 2 * (size_t)n * n;
    const int j0 = it.c0;
    for (int e = threadIdx.x; e < 16 * CH_W; e += 256) {  // 16 rows per item
        const int j = j0 + e / CH_W, d = e % CH_W;
        if (j >= n) continue;
        const int c0j = (j / SBR_B + 1) * SBR_B;  // trailing start of row j's panel
        const int col = j + d;
        double v = 0.0;
        if (d <= SBR_B && col < n) v = (j < nr && col >= c0j) ? P2[(size_t)j * n + col] : mt.A[(size_t)col * n + j];
        band[(size_t)j * CH_W + d] = v;
        if (d == 0) prog[j] = 0;
    }
    for (int jj = 0; jj < 16; ++jj) {
        const int j = j0 + jj;
        if (j >= nr) break;
        const int c0j = (j / SBR_B + 1) * SBR_B;
        for (int col = c0j + threadIdx.x; col < n; col += 256) mt.A[(size_t)j * n + col] = P2[(size_t)j * n + col];
    }
}

__device__ __forceinline__ bool ch_wait(const gint *prog, int need, int *status, int *s_flag) {
    if (threadIdx.x == 0) {
        long long spins = 0;
        int ok = 1;
        while (__hip_atomic_load(prog, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > CH_SPIN_LIMIT || __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                atomicCAS(status, 0, NRG_ERR_HIP);
                break;
            }
        }
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag != 0;
}

// one sweep of one chaser
__device__ __forceinline__ bool ch_sweep(const EigMat &mt, int s, double (*ring)[CH_RS], double *vs, double *dl,
                                         double *wq, double *db, double *sc, int *status, int *s_flag) {
    const int n = mt.n, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    double *const band = sbr_band(mt.ws, n);
    // (global-address pointers: a generic store also counts against the LDS
    // counter and may alias LDS, so the task's next LDS wait or barrier
    // waited for the store to reach memory)
    gint *const prog = (gint *)sbr_prog(mt.ws, n);
    gdouble *const gdg = (gdouble *)vslot(mt.ws, n, V_DG), *const god = (gdouble *)vslot(mt.ws, n, V_OD);
    gdouble *const gA = (gdouble *)mt.A;
    auto slot = [&](int j) { return (j - s - 1 + CH_W) % CH_W; };
    auto L = [&](int j, int d) -> double & { return ring[slot(j)][d]; };
    // (at most b + 1 columns: the loads of a thread are issued together,
    // then stored to LDS; a load and its store in one loop iteration would
    // cost one round trip per entry)
    constexpr int LPT = (SBR_B + 1) * CH_W / CH_NT + 1;
    auto load_cols = [&](int ja, int jb) {
        const int ne = (jb - ja) * CH_W;
        double r[LPT];
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
            const int e = min(tid + CH_NT * u, ne - 1);
            r[u] = ld_dev(band + (size_t)(ja + e / CH_W) * CH_W + e % CH_W);
        }
#pragma unroll
        for (int u = 0; u < LPT; ++u) {
            const int e = tid + CH_NT * u;
            if (e < ne) ring[slot(ja + e / CH_W)][e % CH_W] = r[u];
        }
    };
    auto store_cols = [&](int ja, int jb) {
        for (int e = tid; e < (jb - ja) * CH_W; e += CH_NT) {
            const int j = ja + e / CH_W, d = e % CH_W;
            st_dev(band + (size_t)j * CH_W + d, ring[slot(j)][d]);
        }
    };
    auto signal = [&](int v) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(prog + s, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    if (s > 0 && !ch_wait(prog + s - 1, 3, status, s_flag)) return false;
    load_cols(s, min(s + SBR_B + 1, n));
    __syncthreads();
    for (int k = 0;; ++k) {
        const int c = k == 0 ? s : s + 1 + (k - 1) * SBR_B;
        const int r0 = s + 1 + k * SBR_B, r1 = min(r0 + SBR_B, n), len = r1 - r0;
        const int r2 = min(r1 + SBR_B, n), nb = r2 - r1;
        if (k > 0) {
            if (s > 0 && !ch_wait(prog + s - 1, k + 3, status, s_flag)) return false;
            load_cols(r0, r1);
            __syncthreads();
        }
        // reflector from column c, rows r0 .. r1 - 1
        if (w == 0) {
            const double x = lane < len ? L(c, r0 - c + lane) : 0.0;
            const double sigma = wave_sum(lane >= 1 && lane < len ? x * x : 0.0);
            const double alpha = readlane_d(x, 0);
            double tau = 0.0, beta = alpha, scale = 0.0;
            if (sigma != 0.0) {
                beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
                tau = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            if (lane < SBR_B) vs[lane] = lane == 0 ? 1.0 : lane < len ? x * scale : 0.0;
            if (lane < len) {
                L(c, r0 - c + lane) = lane == 0 ? beta : 0.0;
                // reflector (s, k): row n - 1 - s, columns k b ..
                gA[(size_t)(n - 1 - s) * n + k * SBR_B + lane] = lane == 0 ? tau : x * scale;
            }
            if (lane == 0) {
                sc[0] = tau;
                if (k == 0) {
                    gdg[s] = L(s, 0);
                    god[s] = beta;
                }
            }
        }
        __syncthreads();
        const double tau = sc[0];
        if (tau != 0.0) {
            // dot products: left columns (c, r0), D v, below rows [r1, r2),
            // one wavefront each (three SIMDs at once), 32 terms unrolled
            // with the terms past len skipped by a select (ring entries
            // there may hold anything)
            const int nl = r0 - c - 1;
            if (w == 0 && lane < nl) {
                const int j = c + 1 + lane;
                double t = 0.0;
#pragma unroll
                for (int u = 0; u < SBR_B; ++u) {
                    const double x = L(j, r0 + min(u, len - 1) - j);
                    t = u < len ? fma(vs[u], x, t) : t;
                }
                dl[lane] = t;
            } else if (w == 1 && lane < len) {
                const int t = lane;
                double a = 0.0;
#pragma unroll
                for (int u = 0; u < SBR_B; ++u) {
                    const int uu = min(u, len - 1);
                    const double x = t >= uu ? L(r0 + uu, t - uu) : L(r0 + t, uu - t);
                    a = u < len ? fma(x, vs[u], a) : a;
                }
                wq[t] = tau * a;  // p = tau D v
            } else if (w == 2 && lane < nb) {
                const int t = lane;
                double a = 0.0;
#pragma unroll
                for (int u = 0; u < SBR_B; ++u) {
                    const int uu = min(u, len - 1);
                    const double x = L(r0 + uu, r1 + t - r0 - uu);
                    a = u < len ? fma(x, vs[u], a) : a;
                }
                db[t] = a;
            }
            __syncthreads();
            if (w == 0) {  // q = p - (tau / 2) (v . p) v
                const double pv = wave_sum(lane < len ? vs[lane] * wq[lane] : 0.0);
                const double K = 0.5 * tau * pv;
                if (lane < len) wq[lane] -= K * vs[lane];
            }
            __syncthreads();
            // rows of 32 entries: left columns, D, below rows (u = e mod 32)
            const int nrow = nl + len + nb;
            for (int e = tid; e < nrow * SBR_B; e += CH_NT) {
                const int R = e / SBR_B, u = e % SBR_B;
                if (u >= len) continue;
                if (R < nl) {
                    const int j = c + 1 + R;
                    L(j, r0 + u - j) -= tau * dl[R] * vs[u];
                } else if (R < nl + len) {
                    const int t = R - nl;
                    if (t >= u) L(r0 + u, t - u) -= vs[t] * wq[u] + wq[t] * vs[u];
                } else {
                    const int t = R - nl - len;
                    L(r0 + u, r1 + t - r0 - u) -= tau * db[t] * vs[u];
                }
            }
            __syncthreads();
        }
        if (r1 >= n) {  // last task: the whole window goes back
            if (s == n - 2 && tid == 0) {
                gdg[n - 1] = L(n - 1, 0);
                god[n - 1] = 0.0;
            }
            store_cols(c, n);
            signal(CH_DONE);
            return true;
        }
        store_cols(c, r0);
        signal(k + 1);
    }
}

__global__ __launch_bounds__(CH_NT) void k_sbr_chase(const EigMat *mats, const SbrJob *jobs, const int *tbase, int njobs,
                                                     int *ticket, int *status) {
    __shared__ double ring[CH_W][CH_RS];
    __shared__ double vs[SBR_B], dl[SBR_B], wq[SBR_B], db[SBR_B], sc[4];
    __shared__ int s_t, s_flag;
    const int total = tbase[njobs];
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(ticket, 1);
        __syncthreads();
        const int t = s_t;
        __syncthreads();
        if (t >= total) break;
        int lo = 0, hi = njobs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (tbase[mid] <= t) lo = mid;
            else hi = mid;
        }
        const SbrJob jb = jobs[lo];
        const EigMat mt = mats[jb.mat];
        const int q = t - tbase[lo];
        for (int s = q; s < mt.n - 1; s += jb.P)
            if (!ch_sweep(mt, s, ring, vs, dl, wq, db, sc, status, &s_flag)) return;
        __syncthreads();
    }
}

// Raw buffer access (vector memory instructions: a scalar resource, the
// lane's 32-bit offset and a scalar offset per access, so a run of loads
// from uniform rows needs one address register).  Offsets at or past
// `bytes` read 0 and drop stores (range-checked by the hardware).
typedef unsigned int bt_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, int vo, int so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, int vo, int so, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bt_u32x2, v), r, vo, so, 0);
}

// Two-stage matrices: the eigenvectors of the tridiagonal between Vc (row v
// = vector v) and an entry-major copy Zt in S2 (row p = entry p of every
// vector), in 64 x 64 tiles through LDS (item: matrix, tile ti * nt + tj of
// the source).  The stage-2 back-transformation runs on Zt.
__global__ __launch_bounds__(256) void k_sbr_tr(const EigMat *mats, const BtItem *items, int to_zt) {
    __shared__ double tile[64][65];
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n, nt = (n + 63) / 64;
    const int ti = it.c0 / nt, tj = it.c0 % nt;
    double *const Vc = mt.ws, *const Zt = mt.ws + 2 * (size_t)n * n;
    const double *const src = to_zt ? Vc : Zt;
    double *const dst = to_zt ? Zt : Vc;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    double r[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int row = 64 * ti + ty + 4 * u, col = 64 * tj + tx;
        const double v = src[(size_t)min(row, n - 1) * n + min(col, n - 1)];
        r[u] = (row < n && col < n) ? v : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) tile[ty + 4 * u][tx] = r[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int row = 64 * tj + ty + 4 * u, col = 64 * ti + tx;
        if (row < n && col < n) dst[(size_t)row * n + col] = tile[tx][ty + 4 * u];
    }
}

// ---------------------------------------------------------------------------
// 1g. back-transformation of stage 2: x = Q2 y for every eigenvector y of the
//     tridiagonal (row of Vc).  Q2 = prod over sweep blocks B = [s0, s0 + L)
//     (ascending) of prod_{k descending} prod_{s in B ascending} H(s, k): the
//     same product as the sweeps' order, since H(s, k) and H(s', k') with
//     s < s' commute whenever k' > k and reflectors of one sweep are
//     disjoint.  So for each block (the last first) and k = 0, 1, .. the
//     reflectors H(s, k), s descending, act on the window [s0 + 1 + k b,
//     s0 + 1 + k b + b + L - 1) of every vector.  Lane = vector: the window
//     lives in registers (compile-time indices), moves by b per step (b
//     finished entries stored, b new ones loaded); the step's reflectors are
//     staged in LDS and read by all lanes alike.  BL sweeps per block: the
//     vectors are streamed through once per block.
// ---------------------------------------------------------------------------
// One wavefront per 64 vectors (lane = vector), BL sweeps per block; the
// next step's new window entries are loaded while this step computes, its
// reflectors after it (into the other of two LDS buffers).  Measured on the
// keep-10000 block set: BL = 48 1.14 s, 32 1.30 s, 64 (spills) 1.60 s; the
// window traffic coalesced through an LDS transpose 1.59 s, four wavefronts
// sharing one staging 1.68 s, reflectors a step ahead in registers 1.70 s
// (occupancy): the reflectors' LDS broadcast reads bound a step, not HBM.
constexpr int SBT_BL = 48;  // sweeps per block of the stage-2 back-transformation
// PF: the next step's new window entries are loaded while this step
// computes (measured on the keep-10000 block set: 0.82 s with, 0.85 s
// without; capping the registers at two wavefronts per SIMD spills and
// takes 2.4-3.3 s)
template <int BL, bool PF>
__global__ __launch_bounds__(64) void k_sbr_bt(const EigMat *mats, const BtItem *items) {
    constexpr int WIN = SBR_B + BL - 1;  // window length
    __shared__ double rv[2][BL][SBR_B];
    __shared__ double rt[2][BL];
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n, lane = threadIdx.x;
    const int cv = it.c0 + lane;  // vector (column of Zt)
    const bool act = cv < n;
    // the vectors entry-major (Zt row p = entry p of every vector), so that a
    // window load or store of the 64 lanes is one contiguous 512-byte run
    double *const Zt = mt.ws + 2 * (size_t)n * n;
    // window rows through a buffer resource based at the step's first row:
    // rows past n read 0, and the lanes past n (offset 2^31, beyond the
    // resource's 1 GiB) neither read nor write
    const int vo = act ? cv * 8 : (int)0x80000000u;
    auto rows_from = [&](int base) {
        return buf_rsrc(Zt + (size_t)base * n, (unsigned)min<size_t>((size_t)(n - base) * n * 8, (size_t)1 << 30));
    };
    const int rb = n * 8;  // bytes per row
    const gdouble *const A = (const gdouble *)mt.A;
    const int nsw = n - 1;  // sweeps 0 .. n - 2
    const int nblk = (nsw + BL - 1) / BL;
    double z[WIN], zn[PF ? SBR_B : 1];
    // reflectors H(s, k), s in [s0, s1) (row n-1-s, columns k b ..; tau in
    // place of the unit first element) into LDS buffer buf
    // (all RPL loads issued before their LDS stores, so the staging waits
    // for one round trip; a load and its store in one loop iteration wait
    // for each other, one round trip per entry, and two batches of RPL / 2
    // measured 1.5 % slower)
    constexpr int RPL = BL * SBR_B / 64, RB = RPL;
    static_assert(RPL % RB == 0, "reflector staging batches");
    auto stage_refl = [&](int s0, int s1, int k, int buf) {
        const int base = s0 + 1 + k * SBR_B;
#pragma unroll
        for (int u0 = 0; u0 < RPL; u0 += RB) {
            double r[RB];
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int e = lane + 64 * (u0 + u), o = e / SBR_B, t = e % SBR_B, s = s0 + o;
                const int st = base + o;  // first index of H(s, k)
                const bool ok = s < s1 && st < n && st + t < n;
                const double f = ok ? 1.0 : 0.0;
                r[u] = f * A[ok ? (size_t)(n - 1 - s) * n + k * SBR_B + t : 0];
            }
#pragma unroll
            for (int u = 0; u < RB; ++u) {
                const int e = lane + 64 * (u0 + u), o = e / SBR_B, t = e % SBR_B;
                if (t == 0) {
                    rt[buf][o] = r[u];
                    rv[buf][o][0] = 1.0;
                } else {
                    rv[buf][o][t] = r[u];
                }
            }
        }
    };
    for (int bi = nblk - 1; bi >= 0; --bi) {
        const int s0 = bi * BL, s1 = min(s0 + BL, nsw);
        int base = s0 + 1;
        {
            const __amdgpu_buffer_rsrc_t rs = rows_from(base);
#pragma unroll
            for (int p = 0; p < WIN; ++p) z[p] = buf_ld(rs, vo, p * rb);
        }
        stage_refl(s0, s1, 0, 0);
        wave_lds_sync();
        for (int k = 0, buf = 0; base < n; ++k, base += SBR_B, buf ^= 1) {
            const bool more = base + SBR_B < n;
            const __amdgpu_buffer_rsrc_t rs = rows_from(base);
            if (PF) {
#pragma unroll
                for (int p = 0; p < SBR_B; ++p) zn[PF ? p : 0] = buf_ld(rs, vo, (WIN + p) * rb);
            }
#pragma unroll
            for (int o = BL - 1; o >= 0; --o) {
                const double tau = rt[buf][o];
                if (tau == 0.0) continue;
                // four partial sums: a chain of 8 dependent FMAs instead of 32
                double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
                for (int t = 0; t < SBR_B; t += 4) {
                    d0 = fma(rv[buf][o][t], z[o + t], d0);
                    d1 = fma(rv[buf][o][t + 1], z[o + t + 1], d1);
                    d2 = fma(rv[buf][o][t + 2], z[o + t + 2], d2);
                    d3 = fma(rv[buf][o][t + 3], z[o + t + 3], d3);
                }
                const double d = tau * ((d0 + d1) + (d2 + d3));
#pragma unroll
                for (int t = 0; t < SBR_B; ++t) z[o + t] = fma(-d, rv[buf][o][t], z[o + t]);
            }
            if (more) stage_refl(s0, s1, k + 1, buf ^ 1);
            wave_lds_sync();
            // entries base .. base + b - 1 are final for this block
#pragma unroll
            for (int p = 0; p < SBR_B; ++p) buf_st(rs, vo, p * rb, z[p]);
#pragma unroll
            for (int p = 0; p < WIN - SBR_B; ++p) z[p] = z[p + SBR_B];
#pragma unroll
            for (int p = 0; p < SBR_B; ++p)
                z[WIN - SBR_B + p] = PF ? zn[PF ? p : 0] : more ? buf_ld(rs, vo, (WIN + p) * rb) : 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// 2. divide and conquer on (dg, od)
// Tree: nlev = floor(log2 n) levels, nl = 2^nlev leaves, leaf k = rows
// [k n / nl, (k+1) n / nl) (1 or 2 rows).  The coupling b cut at each leaf
// boundary mid is removed from the two adjacent diagonal entries (Cuppen),
// T = diag(T1, T2) + |b| u u^T with u = e_{mid-1} + sign(b) e_mid.
// ---------------------------------------------------------------------------
__host__ __device__ inline int dc_bound(int k, int n, int nlev) { return (int)(((long long)k * n) >> nlev); }
__host__ __device__ inline int dc_levels(int n) {
    int l = 0;
    while ((2 << l) <= n) ++l;
    return l;  // floor(log2 n)
}

__global__ __launch_bounds__(256) void k_dc_leaf(const EigMat *mats) {
    const EigMat mt = mats[blockIdx.x];
    const int n = mt.n;
    double *ws = mt.ws;
    const double *dg = vslot(ws, n, V_DG), *od = vslot(ws, n, V_OD);
    double *D = vslot(ws, n, V_D), *beta = vslot(ws, n, V_BETA);
    double *Vc = ws;
    if (mt.pad == 1) {
        // reflectors of a team reduction: S2 row j, columns j+1.. -> row j of A
        const double *Y = ws + 2 * (size_t)n * n;
        for (int j = 0; j < n - 2; ++j)
            for (int k = j + 1 + threadIdx.x; k < n; k += blockDim.x) mt.A[(size_t)j * n + k] = Y[(size_t)j * n + k];
    }
    const int nlev = dc_levels(n), nl = 1 << nlev;
    for (int k = threadIdx.x; k < nl; k += blockDim.x) {
        const int a = dc_bound(k, n, nlev), b = dc_bound(k + 1, n, nlev);
        double x = dg[a];
        if (k > 0) {
            x -= fabs(od[a - 1]);
            beta[a] = od[a - 1];
        }
        if (b - a == 1) {
            if (k < nl - 1) x -= fabs(od[a]);
            D[a] = x;
            Vc[(size_t)a * n + a] = 1.0;
        } else {
            double w = dg[a + 1];
            if (k < nl - 1) w -= fabs(od[a + 1]);
            const double y = od[a];
            if (y == 0.0) {
                D[a] = x;
                D[a + 1] = w;
                Vc[(size_t)a * n + a] = 1.0;
                Vc[(size_t)a * n + a + 1] = 0.0;
                Vc[(size_t)(a + 1) * n + a] = 0.0;
                Vc[(size_t)(a + 1) * n + a + 1] = 1.0;
            } else {
                // Jacobi rotation of [[x, y], [y, w]]
                const double th = (w - x) / (2.0 * y);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                D[a] = x - t * y;
                D[a + 1] = w + t * y;
                // eigenvector a = (c, -s), eigenvector a+1 = (s, c)
                Vc[(size_t)a * n + a] = c;
                Vc[(size_t)a * n + a + 1] = -s;
                Vc[(size_t)(a + 1) * n + a] = s;
                Vc[(size_t)(a + 1) * n + a + 1] = c;
            }
        }
    }
}

// Column r of the gathered block of one merge [a, a + K) (h = mid - a) over
// the sorted poles q0 <= q < q1 of one gather segment: row pos[q] of G (in
// S1) is eigenvector perm[q] restricted to its own half, with the deflation
// rotations applied as they come.  A rotation always pairs pole q with the
// last pole before it that was not deflated by a small z (the carried
// value), so the carried element stays in a register until the next new
// pole replaces it; a segment starts at a new pole (or at q = 0), so no
// carry crosses segments.
// Vc and G point at the merge's diagonal block (row and column a), both with
// row stride ld.
__device__ inline void dc_gather_seg(const double *Vc, double *G, int ld, int h, int q0, int q1, int r,
                                     const int *perm, const int *pos, const int *kind, const double *rc,
                                     const double *rs) {
    const size_t n = ld;
    const bool lo = r < h;
    int cr = -1;
    double cv = 0.0;
    constexpr int U = 8;
    for (int qb = q0; qb < q1; qb += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = qb + u;
            v[u] = 0.0;
            if (q < q1) {
                const int i = perm[q];
                if ((i < h) == lo) v[u] = Vc[(size_t)i * n + r];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = qb + u;
            if (q >= q1) break;
            const int k = kind[q] & KIND_MASK;
            if (k == 0) {
                G[(size_t)pos[q] * n + r] = v[u];
            } else if (k == 1 || cr < 0) {  // (a rotation always has a carried pole)
                if (cr >= 0) G[(size_t)pos[cr] * n + r] = cv;
                cr = q;
                cv = v[u];
            } else {
                const double c = rc[q], s = rs[q];
                G[(size_t)pos[cr] * n + r] = c * cv + s * v[u];
                cv = c * v[u] - s * cv;
                cr = q;
            }
        }
    }
    if (cr >= 0) G[(size_t)pos[cr] * n + r] = cv;
}

// Ascending sort of (key[i], idx[i]), i < K, by the total order (key, idx):
// the flip form of Batcher's bitonic network, in which every comparator puts
// the smaller element at the lower index.  Padding up to the power of two P
// is conceptually +inf and never moves, so comparators reaching j >= K are
// skipped.  key/idx are in LDS or in global memory (one workgroup).
__device__ inline void dc_bitonic_sort(double *key, int *idx, int K, int NT) {
    int P = 1;
    while (P < K) P <<= 1;
    const int tid = threadIdx.x;
    for (int B = 2; B <= P; B <<= 1) {
        for (int d = B >> 1; d >= 1; d >>= 1) {
            const bool flip = d == (B >> 1);
            for (int c = tid; c < (P >> 1); c += NT) {
                int i, j;
                if (flip) {
                    const int blk = c / d, t = c - blk * d;
                    i = blk * B + t;
                    j = blk * B + B - 1 - t;
                } else {
                    const int blk = c / d, t = c - blk * d;
                    i = blk * 2 * d + t;
                    j = i + d;
                }
                if (j >= K) continue;
                const double ki = key[i], kj = key[j];
                const int ii = idx[i], ij = idx[j];
                if (kj < ki || (kj == ki && ij < ii)) {
                    key[i] = kj;
                    key[j] = ki;
                    idx[i] = ij;
                    idx[j] = ii;
                }
            }
            __syncthreads();
        }
    }
}

// Block-wide exclusive scans of one int per thread (sum, or max with
// identity -1); every thread also receives the total.
template <bool MAX>
__device__ inline int dc_block_scan(int x, int *sred, int &total);

// The eigenvalues of a merge are the two halves' outputs, each of them
// ascending roots followed by ascending deflated values: a few ascending
// runs.  When key[0, K) has at most DEFL_RUNS of them, this sorts it into
// skey / idx in the order of dc_bitonic_sort ((key, index) ascending, a total
// order, so the permutation is the same): the rank of element i is its offset
// in its own run plus, for every other run, the number of that run's elements
// before it (key <= x in earlier runs, key < x in later ones; binary
// searches).  key and skey must differ; returns false, writing nothing, when
// there are more runs.
constexpr int DEFL_RUNS = 16;
__device__ inline bool dc_run_sort(const double *key, double *skey, int *idx, int K, int *runs, int *sred) {
    const int tid = threadIdx.x;
    const int per = (K + DEFL_NT - 1) / DEFL_NT, q0 = min(K, tid * per), q1 = min(K, q0 + per);
    int nb = 0;
    for (int q = max(q0, 1); q < q1; ++q) nb += key[q] < key[q - 1];
    int total;
    int base = 1 + dc_block_scan<false>(nb, sred, total);
    const int nr = total + 1;
    if (nr > DEFL_RUNS) return false;
    for (int q = max(q0, 1); q < q1; ++q)
        if (key[q] < key[q - 1]) runs[base++] = q;
    if (tid == 0) {
        runs[0] = 0;
        runs[nr] = K;
    }
    __syncthreads();
    for (int i = tid; i < K; i += DEFL_NT) {
        const double x = key[i];
        int rank = 0;
        for (int r = 0; r < nr; ++r) {
            const int lo = runs[r], hi = runs[r + 1];
            if (i >= lo && i < hi) {
                rank += i - lo;
                continue;
            }
            int l = lo, h = hi;
            if (hi <= i) {
                while (l < h) {
                    const int m = (l + h) >> 1;
                    if (key[m] <= x) l = m + 1;
                    else h = m;
                }
            } else {
                while (l < h) {
                    const int m = (l + h) >> 1;
                    if (key[m] < x) l = m + 1;
                    else h = m;
                }
            }
            rank += l - lo;
        }
        idx[rank] = i;
    }
    __syncthreads();
    for (int q = tid; q < K; q += DEFL_NT) skey[q] = key[idx[q]];
    __syncthreads();
    return true;
}

template <bool MAX>
__device__ inline int dc_block_scan(int x, int *sred, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc = MAX ? max(inc, y) : inc + y;
    }
    __syncthreads();
    if (lane == 63) sred[w] = inc;
    __syncthreads();
    int base = MAX ? -1 : 0, all = MAX ? -1 : 0;
#pragma unroll
    for (int k = 0; k < DEFL_NT / 64; ++k) {
        const int v = sred[k];
        if (k < w) base = MAX ? max(base, v) : base + v;
        all = MAX ? max(all, v) : all + v;
    }
    total = all;
    int ex = __shfl_up(inc, 1, 64);
    if (lane == 0) ex = MAX ? -1 : 0;
    return MAX ? max(base, ex) : base + ex;
}

// Deflation test of sorted pole q against the previous undeflated pole p
// (the same arithmetic for the parallel first pass and for rotation chains).
__device__ inline bool dc_close(double zp, double zq, double dp, double dq, double tol, double &c, double &s,
                                double &tau) {
    s = zp;
    c = zq;
    tau = hypot(c, s);
    const double t = dq - dp;
    c /= tau;
    s = -s / tau;
    return fabs(t * c * s) <= tol;
}

// Sort, z, deflation and gather segments of one merge [a, b) with boundary
// mid (the gather itself runs in k_dc_gather).
// Deflation follows the sequential rule of LAPACK dlaed2: a pole with a
// small z component is deflated; otherwise, when it is close to the previous
// undeflated pole, a Givens rotation moves that pole's z into it and the
// previous pole is deflated.  The test of pole q depends on earlier
// decisions only through a rotation into its predecessor, so all tests are
// made in parallel and one thread re-walks only candidates and the poles
// that follow a rotation.
template <bool IN_LDS>
__global__ __launch_bounds__(DEFL_NT) void k_dc_deflate(const DcMerge *merges, nrg_gemm_out *outs,
                                                      nrg_gemm_term *terms) {
    extern __shared__ double lds[];
    __shared__ double red[DEFL_NT / 64];
    __shared__ int sred[DEFL_NT / 64];
    __shared__ int s_info[2];
    const DcMerge mg = merges[blockIdx.x];
    const int n = mg.n, a = mg.a, h = mg.mid - mg.a, K = mg.b - mg.a;
    double *ws = mg.ws;
    const double *Vc = ws;
    double *D = vslot(ws, n, V_D) + a;
    const double beta = vslot(ws, n, V_BETA)[mg.mid];
    const double rho = 2.0 * fabs(beta);
    const double sg = beta >= 0.0 ? 1.0 : -1.0;
    const double rs2 = 0.70710678118654752440;
    const int tid = threadIdx.x;
    double *ds, *zs;
    int *perm, *pos, *prev, *cand;
    if (IN_LDS) {
        ds = lds;
        zs = lds + K;
        perm = reinterpret_cast<int *>(lds + 2 * K);
        pos = perm + K;
        prev = pos + K;
        cand = prev + K;
    } else {
        ds = vslot(ws, n, V_SDS) + a;
        zs = vslot(ws, n, V_SZS) + a;
        perm = islot(ws, n, V_PERM) + a;
        pos = islot(ws, n, V_POS) + a;
        prev = islot(ws, n, V_ORG) + a;  // free until the secular stage
        cand = islot(ws, n, V_TS) + a;
    }
    int *kind = islot(ws, n, V_KIND) + a;
    double *rc = vslot(ws, n, V_RC) + a, *rsn = vslot(ws, n, V_RS) + a;
    __shared__ int s_runs[DEFL_RUNS + 1];
    // unsorted keys: an LDS copy in zs (written only after the sort)
    const double *key = D;
    if (IN_LDS) {
        for (int i = tid; i < K; i += DEFL_NT) zs[i] = D[i];
        __syncthreads();
        key = zs;
    }
    if (!dc_run_sort(key, ds, perm, K, s_runs, sred)) {
        for (int i = tid; i < K; i += DEFL_NT) {
            ds[i] = D[i];
            perm[i] = i;
        }
        __syncthreads();
        dc_bitonic_sort(ds, perm, K, DEFL_NT);
    }
    double md = 0.0, mz = 0.0;
    for (int q = tid; q < K; q += DEFL_NT) {
        const int i = perm[q];
        const size_t c = (size_t)(a + i);
        const double z = (i < h) ? Vc[c * n + (mg.mid - 1)] * rs2 : sg * Vc[c * n + mg.mid] * rs2;
        zs[q] = z;
        md = fmax(md, fabs(ds[q]));
        mz = fmax(mz, fabs(z));
    }
    md = block_max<DEFL_NT>(md, red);
    mz = block_max<DEFL_NT>(mz, red);
    const double tol = 8.0 * kEps * fmax(md, rho * mz);
    // previous pole with a large z: exclusive max-scan over contiguous
    // segments of the sorted order
    const int per = (K + DEFL_NT - 1) / DEFL_NT, q0 = min(K, tid * per), q1 = min(K, q0 + per);
    {
        int m = -1;
        for (int q = q0; q < q1; ++q)
            if (!(rho * fabs(zs[q]) <= tol)) m = q;
        int tot;
        int run = dc_block_scan<true>(m, sred, tot);
        for (int q = q0; q < q1; ++q) {
            prev[q] = run;
            if (!(rho * fabs(zs[q]) <= tol)) run = q;
        }
    }
    __syncthreads();
    // parallel tests with the unrotated values
    int ncand_local = 0;
    for (int q = q0; q < q1; ++q) {
        int k = 0;
        if (!(rho * fabs(zs[q]) <= tol)) {
            k = 1;
            const int p = prev[q];
            double c, s, tau;
            if (p >= 0 && dc_close(zs[p], zs[q], ds[p], ds[q], tol, c, s, tau)) {
                k |= KIND_CAND;
                ++ncand_local;
            }
        }
        kind[q] = k;
    }
    int ncand;
    int cbase = dc_block_scan<false>(ncand_local, sred, ncand);
    for (int q = q0; q < q1; ++q)
        if (kind[q] & KIND_CAND) cand[cbase++] = q;
    // the next large-z pole after each large-z pole (pos is free until the
    // positions below): the walk leaves the poles after its current one as
    // they are, so these successors stay valid while it runs
    for (int q = q0; q < q1; ++q) pos[q] = -1;
    __syncthreads();
    for (int q = q0; q < q1; ++q)
        if (!(rho * fabs(zs[q]) <= tol) && prev[q] >= 0) pos[prev[q]] = q;
    __syncthreads();
    if (tid == 0) {
        int ci = 0, chain = -1, nrot = 0;
        for (;;) {
            int q;
            if (chain >= 0) {
                q = chain;
                chain = -1;
                while (ci < ncand && cand[ci] <= q) ++ci;
            } else if (ci < ncand) {
                q = cand[ci++];
            } else {
                break;
            }
            const int p = prev[q];
            double c, s, tau;
            int k = kind[q] & ~KIND_CAND;
            if (dc_close(zs[p], zs[q], ds[p], ds[q], tol, c, s, tau)) {
                zs[q] = tau;
                zs[p] = 0.0;
                k = (k & ~KIND_MASK) | 2;
                rc[q] = c;
                rsn[q] = s;
                ++nrot;
                const double tt = ds[p] * c * c + ds[q] * s * s;
                ds[q] = ds[p] * s * s + ds[q] * c * c;
                ds[p] = tt;
                kind[p] |= KIND_SRC;
                // the next large-z pole is tested against the rotated values
                const int nx = pos[q];
                if (nx >= 0) chain = nx;
            }
            kind[q] = k;
        }
        s_info[1] = nrot;
    }
    __syncthreads();
    // positions: undeflated poles first (ascending, as they stay sorted),
    // then the deflated ones in sorted order
    {
        int und = 0;
        for (int q = q0; q < q1; ++q) {
            const int k = kind[q];
            und += (k & KIND_MASK) != 0 && !(k & KIND_SRC);
        }
        int kp;
        int ub = dc_block_scan<false>(und, sred, kp);
        int db = kp + (q0 - (ub));  // deflated before q0 = q0 - undeflated before q0
        for (int q = q0; q < q1; ++q) {
            const int k = kind[q] & ~KIND_CAND;
            kind[q] = k;
            if ((k & KIND_MASK) != 0 && !(k & KIND_SRC)) pos[q] = ub++;
            else pos[q] = db++;
        }
        if (tid == 0) {
            s_info[0] = kp;
            int *meta = islot(ws, n, V_META) + a;
            meta[0] = kp;
            meta[1] = s_info[1];
            vslot(ws, n, V_RHO)[a] = rho;
            outs[mg.gemm].m = kp;
            terms[mg.gemm].k = kp;
        }
    }
    __syncthreads();
    const int kp = s_info[0];
    double *dd = vslot(ws, n, V_DD) + a, *zz = vslot(ws, n, V_ZZ) + a;
    for (int q = tid; q < K; q += DEFL_NT) {
        const int p = pos[q];
        if (p < kp) {
            dd[p] = ds[q];
            zz[p] = zs[q];
        } else {
            D[p] = ds[q];  // deflated eigenvalue, final for this merge
        }
    }
    if (IN_LDS) {
        // the gather reads the permutation from global memory
        int *gperm = islot(ws, n, V_PERM) + a, *gpos = islot(ws, n, V_POS) + a;
        for (int q = tid; q < K; q += DEFL_NT) {
            gperm[q] = perm[q];
            gpos[q] = pos[q];
        }
    }
    // gather segments: q = 0 and the first new pole (kind 1) of every block
    // of SEG_L sorted poles (cand is free again)
    const int nblk = (K + SEG_L - 1) / SEG_L;
    for (int i = tid; i < nblk; i += DEFL_NT) {
        int f = -1;
        for (int q = max(1, i * SEG_L); q < min(K, (i + 1) * SEG_L); ++q)
            if ((kind[q] & KIND_MASK) == 1) {
                f = q;
                break;
            }
        cand[i] = f;
    }
    __syncthreads();
    {
        const int bper = (nblk + DEFL_NT - 1) / DEFL_NT, b0 = min(nblk, tid * bper), b1 = min(nblk, b0 + bper);
        int cnt = 0;
        for (int b = b0; b < b1; ++b) cnt += cand[b] >= 0;
        int nseg;
        int sb = 1 + dc_block_scan<false>(cnt, sred, nseg);
        int *seg = islot(ws, n, V_SEG) + a;
        for (int b = b0; b < b1; ++b)
            if (cand[b] >= 0) seg[sb++] = cand[b];
        if (tid == 0) {
            seg[0] = 0;
            islot(ws, n, V_NSEG)[a] = nseg + 1;
        }
    }
}

struct GatherItem {
    int merge, s0, c0, pad;
};

// Gather: wavefront w of item (merge, s0, c0) takes segment s0 + w, lane l
// column c0 + l (the deflation rotations mix rows only, so columns are
// independent; segments are independent by construction).
__global__ __launch_bounds__(256) void k_dc_gather(const DcMerge *merges, const GatherItem *items) {
    const GatherItem it = items[blockIdx.x];
    const DcMerge mg = merges[it.merge];
    const int n = mg.n, a = mg.a, K = mg.b - mg.a;
    double *ws = mg.ws;
    const int nseg = islot(ws, n, V_NSEG)[a];
    const int sgi = it.s0 + (threadIdx.x >> 6), r = it.c0 + (threadIdx.x & 63);
    if (sgi >= nseg || r >= K) return;
    const int *seg = islot(ws, n, V_SEG) + a;
    const int q0 = max(0, seg[sgi]), q1 = min(K, sgi + 1 < nseg ? seg[sgi + 1] : K);
    dc_gather_seg(ws + (size_t)a * n + a, ws + (size_t)n * n + (size_t)a * n + a, n, mg.mid - a, q0, q1, r,
                  islot(ws, n, V_PERM) + a, islot(ws, n, V_POS) + a,
                  islot(ws, n, V_KIND) + a, vslot(ws, n, V_RC) + a, vslot(ws, n, V_RS) + a);
}

// Rows of the deflated poles (positions kp..K-1 of G) are final for this
// merge: copied back into Vc once the whole gather has read Vc.
__device__ inline void dc_copy_deflated(const DcMerge &mg, int j0) {
    const int n = mg.n, a = mg.a, K = mg.b - mg.a;
    double *ws = mg.ws;
    const int kp = islot(ws, n, V_META)[a];
    const int p0 = kp + j0, p1 = min(K, p0 + COPY_ROWS);
    const double *G = ws + (size_t)n * n + (size_t)a * n + a;
    double *Vc = ws + (size_t)a * n + a;
    for (int p = p0; p < p1; ++p)
        for (int c = threadIdx.x; c < K; c += blockDim.x) Vc[(size_t)p * n + c] = G[(size_t)p * n + c];
}

// Root j of 1/rho + sum_i zz_i^2 / (dd_i - lambda) = 0 (dd ascending,
// rho > 0), computed by one wavefront (lanes share the sums over i).  The
// root is found as tau = lambda - dd[org] where org is the nearer pole
// (decided at the middle of the interval), by a bracketed two-pole rational
// model: psi (poles <= j) and phi (poles > j) are each replaced by
// c + s / (pole - x) matching value and slope at the current point; the
// model's root (a quadratic) is taken if it lies inside the bracket,
// otherwise the bracket is bisected.
// W lanes share the sums over i: a wavefront (64; lane = 0..63) or a row of
// 16 lanes (16; lane = 0..15 within the row, every lane of the row active).
template <int W = 64>
__device__ inline double lanes_sum(double v) {
    if constexpr (W == 64) return wave_sum(v);
    else return row_sum(v);
}
template <int W = 64>
__device__ inline void secular_root(const double *dd, const double *zz, int K, double rho, int j, int lane,
                                    int &org_out, double &tau_out) {
    static_assert(W == 64 || W == 16, "a wavefront or a row of 16 lanes");
    const int i0 = lane, istep = W;
    const double irho = 1.0 / rho;
    if (K == 1) {
        org_out = 0;
        tau_out = rho * zz[0] * zz[0];
        return;
    }
    const bool last = (j == K - 1);
    double lo, hi;
    int org;
    if (!last) {
        const double dj = dd[j];
        const double mid = 0.5 * (dd[j + 1] - dj);
        double f = 0.0;
        for (int i = i0; i < K; i += istep) f += zz[i] * zz[i] / ((dd[i] - dj) - mid);
        f = irho + lanes_sum<W>(f);
        if (f >= 0.0) {
            org = j;
            lo = 0.0;
            hi = mid;
        } else {
            org = j + 1;
            lo = -mid;
            hi = 0.0;
        }
    } else {
        double zt = 0.0;
        for (int i = i0; i < K; i += istep) zt += zz[i] * zz[i];
        zt = lanes_sum<W>(zt);
        org = j;
        lo = 0.0;
        hi = rho * zt;
    }
    const double dorg = dd[org];
    const double dk = dd[j] - dorg;
    const double dk1 = last ? 0.0 : dd[j + 1] - dorg;
    double tau = 0.5 * (lo + hi);
    for (int it = 0; it < 64; ++it) {
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
        for (int i = i0; i < K; i += istep) {
            const double r = 1.0 / ((dd[i] - dorg) - tau);
            const double t = zz[i] * zz[i] * r;
            if (i <= j) {
                psi += t;
                dpsi += t * r;
            } else {
                phi += t;
                dphi += t * r;
            }
        }
        psi = lanes_sum<W>(psi);
        dpsi = lanes_sum<W>(dpsi);
        phi = lanes_sum<W>(phi);
        dphi = lanes_sum<W>(dphi);
        const double f = irho + psi + phi;
        const double err = 8.0 * (fabs(psi) + fabs(phi) + irho) + fabs(tau) * (dpsi + dphi);
        if (f < 0.0) lo = tau;
        else hi = tau;
        if (fabs(f) <= kEps * err || hi - lo <= 4.0 * kEps * fmax(fabs(lo), fabs(hi))) break;
        const double ek = dk - tau;
        const double bk = ek * ek * dpsi;
        double C = irho + psi - ek * dpsi;
        double nt;
        if (!last) {
            const double ek1 = dk1 - tau;
            const double e1 = ek1 * ek1 * dphi;
            C += phi - ek1 * dphi;
            // C (dk - x)(dk1 - x) + bk (dk1 - x) + e1 (dk - x) = 0, one of dk, dk1 is 0
            const double a2 = C;
            const double a1 = -(C * (dk + dk1) + bk + e1);
            const double a0 = bk * dk1 + e1 * dk;
            nt = lo;  // invalid unless set below
            if (a2 == 0.0) {
                if (a1 != 0.0) nt = -a0 / a1;
            } else {
                const double disc = fmax(a1 * a1 - 4.0 * a2 * a0, 0.0);
                const double q = -0.5 * (a1 + copysign(sqrt(disc), a1));
                const double r1 = q / a2;
                const double r2 = (q != 0.0) ? a0 / q : lo;
                nt = (r2 > lo && r2 < hi) ? r2 : r1;
            }
        } else {
            nt = (C > 0.0) ? bk / C : lo;
        }
        if (!(nt > lo && nt < hi)) nt = 0.5 * (lo + hi);
        tau = nt;
    }
    org_out = org;
    tau_out = tau;
}

// One wavefront per root: the root, then its row of Delta_ji = (dd_i -
// dd_org(j)) - tau_j.
// Workgroups past nsec copy deflated rows back (items of dc_copy_deflated).
__global__ __launch_bounds__(256) void k_dc_secular(const DcMerge *merges, const DcChunk *chunks, int nsec,
                                                    const DcChunk *copies) {
    if ((int)blockIdx.x >= nsec) {
        const DcChunk cp = copies[blockIdx.x - nsec];
        dc_copy_deflated(merges[cp.merge], cp.j0);
        return;
    }
    const DcChunk ck = chunks[blockIdx.x];
    const DcMerge mg = merges[ck.merge];
    const int n = mg.n, a = mg.a;
    double *ws = mg.ws;
    const int kp = islot(ws, n, V_META)[a];
    const int j = ck.j0 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (j >= kp) return;
    const double rho = vslot(ws, n, V_RHO)[a];
    const double *dd = vslot(ws, n, V_DD) + a, *zz = vslot(ws, n, V_ZZ) + a;
    int org;
    double tau;
    secular_root(dd, zz, kp, rho, j, lane, org, tau);
    const double dorg = dd[org];
    if (lane == 0) {
        vslot(ws, n, V_TS)[a + j] = tau;
        islot(ws, n, V_ORG)[a + j] = org;
        vslot(ws, n, V_D)[a + j] = dorg + tau;
    }
    double *row = ws + 2 * (size_t)n * n + (size_t)(a + j) * n + a;
    for (int i = lane; i < kp; i += 64) row[i] = (dd[i] - dorg) - tau;
}

// z-hat_i^2 = -Delta_ii prod_{j != i} Delta_ji / (dd_i - dd_j) / rho: lane l
// takes i = j0 + l (loads of a row of Delta coalesced over i), wavefront w
// the factors j = w mod ZHAT_NW; the partial products meet in LDS.
constexpr int ZHAT_NW = 16;
__global__ __launch_bounds__(64 * ZHAT_NW) void k_dc_zhat(const DcMerge *merges, const DcChunk *chunks) {
    __shared__ double part[ZHAT_NW][64];
    const DcChunk ck = chunks[blockIdx.x];
    const DcMerge mg = merges[ck.merge];
    const int n = mg.n, a = mg.a;
    double *ws = mg.ws;
    const int kp = islot(ws, n, V_META)[a];
    if (ck.j0 >= kp) return;  // the whole workgroup
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = ck.j0 + lane;
    const double *dd = vslot(ws, n, V_DD) + a;
    const double *S2 = ws + 2 * (size_t)n * n + (size_t)a * n + a;
    double w = 1.0;
    if (i < kp) {
        const double di = dd[i];
        for (int j = wave; j < kp; j += ZHAT_NW) w *= j == i ? S2[(size_t)i * n + i] : S2[(size_t)j * n + i] / (di - dd[j]);
    }
    part[wave][lane] = w;
    __syncthreads();
    if (wave == 0 && i < kp) {
        double p = part[0][lane];
        for (int g = 1; g < ZHAT_NW; ++g) p *= part[g][lane];
        const double rho = vslot(ws, n, V_RHO)[a];
        vslot(ws, n, V_ZH)[a + i] = copysign(sqrt(fmax(-p / rho, 0.0)), vslot(ws, n, V_ZZ)[a + i]);
    }
}

// Row j of S2 <- z-hat / Delta_j, normalised; one wavefront per row.
__global__ __launch_bounds__(256) void k_dc_unorm(const DcMerge *merges, const DcChunk *chunks) {
    const DcChunk ck = chunks[blockIdx.x];
    const DcMerge mg = merges[ck.merge];
    const int n = mg.n, a = mg.a;
    double *ws = mg.ws;
    const int kp = islot(ws, n, V_META)[a];
    const int j = ck.j0 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (j >= kp) return;
    const double *zh = vslot(ws, n, V_ZH) + a;
    double *row = ws + 2 * (size_t)n * n + (size_t)(a + j) * n + a;
    double s = 0.0;
    for (int i = lane; i < kp; i += 64) {
        const double u = zh[i] / row[i];
        s += u * u;
    }
    s = wave_sum(s);
    const double inv = 1.0 / sqrt(s);
    for (int i = lane; i < kp; i += 64) row[i] = zh[i] / row[i] * inv;
}

// ---------------------------------------------------------------------------
// 2b. the lowest levels in one launch.  One workgroup per subtree of the
//     tree at level DC_SMALL_LEV (2^5 leaves of 1 or 2 rows: S <= 64 rows)
//     runs levels 1..5 of that subtree in LDS, with block barriers between
//     the stages instead of one launch per stage and level, and writes the
//     subtree's eigenvalues and S x S eigenvector block back at the end.
//     Every merge has K <= 64 poles, so one wavefront deflates a merge with
//     lane q holding sorted pole q.  The arithmetic is that of the
//     level-by-level kernels: sort by (value, index), z, the dlaed2
//     deflation rule (parallel tests, then the candidates and rotation
//     chains in order), the gather with the rotations, secular roots,
//     z-hat, normalisation and the merge product U G.
// ---------------------------------------------------------------------------
constexpr int DC_SMALL_LEV = 5;
constexpr int DC_SMALL_S = 64;   // rows of a subtree, at most
constexpr int DC_SMALL_LD = 65;  // LDS row stride of the blocks (doubles)
constexpr int DC_SMALL_NT = 1024, DC_SMALL_NW = DC_SMALL_NT / 64;

struct DcSmall {
    double V[DC_SMALL_S * DC_SMALL_LD];  // eigenvectors (rows), block-diagonal by merge
    double G[DC_SMALL_S * DC_SMALL_LD];  // gathered rows of the current merges
    double U[DC_SMALL_S * DC_SMALL_LD];  // Delta, then the eigenvectors of D + rho z z^T (rows)
    double D[DC_SMALL_S], DD[DC_SMALL_S], ZZ[DC_SMALL_S], ZH[DC_SMALL_S];
    double RC[DC_SMALL_S], RS[DC_SMALL_S], BETA[DC_SMALL_S], RHO[DC_SMALL_S];
    int PERM[DC_SMALL_S], POS[DC_SMALL_S], KIND[DC_SMALL_S], ROWM[DC_SMALL_S];
    int MA[DC_SMALL_S], MB[DC_SMALL_S], KP[DC_SMALL_S];
};

// local bounds of merge m of level L of subtree k0 (at level ls) starting at
// global row a0
__device__ inline void dc_small_bounds(int n, int nlev, int ls, int k0, int a0, int L, int m, int &a, int &mid,
                                       int &b) {
    const int k = (k0 << (ls - L)) + m;
    a = dc_bound(k << L, n, nlev) - a0;
    mid = dc_bound((k << L) + (1 << (L - 1)), n, nlev) - a0;
    b = dc_bound((k + 1) << L, n, nlev) - a0;
}

// Deflation of local merge m = [a, b) (K = b - a <= 64) by one wavefront.
__device__ inline void dc_small_deflate(DcSmall &sm, int m, int a, int mid, int b, int lane) {
    const int K = b - a, h = mid - a;
    const double beta = sm.BETA[mid];
    const double rho = 2.0 * fabs(beta);
    const double sg = beta >= 0.0 ? 1.0 : -1.0;
    const double rs2 = 0.70710678118654752440;
    const bool act = lane < K;
    const double di = act ? sm.D[a + lane] : 0.0;
    // rank of pole lane in the order (value, index); pm: the pole of rank lane
    int rk = lane;
    if (act) {
        rk = 0;
        for (int j = 0; j < K; ++j) {
            const double dj = readlane_d(di, j);
            rk += (dj < di) || (dj == di && j < lane);
        }
    }
    int pm = lane;
    for (int j = 0; j < K; ++j)
        if (__builtin_amdgcn_readlane(rk, j) == lane) pm = j;
    double ds = __shfl(di, pm, 64), zs = 0.0;
    if (act) {
        const double *row = sm.V + (a + pm) * DC_SMALL_LD;
        zs = pm < h ? row[mid - 1] * rs2 : sg * row[mid] * rs2;
    }
    const double md = wave_max(fabs(ds)), mz = wave_max(fabs(zs));
    const double tol = 8.0 * kEps * fmax(md, rho * mz);
    const bool large = act && !(rho * fabs(zs) <= tol);
    const unsigned long long Lm = __ballot(large);
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long Lb = Lm & below;
    const int prev = Lb ? 63 - __builtin_clzll(Lb) : -1;
    // tests with the unrotated values
    const int src = prev < 0 ? lane : prev;
    const double zp = __shfl(zs, src, 64), dp = __shfl(ds, src, 64);
    bool cand = false;
    if (large && prev >= 0) {
        double c, s, tau;
        cand = dc_close(zp, zs, dp, ds, tol, c, s, tau);
    }
    unsigned long long Cm = __ballot(cand);
    int kind = large ? 1 : 0;
    double rcv = 0.0, rsv = 0.0;
    int chain = -1;
    for (;;) {  // wave-uniform walk over the candidates and rotation chains
        int q;
        if (chain >= 0) {
            q = chain;
            chain = -1;
            Cm &= ~((2ull << q) - 1ull);  // candidates up to q are settled
        } else if (Cm) {
            q = __builtin_ctzll(Cm);
            Cm &= Cm - 1ull;
        } else {
            break;
        }
        const int p = __builtin_amdgcn_readlane(prev, q);
        const double zpq = readlane_d(zs, p), zqq = readlane_d(zs, q);
        const double dpq = readlane_d(ds, p), dqq = readlane_d(ds, q);
        double c, s, tau;
        if (dc_close(zpq, zqq, dpq, dqq, tol, c, s, tau)) {
            const double tp = dpq * c * c + dqq * s * s, tq = dpq * s * s + dqq * c * c;
            if (lane == q) {
                zs = tau;
                ds = tq;
                kind = 2;
                rcv = c;
                rsv = s;
            }
            if (lane == p) {
                zs = 0.0;
                ds = tp;
                kind |= KIND_SRC;
            }
            // the next large-z pole is tested against the rotated values
            const unsigned long long above = Lm & ~((2ull << q) - 1ull);
            if (above) chain = __builtin_ctzll(above);
        }
    }
    // positions: undeflated poles first, then the deflated ones, both in
    // sorted order
    const bool und = (kind & KIND_MASK) != 0 && !(kind & KIND_SRC);
    const unsigned long long Um = __ballot(und);
    const int kp = __popcll(Um), ub = __popcll(Um & below);
    const int ps = und ? ub : kp + (lane - ub);
    if (act) {
        sm.PERM[a + lane] = pm;
        sm.POS[a + lane] = ps;
        sm.KIND[a + lane] = kind;
        sm.RC[a + lane] = rcv;
        sm.RS[a + lane] = rsv;
        sm.ROWM[a + lane] = m;
        if (ps < kp) {
            sm.DD[a + ps] = ds;
            sm.ZZ[a + ps] = zs;
        } else {
            sm.D[a + ps] = ds;  // deflated eigenvalue, final for this merge
        }
    }
    if (lane == 0) {
        sm.MA[m] = a;
        sm.MB[m] = b;
        sm.KP[m] = kp;
        sm.RHO[m] = rho;
    }
}

__global__ __launch_bounds__(DC_SMALL_NT) void k_dc_small(const EigMat *mats, const BtItem *items) {
    extern __shared__ double lds[];
    DcSmall &sm = *reinterpret_cast<DcSmall *>(lds);
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n, nlev = dc_levels(n), ls = min(DC_SMALL_LEV, nlev), k0 = it.c0;
    double *ws = mt.ws;
    double *Vg = ws, *Dg = vslot(ws, n, V_D);
    const double *betag = vslot(ws, n, V_BETA);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int a0 = dc_bound(k0 << ls, n, nlev), S = dc_bound((k0 + 1) << ls, n, nlev) - a0;
    constexpr int LD = DC_SMALL_LD;
    // the leaves' eigenpairs (1x1 and 2x2 blocks on the diagonal)
    for (int e = tid; e < S * S; e += DC_SMALL_NT) {
        const int r = e / S, c = e - r * S;
        sm.V[r * LD + c] = 0.0;
    }
    __syncthreads();
    for (int l = tid; l < (1 << ls); l += DC_SMALL_NT) {
        const int la = dc_bound((k0 << ls) + l, n, nlev), lb = dc_bound((k0 << ls) + l + 1, n, nlev);
        for (int r = la; r < lb; ++r)
            for (int c = la; c < lb; ++c) sm.V[(r - a0) * LD + (c - a0)] = Vg[(size_t)r * n + c];
    }
    for (int i = tid; i < S; i += DC_SMALL_NT) {
        sm.D[i] = Dg[a0 + i];
        sm.BETA[i] = betag[a0 + i];
    }
    __syncthreads();
    for (int L = 1; L <= ls; ++L) {
        const int nm = 1 << (ls - L);
        int a, mid, b;
        // deflation, then the gather (lane = column) by the same wavefront,
        // one wavefront per merge (LDS operations of a wavefront are in order)
        for (int m = wave; m < nm; m += DC_SMALL_NW) {
            dc_small_bounds(n, nlev, ls, k0, a0, L, m, a, mid, b);
            dc_small_deflate(sm, m, a, mid, b, lane);
            const int K = b - a;
            if (lane < K)
                dc_gather_seg(sm.V + a * LD + a, sm.G + a * LD + a, LD, mid - a, 0, K, lane, sm.PERM + a,
                              sm.POS + a, sm.KIND + a, sm.RC + a, sm.RS + a);
        }
        __syncthreads();
        // root slot = row of 16 lanes (64 slots, and at most 64 roots per
        // level): slot t takes root t of the level, merges in order
        const int slot = tid >> 4, sl = tid & 15;
        int mm = -1, jr = 0;
        {
            int base = 0;
            for (int m = 0; m < nm; ++m) {
                const int kp = sm.KP[m];
                if (slot >= base && slot < base + kp) {
                    mm = m;
                    jr = slot - base;
                }
                base += kp;
            }
        }
        const int ra = mm >= 0 ? sm.MA[mm] : 0, rkp = mm >= 0 ? sm.KP[mm] : 0;
        const double *rdd = sm.DD + ra;
        // secular root jr, and row jr of Delta
        if (mm >= 0) {
            int org;
            double tau;
            secular_root<16>(rdd, sm.ZZ + ra, rkp, sm.RHO[mm], jr, sl, org, tau);
            const double dorg = rdd[org];
            if (sl == 0) sm.D[ra + jr] = dorg + tau;
            double *row = sm.U + (ra + jr) * LD + ra;
            for (int i = sl; i < rkp; i += 16) row[i] = (rdd[i] - dorg) - tau;
        }
        __syncthreads();
        // z-hat_i^2 = -Delta_ii prod_{j != i} Delta_ji / (dd_i - dd_j) / rho, i = jr
        if (mm >= 0) {
            const double *Dl = sm.U + ra * LD + ra;
            const double di = rdd[jr];
            double w = 1.0;
            for (int j = sl; j < rkp; j += 16)
                w *= j == jr ? Dl[jr * LD + jr] : Dl[j * LD + jr] / (di - rdd[j]);
            w = row_prod(w);
            if (sl == 0) sm.ZH[ra + jr] = copysign(sqrt(fmax(-w / sm.RHO[mm], 0.0)), sm.ZZ[ra + jr]);
        }
        __syncthreads();
        // row jr of U = z-hat / Delta, normalised
        if (mm >= 0) {
            const double *zh = sm.ZH + ra;
            double *row = sm.U + (ra + jr) * LD + ra;
            double x[4], ss = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = sl + 16 * q;
                x[q] = i < rkp ? zh[i] / row[i] : 0.0;
                ss += x[q] * x[q];
            }
            const double inv = 1.0 / sqrt(row_sum(ss));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = sl + 16 * q;
                if (i < rkp) row[i] = x[q] * inv;
            }
        }
        __syncthreads();
        // merge product: row r of merge [a, b) <- (U G)_r for r - a < K',
        // the gathered row (a deflated pole) otherwise; 16 threads per row
        {
            const int r = tid >> 4, c0 = tid & 15;
            if (r < S) {
                const int m = sm.ROWM[r];
                a = sm.MA[m];
                const int K = sm.MB[m] - a, kp = sm.KP[m], j = r - a;
                int col[4];  // columns of this thread, clamped into the merge
#pragma unroll
                for (int q = 0; q < 4; ++q) col[q] = a + min(c0 + 16 * q, K - 1);
                double acc[4] = {0.0, 0.0, 0.0, 0.0};
                if (j < kp) {
                    const double *u = sm.U + r * LD + a;
                    const double *g = sm.G + a * LD;
                    for (int i = 0; i < kp; ++i) {
                        const double ui = u[i];
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[q] += ui * g[i * LD + col[q]];
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = sm.G[r * LD + col[q]];
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (c0 + 16 * q < K) sm.V[r * LD + col[q]] = acc[q];
            }
        }
        __syncthreads();
    }
    // eigenvalues and the S x S eigenvector block of the subtree
    for (int e = tid; e < S * S; e += DC_SMALL_NT) {
        const int r = e / S, c = e - r * S;
        Vg[(size_t)(a0 + r) * n + a0 + c] = sm.V[r * LD + c];
    }
    for (int i = tid; i < S; i += DC_SMALL_NT) Dg[a0 + i] = sm.D[i];
}

// Stable ascending rank of the final eigenvalues; W[rank] = D.  Lane l takes
// eigenvalue c0 + l, wavefront w counts over j = w mod RANK_NW; the partial
// counts meet in LDS.
constexpr int RANK_NW = 16;
__global__ __launch_bounds__(64 * RANK_NW) void k_dc_rank(const EigMat *mats, const BtItem *items) {
    __shared__ int part[RANK_NW][64];
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = it.c0 + lane;
    const double *D = vslot(mt.ws, n, V_D);
    const double di = D[min(i, n - 1)];
    int rk = 0;
    for (int j = wave; j < n; j += RANK_NW) {
        const double dj = D[j];
        rk += (dj < di) || (dj == di && j < i);
    }
    part[wave][lane] = rk;
    __syncthreads();
    if (wave == 0 && i < n) {
        int r = 0;
        for (int g = 0; g < RANK_NW; ++g) r += part[g][lane];
        islot(mt.ws, n, V_RANK)[i] = r;
        mt.W[r] = di;
    }
}

// ---------------------------------------------------------------------------
// 3. back-transformation: x_c <- H_0 H_1 ... H_{n-3} x_c for every eigenvector
//    x_c of the tridiagonal (row c of Vc); one wavefront per eigenvector, the
//    vector in registers (NS x 64 elements), reflectors read from A.
//    Output: row rank(c) of S1.
// ---------------------------------------------------------------------------
// Calls f(integral_constant<P>) with P the largest multiple of PS not above
// the run-time p (0 <= p < NS).
template <int P, int PS, int NS, class F>
__device__ __forceinline__ void bt_dispatch(int p, F &f) {
    if constexpr (P + PS >= NS) {
        f(std::integral_constant<int, P>{});
    } else {
        if (p < P + PS) f(std::integral_constant<int, P>{});
        else bt_dispatch<P + PS, PS, NS>(p, f);
    }
}

// BT_NT threads: eight eigenvectors per workgroup, one wavefront each; the
// reflectors are staged BT_RB at a time in LDS (read once per workgroup
// instead of once per eigenvector) and applied from there.
constexpr int BT_NT = 512, BT_RB = 8;
template <int NS>
__global__ __launch_bounds__(BT_NT) void k_eig_bt(const EigMat *mats, const BtItem *items) {
    __shared__ double refl[BT_RB][64 * NS];
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n;
    const int c = it.c0 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const bool act = c < n;  // the same in the whole wavefront
    double *ws = mt.ws;
    const double *tau = vslot(ws, n, V_TAU);
    const double *A = mt.A;
    double x[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int r = 64 * k + lane;
        x[k] = (act && r < n) ? ws[(size_t)c * n + r] : 0.0;
    }
    // batches of reflectors j0, j0-1, .., j0-nb+1; the next batch is loaded
    // into registers while the current one is applied
    constexpr int LQ = BT_RB * 64 * NS / BT_NT;
    static_assert(LQ * BT_NT == BT_RB * 64 * NS, "whole batches per thread");
    double nxt[LQ];
    auto load_batch = [&](int j0) {
        const int nb = min(BT_RB, j0 + 1);
#pragma unroll
        for (int q = 0; q < LQ; ++q) {
            const int e = threadIdx.x + q * BT_NT;
            const int b = e / (64 * NS), r = e - b * 64 * NS, j = j0 - b;
            const bool ok = b < nb && r > j && r < n;
            const double v = A[ok ? (size_t)j * n + r : 0];
            nxt[q] = ok ? v : 0.0;
        }
    };
    if (n >= 3) load_batch(n - 3);
    for (int j0 = n - 3; j0 >= 0; j0 -= BT_RB) {
        const int nb = min(BT_RB, j0 + 1);  // reflectors j0, j0-1, .., j0-nb+1
        __syncthreads();                    // the previous batch has been applied
#pragma unroll
        for (int q = 0; q < LQ; ++q) {
            const int e = threadIdx.x + q * BT_NT;
            refl[e / (64 * NS)][e % (64 * NS)] = nxt[q];
        }
        __syncthreads();
        if (j0 - BT_RB >= 0) load_batch(j0 - BT_RB);
        if (!act) continue;
        // reflectors j >= j0 - nb + 1 are zero in rows <= j: registers below
        // P are untouched (their products are exact zeros), so their work
        // is skipped
        auto apply = [&](auto Pc) {
            constexpr int P = decltype(Pc)::value;
            for (int b = 0; b < nb; ++b) {
                const double t = tau[j0 - b];
                if (t == 0.0) continue;
                double vr[NS];
                double dot = 0.0;
#pragma unroll
                for (int k = P; k < NS; ++k) {
                    vr[k] = refl[b][64 * k + lane];
                    dot += vr[k] * x[k];
                }
                dot = wave_sum(dot) * t;
#pragma unroll
                for (int k = P; k < NS; ++k) x[k] -= dot * vr[k];
            }
            return true;
        };
        bt_dispatch<0, (NS + 3) / 4, NS>((j0 - nb + 2) >> 6, apply);
    }
    if (!act) return;
    const int rk = islot(ws, n, V_RANK)[c];
    double *out = ws + (size_t)n * n + (size_t)rk * n;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int r = 64 * k + lane;
        if (r < n) out[r] = x[k];
    }
}

// ---------------------------------------------------------------------------
// 3b. blocked back-transformation for larger matrices (compact WY, GEMMs):
//     with Q = H_0 H_1 ... H_{n-3}, the eigenvectors (rows of Z) become
//     Z Q^T = Z H_{n-3} ... H_0.  A block of NB reflectors [j0, j1) is
//     H_{j0} ... H_{j1-1} = I - Y T Y^T (forward WY, T upper triangular), so
//     Z <- Z (I - Y T^T Y^T) = Z - (Z M^T) Y^T with M = T Y^T; blocks are
//     applied from the last to the first.  Y^T (zero-padded reflector rows)
//     lives in S1; G = Y^T Y, T, M and W = Z M^T in S2.
// ---------------------------------------------------------------------------
constexpr int WY_NB = 128;
constexpr int WY_BB = 2 * WY_NB;  // reflectors per compact-WY block, at least: T from chunks of WY_NB
constexpr int WY_BB_MAX = 4 * WY_NB;  // and at most, where the workspace holds its scratch
// reflectors per WY block of a matrix of dimension n: the largest of
// WY_BB_MAX, WY_BB whose G, T, M and W (2 B (B + n) doubles) fit S2 (n^2)
inline int wy_bb(int n) {
    return 2 * (size_t)WY_BB_MAX * (WY_BB_MAX + n) <= (size_t)n * n ? WY_BB_MAX : WY_BB;
}
constexpr int BT_WY_MIN = 1025;   // register kernel below, WY from here
constexpr int EIG_MAX_N = 32768;

struct WyItem {
    double *ws;
    int n, j0, nb, pad;
    int go;  // T factor of one chunk: offset of the chunk in its block's G and T
    int ld;  // reflectors per block of the matrix: leading dimension of G and T
};

// Y^T: row j = reflector j (columns j+1.., first element included), zero
// elsewhere; rows n-2, n-1 are zero.  Two-stage matrices: reflector j of
// stage 1 at columns j + SBR_B.. (its unit first element written here).
__global__ __launch_bounds__(256) void k_wy_extract(const EigMat *mats, const BtItem *items) {
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n;
    double *Yt = mt.ws + (size_t)n * n;
    const int j = it.c0;
    const double *src = mt.A + (size_t)j * n;
    double *dst = Yt + (size_t)j * n;
    if (mt.pad == 2) {  // two-stage: v_j at rows j + SBR_B.. (unit first element)
        const bool on = j < sbr_nrefl(n);
        for (int r = threadIdx.x; r < n; r += 256)
            dst[r] = (on && r >= j + SBR_B) ? (r == j + SBR_B ? 1.0 : src[r]) : 0.0;
        return;
    }
    for (int r = threadIdx.x; r < n; r += 256) dst[r] = (j <= n - 3 && r > j) ? src[r] : 0.0;
}

// T of the block (dlarft forward / columnwise): T_kk = tau_k,
// T[0:k, k] = -tau_k T[0:k, 0:k] G[0:k, k], with G = Y^T Y of the block.
// Row i of T needs only its own earlier entries and G, so the four lanes
// 4 i .. 4 i + 3 of one wavefront form row i alone (lane q sums the terms
// l = q mod 4, the quad adds its four partial sums; no barriers).  LDS holds
// the strict upper triangle of G by columns (column k: G[0:k, k] at
// k (k - 1) / 2) and the rows of T packed (row i: T[i, i:] at
// i NB - i (i - 1) / 2): 2 x 8 K doubles for NB = 128.
constexpr int WY_TRI = WY_NB * (WY_NB + 1) / 2;
constexpr int WY_TF_NT = 4 * WY_NB;
__global__ __launch_bounds__(WY_TF_NT) void k_wy_tfactor(const WyItem *items) {
    __shared__ double sG[WY_TRI];
    __shared__ double sT[WY_TRI];
    const WyItem it = items[blockIdx.x];
    const int n = it.n, nb = it.nb, tid = threadIdx.x, i = tid >> 2, q = tid & 3;
    double *S2 = it.ws + 2 * (size_t)n * n;
    // this chunk's diagonal blocks of the block's G and T (ld x ld)
    const int ld = it.ld;
    const double *G = S2 + (size_t)it.go * (ld + 1);
    double *T = S2 + (size_t)ld * ld + (size_t)it.go * (ld + 1);
    const double *tau = vslot(it.ws, n, V_TAU) + it.j0;
    // G's strict upper triangle, column k from row 0: loads issued before
    // the LDS stores (a load and its store in one iteration serialise)
    constexpr int GB = (WY_TRI + WY_TF_NT - 1) / WY_TF_NT;
    {
        double g[GB];
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            const int e = u * WY_TF_NT + tid;
            // column k holds entries k (k - 1) / 2 .. k (k + 1) / 2 - 1
            int k = (int)((1.0 + sqrt(1.0 + 8.0 * (double)e)) * 0.5);
            while (k * (k - 1) / 2 > e) --k;
            while ((k + 1) * k / 2 <= e) ++k;
            const int l = e - k * (k - 1) / 2;
            const double v = G[(size_t)min(l, WY_NB - 1) * ld + min(k, WY_NB - 1)];
            g[u] = (e < WY_TRI && k < nb) ? v : 0.0;
        }
#pragma unroll
        for (int u = 0; u < GB; ++u) {
            const int e = u * WY_TF_NT + tid;
            if (e < WY_TRI) sG[e] = g[u];
        }
    }
    const int ri = i * WY_NB - i * (i - 1) / 2 - i;  // T[i, l] at ri + l
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        const double *gk = sG + k * (k - 1) / 2;
        double t0 = 0.0, t1 = 0.0;
        int l = q;
        for (; l + 4 < k; l += 8) {
            t0 += l >= i ? sT[ri + max(l, i)] * gk[l] : 0.0;
            t1 += l + 4 >= i ? sT[ri + max(l + 4, i)] * gk[l + 4] : 0.0;
        }
        if (l < k) t0 += l >= i ? sT[ri + max(l, i)] * gk[l] : 0.0;
        double t = t0 + t1;
        t += dpp_d<0xB1>(t);  // quad_perm [1, 0, 3, 2]
        t += dpp_d<0x4E>(t);  // quad_perm [2, 3, 0, 1]
        if (q == 0) {
            const double tk = tau[k];
            if (i < k) sT[ri + k] = -tk * t;
            if (i == k) sT[ri + k] = tk;
        }
    }
    if (i < nb) {
        for (int k = q; k < nb; k += 4) T[(size_t)i * ld + k] = k >= i ? sT[ri + k] : 0.0;
        // a later chunk's rows of T are zero left of its diagonal block
        for (int k = q; k < it.go; k += 4) T[(size_t)i * ld + k - it.go] = 0.0;
    }
}

// Row c of Z (the back-transformed eigenvector c) -> row rank(c) of A.
__global__ __launch_bounds__(256) void k_eig_copy_perm(const EigMat *mats, const BtItem *items) {
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n, c = it.c0;
    const int rk = islot(mt.ws, n, V_RANK)[c];
    const double *src = mt.ws + (size_t)c * n;
    double *dst = mt.A + (size_t)rk * n;
    for (int r = threadIdx.x; r < n; r += 256) dst[r] = src[r];
}

__global__ __launch_bounds__(256) void k_eig_copy(const EigMat *mats, const BtItem *items) {
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const size_t nn = (size_t)mt.n * mt.n;
    const double *src = mt.ws + nn;
    const size_t e0 = (size_t)it.c0 * 256 * 16;
    const size_t e1 = std::min(nn, e0 + 256 * 16);
    for (size_t e = e0 + threadIdx.x; e < e1; e += 256) mt.A[e] = src[e];
}

constexpr int kBtSlots[] = {1, 2, 4, 8, 12, 16};  // 64 x 16 columns >= BT_WY_MIN - 1
constexpr int kBtVariants = sizeof(kBtSlots) / sizeof(kBtSlots[0]);
static_assert(64 * kBtSlots[kBtVariants - 1] >= BT_WY_MIN - 1, "register back-transformation covers every n below WY");

template <int NS>
static void launch_bt(hipStream_t s, int nitems, const EigMat *mats, const BtItem *items) {
    hipLaunchKernelGGL(k_eig_bt<NS>, dim3(nitems), dim3(BT_NT), 0, s, mats, items);
}
static void launch_bt_variant(int v, hipStream_t s, int nitems, const EigMat *mats, const BtItem *items) {
    switch (v) {
        case 0: launch_bt<1>(s, nitems, mats, items); break;
        case 1: launch_bt<2>(s, nitems, mats, items); break;
        case 2: launch_bt<4>(s, nitems, mats, items); break;
        case 3: launch_bt<8>(s, nitems, mats, items); break;
        case 4: launch_bt<12>(s, nitems, mats, items); break;
        default: launch_bt<16>(s, nitems, mats, items); break;
    }
}

// descriptor arena assembled on the host and staged in one copy
struct HostArena {
    std::vector<char> buf;
    template <class T>
    size_t put(const std::vector<T> &v) {
        size_t off = (buf.size() + 255) & ~(size_t)255;
        buf.resize(off + sizeof(T) * v.size());
        if (!v.empty()) memcpy(buf.data() + off, v.data(), sizeof(T) * v.size());
        return off;
    }
};

extern "C" nrg_status nrg_diag_batch(nrg_ctx *c, const nrg_eig_task *tasks_in, int ntasks) {
    if (ntasks <= 0) return NRG_OK;
    // largest first: the LDS-sized tail of the tridiagonalisation is contiguous
    std::vector<int> by_size(ntasks);
    for (int i = 0; i < ntasks; ++i) by_size[i] = i;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) { return tasks_in[a].n > tasks_in[b].n; });
    std::vector<nrg_eig_task> tasks(ntasks);
    for (int i = 0; i < ntasks; ++i) tasks[i] = tasks_in[by_size[i]];
    // eigenvalues for nrg_diag_values, concatenated in the caller's task order
    std::vector<long long> voff_in(ntasks + 1, 0);
    for (int i = 0; i < ntasks; ++i) voff_in[i + 1] = voff_in[i] + std::max(0, (int)tasks_in[i].n);
    std::vector<long long> voff(ntasks);
    for (int i = 0; i < ntasks; ++i) voff[i] = voff_in[by_size[i]];
    const size_t nvals = (size_t)voff_in[ntasks];
    if (nvals > c->hval_doubles) {  // (no queued copy may still write the old buffer)
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->hval) HIPCHK(c, hipHostFree(c->hval));
        c->hval = nullptr;
        const size_t nd = std::max<size_t>(nvals * 2, 1 << 14);
        HIPCHK(c, hipHostMalloc((void **)&c->hval, nd * sizeof(double), hipHostMallocDefault));
        c->hval_doubles = nd;
    }
    size_t tot = 0;
    std::vector<size_t> off(ntasks);
    for (int i = 0; i < ntasks; ++i) {
        if (tasks[i].n <= 0 || !tasks[i].A || !tasks[i].W) return fail(c, NRG_ERR_ARG, "diag: bad task");
        if (tasks[i].n > EIG_MAX_N)
            return fail(c, NRG_ERR_ARG, "diag: matrix dimension above " + std::to_string(EIG_MAX_N));
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
    // device pointers inside the staged arena are fixed up after staging, so
    // descriptors hold workspace pointers directly
    std::vector<EigMat> mats(ntasks);
    for (int i = 0; i < ntasks; ++i) mats[i] = {tasks[i].A, tasks[i].W, c->ews + off[i], tasks[i].n, 0};
    int nbig = 0;  // matrices with n >= 2 form the prefix
    while (nbig < ntasks && tasks[nbig].n >= 2) ++nbig;

    // tridiagonalisation jobs: ticket ranges, team sizes
    if (c->tri_capacity <= 0) {
        int per_cu = 0, dev = 0, ncu = 0;
        HIPCHK(c, hipGetDevice(&dev));
        HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(c, hipFuncSetAttribute((const void *)k_eig_tri_coop, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)TRI_COOP_LDS));
        if (const char *e = getenv("NRG_BLK_MIN_N")) c->blk_min_n = std::max(3, atoi(e));
        if (const char *e = getenv("NRG_BLK_TDIV")) c->blk_tdiv = std::max(0, atoi(e));
        if (const char *e = getenv("NRG_SBR")) c->sbr = atoi(e) != 0;
        {
            int per_cu_ch = 0;
            HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_ch, k_sbr_chase, CH_NT, 0));
            c->chase_capacity = std::max(1, per_cu_ch * ncu);
        }
        HIPCHK(c, hipFuncSetAttribute((const void *)k_eig_tri_blk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)TRI_COOP_LDS));
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_eig_tri_coop, TRI_NT, TRI_COOP_LDS));
        {
            int per_cu_blk = 0;
            HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_blk, k_eig_tri_blk, TRI_NT, TRI_COOP_LDS));
            per_cu = std::min(per_cu, per_cu_blk);
        }
        c->tri_capacity = std::max(1, per_cu * ncu);
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_eig_tri_reg, REG_NT, 0));
        // at most one team workgroup per CU: two would share its SIMDs
        c->reg_capacity = std::max(1, std::min(per_cu, 1) * ncu);
        HIPCHK(c, hipFuncSetAttribute((const void *)k_dc_small, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)sizeof(DcSmall)));
    }
    // Team sizes, then longest-processing-time assignment of the jobs to the
    // XCD queues (work = estimated time x workgroups; a team stays on one XCD)
    const int tri_grid = c->tri_capacity;
    const int per_xcd = std::max(1, tri_grid / N_XCD);
    const int reg_grid = c->reg_capacity, reg_per_xcd = std::max(1, reg_grid / N_XCD);
    std::vector<int> teamT(ntasks, 1);
    std::vector<double> cost(ntasks);
    std::vector<char> global(ntasks, 0);
    // chains of two-stage jobs on one team: the next job of a task, and
    // whether a task is a chain's continuation (not a queue job)
    std::vector<int> chain_next(ntasks, -1);
    std::vector<char> chained(ntasks, 0);
    std::vector<char> regjob(ntasks, 0), blkjob(ntasks, 0);
    const int blk_min_n = c->blk_min_n, blk_tdiv = c->blk_tdiv;
    for (int i = 0; i < ntasks; ++i) {
        const double n = tasks[i].n;
        if (tasks[i].n <= REG_ONE_N) {  // one workgroup on the side stream, not queued
            cost[i] = 0.5 * n;
            continue;
        }
        // microseconds per step: a register-resident team pays its exchange;
        // a streaming team a barrier plus its share of the matrix
        if (tasks[i].n <= REG_MAX_N && reg_team(tasks[i].n) <= reg_per_xcd) {
            teamT[i] = reg_team(tasks[i].n);
            regjob[i] = 1;
            cost[i] = 2.0 * n;
            continue;
        }
        // (S2 also holds the compact-WY scratch of the back-transformation:
        // G and T of a block, M = T Y^T and W, 2 WY_BB (WY_BB + n) doubles)
        if (tasks[i].n >= blk_min_n && c->sbr && sbr_ws_doubles(tasks[i].n) <= (size_t)tasks[i].n * tasks[i].n &&
            (size_t)tasks[i].n * (CH_W + 1) + 64 <= (size_t)tasks[i].n * tasks[i].n &&
            2 * (size_t)WY_BB * (WY_BB + tasks[i].n) <= (size_t)tasks[i].n * tasks[i].n) {
            blkjob[i] = 2;  // two-stage; team size below, once all blocked jobs are known
            mats[i].pad = 2;
            continue;
        }
        if (tasks[i].n >= blk_min_n && blk_ws_doubles(tasks[i].n) <= (size_t)tasks[i].n * tasks[i].n) {
            blkjob[i] = 1;  // team size below, once all blocked jobs are known
            continue;
        }
        const int nsl = (tasks[i].n - 1 + TEAM_W - 1) / TEAM_W;
        int T;
        if (tasks[i].n > TEAM_GLOBAL_N) {
            T = std::min({nsl, TEAM_GLOBAL_MAX, tri_grid});
            global[i] = 1;
        } else {
            T = std::min({nsl, TEAM_MAX, per_xcd});
        }
        teamT[i] = std::max(T, 2);
        cost[i] = n * (3.0 + 1.1e-4 * n * n / teamT[i]);
        mats[i].pad = 1;  // reflectors to copy from S2
    }
    // Blocked reductions: a team of T takes about o(n) + W(n) / T
    // microseconds: o(n) = n (0.0114 n - 25) of per-column latency that no
    // team size removes (barriers, block sums), and W(n) = 0.266 n^2 +
    // 7e-5 n^3 CU-us of work split over the ranks (17 us per column and
    // owned 64-column block for the column and p, and the half-matrix
    // product).  Fitted to per-phase clock probes on MI355X with the
    // keep-10000 blocks running together.  The team sizes make every blocked
    // job finish by a common time M, the smallest for which all teams fit on
    // the device together (bisection on M), so CUs left over by the smallest
    // sizes go to the largest matrices (a single 13136-row block: 51 -> 103
    // ranks, 4.24 -> 2.73 s); when even the smallest sizes do not fit at
    // once (the keep-10000 iterations) every team has its smallest size and
    // the queues run them in turn.  NRG_BLK_TDIV > 0 instead gives every
    // matrix nt / NRG_BLK_TDIV ranks.
    {
        // The two-stage reduction's stage 1 (blkjob 2) pays ~35 barriers per
        // panel of 32 columns (o(n) ~ 3.5 n us) and W(n) ~ 2.7e-5 n^3 CU-us of
        // tile products (clock probes on MI355X: a 12680-row block with 49
        // ranks among the keep-10000 blocks 1.07 s, a 13136-row block alone
        // with 103 ranks 0.65 s).
        auto blk_o1 = [](double n) { return n * std::max(0.0114 * n - 25.0, 10.0); };
        auto blk_w1 = [](double n) { return 0.266 * n * n + 7e-5 * n * n * n; };
        const bool two = c->sbr != 0;
        auto blk_o = [&](double n) { return two ? 3.5 * n : blk_o1(n); };
        auto blk_w = [&](double n) { return two ? 0.5 * n * n + 2.7e-5 * n * n * n : blk_w1(n); };
        const int tmax = std::min(TEAM_GLOBAL_MAX, tri_grid);
        // between one rank per four and one per two 64-column blocks: smaller
        // teams measured slower than the model in a full mix (their ranks'
        // per-block work grows; two-stage floors of one rank per six blocks:
        // keep-10000 blocks 1.44 -> 1.76 s), larger ones were not measured
        auto team_for = [&](double n, double M) {
            const int nt = ((int)n + 63) / 64;
            const int cap = std::max(2, std::min(tmax, nt / 2)),
                      floor_ = std::min(cap, std::max(2, nt / 4));
            const double room = M - blk_o(n);
            if (room <= 0.0) return cap;
            return (int)std::min<double>(cap, std::max<double>(floor_, std::ceil(blk_w(n) / room)));
        };
        std::vector<int> bj;
        for (int i = 0; i < ntasks; ++i)
            if (blkjob[i]) bj.push_back(i);
        double M = 0.0;
        if (!bj.empty() && blk_tdiv <= 0) {
            double lo = 0.0, hi = 0.0;
            for (int i : bj) {
                lo = std::max(lo, blk_o(tasks[i].n));
                hi = std::max(hi, blk_o(tasks[i].n) + blk_w(tasks[i].n) / 2.0);
            }
            long floors = 0;
            for (int i : bj) floors += team_for(tasks[i].n, 1e300);
            if (floors <= tri_grid) {
                for (int it = 0; it < 60; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    long sum = 0;
                    for (int i : bj) sum += team_for(tasks[i].n, mid);
                    (sum <= tri_grid ? hi : lo) = mid;
                }
                M = hi;
            } else if (two) {
                // the teams run in turn from the queues: size them for the
                // average load of the device (total work over the grid), so
                // the largest matrices are not the critical path
                double wsum = 0.0;
                for (int i : bj) wsum += blk_w(tasks[i].n);
                M = std::max(lo, wsum / tri_grid);
            } else {
                M = hi;
            }
        }
        for (int i : bj) {
            const double n = tasks[i].n;
            const int nt = (tasks[i].n + 63) / 64;
            int T = blk_tdiv > 0 ? std::max(2, nt / blk_tdiv) : team_for(n, M);
            T = std::min(T, tmax);
            if (blkjob[i] == 2) {  // a rank holds at most SBR_MAXQ blocks of panel rows
                while (T < tmax && sbr_owned_max(nt, T) > SBR_MAXQ) ++T;
                if (sbr_owned_max(nt, T) > SBR_MAXQ)
                    return fail(c, NRG_ERR_ARG, "diag: two-stage team too small for dimension " + std::to_string(tasks[i].n));
            }
            if (T > std::min(TEAM_MAX, per_xcd)) global[i] = 1;
            teamT[i] = T;
            cost[i] = blk_o(n) + blk_w(n) / T;
        }
        // When the two-stage jobs do not all fit on the device at once, a
        // team drawn from the ticket queue would assemble from workgroups
        // that finish earlier jobs at different times, the first ones idle
        // until the last arrives (clock probes: a third of the launch's
        // CU time).  Instead the jobs form chains on fixed teams: the
        // largest that fit at once start the chains (their team sizes), and
        // every further job, largest first, is appended to the chain that
        // would finish it first, run by that chain's team.
        std::vector<int> sj;
        long sumT = 0;
        for (int i : bj)
            if (blkjob[i] == 2) {
                sj.push_back(i);
                sumT += teamT[i];
            }
        if (sumT > tri_grid && sj.size() > 1) {
            std::stable_sort(sj.begin(), sj.end(), [&](int a, int b) { return cost[a] > cost[b]; });
            struct Lane {
                int T, last;
                double fin;
            };
            std::vector<Lane> lanes;
            int used = 0;
            size_t k = 0;
            for (; k < sj.size() && used + teamT[sj[k]] <= tri_grid; ++k) {
                lanes.push_back({teamT[sj[k]], sj[k], cost[sj[k]]});
                global[sj[k]] = 1;
                used += teamT[sj[k]];
            }
            // the workgroups left over start one more chain, with the largest
            // job whose panel rows they can hold
            if (tri_grid - used >= 2) {
                for (size_t m = k; m < sj.size(); ++m) {
                    const int i = sj[m], nt = (tasks[i].n + 63) / 64;
                    const int T = std::min(teamT[i], tri_grid - used);
                    if (sbr_owned_max(nt, T) > SBR_MAXQ) continue;
                    teamT[i] = T;
                    cost[i] = blk_o(tasks[i].n) + blk_w(tasks[i].n) / T;
                    lanes.push_back({T, i, cost[i]});
                    global[i] = 1;
                    used += T;
                    sj.erase(sj.begin() + m);
                    break;
                }
            }
            for (; k < sj.size(); ++k) {
                const int i = sj[k], nt = (tasks[i].n + 63) / 64;
                const double n = tasks[i].n;
                int best = -1;
                double bf = 0.0;
                for (int l = 0; l < (int)lanes.size(); ++l) {
                    if (sbr_owned_max(nt, lanes[l].T) > SBR_MAXQ) continue;
                    const double f = lanes[l].fin + blk_o(n) + blk_w(n) / lanes[l].T;
                    if (best < 0 || f < bf) {
                        best = l;
                        bf = f;
                    }
                }
                if (best < 0) continue;  // no chain's team suits it: it stays a queue job
                chain_next[lanes[best].last] = i;
                chained[i] = 1;
                teamT[i] = lanes[best].T;
                lanes[best].last = i;
                lanes[best].fin = bf;
            }
        }
    }
    std::vector<int> order(ntasks);
    for (int i = 0; i < ntasks; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int> gq;
    std::vector<std::vector<int>> queue(N_XCD), rqueue(N_XCD);
    std::vector<double> load(N_XCD, 0.0), rload(N_XCD, 0.0);
    for (int i : order)
        if (global[i] && !chained[i]) {
            gq.push_back(i);
            for (int x = 0; x < N_XCD; ++x) load[x] += cost[i] * teamT[i] / tri_grid;
        }
    std::vector<int> ones;  // one-workgroup matrices
    for (int i : order) {
        if (global[i] || chained[i]) continue;
        if (tasks[i].n <= REG_ONE_N) {
            ones.push_back(i);
            continue;
        }
        std::vector<double> &ld = regjob[i] ? rload : load;
        const int x = (int)(std::min_element(ld.begin(), ld.end()) - ld.begin());
        (regjob[i] ? rqueue : queue)[x].push_back(i);
        ld[x] += cost[i] * teamT[i] / (regjob[i] ? reg_per_xcd : per_xcd);
    }
    // job lists: [global queue, queue of XCD 0, ..., queue of XCD 7]
    auto job_list = [&](const std::vector<int> &g, const std::vector<std::vector<int>> &q, std::vector<TriJob> &jobs,
                        std::vector<int> &tbase, std::vector<int> &qjob) {
        tbase.assign(1, 0);
        qjob.assign(1, 0);
        for (int i : g) {
            jobs.push_back({i, teamT[i], blkjob[i], -1});
            tbase.push_back(tbase.back() + teamT[i]);
        }
        qjob.push_back((int)jobs.size());
        for (int x = 0; x < N_XCD; ++x) {
            for (int i : q[x]) {
                jobs.push_back({i, teamT[i], blkjob[i], -1});
                tbase.push_back(tbase.back() + teamT[i]);
            }
            qjob.push_back((int)jobs.size());
        }
    };
    std::vector<TriJob> jobs, rjobs;
    std::vector<int> tbase, qjob, rtbase, rqjob;
    job_list(gq, queue, jobs, tbase, qjob);
    job_list({}, rqueue, rjobs, rtbase, rqjob);
    // chain continuations after the queues (reached only through `next`)
    {
        std::vector<int> jof(ntasks, -1);
        for (int j = 0; j < (int)jobs.size(); ++j) jof[jobs[j].mat] = j;
        for (int i = 0; i < ntasks; ++i)
            if (chained[i]) {
                jof[i] = (int)jobs.size();
                jobs.push_back({i, teamT[i], blkjob[i], -1});
            }
        for (int j = 0; j < (int)jobs.size(); ++j)
            if (chain_next[jobs[j].mat] >= 0) jobs[j].next = jof[chain_next[jobs[j].mat]];
    }
    // counters: tickets + one barrier per job, register teams first; barrier
    // counters a cache line apart (teams polling neighbouring counters would
    // contend for one line)
    // (the last counter: the bulge chasers' ticket)
    const size_t creg = N_XCD + 1 + (size_t)BAR_STRIDE * rjobs.size(),
                 ncnt = creg + N_XCD + 1 + (size_t)BAR_STRIDE * jobs.size() + 1;
    if (ncnt > c->cnt_ints) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->cnt) HIPCHK(c, hipFree(c->cnt));
        c->cnt = nullptr;
        c->cnt_ints = std::max<size_t>(ncnt * 2, 1024);
        HIPCHK(c, hipMalloc(&c->cnt, c->cnt_ints * sizeof(int)));
    }
    // The tridiagonalisation is staged and launched before the rest of the
    // plan is built, so the host builds that while it runs.  Its slot of the
    // staging ring (the matrix descriptors, which the later stages read too)
    // is next written by a copy queued on this stream after all of them.
    // two-stage matrices: band extraction items (16 columns each) and the
    // chaser teams (largest first; a team never exceeds the co-resident grid)
    std::vector<BtItem> sbitems;
    std::vector<SbrJob> sjobs;
    std::vector<int> stbase(1, 0);
    for (int i = 0; i < ntasks; ++i) {
        if (blkjob[i] != 2) continue;
        const int n = tasks[i].n;
        for (int j = 0; j < n; j += 16) sbitems.push_back({i, j});
        const int P = std::max(1, std::min((n - 1) / (3 * SBR_B) + 1, c->chase_capacity));
        sjobs.push_back({i, P});
        stbase.push_back(stbase.back() + P);
    }
    HostArena ar1;
    const size_t o_jobs = ar1.put(jobs), o_tb = ar1.put(tbase), o_qj = ar1.put(qjob), o_ones = ar1.put(ones);
    const size_t o_sbitems = ar1.put(sbitems), o_sjobs = ar1.put(sjobs), o_stbase = ar1.put(stbase);
    const size_t o_rjobs = ar1.put(rjobs), o_rtb = ar1.put(rtbase), o_rqj = ar1.put(rqjob);
    const size_t o_mats = ar1.put(mats);
    void *d1;
    {
        const nrg_status s1 = stage(c, ar1.buf.data(), ar1.buf.size(), &d1);
        if (s1) return s1;
    }
    char *db1 = (char *)d1;
    const EigMat *d_mats = (const EigMat *)(db1 + o_mats);
    const TriJob *d_jobs = (const TriJob *)(db1 + o_jobs);
    const int *d_tb = (const int *)(db1 + o_tb);
    const int *d_qj = (const int *)(db1 + o_qj);

    // 1. tridiagonalisation: matrices up to REG_ONE_N one workgroup each,
    //    then one persistent launch for the register-resident teams and one
    //    for the streaming teams
    {
        double fl1 = 0, by1 = 0, flr = 0, byr = 0, flt = 0, byt = 0;
        for (int i = 0; i < ntasks; ++i) {
            const double n = tasks[i].n;
            // algorithmic work: 4/3 n^3 flops; the trailing matrix is read once
            // per column (sum_j 8 (n-j)^2 bytes)
            double &fl = tasks[i].n <= REG_ONE_N ? fl1 : regjob[i] ? flr : flt;
            double &by = tasks[i].n <= REG_ONE_N ? by1 : regjob[i] ? byr : byt;
            fl += 4.0 / 3.0 * n * n * n;
            // the blocked product reads half the trailing matrix per column;
            // the two-stage reduction, per panel of 32 columns of a trailing
            // matrix of order m, reads it for the product (8 m^2 bytes) and
            // reads and writes its lower triangle in the update (8 m^2):
            // sum over panels 16 m^2 = n^3 / 6 bytes
            by += (blkjob[i] == 2 ? 1.0 / 6.0 : blkjob[i] ? 4.0 / 3.0 : 8.0 / 3.0) * n * n * n;
        }
        bool any_blk = false;
        for (int i = 0; i < ntasks; ++i) any_blk = any_blk || blkjob[i];
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, c->stream, c->cnt, (int)ncnt);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->dstatus, 1);
        HIPCHK(c, hipGetLastError());
        // With register teams, the one-workgroup matrices are drained by that
        // launch's workgroups without a team ticket (a separate launch would
        // need free CUs, and its workgroups wait when the dispatcher places
        // them where teams hold every CU).  Without teams they run in their
        // own launch on the side stream, beside the streaming teams.
        const bool ones_in_reg = !rjobs.empty();
        const bool ones_side = !ones.empty() && !ones_in_reg;
        if (ones_side) HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        // grids <= co-resident capacity (one workgroup per CU at these sizes)
        // and nothing else runs on the stream, so every team assembles; the
        // barrier's bounded wait reports an error otherwise
        if (!rjobs.empty()) {
            LaunchTimer lt(K_EIG_TRI_REG, c->stream, flr + fl1, byr + by1);
            hipLaunchKernelGGL(k_eig_tri_reg, dim3(reg_grid), dim3(REG_NT), 0, c->stream, d_mats,
                               (const TriJob *)(db1 + o_rjobs), (const int *)(db1 + o_rtb), (const int *)(db1 + o_rqj),
                               c->cnt, c->dstatus, (const int *)(db1 + o_ones), (int)ones.size());
        }
        if (!jobs.empty()) {
            LaunchTimer lt(any_blk ? K_EIG_TRI_BLK : K_EIG_TRI, c->stream, flt, byt);
            hipLaunchKernelGGL(any_blk ? k_eig_tri_blk : k_eig_tri_coop, dim3(tri_grid), dim3(TRI_NT), TRI_COOP_LDS,
                               c->stream, d_mats, d_jobs, d_tb, d_qj, c->cnt + creg, c->dstatus);
        }
        if (ones_side) {
            HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
            {
                LaunchTimer lt(K_EIG_TRI1, c->side, fl1, by1);
                hipLaunchKernelGGL(k_eig_tri_one, dim3(ones.size()), dim3(TRI_NT), 0, c->side, d_mats,
                                   (const int *)(db1 + o_ones), c->dstatus);
            }
            HIPCHK(c, hipEventRecord(c->ev_join, c->side));
            HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
        }
        // two-stage matrices: band -> tridiagonal by bulge chasing (grid <=
        // co-resident capacity; every team fits it, so every chaser's
        // predecessor runs)
        if (!sjobs.empty()) {
            hipLaunchKernelGGL(k_sbr_band, dim3(sbitems.size()), dim3(256), 0, c->stream, d_mats,
                               (const BtItem *)(db1 + o_sbitems));
            HIPCHK(c, hipGetLastError());
            double flc = 0.0;
            for (const SbrJob &jb : sjobs) flc += 6.0 * SBR_B * (double)tasks[jb.mat].n * tasks[jb.mat].n;
            LaunchTimer lt(K_SBR_CHASE, c->stream, flc, 0.0);
            const int grid = std::min(c->chase_capacity, stbase.back());
            hipLaunchKernelGGL(k_sbr_chase, dim3(grid), dim3(CH_NT), 0, c->stream, d_mats,
                               (const SbrJob *)(db1 + o_sjobs), (const int *)(db1 + o_stbase), (int)sjobs.size(),
                               c->cnt + ncnt - 1, c->dstatus);
        }
    }
    HIPCHK(c, hipGetLastError());

    // divide-and-conquer tree, level by level over all matrices
    int maxlev = 0;
    for (int i = 0; i < nbig; ++i) maxlev = std::max(maxlev, dc_levels(tasks[i].n));
    std::vector<DcMerge> merges;
    std::vector<DcChunk> schunks, uchunks, copies;  // copies: deflated-row chunks (upper bound)
    std::vector<GatherItem> gathers;
    std::vector<nrg_gemm_out> gouts;
    std::vector<nrg_gemm_term> gterms;
    std::vector<GemmItem> gitems, gred;  // tiles (or slices), reduce items of split tiles
    std::vector<int> gks;                 // k-steps bound of each merge output
    int gslots = 0;                       // partial tiles needed by the largest launch
    struct Level {
        int m0, m1, s0, s1, u0, u1, g0, g1, x0, x1, c0, c1, r0, r1, maxK;
        double flops;
        bool mfma;
    };
    std::vector<Level> levels(maxlev + 1);
    // levels 1..DC_SMALL_LEV: one k_dc_small workgroup per subtree
    std::vector<BtItem> small;
    for (int i = 0; i < nbig; ++i) {
        const int nlev = dc_levels(tasks[i].n), ls = std::min(DC_SMALL_LEV, nlev);
        for (int k = 0; k < (1 << (nlev - ls)); ++k) small.push_back({i, k});
    }
    for (int L = DC_SMALL_LEV + 1; L <= maxlev; ++L) {
        Level lv{};
        lv.m0 = (int)merges.size();
        lv.s0 = (int)schunks.size();
        lv.u0 = (int)uchunks.size();
        lv.g0 = (int)gitems.size();
        lv.x0 = (int)gathers.size();
        lv.c0 = (int)copies.size();
        const size_t lo0 = gouts.size();
        for (int i = 0; i < nbig; ++i) {
            const int n = tasks[i].n, nlev = dc_levels(n);
            if (L > nlev) continue;
            double *ws = c->ews + off[i];
            for (int k = 0; k < (1 << (nlev - L)); ++k) {
                const int a = dc_bound(k << L, n, nlev), mid = dc_bound((k << L) + (1 << (L - 1)), n, nlev),
                          b = dc_bound((k + 1) << L, n, nlev);
                const int K = b - a, mi = (int)merges.size(), g = (int)gouts.size();
                merges.push_back({ws, n, a, mid, b, g, 0});
                nrg_gemm_out o{};
                o.C = ws + (size_t)a * n + a;
                o.m = K;
                o.n = K;
                o.ldc = n;
                o.term_begin = g;
                o.term_count = 1;
                o.beta = 0.0;
                gouts.push_back(o);
                gks.push_back((K + GK - 1) / GK);
                nrg_gemm_term t{};
                t.A = ws + 2 * (size_t)n * n + (size_t)a * n + a;
                t.B = ws + (size_t)n * n + (size_t)a * n + a;
                t.s = nullptr;
                t.alpha = 1.0;
                t.k = K;
                t.lda = n;
                t.ldb = n;
                gterms.push_back(t);
                for (int j0 = 0; j0 < K; j0 += SEC_NT) schunks.push_back({mi, j0});
                for (int j0 = 0; j0 < K; j0 += 4) uchunks.push_back({mi, j0});
                // gather: at most ceil(K / SEG_L) + 1 segments, four per workgroup
                const int nseg_max = (K + SEG_L - 1) / SEG_L + 1;
                for (int s0 = 0; s0 < nseg_max; s0 += 4)
                    for (int c0 = 0; c0 < K; c0 += 64) gathers.push_back({mi, s0, c0, 0});
                for (int j0 = 0; j0 < K; j0 += COPY_ROWS) copies.push_back({mi, j0});
                lv.maxK = std::max(lv.maxK, K);
                lv.flops += 2.0 * K * K * K;
            }
        }
        lv.m1 = (int)merges.size();
        lv.s1 = (int)schunks.size();
        lv.u1 = (int)uchunks.size();
        lv.r0 = (int)gred.size();
        lv.mfma = gemm_plan(gouts, lo0, gks, gitems, gred, gslots, c->gemm_mfma_min_tiles);
        lv.r1 = (int)gred.size();
        lv.g1 = (int)gitems.size();
        lv.x1 = (int)gathers.size();
        lv.c1 = (int)copies.size();
        levels[L] = lv;
    }
    // back-transformation items by register-slot variant; rank / copy items
    std::vector<BtItem> bt[kBtVariants];
    std::vector<BtItem> rk, cp, wyrows;  // wyrows: one row per item (extract, permuted copy)
    std::vector<int> wymat;              // matrices back-transformed by WY blocks
    std::vector<BtItem> sbt;             // two-stage: 64 eigenvectors per item (stage 2's reflectors)
    std::vector<BtItem> str;             // two-stage: 64 x 64 tiles of Vc (entry-major copy and back)
    int wy_blocks = 0;
    for (int i = 0; i < nbig; ++i) {
        const int n = tasks[i].n, ns = (n + 63) / 64;
        if (mats[i].pad == 2) {
            for (int c0 = 0; c0 < n; c0 += 64) sbt.push_back({i, c0});
            for (int t = 0; t < ns * ns; ++t) str.push_back({i, t});
            wymat.push_back(i);
            wy_blocks = std::max(wy_blocks, (sbr_nrefl(n) + wy_bb(n) - 1) / wy_bb(n));
            for (int r = 0; r < n; ++r) wyrows.push_back({i, r});
        } else if (n >= BT_WY_MIN) {
            wymat.push_back(i);
            wy_blocks = std::max(wy_blocks, (n - 2 + wy_bb(n) - 1) / wy_bb(n));
            for (int r = 0; r < n; ++r) wyrows.push_back({i, r});
        } else {
            int v = 0;
            while (kBtSlots[v] < ns) ++v;
            for (int c0 = 0; c0 < n; c0 += BT_NT / 64) bt[v].push_back({i, c0});
            const size_t nn = (size_t)n * n;
            for (size_t e = 0; e < nn; e += 256 * 16) cp.push_back({i, (int)(e / (256 * 16))});
        }
        for (int c0 = 0; c0 < n; c0 += 64) rk.push_back({i, c0});
    }
    // WY blocks, counted from the last: block q of matrix i covers reflectors
    // [max(0, j1 - B), j1) with j1 = NR - q B, B = wy_bb(n): NR = n - 2
    // reflectors starting at row j + 1 (one-stage), or sbr_nrefl(n) starting
    // at row j + SBR_B (two-stage; WyItem.pad holds that offset).  T of a
    // block by dlarft's block recurrence over chunks of WY_NB: the diagonal
    // blocks by k_wy_tfactor, then chunk c's column above them,
    // T[0:c, c] = -T[0:c, 0:c] (Y[0:c]^T Y_c) T_cc, by two GEMM stages (P_c =
    // G[0:c, c] T_cc, then the product with the part of T already built).
    std::vector<WyItem> wyitems, wyblk;  // T-factor chunks; whole blocks
    std::vector<nrg_gemm_out> wyouts;
    std::vector<nrg_gemm_term> wyterms;
    std::vector<GemmItem> wygi, wyred;
    std::vector<int> wyks;
    constexpr int WY_NCH = WY_BB_MAX / WY_NB;
    constexpr int WY_STAGES = 2 * WY_NCH + 2;  // G, (P_c, T_c) for c = 1 .., M, W, Z
    struct WyLevel {
        int w0, w1;
        int g[WY_STAGES][2];  // gemm item ranges per stage
        int r[WY_STAGES][2];  // reduce item ranges
        bool mfma[WY_STAGES];
        double flops;
    };
    std::vector<WyLevel> wylev(wy_blocks);
    for (int q = 0; q < wy_blocks; ++q) {
        WyLevel L{};
        L.w0 = (int)wyitems.size();
        const size_t b0 = wyblk.size();
        for (int i : wymat) {
            const int n = tasks[i].n, two = mats[i].pad == 2, bb = wy_bb(n);
            const int j1 = (two ? sbr_nrefl(n) : n - 2) - q * bb;
            if (j1 <= 0) continue;
            const int j0 = std::max(0, j1 - bb), nb = j1 - j0;
            wyblk.push_back({c->ews + off[i], n, j0, nb, two ? SBR_B : 1, 0, bb});
            for (int g0 = 0; g0 < nb; g0 += WY_NB)
                wyitems.push_back({c->ews + off[i], n, j0 + g0, std::min(WY_NB, nb - g0), two ? SBR_B : 1, g0, bb});
        }
        L.w1 = (int)wyitems.size();
        for (int stage = 0; stage < WY_STAGES; ++stage) {
            L.g[stage][0] = (int)wygi.size();
            const size_t so0 = wyouts.size();
            for (size_t w = b0; w < wyblk.size(); ++w) {
                const WyItem &it = wyblk[w];
                const int n = it.n, j0 = it.j0, nb = it.nb, ro = it.pad, np = n - j0 - ro, ld = it.ld;
                // chunk stages: chunk ch = stage / 2 + ... (stages 1, 2: chunk 1; 3, 4: chunk 2; ..)
                const int ch = (stage + 1) / 2, c0 = ch * WY_NB, nbc = std::min(WY_NB, nb - c0);
                if (stage >= 1 && stage < WY_STAGES - 3 && nbc <= 0) continue;
                double *ws = it.ws, *Yt = ws + (size_t)n * n, *S2 = ws + 2 * (size_t)n * n;
                double *G = S2, *T = S2 + (size_t)ld * ld, *M = S2 + 2 * (size_t)ld * ld, *W = M + (size_t)ld * n;
                double *P = W;  // P_c (c0 x nbc, ld WY_NB), before W is computed
                const double *Yb = Yt + (size_t)j0 * n + j0 + ro;  // block rows, columns j0+ro..
                nrg_gemm_out o{};
                nrg_gemm_term t{};
                t.alpha = 1.0;
                if (stage == 0) {  // G = Yb Yb^T
                    o.C = G; o.m = nb; o.n = nb; o.ldc = ld;
                    t.A = Yb; t.lda = n; t.B = Yb; t.ldb = n; t.transB = 1; t.k = np;
                } else if (stage < WY_STAGES - 3 && (stage & 1)) {  // P_c = G[0:c0, c0:] T_cc
                    o.C = P; o.m = c0; o.n = nbc; o.ldc = WY_NB;
                    t.A = G + c0; t.lda = ld; t.B = T + (size_t)c0 * ld + c0; t.ldb = ld; t.k = nbc;
                } else if (stage < WY_STAGES - 3) {  // T[0:c0, c0:] = -T[0:c0, 0:c0] P_c
                    o.C = T + c0; o.m = c0; o.n = nbc; o.ldc = ld;
                    t.A = T; t.lda = ld; t.B = P; t.ldb = WY_NB; t.k = c0; t.alpha = -1.0;
                } else if (stage == WY_STAGES - 3) {  // M = T Yb
                    o.C = M; o.m = nb; o.n = np; o.ldc = n;
                    t.A = T; t.lda = ld; t.B = Yb; t.ldb = n; t.k = nb;
                } else if (stage == WY_STAGES - 2) {  // W = Z[:, j0+ro:] M^T
                    o.C = W; o.m = n; o.n = nb; o.ldc = ld;
                    t.A = ws + j0 + ro; t.lda = n; t.B = M; t.ldb = n; t.transB = 1; t.k = np;
                } else {  // Z[:, j0+ro:] -= W Yb
                    o.C = ws + j0 + ro; o.m = n; o.n = np; o.ldc = n; o.beta = 1.0;
                    t.A = W; t.lda = ld; t.B = Yb; t.ldb = n; t.k = nb; t.alpha = -1.0;
                }
                o.term_begin = (int)wyterms.size();
                o.term_count = 1;
                wyouts.push_back(o);
                wyterms.push_back(t);
                wyks.push_back((t.k + GK - 1) / GK);
                if (stage >= WY_STAGES - 2) L.flops += 2.0 * n * (double)nb * np;
            }
            L.r[stage][0] = (int)wyred.size();
            L.mfma[stage] = gemm_plan(wyouts, so0, wyks, wygi, wyred, gslots, c->gemm_mfma_min_tiles);
            L.r[stage][1] = (int)wyred.size();
            L.g[stage][1] = (int)wygi.size();
        }
        wylev[q] = L;
    }
    HostArena ar;
    const size_t o_small = ar.put(small), o_voff = ar.put(voff);
    const size_t o_merges = ar.put(merges), o_sch = ar.put(schunks), o_uch = ar.put(uchunks),
                 o_xch = ar.put(gathers), o_cpy = ar.put(copies), o_gouts = ar.put(gouts), o_gterms = ar.put(gterms), o_gitems = ar.put(gitems), o_gred = ar.put(gred), o_rk = ar.put(rk),
                 o_cp = ar.put(cp);
    size_t o_bt[kBtVariants];
    for (int v = 0; v < kBtVariants; ++v) o_bt[v] = ar.put(bt[v]);
    const size_t o_sbt = ar.put(sbt), o_str = ar.put(str);
    const size_t o_wyrows = ar.put(wyrows), o_wyitems = ar.put(wyitems), o_wyouts = ar.put(wyouts),
                 o_wyterms = ar.put(wyterms), o_wygi = ar.put(wygi), o_wyred = ar.put(wyred);
    if (gslots) {
        nrg_status sr = gemm_reserve(c, gslots);
        if (sr) return sr;
    }
    void *d;
    nrg_status s = stage(c, ar.buf.data(), ar.buf.size(), &d);
    if (s) return s;
    char *db = (char *)d;
    const DcMerge *d_merges = (const DcMerge *)(db + o_merges);
    const DcChunk *d_sch = (const DcChunk *)(db + o_sch), *d_uch = (const DcChunk *)(db + o_uch),
                  *d_cpy = (const DcChunk *)(db + o_cpy);
    const GatherItem *d_xch = (const GatherItem *)(db + o_xch);
    nrg_gemm_out *d_gouts = (nrg_gemm_out *)(db + o_gouts);
    nrg_gemm_term *d_gterms = (nrg_gemm_term *)(db + o_gterms);
    const GemmItem *d_gitems = (const GemmItem *)(db + o_gitems), *d_gred = (const GemmItem *)(db + o_gred);

    HIPCHK(c, hipGetLastError());
    // once the eigenvalues are final: the status word (set only by the
    // tridiagonalisation teams) and the eigenvalues go to pinned memory, and
    // ev_values marks them for nrg_diag_values, before the eigenvectors
    auto values_out = [&]() -> nrg_status {
        if (!c->status_pending) *c->hstat = 0;  // nothing queued writes it now
        hipLaunchKernelGGL(k_status_or, dim3(1), dim3(64), 0, c->stream, c->dstatus, c->hstat);
        c->status_pending = true;
        hipLaunchKernelGGL(k_values_out, dim3(ntasks), dim3(256), 0, c->stream, d_mats,
                           (const long long *)(db + o_voff), c->hval);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(c->ev_values, c->stream));
        c->values_pending = true;
        c->values_count = nvals;
        return NRG_OK;
    };
    if (nbig == 0) return values_out();
    // 2. divide and conquer
    {
        LaunchTimer lt(K_DC_LEAF, c->stream, 0.0, 0.0);
        hipLaunchKernelGGL(k_dc_leaf, dim3(nbig), dim3(256), 0, c->stream, d_mats);
    }
    HIPCHK(c, hipGetLastError());
    {
        LaunchTimer lt(K_DC_SMALL, c->stream, 0.0, 0.0);
        hipLaunchKernelGGL(k_dc_small, dim3(small.size()), dim3(DC_SMALL_NT), sizeof(DcSmall), c->stream, d_mats,
                           (const BtItem *)(db + o_small));
    }
    HIPCHK(c, hipGetLastError());
    for (int L = 1; L <= maxlev; ++L) {
        const Level &lv = levels[L];
        const int nm = lv.m1 - lv.m0;
        if (nm <= 0) continue;
        {
            LaunchTimer lt(K_DC_DEFLATE, c->stream, 0.0, 0.0);
            if (lv.maxK <= DEFL_LDS_K)
                hipLaunchKernelGGL(k_dc_deflate<true>, dim3(nm), dim3(DEFL_NT), 32 * (size_t)lv.maxK, c->stream,
                                   d_merges + lv.m0, d_gouts, d_gterms);
            else
                hipLaunchKernelGGL(k_dc_deflate<false>, dim3(nm), dim3(DEFL_NT), 0, c->stream, d_merges + lv.m0,
                                   d_gouts, d_gterms);
        }
        HIPCHK(c, hipGetLastError());
        {
            LaunchTimer lt(K_DC_DEFLATE, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_dc_gather, dim3(lv.x1 - lv.x0), dim3(256), 0, c->stream, d_merges, d_xch + lv.x0);
        }
        HIPCHK(c, hipGetLastError());
        {
            // the secular equations, and the copy of deflated rows once the
            // gather has finished reading Vc
            LaunchTimer lt(K_DC_SECULAR, c->stream, 0.0, 0.0);
            const int nsec = lv.u1 - lv.u0;
            hipLaunchKernelGGL(k_dc_secular, dim3(nsec + lv.c1 - lv.c0), dim3(256), 0, c->stream, d_merges,
                               d_uch + lv.u0, nsec, d_cpy + lv.c0);
        }
        {
            LaunchTimer lt(K_DC_ZHAT, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_dc_zhat, dim3(lv.s1 - lv.s0), dim3(64 * ZHAT_NW), 0, c->stream, d_merges, d_sch + lv.s0);
        }
        {
            LaunchTimer lt(K_DC_UNORM, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_dc_unorm, dim3(lv.u1 - lv.u0), dim3(256), 0, c->stream, d_merges, d_uch + lv.u0);
        }
        HIPCHK(c, hipGetLastError());
        s = gemm_launch_dev(c, K_GEMM_DC, d_gouts, d_gterms, d_gitems + lv.g0, lv.g1 - lv.g0, d_gred + lv.r0,
                            lv.r1 - lv.r0, lv.flops, 0.0, lv.mfma);
        if (s) return s;
    }
    // 3. sort, back-transformation, copy into A
    {
        LaunchTimer lt(K_DC_RANK, c->stream, 0.0, 0.0);
        hipLaunchKernelGGL(k_dc_rank, dim3(rk.size()), dim3(64 * RANK_NW), 0, c->stream, d_mats, (const BtItem *)(db + o_rk));
    }
    HIPCHK(c, hipGetLastError());
    s = values_out();
    if (s) return s;
    {
        double fl = 0, by = 0;
        for (int i = 0; i < nbig; ++i) {
            const double n = tasks[i].n;
            if (tasks[i].n >= BT_WY_MIN) continue;
            fl += 2.0 * n * n * n;
            by += 8.0 * 2.0 * n * n;
        }
        LaunchTimer lt(K_EIG_BT, c->stream, fl, by);
        // the variants work on different matrices: the largest runs on the
        // stream, the others beside it on the side stream
        int vmax = -1;
        for (int v = 0; v < kBtVariants; ++v)
            if (!bt[v].empty()) vmax = v;
        bool forked = false;
        for (int v = 0; v < kBtVariants; ++v) {
            if (bt[v].empty() || v == vmax) continue;
            if (!forked) {
                HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
                HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
                forked = true;
            }
            launch_bt_variant(v, c->side, (int)bt[v].size(), d_mats, (const BtItem *)(db + o_bt[v]));
        }
        if (vmax >= 0) launch_bt_variant(vmax, c->stream, (int)bt[vmax].size(), d_mats, (const BtItem *)(db + o_bt[vmax]));
        if (forked) {
            HIPCHK(c, hipEventRecord(c->ev_join, c->side));
            HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
        }
    }
    HIPCHK(c, hipGetLastError());
    if (!cp.empty()) {
        double by = 0;
        for (int i = 0; i < nbig; ++i)
            if (tasks[i].n < BT_WY_MIN) by += 16.0 * tasks[i].n * tasks[i].n;
        LaunchTimer lt(K_EIG_COPY, c->stream, 0.0, by);
        hipLaunchKernelGGL(k_eig_copy, dim3(cp.size()), dim3(256), 0, c->stream, d_mats, (const BtItem *)(db + o_cp));
    }
    HIPCHK(c, hipGetLastError());
    if (!sbt.empty()) {  // stage 2's reflectors on the eigenvectors of the tridiagonal
        double fl = 0.0;
        for (int i = 0; i < nbig; ++i)
            if (mats[i].pad == 2) fl += 2.0 * (double)tasks[i].n * tasks[i].n * tasks[i].n;
        double by = 0.0;
        for (const BtItem &b : sbt)
            if (b.c0 == 0) by += 16.0 * tasks[b.mat].n * (double)tasks[b.mat].n;
        {
            LaunchTimer lt(K_EIG_COPY, c->stream, 0.0, by);
            hipLaunchKernelGGL(k_sbr_tr, dim3(str.size()), dim3(256), 0, c->stream, d_mats, (const BtItem *)(db + o_str), 1);
        }
        {
            LaunchTimer lt(K_SBR_BT, c->stream, fl, 0.0);
            hipLaunchKernelGGL((k_sbr_bt<SBT_BL, true>), dim3(sbt.size()), dim3(64), 0, c->stream, d_mats, (const BtItem *)(db + o_sbt));
        }
        {
            LaunchTimer lt(K_EIG_COPY, c->stream, 0.0, by);
            hipLaunchKernelGGL(k_sbr_tr, dim3(str.size()), dim3(256), 0, c->stream, d_mats, (const BtItem *)(db + o_str), 0);
        }
    }
    HIPCHK(c, hipGetLastError());
    if (!wymat.empty()) {
        const BtItem *d_wyrows = (const BtItem *)(db + o_wyrows);
        const WyItem *d_wyitems = (const WyItem *)(db + o_wyitems);
        const nrg_gemm_out *d_wyouts = (const nrg_gemm_out *)(db + o_wyouts);
        const nrg_gemm_term *d_wyterms = (const nrg_gemm_term *)(db + o_wyterms);
        const GemmItem *d_wygi = (const GemmItem *)(db + o_wygi), *d_wyred = (const GemmItem *)(db + o_wyred);
        auto wy_gemm = [&](const WyLevel &L, int stage, double fl) {
            return gemm_launch_dev(c, K_GEMM_BT, d_wyouts, d_wyterms, d_wygi + L.g[stage][0],
                                   L.g[stage][1] - L.g[stage][0], d_wyred + L.r[stage][0],
                                   L.r[stage][1] - L.r[stage][0], fl, 0.0, L.mfma[stage]);
        };
        {
            LaunchTimer lt(K_EIG_BT, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_wy_extract, dim3(wyrows.size()), dim3(256), 0, c->stream, d_mats, d_wyrows);
        }
        HIPCHK(c, hipGetLastError());
        for (int q = 0; q < wy_blocks; ++q) {
            const WyLevel &L = wylev[q];
            s = wy_gemm(L, 0, 0.0);
            if (s) return s;
            {
                LaunchTimer lt(K_EIG_BT, c->stream, 0.0, 0.0);
                hipLaunchKernelGGL(k_wy_tfactor, dim3(L.w1 - L.w0), dim3(WY_TF_NT), 0, c->stream, d_wyitems + L.w0);
            }
            HIPCHK(c, hipGetLastError());
            for (int stage = 1; stage < WY_STAGES; ++stage) {
                s = wy_gemm(L, stage, stage >= WY_STAGES - 2 ? 0.5 * L.flops : 0.0);
                if (s) return s;
            }
        }
        {
            LaunchTimer lt(K_EIG_COPY, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_eig_copy_perm, dim3(wyrows.size()), dim3(256), 0, c->stream, d_mats, d_wyrows);
        }
        HIPCHK(c, hipGetLastError());
    }
    return NRG_OK;
}
