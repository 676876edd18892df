// nrg_eig.hip - batched FP64 symmetric eigensolver (replaces the dsyev call
// of diagonalise(), src/diag.h:54-89 and 297-353).
//
// All matrices of one NRG iteration are solved together, stage by stage:
//   k_eig_tri    Householder tridiagonalisation, one 1024-thread workgroup per
//                matrix; step j's rank-2 update is fused with step j+1's
//                matrix-vector product (one pass over the trailing matrix per
//                step).  Reflectors stay in the rows of A.
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
//   k_eig_bt     back-transformation: one wavefront per eigenvector applies
//                the reflectors H_{n-3} .. H_0 with the vector held in
//                registers, and writes it as row rank(c) of the result.
// Output convention as diagonalise(): eigenvalues ascending in W, row r of A
// = eigenvector r.
#include "nrg_internal.h"

constexpr int TRI_NT = 1024;
constexpr int TRI_NW = TRI_NT / 64;
constexpr int TRI_LDS_N = 1344;  // 5 vectors x 1344 x 8 B = 53 KB dynamic LDS
constexpr int DEFL_NT = 256;
constexpr int DEFL_LDS_K = 2048;  // 2 x K doubles + 2 x K ints + K doubles = 32 B x K
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
    V_PV,       // p = tau A v
    V_VA,       // tridiagonalisation vectors (global variant)
    V_VB,
    V_WV,
    V_DGT,
    V_ODT,
    V_DG,    // diagonal of T (after tridiagonalisation)
    V_OD,    // off-diagonal of T
    V_D,     // current eigenvalues (per position)
    V_BETA,  // coupling cut at a merge boundary, stored at index mid
    V_DD,    // sorted non-deflated poles of a merge (at a..a+K')
    V_ZZ,    // their z components
    V_ZH,    // z-hat
    V_TS,    // secular root offsets
    V_ORG,   // int: origin pole of each root
    V_META,  // int: [a] = K', [a+1] = number of rotations
    V_RHO,   // rho of the merge at [a]
    V_RANK,  // int: final rank of each eigenvalue
    V_RP,    // int: rotation pairs (sorted indices)
    V_RQ,
    V_RC,
    V_RS,
    V_SDS,   // deflation arrays (global variant)
    V_SZS,
    V_PERM,
    V_POS,
    V_PART,  // TRI_NW vectors of row-group partial sums
    V_COUNT = V_PART + TRI_NW
};

__host__ __device__ inline size_t eig_ws_doubles(int n) { return 3 * (size_t)n * n + (size_t)V_COUNT * n; }
__device__ inline double *vslot(double *ws, int n, int s) { return ws + 3 * (size_t)n * n + (size_t)s * n; }
__device__ inline int *islot(double *ws, int n, int s) { return reinterpret_cast<int *>(vslot(ws, n, s)); }

struct EigMat {
    double *A;
    double *W;
    double *ws;
    int n;
    int pad;  // 1: reduced by a team (reflectors in S2)
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

// Decomposition of a pass over an (mr x mc) block whose threads own columns:
// waves take 64-column chunks, and when there are fewer chunks than waves the
// rows are split into R contiguous groups so every wave keeps loads in flight.
struct PassMap {
    int ncc, R, rpg;
};
__device__ inline PassMap pass_map(int mr, int mc) {
    PassMap p;
    p.ncc = (mc + 63) / 64;
    p.R = (p.ncc >= TRI_NW) ? 1 : TRI_NW / p.ncc;
    if (p.R > mr) p.R = mr > 0 ? mr : 1;
    p.rpg = (mr + p.R - 1) / p.R;
    return p;
}

// ---------------------------------------------------------------------------
// 1. tridiagonalisation: A <- H_j A H_j, j = 0..n-3, H_j = I - tau_j v_j v_j^T
//    acting on rows/columns j+1..n-1.  v_j is left in row j of A (columns
//    j+1..n-1, first element included), tau_j in V_TAU, T in V_DG / V_OD.
// IN_LDS selects at compile time where the work vectors live, so that the
// LDS variant addresses them with ds_* instructions.
// ---------------------------------------------------------------------------
template <bool IN_LDS>
__device__ __forceinline__ void tri_core(const EigMat &mt, double *red) {
    extern __shared__ double lds[];
    const int n = mt.n;
    double *__restrict__ A = mt.A;
    const int tid = threadIdx.x;
    if (n <= 0) return;
    if (n == 1) {
        if (tid == 0) {
            mt.W[0] = A[0];
            A[0] = 1.0;
        }
        return;
    }
    double *ws = mt.ws;
    double *vecs = IN_LDS ? lds : vslot(ws, n, V_VA);
    double *va = vecs;          // reflector j (length m)
    double *vb = vecs + n;      // reflector j+1
    double *wv = vecs + 2 * n;  // w = p - K v
    double *dg = vecs + 3 * n;
    double *od = vecs + 4 * n;
    double *tau = vslot(ws, n, V_TAU);
    double *pv = vslot(ws, n, V_PV);
    double *part = vslot(ws, n, V_PART);
    const int wave = tid >> 6, lane = tid & 63;

    bool p_ready = false;
    for (int j = 0; j < n - 2; ++j) {
        const int m = n - j - 1;
        double *xr = A + (size_t)j * n + j + 1;  // row j right of the diagonal (final)
        double *A22 = A + (size_t)(j + 1) * n + (j + 1);
        // reflector of step j (may have been prepared by the previous fused pass)
        if (!p_ready) {
            double psum = 0.0;
            for (int k = 1 + tid; k < m; k += TRI_NT) psum += xr[k] * xr[k];
            const double sigma = tri_block_sum(psum, red);
            const double x0 = xr[0];
            double t = 0.0;
            if (sigma != 0.0) {
                const double alpha = -copysign(sqrt(x0 * x0 + sigma), x0);
                const double v0 = x0 - alpha;
                t = 2.0 / (v0 * v0 + sigma);
                for (int k = tid; k < m; k += TRI_NT) va[k] = (k == 0) ? v0 : xr[k];
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
                for (int ch = (pm.ncc >= TRI_NW ? wave : wave % pm.ncc); ch < pm.ncc;
                     ch += (pm.ncc >= TRI_NW ? TRI_NW : pm.ncc * TRI_NW)) {
                    const int g = (pm.ncc >= TRI_NW) ? 0 : wave / pm.ncc;
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
                for (int c = tid; c < m; c += TRI_NT) {
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
        for (int k = tid; k < m; k += TRI_NT) pp += va[k] * pv[k];
        const double K = 0.5 * t * tri_block_sum(pp, red);
        for (int k = tid; k < m; k += TRI_NT) wv[k] = pv[k] - K * va[k];
        __syncthreads();
        // row 0 of A22 (row j+1 of A) first: it carries the next reflector
        for (int c = tid; c < m; c += TRI_NT) A22[c] -= va[0] * wv[c] + wv[0] * va[c];
        __syncthreads();
        const bool has_next = (j + 1) < n - 2;
        double tn = 0.0;
        if (has_next) {
            const int mn = m - 1;
            double *xn = A22 + 1;  // row j+1, columns j+2..
            double psum = 0.0;
            for (int k = 1 + tid; k < mn; k += TRI_NT) psum += xn[k] * xn[k];
            const double sigma = tri_block_sum(psum, red);
            const double x0 = xn[0];
            if (sigma != 0.0) {
                const double alpha = -copysign(sqrt(x0 * x0 + sigma), x0);
                const double v0 = x0 - alpha;
                tn = 2.0 / (v0 * v0 + sigma);
                for (int k = tid; k < mn; k += TRI_NT) vb[k] = (k == 0) ? v0 : xn[k];
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
            for (int ch = (pm.ncc >= TRI_NW ? wave : wave % pm.ncc); ch < pm.ncc;
                 ch += (pm.ncc >= TRI_NW ? TRI_NW : pm.ncc * TRI_NW)) {
                const int g = (pm.ncc >= TRI_NW) ? 0 : wave / pm.ncc;
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
                for (int c = tid; c < mm; c += TRI_NT) {
                    double acc = 0.0;
                    for (int g = 0; g < pm.R; ++g) acc += part[(size_t)g * n + c];
                    pv[c] = tn * acc;
                }
            }
        }
        __syncthreads();
        if (fuse) {
            // the next reflector becomes current
            for (int k = tid; k < m - 1; k += TRI_NT) va[k] = vb[k];
            __syncthreads();
        }
        p_ready = fuse;
    }
    if (tid == 0) {
        od[n - 2] = A[(size_t)(n - 2) * n + n - 1];
        od[n - 1] = 0.0;
    }
    for (int i = tid; i < n; i += TRI_NT) dg[i] = A[(size_t)i * n + i];
    __syncthreads();
    double *gdg = vslot(ws, n, V_DG), *god = vslot(ws, n, V_OD);
    for (int i = tid; i < n; i += TRI_NT) {
        gdg[i] = dg[i];
        god[i] = od[i];
    }
}

__global__ void k_fill_int(int *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

__global__ void k_copy_int(const int *src, int *dst) {
    if (threadIdx.x == 0) *dst = *src;
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
constexpr size_t TRI_COOP_LDS = 3 * (size_t)TEAM_LDS_N * sizeof(double);  // >= 5 x TRI_LDS_N doubles
static_assert(TRI_COOP_LDS >= 5 * (size_t)TRI_LDS_N * sizeof(double), "one-workgroup vectors must fit");
constexpr int TEAM_MIN_N = 128;              // smaller matrices: one workgroup
constexpr int TEAM_MAX = 32;                 // workgroups of one XCD
constexpr int TEAM_GLOBAL_N = 1024;          // larger matrices: teams across XCDs
constexpr int TEAM_GLOBAL_MAX = 128;
constexpr int N_XCD = 8;                     // workgroup b runs on XCD b mod 8
constexpr long long TEAM_SPIN_LIMIT = 1LL << 22;  // polls (~1 us each): a few seconds

struct TriJob {
    int mat, T;
};

__device__ inline void st_dev(double *p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline double ld_dev(const double *p) {
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Team barrier with a bounded wait: returns false (and flags the status) if
// the team did not arrive, so that every workgroup still terminates.
__device__ inline bool team_barrier(int *cnt, int target, int *status, int *s_flag) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's device-scope stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        long long spins = 0;
        int ok = 1;
        while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > TEAM_SPIN_LIMIT || __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
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

// One persistent launch for the team matrices of a batch.  Every workgroup
// first draws tickets from the global queue (teams of the largest matrices,
// spread over XCDs), then workgroups of XCD x (blockIdx mod 8) draw from queue
// x; a job of T tickets is a team of T workgroups.  Tickets are handed out in
// job order and a workgroup finishes its job before it draws again, so every
// team assembles.
// Matrices below TEAM_MIN_N: one workgroup each.
__global__ __launch_bounds__(TRI_NT) void k_eig_tri_one(const EigMat *mats, const int *idx) {
    __shared__ double red[TRI_NW];
    const EigMat mt = mats[idx[blockIdx.x]];
    tri_core<true>(mt, red);
}

__global__ __launch_bounds__(TRI_NT) void k_eig_tri_coop(const EigMat *mats, const TriJob *jobs, const int *tbase,
                                                         const int *qjob, int *counters, int *status) {
    __shared__ double red[TRI_NW];
    __shared__ int s_ticket, s_flag;
    const int x = blockIdx.x % N_XCD;
    int *bars = counters + N_XCD + 1;     // counters[9 + job]: team barrier
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
        const TriJob jb = jobs[lo];
        const EigMat mt = mats[jb.mat];
        const int rank = t0 + t - tbase[lo];
        if (mt.n <= TEAM_LDS_N) tri_team<true>(mt, rank, jb.T, bars + lo, status, red, &s_flag);
        else tri_team<false>(mt, rank, jb.T, bars + lo, status, red, &s_flag);
        __syncthreads();
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
    if (mt.pad) {
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

// Sort, z, deflation and gather of one merge [a, b) with boundary mid.
template <bool IN_LDS>
__global__ __launch_bounds__(DEFL_NT) void k_dc_deflate(const DcMerge *merges, nrg_gemm_out *outs,
                                                      nrg_gemm_term *terms) {
    extern __shared__ double lds[];
    __shared__ double red[DEFL_NT / 64];
    __shared__ int s_info[2];
    const DcMerge mg = merges[blockIdx.x];
    const int n = mg.n, a = mg.a, h = mg.mid - mg.a, K = mg.b - mg.a;
    double *ws = mg.ws;
    double *Vc = ws, *S1 = ws + (size_t)n * n;
    double *D = vslot(ws, n, V_D) + a;
    const double beta = vslot(ws, n, V_BETA)[mg.mid];
    const double rho = 2.0 * fabs(beta);
    const double sg = beta >= 0.0 ? 1.0 : -1.0;
    const double rs2 = 0.70710678118654752440;
    const int tid = threadIdx.x;
    double *draw, *ds, *zs;
    int *perm, *pos;
    if (IN_LDS) {
        draw = lds;
        ds = lds + K;
        zs = lds + 2 * K;
        perm = reinterpret_cast<int *>(lds + 3 * K);
        pos = perm + K;
    } else {
        draw = vslot(ws, n, V_TS) + a;  // free until the secular stage
        ds = vslot(ws, n, V_SDS) + a;
        zs = vslot(ws, n, V_SZS) + a;
        perm = islot(ws, n, V_PERM) + a;
        pos = islot(ws, n, V_POS) + a;
    }
    for (int i = tid; i < K; i += DEFL_NT) draw[i] = D[i];
    __syncthreads();
    // stable ascending rank of the poles
    for (int i = tid; i < K; i += DEFL_NT) {
        const double di = draw[i];
        int rk = 0;
        for (int j = 0; j < K; ++j) {
            const double dj = draw[j];
            rk += (dj < di) || (dj == di && j < i);
        }
        perm[rk] = i;
    }
    __syncthreads();
    double md = 0.0, mz = 0.0;
    for (int q = tid; q < K; q += DEFL_NT) {
        const int i = perm[q];
        const size_t c = (size_t)(a + i);
        const double z = (i < h) ? Vc[c * n + (mg.mid - 1)] * rs2 : sg * Vc[c * n + mg.mid] * rs2;
        ds[q] = draw[i];
        zs[q] = z;
        md = fmax(md, fabs(draw[i]));
        mz = fmax(mz, fabs(z));
    }
    md = block_max<DEFL_NT>(md, red);
    mz = block_max<DEFL_NT>(mz, red);
    const double tol = 8.0 * kEps * fmax(md, rho * mz);
    int *rp = islot(ws, n, V_RP) + a, *rq = islot(ws, n, V_RQ) + a;
    double *rc = vslot(ws, n, V_RC) + a, *rsn = vslot(ws, n, V_RS) + a;
    if (tid == 0) {
        // pos[q] = -1 marks a deflated pole
        int pj = -1, nrot = 0;
        for (int q = 0; q < K; ++q) {
            pos[q] = 0;
            if (rho * fabs(zs[q]) <= tol) {
                pos[q] = -1;
                continue;
            }
            if (pj < 0) {
                pj = q;
                continue;
            }
            double s = zs[pj], c = zs[q];
            const double tau = hypot(c, s);
            const double t = ds[q] - ds[pj];
            c /= tau;
            s = -s / tau;
            if (fabs(t * c * s) <= tol) {
                // close poles: rotate the pair so that pj's z component vanishes
                zs[q] = tau;
                zs[pj] = 0.0;
                rp[nrot] = pj;
                rq[nrot] = q;
                rc[nrot] = c;
                rsn[nrot] = s;
                ++nrot;
                const double tt = ds[pj] * c * c + ds[q] * s * s;
                ds[q] = ds[pj] * s * s + ds[q] * c * c;
                ds[pj] = tt;
                pos[pj] = -1;
            }
            pj = q;
        }
        // positions: non-deflated first (ascending, as they stay sorted),
        // then the deflated ones
        int kp = 0;
        for (int q = 0; q < K; ++q)
            if (pos[q] >= 0) pos[q] = kp++;
        int kd = kp;
        for (int q = 0; q < K; ++q)
            if (pos[q] < 0) pos[q] = kd++;
        s_info[0] = kp;
        s_info[1] = nrot;
        int *meta = islot(ws, n, V_META) + a;
        meta[0] = kp;
        meta[1] = nrot;
        vslot(ws, n, V_RHO)[a] = rho;
        outs[mg.gemm].m = kp;
        terms[mg.gemm].k = kp;
    }
    __syncthreads();
    const int kp = s_info[0], nrot = s_info[1];
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
    // gather: row pos(q) of the block of S1 = eigenvector perm[q] restricted to
    // its own half (zero in the other half)
    double *G = S1 + (size_t)a * n + a;
    const int wave = tid >> 6, lane = tid & 63;
    for (int q = wave; q < K; q += DEFL_NT / 64) {
        const int i = perm[q];
        const double *src = Vc + (size_t)(a + i) * n + a;
        double *dst = G + (size_t)pos[q] * n;
        const int r0 = (i < h) ? 0 : h, r1 = (i < h) ? h : K;
        for (int r = lane; r < K; r += 64) dst[r] = (r >= r0 && r < r1) ? src[r] : 0.0;
    }
    __syncthreads();
    // deflation rotations, in order; each thread owns element r of every row
    if (nrot > 0) {
        for (int r = tid; r < K; r += DEFL_NT) {
            for (int t = 0; t < nrot; ++t) {
                double *x = G + (size_t)pos[rp[t]] * n + r;
                double *y = G + (size_t)pos[rq[t]] * n + r;
                const double c = rc[t], s = rsn[t];
                const double xv = *x, yv = *y;
                *x = c * xv + s * yv;
                *y = c * yv - s * xv;
            }
        }
    }
    __syncthreads();
    // deflated eigenvectors are final for this merge
    for (int p = kp + wave; p < K; p += DEFL_NT / 64) {
        const double *src = G + (size_t)p * n;
        double *dst = Vc + (size_t)(a + p) * n + a;
        for (int r = lane; r < K; r += 64) dst[r] = src[r];
    }
}

// Root j of 1/rho + sum_i zz_i^2 / (dd_i - lambda) = 0 (dd ascending,
// rho > 0), computed by one wavefront (lanes share the sums over i).  The
// root is found as tau = lambda - dd[org] where org is the nearer pole
// (decided at the middle of the interval), by a bracketed two-pole rational
// model: psi (poles <= j) and phi (poles > j) are each replaced by
// c + s / (pole - x) matching value and slope at the current point; the
// model's root (a quadratic) is taken if it lies inside the bracket,
// otherwise the bracket is bisected.
__device__ inline void secular_root(const double *dd, const double *zz, int K, double rho, int j, int lane,
                                    int &org_out, double &tau_out) {
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
        for (int i = lane; i < K; i += 64) f += zz[i] * zz[i] / ((dd[i] - dj) - mid);
        f = irho + wave_sum(f);
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
        for (int i = lane; i < K; i += 64) zt += zz[i] * zz[i];
        zt = wave_sum(zt);
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
        for (int i = lane; i < K; i += 64) {
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
        psi = wave_sum(psi);
        dpsi = wave_sum(dpsi);
        phi = wave_sum(phi);
        dphi = wave_sum(dphi);
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
__global__ __launch_bounds__(256) void k_dc_secular(const DcMerge *merges, const DcChunk *chunks) {
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

// z-hat_i^2 = -Delta_ii prod_{j != i} Delta_ji / (dd_i - dd_j) / rho
__global__ __launch_bounds__(SEC_NT) void k_dc_zhat(const DcMerge *merges, const DcChunk *chunks) {
    const DcChunk ck = chunks[blockIdx.x];
    const DcMerge mg = merges[ck.merge];
    const int n = mg.n, a = mg.a;
    double *ws = mg.ws;
    const int kp = islot(ws, n, V_META)[a];
    const int i = ck.j0 + threadIdx.x;
    if (i >= kp) return;
    const double rho = vslot(ws, n, V_RHO)[a];
    const double *dd = vslot(ws, n, V_DD) + a, *zz = vslot(ws, n, V_ZZ) + a;
    const double *S2 = ws + 2 * (size_t)n * n + (size_t)a * n + a;
    const double di = dd[i];
    double w = S2[(size_t)i * n + i];
    for (int j = 0; j < kp; ++j) {
        if (j == i) continue;
        w *= S2[(size_t)j * n + i] / (di - dd[j]);
    }
    vslot(ws, n, V_ZH)[a + i] = copysign(sqrt(fmax(-w / rho, 0.0)), zz[i]);
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

// Stable ascending rank of the final eigenvalues; W[rank] = D.
__global__ __launch_bounds__(256) void k_dc_rank(const EigMat *mats, const BtItem *items) {
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n;
    const int i = it.c0 + threadIdx.x;
    if (i >= n) return;
    const double *D = vslot(mt.ws, n, V_D);
    const double di = D[i];
    int rk = 0;
    for (int j = 0; j < n; ++j) {
        const double dj = D[j];
        rk += (dj < di) || (dj == di && j < i);
    }
    islot(mt.ws, n, V_RANK)[i] = rk;
    mt.W[rk] = di;
}

// ---------------------------------------------------------------------------
// 3. back-transformation: x_c <- H_0 H_1 ... H_{n-3} x_c for every eigenvector
//    x_c of the tridiagonal (row c of Vc); one wavefront per eigenvector, the
//    vector in registers (NS x 64 elements), reflectors read from A.
//    Output: row rank(c) of S1.
// ---------------------------------------------------------------------------
template <int NS>
__global__ __launch_bounds__(256) void k_eig_bt(const EigMat *mats, const BtItem *items) {
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n;
    const int c = it.c0 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= n) return;
    double *ws = mt.ws;
    const double *xc = ws + (size_t)c * n;
    const double *tau = vslot(ws, n, V_TAU);
    const double *A = mt.A;
    double x[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int r = 64 * k + lane;
        x[k] = (r < n) ? xc[r] : 0.0;
    }
    for (int j = n - 3; j >= 0; --j) {
        const double t = tau[j];
        if (t == 0.0) continue;
        const double *v = A + (size_t)j * n;
        double vr[NS];
        double dot = 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int r = 64 * k + lane;
            vr[k] = (r > j && r < n) ? v[r] : 0.0;
            dot += vr[k] * x[k];
        }
        dot = wave_sum(dot) * t;
#pragma unroll
        for (int k = 0; k < NS; ++k) x[k] -= dot * vr[k];
    }
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
constexpr int WY_NB = 64;
constexpr int BT_WY_MIN = 1025;   // register kernel below, WY from here
constexpr int EIG_MAX_N = 32768;

struct WyItem {
    double *ws;
    int n, j0, nb, pad;
};

// Y^T: row j = reflector j (columns j+1.., first element included), zero
// elsewhere; rows n-2, n-1 are zero.
__global__ __launch_bounds__(256) void k_wy_extract(const EigMat *mats, const BtItem *items) {
    const BtItem it = items[blockIdx.x];
    const EigMat mt = mats[it.mat];
    const int n = mt.n;
    double *Yt = mt.ws + (size_t)n * n;
    const int j = it.c0;
    const double *src = mt.A + (size_t)j * n;
    double *dst = Yt + (size_t)j * n;
    for (int r = threadIdx.x; r < n; r += 256) dst[r] = (j <= n - 3 && r > j) ? src[r] : 0.0;
}

// T of the block (dlarft forward / columnwise): T_kk = tau_k,
// T[0:k, k] = -tau_k T[0:k, 0:k] G[0:k, k], with G = Y^T Y of the block.
__global__ __launch_bounds__(WY_NB) void k_wy_tfactor(const WyItem *items) {
    __shared__ double sG[WY_NB][WY_NB + 1];
    __shared__ double sT[WY_NB][WY_NB + 1];
    const WyItem it = items[blockIdx.x];
    const int n = it.n, nb = it.nb, i = threadIdx.x;
    double *S2 = it.ws + 2 * (size_t)n * n;
    const double *G = S2;
    double *T = S2 + WY_NB * WY_NB;
    const double *tau = vslot(it.ws, n, V_TAU) + it.j0;
    for (int k = 0; k < nb; ++k) {
        sG[i][k] = (i < nb) ? G[i * WY_NB + k] : 0.0;
        sT[i][k] = 0.0;
    }
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        const double tk = tau[k];
        double t = 0.0;
        if (i < k)
            for (int l = i; l < k; ++l) t += sT[i][l] * sG[l][k];
        __syncthreads();
        if (i < k) sT[i][k] = -tk * t;
        if (i == k) sT[k][k] = tk;
        __syncthreads();
    }
    if (i < nb)
        for (int k = 0; k < nb; ++k) T[i * WY_NB + k] = sT[i][k];
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

constexpr int kBtSlots[] = {1, 2, 4, 8, 12, 16, 24, 32, 48, 64};
constexpr int kBtVariants = sizeof(kBtSlots) / sizeof(kBtSlots[0]);

template <int NS>
static void launch_bt(hipStream_t s, int nitems, const EigMat *mats, const BtItem *items) {
    hipLaunchKernelGGL(k_eig_bt<NS>, dim3(nitems), dim3(256), 0, s, mats, items);
}
static void launch_bt_variant(int v, hipStream_t s, int nitems, const EigMat *mats, const BtItem *items) {
    switch (v) {
        case 0: launch_bt<1>(s, nitems, mats, items); break;
        case 1: launch_bt<2>(s, nitems, mats, items); break;
        case 2: launch_bt<4>(s, nitems, mats, items); break;
        case 3: launch_bt<8>(s, nitems, mats, items); break;
        case 4: launch_bt<12>(s, nitems, mats, items); break;
        case 5: launch_bt<16>(s, nitems, mats, items); break;
        case 6: launch_bt<24>(s, nitems, mats, items); break;
        case 7: launch_bt<32>(s, nitems, mats, items); break;
        case 8: launch_bt<48>(s, nitems, mats, items); break;
        default: launch_bt<64>(s, nitems, mats, items); break;
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
    std::vector<nrg_eig_task> tasks(tasks_in, tasks_in + ntasks);
    std::stable_sort(tasks.begin(), tasks.end(), [](const nrg_eig_task &a, const nrg_eig_task &b) { return a.n > b.n; });
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

    // divide-and-conquer tree, level by level over all matrices
    int maxlev = 0;
    for (int i = 0; i < nbig; ++i) maxlev = std::max(maxlev, dc_levels(tasks[i].n));
    std::vector<DcMerge> merges;
    std::vector<DcChunk> schunks, uchunks;
    std::vector<nrg_gemm_out> gouts;
    std::vector<nrg_gemm_term> gterms;
    std::vector<GemmItem> gitems;
    struct Level {
        int m0, m1, s0, s1, u0, u1, g0, g1, maxK;
        double flops;
    };
    std::vector<Level> levels(maxlev + 1);
    for (int L = 1; L <= maxlev; ++L) {
        Level lv{};
        lv.m0 = (int)merges.size();
        lv.s0 = (int)schunks.size();
        lv.u0 = (int)uchunks.size();
        lv.g0 = (int)gitems.size();
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
                for (int tm = 0; tm * GT < K; ++tm)
                    for (int tn = 0; tn * GT < K; ++tn) gitems.push_back({g, tm, tn, 0});
                lv.maxK = std::max(lv.maxK, K);
                lv.flops += 2.0 * K * K * K;
            }
        }
        lv.m1 = (int)merges.size();
        lv.s1 = (int)schunks.size();
        lv.u1 = (int)uchunks.size();
        lv.g1 = (int)gitems.size();
        levels[L] = lv;
    }
    // back-transformation items by register-slot variant; rank / copy items
    std::vector<BtItem> bt[kBtVariants];
    std::vector<BtItem> rk, cp, wyrows;  // wyrows: one row per item (extract, permuted copy)
    std::vector<int> wymat;              // matrices back-transformed by WY blocks
    int wy_blocks = 0;
    for (int i = 0; i < nbig; ++i) {
        const int n = tasks[i].n, ns = (n + 63) / 64;
        if (n >= BT_WY_MIN) {
            wymat.push_back(i);
            wy_blocks = std::max(wy_blocks, (n - 2 + WY_NB - 1) / WY_NB);
            for (int r = 0; r < n; ++r) wyrows.push_back({i, r});
        } else {
            int v = 0;
            while (kBtSlots[v] < ns) ++v;
            for (int c0 = 0; c0 < n; c0 += 4) bt[v].push_back({i, c0});
            const size_t nn = (size_t)n * n;
            for (size_t e = 0; e < nn; e += 256 * 16) cp.push_back({i, (int)(e / (256 * 16))});
        }
        for (int c0 = 0; c0 < n; c0 += 256) rk.push_back({i, c0});
    }
    // WY blocks, counted from the last: block q of matrix i covers reflectors
    // [max(0, j1 - WY_NB), j1) with j1 = n - 2 - q WY_NB
    std::vector<WyItem> wyitems;
    std::vector<nrg_gemm_out> wyouts;
    std::vector<nrg_gemm_term> wyterms;
    std::vector<GemmItem> wygi;
    struct WyLevel {
        int w0, w1;
        int g[4][2];  // gemm item ranges of G, M, W, Z
        double flops;
    };
    std::vector<WyLevel> wylev(wy_blocks);
    for (int q = 0; q < wy_blocks; ++q) {
        WyLevel L{};
        L.w0 = (int)wyitems.size();
        std::vector<int> act;
        for (int i : wymat) {
            const int n = tasks[i].n, j1 = n - 2 - q * WY_NB;
            if (j1 <= 0) continue;
            const int j0 = std::max(0, j1 - WY_NB);
            wyitems.push_back({c->ews + off[i], n, j0, j1 - j0, 0});
            act.push_back(i);
        }
        L.w1 = (int)wyitems.size();
        for (int stage = 0; stage < 4; ++stage) {
            L.g[stage][0] = (int)wygi.size();
            for (int w = L.w0; w < L.w1; ++w) {
                const WyItem &it = wyitems[w];
                const int n = it.n, j0 = it.j0, nb = it.nb, np = n - j0 - 1;
                double *ws = it.ws, *Yt = ws + (size_t)n * n, *S2 = ws + 2 * (size_t)n * n;
                double *G = S2, *T = S2 + WY_NB * WY_NB, *M = S2 + 2 * WY_NB * WY_NB, *W = M + (size_t)WY_NB * n;
                const double *Yb = Yt + (size_t)j0 * n + j0 + 1;  // block rows, columns j0+1..
                nrg_gemm_out o{};
                nrg_gemm_term t{};
                t.alpha = 1.0;
                if (stage == 0) {  // G = Yb Yb^T
                    o.C = G; o.m = nb; o.n = nb; o.ldc = WY_NB;
                    t.A = Yb; t.lda = n; t.B = Yb; t.ldb = n; t.transB = 1; t.k = np;
                } else if (stage == 1) {  // M = T Yb
                    o.C = M; o.m = nb; o.n = np; o.ldc = n;
                    t.A = T; t.lda = WY_NB; t.B = Yb; t.ldb = n; t.k = nb;
                } else if (stage == 2) {  // W = Z[:, j0+1:] M^T
                    o.C = W; o.m = n; o.n = nb; o.ldc = WY_NB;
                    t.A = ws + j0 + 1; t.lda = n; t.B = M; t.ldb = n; t.transB = 1; t.k = np;
                } else {  // Z[:, j0+1:] -= W Yb
                    o.C = ws + j0 + 1; o.m = n; o.n = np; o.ldc = n; o.beta = 1.0;
                    t.A = W; t.lda = WY_NB; t.B = Yb; t.ldb = n; t.k = nb; t.alpha = -1.0;
                }
                o.term_begin = (int)wyterms.size();
                o.term_count = 1;
                const int g = (int)wyouts.size();
                wyouts.push_back(o);
                wyterms.push_back(t);
                for (int tm = 0; tm * GT < o.m; ++tm)
                    for (int tn = 0; tn * GT < o.n; ++tn) wygi.push_back({g, tm, tn, 0});
                if (stage >= 2) L.flops += 2.0 * n * (double)nb * np;
            }
            L.g[stage][1] = (int)wygi.size();
        }
        wylev[q] = L;
    }
    // tridiagonalisation jobs: ticket ranges, team sizes
    if (c->tri_capacity <= 0) {
        int per_cu = 0, dev = 0, ncu = 0;
        HIPCHK(c, hipGetDevice(&dev));
        HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(c, hipFuncSetAttribute((const void *)k_eig_tri_coop, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)TRI_COOP_LDS));
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_eig_tri_coop, TRI_NT, TRI_COOP_LDS));
        c->tri_capacity = std::max(1, per_cu * ncu);
    }
    // Team sizes, then longest-processing-time assignment of the jobs to the
    // XCD queues (work = estimated time x workgroups; a team stays on one XCD)
    const int tri_grid = c->tri_capacity;
    const int per_xcd = std::max(1, tri_grid / N_XCD);
    std::vector<int> teamT(ntasks, 1);
    std::vector<double> cost(ntasks);
    std::vector<char> global(ntasks, 0);
    for (int i = 0; i < ntasks; ++i) {
        const double n = tasks[i].n;
        const int nsl = (tasks[i].n - 1 + TEAM_W - 1) / TEAM_W;
        int T = 1;
        if (tasks[i].n > TEAM_GLOBAL_N) {
            T = std::min({nsl, TEAM_GLOBAL_MAX, tri_grid});
            global[i] = 1;
        } else if (tasks[i].n >= TEAM_MIN_N) {
            T = std::min({nsl, TEAM_MAX, per_xcd});
        }
        if (T < 2) T = 2;  // team jobs only; matrices below TEAM_MIN_N are not queued
        teamT[i] = T;
        // microseconds: one workgroup streams the trailing matrix per step;
        // a team pays a barrier per step plus its share of the matrix
        cost[i] = (T == 1) ? 0.047 * n * n : n * (3.0 + 1.1e-4 * n * n / T);
        if (tasks[i].n >= TEAM_MIN_N) mats[i].pad = 1;  // reflectors to copy from S2
    }
    std::vector<int> order(ntasks);
    for (int i = 0; i < ntasks; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<int> gq;
    std::vector<std::vector<int>> queue(N_XCD);
    std::vector<double> load(N_XCD, 0.0);
    for (int i : order)
        if (global[i]) {
            gq.push_back(i);
            for (int x = 0; x < N_XCD; ++x) load[x] += cost[i] * teamT[i] / tri_grid;
        }
    std::vector<int> ones;  // one-workgroup matrices
    for (int i : order) {
        if (global[i]) continue;
        if (tasks[i].n < TEAM_MIN_N) {
            ones.push_back(i);
            continue;
        }
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        queue[x].push_back(i);
        load[x] += cost[i] * teamT[i] / per_xcd;
    }
    std::vector<TriJob> jobs;
    std::vector<int> tbase(1, 0), qjob(1, 0);
    for (int i : gq) {
        jobs.push_back({i, teamT[i]});
        tbase.push_back(tbase.back() + teamT[i]);
    }
    qjob.push_back((int)jobs.size());
    for (int x = 0; x < N_XCD; ++x) {
        for (int i : queue[x]) {
            jobs.push_back({i, teamT[i]});
            tbase.push_back(tbase.back() + teamT[i]);
        }
        qjob.push_back((int)jobs.size());
    }
    const size_t ncnt = N_XCD + 1 + (size_t)ntasks;  // ticket counters + one barrier per job
    if (ncnt > c->cnt_ints) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->cnt) HIPCHK(c, hipFree(c->cnt));
        c->cnt = nullptr;
        c->cnt_ints = std::max<size_t>(ncnt * 2, 1024);
        HIPCHK(c, hipMalloc(&c->cnt, c->cnt_ints * sizeof(int)));
    }
    HostArena ar;
    const size_t o_jobs = ar.put(jobs), o_tb = ar.put(tbase), o_qj = ar.put(qjob), o_ones = ar.put(ones);
    const size_t o_mats = ar.put(mats), o_merges = ar.put(merges), o_sch = ar.put(schunks), o_uch = ar.put(uchunks),
                 o_gouts = ar.put(gouts), o_gterms = ar.put(gterms), o_gitems = ar.put(gitems), o_rk = ar.put(rk),
                 o_cp = ar.put(cp);
    size_t o_bt[kBtVariants];
    for (int v = 0; v < kBtVariants; ++v) o_bt[v] = ar.put(bt[v]);
    const size_t o_wyrows = ar.put(wyrows), o_wyitems = ar.put(wyitems), o_wyouts = ar.put(wyouts),
                 o_wyterms = ar.put(wyterms), o_wygi = ar.put(wygi);
    void *d;
    nrg_status s = stage(c, ar.buf.data(), ar.buf.size(), &d);
    if (s) return s;
    char *db = (char *)d;
    const EigMat *d_mats = (const EigMat *)(db + o_mats);
    const TriJob *d_jobs = (const TriJob *)(db + o_jobs);
    const int *d_tb = (const int *)(db + o_tb);
    const int *d_qj = (const int *)(db + o_qj);
    const DcMerge *d_merges = (const DcMerge *)(db + o_merges);
    const DcChunk *d_sch = (const DcChunk *)(db + o_sch), *d_uch = (const DcChunk *)(db + o_uch);
    nrg_gemm_out *d_gouts = (nrg_gemm_out *)(db + o_gouts);
    nrg_gemm_term *d_gterms = (nrg_gemm_term *)(db + o_gterms);
    const GemmItem *d_gitems = (const GemmItem *)(db + o_gitems);

    // 1. tridiagonalisation: matrices below TEAM_MIN_N one workgroup each,
    //    then one persistent launch for the teams
    {
        double fl1 = 0, by1 = 0, flt = 0, byt = 0;
        for (int i = 0; i < ntasks; ++i) {
            const double n = tasks[i].n;
            // algorithmic work: 4/3 n^3 flops; the trailing matrix is read once
            // per column (sum_j 8 (n-j)^2 bytes)
            (tasks[i].n < TEAM_MIN_N ? fl1 : flt) += 4.0 / 3.0 * n * n * n;
            (tasks[i].n < TEAM_MIN_N ? by1 : byt) += 8.0 / 3.0 * n * n * n;
        }
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, c->stream, c->cnt, (int)ncnt);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->dstatus, 1);
        HIPCHK(c, hipGetLastError());
        // the one-workgroup matrices run beside the teams on the side stream;
        // they wait for nothing, so team workgroups they delay still start
        if (!ones.empty()) {
            HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
            HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
            {
                LaunchTimer lt(K_EIG_TRI1, c->side, fl1, by1);
                hipLaunchKernelGGL(k_eig_tri_one, dim3(ones.size()), dim3(TRI_NT),
                                   5 * (size_t)TRI_LDS_N * sizeof(double), c->side, d_mats, (const int *)(db + o_ones));
            }
            HIPCHK(c, hipEventRecord(c->ev_join, c->side));
        }
        // grid <= co-resident capacity (one workgroup per CU at this size) and
        // nothing else runs on the stream, so every team assembles; the
        // barrier's bounded wait reports an error otherwise
        if (!jobs.empty()) {
            LaunchTimer lt(K_EIG_TRI, c->stream, flt, byt);
            hipLaunchKernelGGL(k_eig_tri_coop, dim3(tri_grid), dim3(TRI_NT), TRI_COOP_LDS, c->stream, d_mats, d_jobs,
                               d_tb, d_qj, c->cnt, c->dstatus);
        }
        if (!ones.empty()) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    }
    HIPCHK(c, hipGetLastError());
    if (nbig == 0) return NRG_OK;
    // 2. divide and conquer
    {
        LaunchTimer lt(K_DC_LEAF, c->stream, 0.0, 0.0);
        hipLaunchKernelGGL(k_dc_leaf, dim3(nbig), dim3(256), 0, c->stream, d_mats);
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
            LaunchTimer lt(K_DC_SECULAR, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_dc_secular, dim3(lv.u1 - lv.u0), dim3(256), 0, c->stream, d_merges, d_uch + lv.u0);
        }
        {
            LaunchTimer lt(K_DC_ZHAT, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_dc_zhat, dim3(lv.s1 - lv.s0), dim3(SEC_NT), 0, c->stream, d_merges, d_sch + lv.s0);
        }
        {
            LaunchTimer lt(K_DC_UNORM, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_dc_unorm, dim3(lv.u1 - lv.u0), dim3(256), 0, c->stream, d_merges, d_uch + lv.u0);
        }
        HIPCHK(c, hipGetLastError());
        s = gemm_launch_dev(c, K_GEMM_DC, d_gouts, d_gterms, d_gitems + lv.g0, lv.g1 - lv.g0, lv.flops, 0.0);
        if (s) return s;
    }
    // 3. sort, back-transformation, copy into A
    {
        LaunchTimer lt(K_DC_RANK, c->stream, 0.0, 0.0);
        hipLaunchKernelGGL(k_dc_rank, dim3(rk.size()), dim3(256), 0, c->stream, d_mats, (const BtItem *)(db + o_rk));
    }
    HIPCHK(c, hipGetLastError());
    {
        double fl = 0, by = 0;
        for (int i = 0; i < nbig; ++i) {
            const double n = tasks[i].n;
            if (tasks[i].n >= BT_WY_MIN) continue;
            fl += 2.0 * n * n * n;
            by += 8.0 * 2.0 * n * n;
        }
        LaunchTimer lt(K_EIG_BT, c->stream, fl, by);
        for (int v = 0; v < kBtVariants; ++v)
            if (!bt[v].empty())
                launch_bt_variant(v, c->stream, (int)bt[v].size(), d_mats, (const BtItem *)(db + o_bt[v]));
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
    if (!wymat.empty()) {
        const BtItem *d_wyrows = (const BtItem *)(db + o_wyrows);
        const WyItem *d_wyitems = (const WyItem *)(db + o_wyitems);
        const nrg_gemm_out *d_wyouts = (const nrg_gemm_out *)(db + o_wyouts);
        const nrg_gemm_term *d_wyterms = (const nrg_gemm_term *)(db + o_wyterms);
        const GemmItem *d_wygi = (const GemmItem *)(db + o_wygi);
        {
            LaunchTimer lt(K_EIG_BT, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_wy_extract, dim3(wyrows.size()), dim3(256), 0, c->stream, d_mats, d_wyrows);
        }
        HIPCHK(c, hipGetLastError());
        for (int q = 0; q < wy_blocks; ++q) {
            const WyLevel &L = wylev[q];
            s = gemm_launch_dev(c, K_GEMM_BT, d_wyouts, d_wyterms, d_wygi + L.g[0][0], L.g[0][1] - L.g[0][0], 0.0, 0.0);
            if (s) return s;
            {
                LaunchTimer lt(K_EIG_BT, c->stream, 0.0, 0.0);
                hipLaunchKernelGGL(k_wy_tfactor, dim3(L.w1 - L.w0), dim3(WY_NB), 0, c->stream, d_wyitems + L.w0);
            }
            HIPCHK(c, hipGetLastError());
            s = gemm_launch_dev(c, K_GEMM_BT, d_wyouts, d_wyterms, d_wygi + L.g[1][0], L.g[1][1] - L.g[1][0], 0.0, 0.0);
            if (s) return s;
            s = gemm_launch_dev(c, K_GEMM_BT, d_wyouts, d_wyterms, d_wygi + L.g[2][0], L.g[2][1] - L.g[2][0],
                                0.5 * L.flops, 0.0);
            if (s) return s;
            s = gemm_launch_dev(c, K_GEMM_BT, d_wyouts, d_wyterms, d_wygi + L.g[3][0], L.g[3][1] - L.g[3][0],
                                0.5 * L.flops, 0.0);
            if (s) return s;
        }
        {
            LaunchTimer lt(K_EIG_COPY, c->stream, 0.0, 0.0);
            hipLaunchKernelGGL(k_eig_copy_perm, dim3(wyrows.size()), dim3(256), 0, c->stream, d_mats, d_wyrows);
        }
        HIPCHK(c, hipGetLastError());
    }
    // the status word reaches the host through pinned memory (no copy)
    hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(64), 0, c->stream, c->dstatus, c->hstat);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (*c->hstat) return fail(c, NRG_ERR_HIP, "diag: a tridiagonalisation team did not assemble in time");
    return NRG_OK;
}
