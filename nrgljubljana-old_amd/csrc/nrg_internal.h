// nrg_internal.h - state and helpers shared by the translation units of
// libnrg_gpu.so (context, error reporting, launch accounting, descriptor
// staging, wave/block reductions, device-side grouped GEMM launches).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nrg_gpu.h"

struct nrg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // side stream (default priority) for work that overlaps the main stream
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_wait = nullptr;  // recorded by wait_stream
    std::string err;
    // staging ring for descriptor arrays: slot k (device buffer, pinned host
    // buffer) is reused once the event recorded when the next slot was taken
    // has completed, so the host runs up to NSTAGE launches ahead of the GPU
    static constexpr int NSTAGE = 32;
    void *desc[NSTAGE] = {};
    void *hdesc[NSTAGE] = {};
    size_t desc_bytes[NSTAGE] = {};
    hipEvent_t desc_ev[NSTAGE] = {};
    bool desc_busy[NSTAGE] = {};
    int desc_cur = -1;
    // eigensolver workspace
    double *ews = nullptr;
    size_t ews_doubles = 0;
    int *dstatus = nullptr;
    // tridiagonalisation: ticket / barrier counters, co-resident workgroups
    int *cnt = nullptr;
    size_t cnt_ints = 0;
    int tri_capacity = 0;
    int reg_capacity = 0;
    // matrices from this dimension up use the blocked tridiagonalisation
    int blk_min_n = 1024;
    // blocked teams: sized by the makespan model (0), or one workgroup per
    // blk_tdiv 64-column blocks
    int blk_tdiv = 0;
    // large matrices: two-stage reduction (band, then bulge chasing) when
    // set, else the one-stage blocked reduction
    int sbr = 1;
    // co-resident bulge-chasing workgroups
    int chase_capacity = 0;
    // grouped GEMM: launches with at least this many 64 x 64 output tiles run
    // on FP64 MFMA (NRG_GEMM_MFMA_MIN_TILES at creation)
    int gemm_mfma_min_tiles = 256;
    // grouped GEMM: partial tiles of split outputs
    double *gpart = nullptr;
    size_t gpart_tiles = 0;
    // pinned host memory: gathers (written by kernels directly), the
    // eigensolver status
    int *hstat = nullptr;
    // a queued eigensolve's status word has not been read yet (wait_stream
    // reads it once the stream has drained)
    bool status_pending = false;
    // eigenvalues of the last eigensolve, copied to pinned memory as soon as
    // they are final (before the eigenvectors); ev_values follows that copy
    double *hval = nullptr;
    size_t hval_doubles = 0;
    hipEvent_t ev_values = nullptr;
    bool values_pending = false;
    size_t values_count = 0;
    double *hpin = nullptr;
    size_t hpin_doubles = 0;
    // device address of the no-allocation scratch array on this context's
    // device (hipGetSymbolAddress at creation)
    void *scratch = nullptr;
};

inline nrg_status fail(nrg_ctx *c, nrg_status code, const std::string &msg) {
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
enum KernelId {
    K_ZERO,
    K_ASSEMBLE,
    K_EIG_TRI,
    K_EIG_TRI1,
    K_EIG_TRI_REG,
    K_EIG_TRI_BLK,
    K_DC_LEAF,
    K_DC_SMALL,
    K_DC_DEFLATE,
    K_DC_SECULAR,
    K_DC_ZHAT,
    K_DC_UNORM,
    K_GEMM_DC,
    K_GEMM_BT,
    K_DC_RANK,
    K_EIG_BT,
    K_EIG_COPY,
    K_GEMM,
    K_SPEC,
    K_XFER,
    K_SBR_CHASE,
    K_SBR_BT,
    K_COUNT
};
extern const char *kKernelNames[K_COUNT];

struct PendingLaunch {
    int kid;
    hipEvent_t a, b;
    double flops, bytes;
};
extern bool g_timing;
extern int g_timing_only;  // kernel id timed alone, -1: all
extern nrg_kernel_stat g_stats[K_COUNT];
extern std::vector<PendingLaunch> g_pending;
void timing_fold();

struct LaunchTimer {
    bool on;
    PendingLaunch p;
    hipStream_t s;
    LaunchTimer(int kid, hipStream_t st, double flops, double bytes)
        : on(g_timing && (g_timing_only < 0 || g_timing_only == kid)), s(st) {
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

// Waits for an event / for everything queued on the context's stream by
// polling (the runtime's blocking waits may sleep well past completion,
// which leaves the GPU idle until the host wakes up).  wait_stream also
// reports the status of eigensolves queued since the last wait.
nrg_status wait_event(nrg_ctx *c, hipEvent_t e);
nrg_status wait_stream(nrg_ctx *c);

// Copies `bytes` of host descriptors into the next slot of the context's
// staging ring (waiting only if that slot's previous use is still queued);
// the device copy is valid for the work queued until the next stage call.
nrg_status stage(nrg_ctx *c, const void *host, size_t bytes, void **dev);

// ---------------------------------------------------------------------------
// reductions (wave64)
// ---------------------------------------------------------------------------
// DPP move of a double within rows of 16 lanes (two 32-bit moves on the VALU)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// Lane `lane` (the same in every lane) of a double, as a scalar
__device__ __forceinline__ double readlane_d(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// Sum over the 64 lanes of a wavefront (called with every lane active): pairs
// within quads and rows by DPP (quad_perm [1,0,3,2], [2,3,0,1], row half
// mirror, row mirror), then the four row sums read as scalars and added in a
// fixed order, so every lane returns the same bits.  No LDS traffic, unlike
// shuffles, which all go through the CU's LDS unit.
// Sum over each row of 16 lanes (every lane of the row active), the same
// bits in all 16 lanes: the row-local DPP steps of wave_sum.
__device__ inline double row_sum(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return v;
}
__device__ inline double row_prod(double v) {
    v *= dpp_d<0xB1>(v);
    v *= dpp_d<0x4E>(v);
    v *= dpp_d<0x141>(v);
    v *= dpp_d<0x140>(v);
    return v;
}
__device__ inline double wave_sum(double v) {
    v = row_sum(v);
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
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
template <int NT>
__device__ inline double block_max(double v, double *red) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
    for (int k = 1; k < NT / 64; ++k) s = fmax(s, red[k]);
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// grouped GEMM work items: one GT x GT output tile, or slice s of ns of its
// (term, k) sequence.  Slices write partial tiles (slot part + s of the
// context's partial buffer); one reduce item per split tile then adds them
// in slice order into C, so results do not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int GT = 32, GK = 16;
struct GemmItem {
    int out, tm, tn;
    int s, ns, part;  // ns <= 1: the tile is written directly
};
// Launch with few tiles: tiles are split until about GEMM_SPLIT_TARGET work
// items exist, each slice keeping at least GEMM_SLICE_MIN k-steps.
constexpr int GEMM_SPLIT_MIN = 384, GEMM_SPLIT_TARGET = 1024, GEMM_SLICE_MIN = 8;
// Splits the tile items [first, end) of one launch when there are few of
// them; ksteps(o) bounds the k-steps of output o.  Appends one reduce item
// per split tile; returns the partial slots used (numbered from 0).
int gemm_split(std::vector<GemmItem> &items, size_t first, const std::vector<int> &ksteps,
               std::vector<GemmItem> &reduce);
// k-steps of every output (sum over its terms of ceil(k / GK))
std::vector<int> gemm_ksteps(const nrg_gemm_out *outs, int nouts, const nrg_gemm_term *terms);
// Makes room for `slots` partial tiles (synchronises the stream if it grows).
nrg_status gemm_reserve(nrg_ctx *c, int slots);
// Launches the grouped GEMM on descriptors that already live in device
// memory (outputs may have m / k patched by earlier kernels on the stream:
// tiles past the patched m exit), then the reduction of split tiles.
nrg_status gemm_launch_dev(nrg_ctx *c, int kid, const nrg_gemm_out *d_outs, const nrg_gemm_term *d_terms,
                           const GemmItem *d_items, int nitems, const GemmItem *d_reduce, int nreduce,
                           double flops, double bytes, bool mfma = false, double *part = nullptr);
// Launches with at least `min_tiles` tiles of GTM x GTM (the context's
// gemm_mfma_min_tiles) run on FP64 MFMA (k_gemm_mfma, tiles never split);
// smaller ones on the 32 x 32 FMA tile (k_gemm, split when few).  gemm_plan
// appends the work items of the outputs [o0, outs.size()) in the chosen
// tiling (and the reduce items of split tiles, raising `slots` to the
// partial slots they need) and returns true for the MFMA kernel.
constexpr int GTM = 64;
bool gemm_plan(const std::vector<nrg_gemm_out> &outs, size_t o0, const std::vector<int> &ksteps,
               std::vector<GemmItem> &items, std::vector<GemmItem> &reduce, int &slots, int min_tiles);
