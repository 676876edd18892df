// nrg_host.h - host side of the MI355X NRG engine (drop-in for the `nrg`
// binary of NRG Ljubljana).  Only small vectors and descriptors live here;
// every matrix of the path is a device buffer driven through nrg_gpu.h.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "nrg_gpu.h"

namespace nrg {

// Invariant subspace label: (Q, SS) for QS, (Q, SSZ) for QSZ.  std::pair's
// lexicographic order is the order of Invar::operator< (src/invar.cc:198-200),
// which fixes every loop order of the reference.
using Invar = std::pair<int, int>;
using Twoinvar = std::pair<Invar, Invar>;

// Errors end the solve, not the process: die() throws Error, which the entry
// points (nrg_main, the serving workers) catch and report with status 1, the
// exit status of the reference's my_error/exit1 (src/debug.h, portabil.h).
struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};
// A requested stop (forcestop, stopafter): the reference prints the message
// and exits with status 1 (exit1, src/nrg.cc:3113-3114, 3453, 3462).
struct Stop : Error {
    explicit Stop(const std::string &m) : Error(m) {}
};
[[noreturn]] void die(const std::string &msg);
// Size thresholds measured on the device by nrgamd/autotune.py, read from
// `<tables_dir>/tune.txt` ("NAME value" lines, # comments) into environment
// variables of the same names that are not set already (the library reads
// them when a context is created or first eigensolves).  A missing file
// leaves the built-in defaults.
void load_tuning(const std::string &tables_dir);

// ---- parameters (src/param.cc defaults, src/misc.h:71-124 parser) --------
struct Params {
    std::map<std::string, std::string> raw;
    double Lambda = 1.0, z = 1.0, bandrescale = 1.0, keepenergy = -1.0, T = -999.0, betabar = 1.0;
    std::string discretization = "Y", strategy = "kept", specd;
    int keep = 100, keepmin = 0, bins = 100, safeguardmax = 200, Nmax = 0, Ninit = 0;
    int width_td = 16, prec_td = 10, width_custom = 16, prec_custom = 10, prec_xy = 10;
    bool dmnrg = false, fdm = false, dm = false, savebins = false, broaden = true, NN2even = true,
         NN2avg = false, NN1 = false, lastall = false, lastalloverride = false, calc0 = true,
         dumpenergies = false, dumpsubspaces = false, fdmexpv = false;
    int fdmexpvn = 0;       // iteration of the FDM expectation values (src/param.cc:428)
    int forcestop = -1;     // stop at the start of this iteration of the first pass (src/param.cc:558)
    std::string stopafter;  // "nrg": after the first pass, "rho": after the density matrices (src/param.cc:560-563)
    int dumpannotated = 0;  // number of lowest states listed per iteration in annotated.dat (src/param.cc:338)
    bool dumpabs = false, dumpscaled = true;
    int dumpprecision = 8;
    bool dumpgroups = true;
    double grouptol = 1e-6;
    double goodE = 2.0, NNtanh = 0.0, safeguard = 1e-5, fixeps = 1e-15, emin = -1.0, emax = -1.0,
           discard_trim = 1e-16, discard_immediately = 1e-16, alpha = 0.3, omega0 = -1.0,
           omega0_ratio = 1.0, broaden_max = 2.5, broaden_min = -99.0, broaden_min_ratio = 3.0,
           broaden_ratio = 1.05;
    void load(const std::string &path);
};

// ---- `data` file (src/nrg.cc:3293-3401, src/read-input.cc:98-144) -------
struct HostMat {
    int rows = 0, cols = 0;
    std::vector<double> v;
};
using HostOp = std::map<Twoinvar, HostMat>;

struct Data {
    std::string sym;
    int channels = 0, Nmax = 0;
    std::map<Invar, std::vector<double>> energies;
    std::vector<HostOp> f;  // per channel
    std::map<std::string, HostOp> ops_s, ops_d;
    double GS = 0.0;
    std::vector<std::vector<double>> xi, zeta;  // per channel, 0..Nmax
    void load(const std::string &path);
};

// ---- coefficient tables (derived by nrgamd/symtab.py) --------------------
struct TabEntry {
    int a, b, ch;
    std::string kind;
    std::vector<double> vals;
    double at(int ss) const;  // value at SS (QS) / constant (QSZ)
};

struct Tables {
    std::string sym;
    int channels = 0, combs = 0, ssmax = 0;
    std::vector<Invar> In, QN;
    std::vector<std::vector<int>> occ;
    std::vector<TabEntry> offdiag, recalcf, doublet;
    void load(const std::string &path);
    int mult(const Invar &I) const { return sym == "QS" ? I.second : 1; }
    bool allowed(const Invar &I) const { return sym == "QS" ? I.second > 0 : true; }
    bool triangle(const Invar &a, const Invar &b, const Invar &c) const;
    double specdens_factor(const Invar &Ip, const Invar &I1) const;
};

// ---- device helpers -------------------------------------------------------
// Device memory comes from a pool of size classes (8 per octave): a freed
// block is kept for the next request of its class instead of being returned
// with hipFree, which would synchronise the device.  All work is on one
// stream, so a block reused by a later launch is ordered after its previous
// users.
struct Gpu {
    nrg_ctx *ctx = nullptr;
    int device = 0;
    std::multimap<size_t, double *> pool_free;
    std::map<double *, size_t> pool_size;
    explicit Gpu(int device);
    ~Gpu();
    void check(nrg_status s, const char *what) const;
    double *alloc(size_t n);
    void free(void *p);
    void h2d(double *d, const double *h, size_t n);
    void d2h(double *h, const double *d, size_t n);
    // one transfer for many (strided) device vectors, host side back to back
    void gather(const std::vector<nrg_strided> &items, double *host);
    void scatter(const std::vector<nrg_strided> &items, const double *host);
    // returns the pool's unused blocks to the device (after a spill)
    void trim();
    // bytes held by the pool and not in use
    size_t idle_bytes() const;
};

// A device context with its memory pool, lent to one solve at a time.  On
// release the context (streams, staging, workspaces) and the pool's blocks
// are kept for the next solve in the process (a few are kept; solves in a
// serving loop reuse them), so repeated solves skip stream creation and the
// device-synchronising allocations and frees.
struct GpuLease {
    std::unique_ptr<Gpu> g;
    explicit GpuLease(int device);
    ~GpuLease();
    GpuLease(const GpuLease &) = delete;
    GpuLease &operator=(const GpuLease &) = delete;
};

// Bump allocator over large device chunks, released together.
struct Arena {
    Gpu *gpu = nullptr;
    std::vector<double *> chunks;
    double *cur = nullptr;
    size_t left = 0;
    size_t chunk = size_t(1) << 22;  // doubles (32 MiB); grows to fit
    explicit Arena(Gpu *g) : gpu(g) {}
    ~Arena() { release(); }
    double *alloc(size_t n);
    void release();
};

struct DevMat {
    double *p = nullptr;
    int rows = 0, cols = 0, ld = 0;
};
using DevOp = std::map<Twoinvar, DevMat>;

// One invariant subspace of one iteration, with its (pre-truncation)
// eigenvectors resident on the device for the density-matrix sweeps.
struct Sub {
    Invar I;
    std::vector<Invar> anc;
    std::vector<int> rmax, off;
    int d = 0, kept = 0;
    double *U = nullptr;  // d x d, row r = eigenvector r
    double *W = nullptr;  // device eigenvalues (after shift/fix for use as E_prev)
    double *Eabs = nullptr;  // device absolute energies - GSenergy (pass 2)
    std::vector<double> value, absenergy;
};

struct Iter {
    int N = 0;
    double scale = 0.0, Egs = 0.0;
    std::map<Invar, Sub> subs;
    std::vector<double> sorted;    // all energies (shifted, splittings fixed), ascending
    std::unique_ptr<Arena> arena;  // U, W (one chunk)
    // spill to pinned host memory when the eigenvectors of all iterations do
    // not fit in the device budget (the reference's workdir/unitaryN files)
    double *base = nullptr;  // device chunk while resident
    size_t bytes = 0;
    void *host = nullptr;    // host copy while spilled
    bool host_pinned = true; // pinned (else pageable) host copy
    bool spilled = false;
};

// Shared eigensolver for several solves running in one process (the serving
// loop): a solve announces itself while it is in the NRG pass, and hands each
// iteration's task list to diagonalise(), which returns when the shared
// batched launch that contains it has finished.
struct DiagService {
    virtual ~DiagService() = default;
    virtual void enter() = 0;
    virtual void leave() = 0;
    virtual nrg_status diagonalise(const nrg_eig_task *tasks, int ntasks) = 0;
};

struct Options {
    std::string workdir = ".";
    std::string dump_eigen;  // binary dump of eigenvalues / kept counts
    std::string stats;       // JSON with stage timings
    bool stats_light = false;  // stats without stage timing (no stream syncs between stages)
    std::string tables_dir;
    int device = 0;
    bool quiet = false;  // no per-iteration log lines
    bool plain = false;  // one eigensolver call per block (the unbatched GPU path), for comparison
    double hbm_budget_gb = -1.0;  // device bytes for resident eigenvectors; <0: 60% of free HBM at start
    bool spill_pageable = false;  // spill to pageable host memory from the start (a test of that path)
    std::string iter_log;         // per first-pass iteration: N, seconds since start (stream drained), eigenpairs
    DiagService *diag = nullptr;  // shared eigensolver (serving loop); null: own batched calls
};

int run(const Options &opt);

}  // namespace nrg
