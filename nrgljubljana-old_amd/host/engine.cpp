// engine.cpp - the two-pass NRG driver of the drop-in `nrg`.
//
// Host logic follows the reference driver (src/nrg.cc): subspace
// enumeration, truncation, thermodynamics, recalculation task lists and the
// density-matrix sweeps are decided here on small host vectors; every
// matrix operation is a batched launch through nrg_gpu.h, one launch per
// stage per iteration.  Eigenvectors of every iteration stay on the device
// until the DMNRG pass has consumed them (the reference writes them to
// workdir/unitaryN, src/dmnrg.h:198-304).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <set>
#include <sstream>
#include <thread>
#include <unordered_map>

#include <mutex>

#include "nrg_host.h"

namespace nrg {

// exp(-x) in long double for x >= 0.  Past x = 11400 the result is below
// half the smallest subnormal and rounds to zero, so it is returned without
// calling expl (the Boltzmann factors of all but the last shells are there).
static inline long double expl_neg(long double x) { return x > 11400.0L ? 0.0L : expl(-x); }

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
Gpu::Gpu(int dev) : device(dev) { check(nrg_ctx_create(&ctx, dev), "nrg_ctx_create"); }
size_t Gpu::idle_bytes() const {
    size_t b = 0;
    for (auto &kv : pool_free) b += kv.first * sizeof(double);
    return b;
}

namespace {
std::mutex g_lease_m;
// never destroyed: contexts must not be torn down during static destruction
std::vector<std::unique_ptr<Gpu>> *g_idle = new std::vector<std::unique_ptr<Gpu>>();
constexpr size_t kIdleMax = 2;
}  // namespace

GpuLease::GpuLease(int device) {
    {
        std::lock_guard<std::mutex> lk(g_lease_m);
        for (auto it = g_idle->begin(); it != g_idle->end(); ++it)
            if ((*it)->device == device) {
                g = std::move(*it);
                g_idle->erase(it);
                break;
            }
    }
    if (!g) g.reset(new Gpu(device));
}
GpuLease::~GpuLease() {
    if (!g) return;
    std::unique_ptr<Gpu> drop;
    {
        std::lock_guard<std::mutex> lk(g_lease_m);
        if (g_idle->size() < kIdleMax) g_idle->push_back(std::move(g));
        else drop = std::move(g);
    }
}
Gpu::~Gpu() {
    if (ctx) (void)nrg_sync(ctx);
    for (auto &kv : pool_size) (void)nrg_dev_free(ctx, kv.first);
    nrg_ctx_destroy(ctx);
}
void Gpu::check(nrg_status s, const char *what) const {
    if (s != NRG_OK) die(std::string(what) + " failed: " + (ctx ? nrg_last_error(ctx) : "no GPU context"));
}
static size_t pool_class(size_t n) {
    size_t c = 1 << 12;
    while (c < n) c <<= 1;
    if (c <= (1 << 12)) return c;
    // 8 classes between c/2 and c
    const size_t step = c / 16;
    return (n + step - 1) / step * step;
}
double *Gpu::alloc(size_t n) {
    const size_t sz = pool_class(std::max<size_t>(n, 1));
    auto it = pool_free.find(sz);
    if (it != pool_free.end()) {
        double *p = it->second;
        pool_free.erase(it);
        return p;
    }
    void *p = nullptr;
    check(nrg_dev_alloc(ctx, sz * sizeof(double), &p), "nrg_dev_alloc");
    pool_size[(double *)p] = sz;
    return (double *)p;
}
void Gpu::free(void *p) {
    if (!p) return;
    auto it = pool_size.find((double *)p);
    if (it == pool_size.end()) die("Gpu::free: pointer not from the pool");
    pool_free.insert({it->second, it->first});
}
void Gpu::trim() {
    check(nrg_sync(ctx), "nrg_sync");
    for (auto &kv : pool_free) {
        check(nrg_dev_free(ctx, kv.second), "nrg_dev_free");
        pool_size.erase(kv.second);
    }
    pool_free.clear();
}
void Gpu::gather(const std::vector<nrg_strided> &items, double *host) {
    check(nrg_gather_d2h(ctx, items.data(), (int)items.size(), host), "nrg_gather_d2h");
}
void Gpu::scatter(const std::vector<nrg_strided> &items, const double *host) {
    check(nrg_scatter_h2d(ctx, items.data(), (int)items.size(), host), "nrg_scatter_h2d");
}
void Gpu::h2d(double *d, const double *h, size_t n) { check(nrg_h2d(ctx, d, h, n * sizeof(double)), "nrg_h2d"); }
void Gpu::d2h(double *h, const double *d, size_t n) { check(nrg_d2h(ctx, h, d, n * sizeof(double)), "nrg_d2h"); }

double *Arena::alloc(size_t n) {
    n = (n + 31) & ~size_t(31);  // 256-byte granules
    if (n > left) {
        size_t sz = std::max(chunk, n);
        cur = gpu->alloc(sz);
        chunks.push_back(cur);
        left = sz;
    }
    double *p = cur;
    cur += n;
    left -= n;
    return p;
}
void Arena::release() {
    for (double *p : chunks) gpu->free(p);
    chunks.clear();
    cur = nullptr;
    left = 0;
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
// Ascending sort of (value, location) pairs by value: LSD radix sort on
// order-preserving 64-bit keys (negative values: all bits flipped; others:
// sign bit set), one 8-bit digit per pass, passes on a digit that every key
// shares skipped.  Values are finite.
struct KeyedRef {
    uint64_t key;
    double *ref;
};
static void radix_sort(std::vector<KeyedRef> &a, std::vector<KeyedRef> &tmp) {
    const size_t m = a.size();
    tmp.resize(m);
    for (int shift = 0; shift < 64; shift += 8) {
        size_t cnt[256] = {};
        for (const KeyedRef &e : a) ++cnt[(e.key >> shift) & 255];
        if (cnt[(a[0].key >> shift) & 255] == m) continue;
        size_t off = 0;
        for (size_t &c : cnt) {
            const size_t x = c;
            c = off;
            off += x;
        }
        for (const KeyedRef &e : a) tmp[cnt[(e.key >> shift) & 255]++] = e;
        a.swap(tmp);
    }
}
static inline uint64_t order_key(double x) {
    uint64_t u;
    std::memcpy(&u, &x, sizeof u);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}

static Invar add(const Invar &a, const Invar &b) { return {a.first + b.first, a.second + b.second}; }
static Invar sub(const Invar &a, const Invar &b) { return {a.first - b.first, a.second - b.second}; }

struct GemmBatch {
    std::vector<nrg_gemm_out> outs;
    std::vector<nrg_gemm_term> terms;
    // begin an output; terms appended afterwards belong to it
    void out(double *C, int m, int n, int ldc, double beta = 0.0) {
        nrg_gemm_out o{};
        o.C = C;
        o.m = m;
        o.n = n;
        o.ldc = ldc;
        o.beta = beta;
        o.term_begin = (int)terms.size();
        o.term_count = 0;
        outs.push_back(o);
    }
    void term(const double *A, int lda, bool tA, const double *B, int ldb, bool tB, int k, double alpha,
              const double *s = nullptr) {
        nrg_gemm_term t{};
        t.A = A;
        t.B = B;
        t.s = s;
        t.alpha = alpha;
        t.k = k;
        t.lda = lda;
        t.ldb = ldb;
        t.transA = tA;
        t.transB = tB;
        terms.push_back(t);
        outs.back().term_count++;
    }
    void run(Gpu &g, const char *what) {
        if (outs.empty()) return;
        g.check(nrg_gemm_grouped(g.ctx, outs.data(), (int)outs.size(), terms.data(), (int)terms.size()), what);
        outs.clear();
        terms.clear();
    }
};

// Density matrix of one subspace: dense (dim x dim) or diagonal (vector).
struct Rho {
    int dim = 0;
    bool diag = false;
    double *p = nullptr;  // dense: dim x dim (ld = dim); diag: dim entries
};
using RhoMap = std::map<Invar, Rho>;

struct Stage {
    double t = 0.0;
    long n = 0;
};

// ---------------------------------------------------------------------------
// Engine
// ---------------------------------------------------------------------------
struct Engine {
    Options opt;
    Params P;
    Data D;
    Tables T;
    GpuLease lease;  // before every member that allocates from it
    Gpu &gpu;
    int Nmax = 0, combs = 0, nch = 0;
    std::vector<Iter> iters;
    double GSenergy = 0.0, totalenergy2 = 0.0, scale_last = 0.0;
    // eigenvector residency: budget, resident bytes, the order of later uses
    size_t budget = 0, resident = 0;
    std::vector<int> schedule;
    size_t sched_pos = 0;
    int spills = 0, restores = 0;
    // host memory holding spilled eigenvectors: now, and the most at any time
    size_t host_pinned_bytes = 0, host_pageable_bytes = 0, host_peak_bytes = 0;
    int pass1_N = -1;  // iteration being built by pass 1 (its predecessor holds E_prev)
    std::map<std::string, Stage> stages;
    long long eigenpairs = 0;

    // outputs of pass 1
    std::ofstream ftd, fcustom;
    std::vector<std::string> singlet_names;
    std::ofstream fdump, fenergies;

    // FDM statistics
    std::vector<double> ZnD, wn, wnfactor;
    // density matrices per iteration (device)
    std::vector<RhoMap> rhos, rhofdm;
    RhoMap rho_init;
    std::unique_ptr<Arena> rho_arena;

    // spectra
    struct Spec {
        std::string kind, op1, op2;
        int spin = 0;
        std::vector<double> pos, neg;  // merged bins
    };
    std::vector<Spec> spectra;
    int nbins = 0;
    double log10emin = 0, log10emax = 0;
    std::vector<double> bin_e;
    // histograms of pass 2: one slot of 2 nbins per (shell, spectrum), read
    // back once at the end and merged in the order they were filled
    double *dbins = nullptr;
    struct BinSlot {
        int N;
        bool last;
        size_t spec;
        double scale;
    };
    std::vector<BinSlot> bin_slots;
    size_t bin_capacity = 0;

    explicit Engine(const Options &o) : opt(o), lease(o.device), gpu(*lease.g) {}
    // after an error: queued work drains before the memory goes back to the
    // pool, and spilled eigenvectors release their pinned host memory
    ~Engine() {
        (void)nrg_sync(gpu.ctx);
        for (auto &it : iters) free_host(it);
    }

    double SCALE(int N) const {
        double s = 0.5 * (1.0 + 1.0 / P.Lambda);
        s *= std::pow(P.Lambda, -(N - 1) / 2.0 + 1.0 - P.z);
        return s * P.bandrescale;
    }

    // stage timing (a stream sync per stage) only when timed stats are asked for
    bool stage_timing() const { return !opt.stats.empty() && !opt.stats_light; }
    template <class F>
    void timed(const char *name, F f) {
        if (!stage_timing()) {  // keep the stream running
            f();
            return;
        }
        auto t0 = std::chrono::steady_clock::now();
        f();
        gpu.check(nrg_sync(gpu.ctx), "nrg_sync");
        auto t1 = std::chrono::steady_clock::now();
        stages[name].t += std::chrono::duration<double>(t1 - t0).count();
        stages[name].n++;
    }

    std::string path(const std::string &f) const { return opt.workdir + "/" + f; }

    // ------------------------------------------------------------ set-up
    void setup() {
        P.load(path("param"));
        D.load(path("data"));
        {
            size_t fr = 0, tot = 0;
            gpu.check(nrg_mem_info(gpu.ctx, &fr, &tot), "nrg_mem_info");
            // the pool's idle blocks (kept from an earlier solve) are available too
            fr += gpu.idle_bytes();
            budget = opt.hbm_budget_gb >= 0.0 ? (size_t)(opt.hbm_budget_gb * 1e9) : (size_t)(0.6 * (double)fr);
        }
        nch = D.channels;
        if (P.Nmax == 0) P.Nmax = D.Nmax;
        if ((int)D.xi.at(0).size() - 1 != P.Nmax) die("length of the xi table differs from Nmax");
        Nmax = P.Nmax;
        if (D.sym != "QS" && D.sym != "QSZ") die("unsupported symmetry " + D.sym);
        std::string tdir = opt.tables_dir;
        std::string tf = tdir + "/" + (D.sym == "QS" ? "qs" : "qsz") + "-" + std::to_string(nch) + "ch.tab";
        T.load(tf);
        combs = T.combs;
        if (T.channels != nch) die("table channel count mismatch");
        if (P.dm && P.T <= 0) die("dm requires T > 0");
        // spectra requested by specd (names A-B or A-B-u / A-B-d)
        std::istringstream ss(P.specd);
        std::string tok;
        while (ss >> tok) {
            for (auto &c : tok)
                if (c == ',') c = ' ';
            std::istringstream s2(tok);
            std::string name;
            while (s2 >> name) {
                std::vector<std::string> parts;
                std::string cur;
                for (char c : name) {
                    if (c == '-') {
                        parts.push_back(cur);
                        cur.clear();
                    } else cur += c;
                }
                parts.push_back(cur);
                if (parts.size() < 2) die("bad spectrum name " + name);
                int spin = 0;
                if (parts.size() == 3) spin = parts[2] == "u" ? 1 : (parts[2] == "d" ? -1 : 0);
                if (!D.ops_d.count(parts[0]) || !D.ops_d.count(parts[1])) die("unknown doublet operator in " + name);
                for (const char *k : {"DMNRG", "FDM"}) {
                    if ((std::string(k) == "DMNRG" && !P.dmnrg) || (std::string(k) == "FDM" && !P.fdm)) continue;
                    Spec sp;
                    sp.kind = k;
                    sp.op1 = parts[0];
                    sp.op2 = parts[1];
                    sp.spin = spin;
                    spectra.push_back(sp);
                }
            }
        }
        // Bins::setlimits / loggrid_std (src/bins.h:50-99)
        const double emax = P.emax > 0 ? P.emax : SCALE(0) * std::pow(10.0, 2.0);
        const double emin = P.emin > 0 ? P.emin : SCALE(Nmax) / std::pow(10.0, 2.0);
        log10emin = std::floor(std::log10(emin));
        log10emax = std::ceil(std::log10(emax));
        nbins = (int)((log10emax - log10emin) * P.bins + 1.0);
        bin_e.resize(nbins);
        for (int i = 0; i < nbins; ++i) bin_e[i] = std::pow(10.0, log10emin + (double)i / P.bins);
        for (auto &s : spectra) {
            s.pos.assign(nbins, 0.0);
            s.neg.assign(nbins, 0.0);
        }
    }

    // ------------------------------------------------------------ state
    struct Prev {
        int nr = 0;
        const double *E = nullptr;  // device, >= nr values (shifted)
    };
    struct State {
        std::map<Invar, Prev> prev;
        std::vector<DevOp> f;
        std::map<std::string, DevOp> ops_s, ops_d;
        std::unique_ptr<Arena> data_arena, op_arena, old_op_arena;
    };

    DevOp upload_op(Arena &a, const HostOp &op) {
        DevOp r;
        for (auto &kv : op) {
            DevMat m;
            m.rows = kv.second.rows;
            m.cols = kv.second.cols;
            m.ld = std::max(1, m.cols);
            m.p = a.alloc(std::max<size_t>(1, kv.second.v.size()));
            gpu.h2d(m.p, kv.second.v.data(), kv.second.v.size());
            r[kv.first] = m;
        }
        return r;
    }

    void fresh_state(State &st) {
        st.data_arena.reset(new Arena(&gpu));
        st.data_arena->chunk = size_t(1) << 20;
        st.prev.clear();
        for (auto &kv : D.energies) {
            double *e = st.data_arena->alloc(kv.second.size());
            gpu.h2d(e, kv.second.data(), kv.second.size());
            st.prev[kv.first] = Prev{(int)kv.second.size(), e};
        }
        st.f.clear();
        for (auto &op : D.f) st.f.push_back(upload_op(*st.data_arena, op));
        st.ops_s.clear();
        for (auto &kv : D.ops_s) st.ops_s[kv.first] = upload_op(*st.data_arena, kv.second);
        st.ops_d.clear();
        for (auto &kv : D.ops_d) st.ops_d[kv.first] = upload_op(*st.data_arena, kv.second);
    }

    // ------------------------------------------------------------ tasks
    // nrg_make_subspaces_list + nrg_determine_tasks (src/nrg.cc:2415-2430,
    // 2753-2781) and Rmaxvals::determine_ranges (src/nrg.cc:1479-1487).
    void determine_tasks(const std::map<Invar, Prev> &prev, Iter &it) {
        std::set<Invar> subs;
        for (auto &kv : prev) {
            if (kv.second.nr == 0) continue;
            for (int i = 0; i < combs; ++i) {
                Invar I = sub(kv.first, T.In[i]);
                if (T.allowed(I)) subs.insert(I);
            }
        }
        for (const Invar &I : subs) {
            Sub s;
            s.I = I;
            s.anc.resize(combs);
            s.rmax.assign(combs, 0);
            s.off.assign(combs + 1, 0);
            for (int i = 0; i < combs; ++i) {
                s.anc[i] = add(I, T.In[i]);
                auto p = prev.find(s.anc[i]);
                if (p != prev.end() && T.triangle(I, s.anc[i], T.QN[i])) s.rmax[i] = p->second.nr;
                s.off[i + 1] = s.off[i] + s.rmax[i];
            }
            s.d = s.off[combs];
            if (s.d > 0) it.subs[I] = std::move(s);
        }
    }

    // ------------------------------------------------------------ truncation
    // STAT::cluster_mapping (src/nrg.cc:561): value -> replacement, kept over
    // every iteration of both passes; an entry, once made, is never replaced.
    std::unordered_map<double, double> cluster_map;
    // find_clusters (src/splitting.cc:44-77) over ascending values: a cluster
    // runs while values stay less than fixeps above its first value; when a
    // later value closes it and its members are not all equal, every member
    // unlike the first maps to the first.  The last cluster is never closed.
    void find_clusters(const std::vector<KeyedRef> &e) {
        size_t i0 = 0;
        for (size_t i = 0; i < e.size(); ++i) {
            if (*e[i].ref - *e[i0].ref < P.fixeps) continue;
            for (size_t k = i0 + 1; k < i; ++k)
                if (*e[k].ref != *e[i0].ref) cluster_map.emplace(*e[k].ref, *e[i0].ref);
            i0 = i;
        }
    }
    // find_clusters + fix_splittings (src/splitting.cc:5-17; the sequence of
    // src/nrg.cc:2910-2918): clusters of this iteration's sorted values are
    // added to the map, every value found in the map is replaced, and after
    // a replacement the clusters of the new values are added as well.  One
    // radix sort of (value, location) (two after a replacement); the fixed
    // values, ascending, are kept in it.sorted for the truncation.
    std::vector<KeyedRef> sort_buf, sort_tmp;
    void fix_splittings(Iter &it) {
        std::vector<KeyedRef> &e = sort_buf;
        e.clear();
        for (auto &kv : it.subs)
            for (double &x : kv.second.value) e.push_back({order_key(x), &x});
        if (e.empty()) return;
        radix_sort(e, sort_tmp);
        find_clusters(e);
        bool changed = false;
        if (!cluster_map.empty())
            for (KeyedRef &k : e) {
                auto m = cluster_map.find(*k.ref);
                if (m == cluster_map.end()) continue;
                *k.ref = m->second;
                k.key = order_key(m->second);
                changed = true;
            }
        if (changed) {
            radix_sort(e, sort_tmp);
            find_clusters(e);
        }
        it.sorted.resize(e.size());
        for (size_t i = 0; i < e.size(); ++i) it.sorted[i] = *e[i].ref;
    }

    // highest_retained_energy + nrg_truncate_prepare (src/nrg.cc:1919-1998)
    void truncate_prepare(Iter &it, bool last, bool dmnrgrun) {
        if (it.sorted.empty()) {
            for (auto &kv : it.subs) it.sorted.insert(it.sorted.end(), kv.second.value.begin(), kv.second.value.end());
            std::sort(it.sorted.begin(), it.sorted.end());
        }
        const std::vector<double> &e = it.sorted;
        const size_t total = e.size();
        size_t nrkeep;
        if (P.keepenergy <= 0.0) nrkeep = P.keep;
        else {
            nrkeep = 1 + std::count_if(e.begin(), e.end(), [&](double x) { return x <= P.keepenergy; });
            nrkeep = std::min<size_t>(std::max<size_t>(nrkeep, P.keepmin), P.keep);
        }
        if (P.safeguard > 0.0) {
            int extra = 0;
            while (nrkeep < total && (e[nrkeep] - e[nrkeep - 1]) <= P.safeguard && extra < P.safeguardmax) {
                nrkeep++;
                extra++;
            }
        }
        nrkeep = std::min(std::max<size_t>(nrkeep, 1), total);
        const double Emax = e[nrkeep - 1];
        for (auto &kv : it.subs) {
            Sub &s = kv.second;
            int c = (int)std::count_if(s.value.begin(), s.value.end(), [&](double x) { return x <= Emax; });
            if (P.fdm && last && !P.lastalloverride) c = (int)s.value.size();
            if (dmnrgrun && last && P.lastall) c = (int)s.value.size();
            s.kept = c;
        }
    }

    // ------------------------------------------------------------ TD
    // nrg_calculate_TD (src/nrg.cc:2333-2369) + SymmetryQS/QSZ::calculate_TD
    void td_row(const std::map<Invar, std::vector<double>> &vals, double scale_now) {
        const double rf = P.betabar;
        double Z = 0, E = 0, E2 = 0, a = 0, b = 0, c = 0;
        for (auto &kv : vals) {
            double sZ = 0, sE = 0, sE2 = 0;
            for (double x : kv.second) {
                const double be = rf * x, ex = std::exp(-be);
                sE += be * ex;
                sE2 += be * be * ex;
                sZ += ex;
            }
            const int m = T.mult(kv.first);
            Z += m * sZ;
            E += m * sE;
            E2 += m * sE2;
            const double sumZ = m * sZ;
            const int q = kv.first.first, ss = kv.first.second;
            if (T.sym == "QS") {
                a += sumZ * (ss * ss - 1) / 12.0;
                b += sumZ * q;
                c += sumZ * q * q;
            } else {
                a += sumZ * (ss / 2.0);
                b += sumZ * (ss / 2.0) * (ss / 2.0);
                c += sumZ * q;
            }
        }
        const double Em = E / Z, E2m = E2 / Z;
        std::vector<double> row;
        row.push_back(scale_now / P.betabar);
        row.push_back(a / Z);
        row.push_back(b / Z);
        row.push_back(c / Z);
        row.push_back(Em);
        row.push_back(E2m);
        row.push_back(E2m - Em * Em);
        row.push_back(-std::log(Z));
        row.push_back(Em + std::log(Z));
        ftd << ' ';
        for (double v : row) {
            std::ostringstream s;
            s << std::setw(P.width_td) << std::setprecision(P.prec_td) << v;
            ftd << std::setw(P.width_td) << s.str() << ' ';
        }
        ftd << '\n';
    }

    void open_td() {
        ftd.open(path("td"));
        std::vector<std::string> names = {"T"};
        if (T.sym == "QS") names.insert(names.end(), {"<Sz^2>", "<Q>", "<Q^2>"});
        else names.insert(names.end(), {"<Sz>", "<Sz^2>", "<Q>"});
        names.insert(names.end(), {"<E>", "<E^2>", "C", "F", "S"});
        ftd << '#';
        for (auto &n : names) ftd << std::setw(P.width_td) << n << ' ';
        ftd << '\n';
    }

    // custom: expectation values of singlet operators (src/nrg.cc:946-971,
    // 2218-2235; ExpvOutput src/nrg.cc:1508-1546)
    void open_custom() {
        for (auto &kv : D.ops_s) singlet_names.push_back(kv.first);
        fcustom.open(path("custom"));
        auto fo = [&](auto x) { fcustom << std::setw(P.width_custom) << std::setprecision(P.prec_custom) << x << ' '; };
        fcustom << '#';
        fo(1);
        for (size_t k = 1; k <= singlet_names.size(); ++k) fo(1 + k);
        fcustom << '\n' << '#';
        fo("T");
        for (auto &n : singlet_names) fo(n);
        fcustom << '\n';
    }

    // A row of the custom file (calc_trace_singlet, src/nrg.cc:1204-1222):
    // the diagonals of the singlet operators are collected on the device now
    // (the operators are replaced by the next recalculation) and the rows are
    // written at the end of the pass, so that the pass never waits for them.
    struct CustomRow {
        double Tnow;
        std::map<Invar, std::vector<double>> vals;
        double *dg;  // device copy of the diagonals
        size_t n;
    };
    std::vector<CustomRow> custom_rows;
    std::unique_ptr<Arena> custom_arena;

    void custom_row(double Tnow, const std::map<Invar, std::vector<double>> &vals,
                    const std::map<std::string, DevOp> &ops) {
        std::vector<nrg_strided> items;
        size_t tot = 0;
        for (auto &name : singlet_names) {
            const DevOp &op = ops.at(name);
            for (auto &kv : vals) {
                const int n = (int)kv.second.size();
                if (!n) continue;
                auto it = op.find({kv.first, kv.first});
                if (it == op.end()) die("calc_trace_singlet: missing block");
                items.push_back({it->second.p, n, it->second.ld + 1});
                tot += n;
            }
        }
        if (!custom_arena) custom_arena.reset(new Arena(&gpu));
        double *dg = tot ? custom_arena->alloc(tot) : nullptr;
        if (tot) gpu.check(nrg_gather_dev(gpu.ctx, items.data(), (int)items.size(), dg), "nrg_gather_dev");
        custom_rows.push_back({Tnow, vals, dg, tot});
    }

    void flush_custom() {
        std::vector<nrg_strided> items;
        size_t tot = 0;
        for (auto &r : custom_rows)
            if (r.n) {
                items.push_back({r.dg, (int64_t)r.n, 1});
                tot += r.n;
            }
        std::vector<double> all(tot);
        gpu.gather(items, all.data());
        size_t pos = 0;
        auto fo = [&](auto x) { fcustom << std::setw(P.width_custom) << std::setprecision(P.prec_custom) << x << ' '; };
        for (auto &r : custom_rows) {
            double Z = 0;
            for (auto &kv : r.vals)
                for (double x : kv.second) Z += T.mult(kv.first) * std::exp(-P.betabar * x);
            fcustom << ' ';
            fo(r.Tnow);
            for (size_t k = 0; k < singlet_names.size(); ++k) {
                double tr = 0;
                for (auto &kv : r.vals) {
                    const int n = (int)kv.second.size();
                    if (!n) continue;
                    double s = 0;
                    for (int i = 0; i < n; ++i) s += std::exp(-P.betabar * kv.second[i]) * all[pos + i];
                    pos += n;
                    tr += T.mult(kv.first) * s;
                }
                fo(tr / Z);
            }
            fcustom << '\n';
        }
        custom_rows.clear();
        custom_arena.reset();
    }

    // ------------------------------------------------------------ recalc
    // recalc_f (src/recalc.cc:12-97) driven by Symmetry::recalc_irreduc
    // (src/nrg-recalc-QS.cc.m4:61-103): one grouped GEMM for all pairs.
    std::vector<DevOp> recalc_f(Iter &it, Arena &a) {
        std::vector<DevOp> nf(nch);
        GemmBatch gb;
        for (auto &kp : it.subs) {
            const Sub &Sp = kp.second;
            const Invar Ip = kp.first;
            for (int dir : {+1, -1}) {
                const std::string kind = dir > 0 ? "up" : "down";
                const Invar I1 = T.sym == "QS" ? Invar{Ip.first + 1, Ip.second + dir} : Invar{Ip.first + 1, Ip.second + dir};
                auto f1 = it.subs.find(I1);
                if (f1 == it.subs.end()) continue;
                const Sub &S1 = f1->second;
                if (S1.kept == 0 || Sp.kept == 0) continue;
                for (int ch = 0; ch < nch; ++ch) {
                    DevMat F;
                    F.rows = S1.kept;
                    F.cols = Sp.kept;
                    F.ld = Sp.kept;
                    F.p = a.alloc((size_t)F.rows * F.cols);
                    gb.out(F.p, F.rows, F.cols, F.ld);
                    for (const TabEntry &e : T.recalcf) {
                        if (e.ch != ch || e.kind != kind) continue;
                        const int r1 = S1.rmax[e.a], rp = Sp.rmax[e.b];
                        if (!(r1 > 0 && rp > 0)) continue;
                        if (S1.anc[e.a] != Sp.anc[e.b] || r1 != rp) die("recalc_f: ancestor mismatch");
                        const double fac = e.at(Ip.second);
                        if (!std::isfinite(fac) || std::fabs(fac) >= 1000) die("recalc_f: bad factor");
                        gb.term(S1.U + S1.off[e.a], S1.d, false, Sp.U + Sp.off[e.b], Sp.d, true, r1, fac);
                    }
                    nf[ch][{I1, Ip}] = F;
                }
            }
        }
        gb.run(gpu, "recalc_f");
        return nf;
    }

    // recalc_general (src/recalc.cc:164-281) for the singlet
    // (src/nrg.cc:2112-2142) and doublet (src/nrg-recalc-QS.cc.m4:17-54)
    // tables; `nstate` gives the number of states (kept or all) per subspace.
    struct RTerm {
        int i1, ip;
        Invar IN1, INp;
        double fac;
    };
    DevOp recalc_general_all(Iter &it, const DevOp &cold, bool doublet, bool all_states, Arena &a, Arena &tmp) {
        DevOp cnew;
        GemmBatch st1, st2;
        struct Pend {
            double *C;
            int d1, dp, ldc;
            std::vector<std::pair<std::pair<double *, int>, std::pair<int, double>>> t;  // (T, r1) ; (i1, fac)
        };
        auto nst = [&](const Sub &s) { return all_states ? s.d : s.kept; };
        for (auto &k1 : it.subs) {
            const Invar I1 = k1.first;
            const Sub &S1 = k1.second;
            std::vector<std::pair<Invar, std::vector<RTerm>>> targets;
            if (doublet) {
                for (int dir : {+1, -1}) {
                    const Invar Ip{I1.first - 1, I1.second + dir};
                    if (!it.subs.count(Ip)) continue;
                    std::vector<RTerm> tab;
                    const std::string kind = dir > 0 ? "p" : "m";
                    for (const TabEntry &e : T.doublet) {
                        if (e.kind != kind) continue;
                        const Sub &Sp = it.subs.at(Ip);
                        tab.push_back({e.a, e.b, S1.anc[e.a], Sp.anc[e.b], e.at(I1.second)});
                    }
                    targets.push_back({Ip, tab});
                }
            } else {
                std::vector<RTerm> tab;
                for (int i = 0; i < combs; ++i) tab.push_back({i, i, S1.anc[i], S1.anc[i], 1.0});
                targets.push_back({I1, tab});
            }
            for (auto &tg : targets) {
                const Invar Ip = tg.first;
                const Sub &Sp = it.subs.at(Ip);
                Invar Iop = doublet ? (T.sym == "QS" ? Invar{1, 2} : Invar{1, I1.second - Ip.second})
                                    : (T.sym == "QS" ? Invar{0, 1} : Invar{0, 0});
                if (!T.triangle(I1, Ip, Iop)) continue;
                const int d1 = nst(S1), dp = nst(Sp);
                DevMat C;
                C.rows = d1;
                C.cols = dp;
                C.ld = std::max(1, dp);
                C.p = a.alloc(std::max<size_t>(1, (size_t)d1 * dp));
                cnew[{I1, Ip}] = C;
                if (d1 == 0 || dp == 0) continue;
                st2.out(C.p, d1, dp, C.ld);
                for (const RTerm &t : tg.second) {
                    if (!T.allowed(t.IN1) || !T.allowed(t.INp)) continue;
                    const int r1 = S1.rmax[t.i1], rp = Sp.rmax[t.ip];
                    if (r1 == 0 || rp == 0) continue;
                    if (S1.anc[t.i1] != t.IN1 || Sp.anc[t.ip] != t.INp) die("recalc_general: ancestor mismatch");
                    auto m = cold.find({t.IN1, t.INp});
                    const bool tri = T.triangle(t.IN1, Iop, t.INp);
                    if (!tri && m == cold.end()) continue;
                    if (m == cold.end()) die("recalc_general: cold.count(ININ) != 1");
                    if (!std::isfinite(t.fac) || std::fabs(t.fac) >= 1000) die("recalc_general: bad factor");
                    if (m->second.rows != r1 || m->second.cols != rp) die("recalc_general: block size mismatch");
                    // stage 1: Tm = m (r1 x rp) . Up_blk^T (rp x dp)
                    double *Tm = tmp.alloc((size_t)r1 * dp);
                    st1.out(Tm, r1, dp, dp);
                    st1.term(m->second.p, m->second.ld, false, Sp.U + Sp.off[t.ip], Sp.d, true, rp, 1.0);
                    // stage 2: C += fac U1_blk (d1 x r1) . Tm
                    st2.term(S1.U + S1.off[t.i1], S1.d, false, Tm, dp, false, r1, t.fac);
                }
            }
        }
        st1.run(gpu, "recalc stage 1");
        st2.run(gpu, "recalc stage 2");
        return cnew;
    }

    // ------------------------------------------------------------ residency
    // Order in which iterations' eigenvectors are used after pass 1: the
    // DM-NRG and FDM sweeps go down from the last shell, pass 2 goes up.
    void plan_schedule() {
        schedule.clear();
        sched_pos = 0;
        if (!P.dm) return;
        for (int k = 0; k < (P.dmnrg ? 1 : 0) + (P.fdm ? 1 : 0); ++k)
            for (int N = Nmax - 1; N > P.Ninit; --N) schedule.push_back(N);
        for (int N = P.Ninit; N < Nmax; ++N) schedule.push_back(N);
    }
    size_t next_use(int N) const {
        for (size_t k = sched_pos; k < schedule.size(); ++k)
            if (schedule[k] == N) return k;
        return SIZE_MAX;
    }
    // The host copy is pinned while pinned memory lasts (the runtime refuses
    // it beyond a limit of its own, measured at about 66 GB on the MI355X
    // box; spills stop at kPinnedSpillMax so that the small pinned buffers
    // the library grows later still fit), then ordinary pageable memory,
    // which the runtime copies through staging buffers.
    static constexpr size_t kPinnedSpillMax = (size_t)48 << 30;
    void spill(Iter &it) {
        it.host = nullptr;
        if (host_pinned_bytes + it.bytes > kPinnedSpillMax) pinned_exhausted = true;
        it.host_pinned = !pinned_exhausted && nrg_host_alloc(gpu.ctx, it.bytes, &it.host) == NRG_OK;
        if (!it.host_pinned) {
            pinned_exhausted = true;
            it.host = std::malloc(it.bytes);
            if (!it.host) die("spill: out of host memory for " + std::to_string(it.bytes) + " bytes of eigenvectors");
        }
        (it.host_pinned ? host_pinned_bytes : host_pageable_bytes) += it.bytes;
        host_peak_bytes = std::max(host_peak_bytes, host_pinned_bytes + host_pageable_bytes);
        gpu.check(nrg_d2h(gpu.ctx, it.host, it.base, it.bytes), "nrg_d2h");
        it.arena.reset();
        gpu.trim();
        it.spilled = true;
        resident -= it.bytes;
        ++spills;
    }
    void restore(Iter &it) {
        double *nb = gpu.alloc(it.bytes / 8);
        gpu.check(nrg_h2d(gpu.ctx, nb, it.host, it.bytes), "nrg_h2d");
        free_host(it);
        for (auto &kv : it.subs) {
            kv.second.U = nb + (kv.second.U - it.base);
            kv.second.W = nb + (kv.second.W - it.base);
        }
        it.arena.reset(new Arena(&gpu));
        it.arena->chunks.push_back(nb);  // released with the arena
        it.base = nb;
        it.spilled = false;
        resident += it.bytes;
        ++restores;
    }
    void free_host(Iter &it) {
        if (!it.host) return;
        (it.host_pinned ? host_pinned_bytes : host_pageable_bytes) -= it.bytes;
        if (it.host_pinned) (void)nrg_host_free(gpu.ctx, it.host);
        else std::free(it.host);
        it.host = nullptr;
    }
    bool pinned_exhausted = opt.spill_pageable;  // a pinned allocation failed: pageable from then on
    // Spill resident iterations (the one used last first) until `need` more
    // bytes fit in the budget; iteration `keep` and, in pass 1, its
    // predecessor (E_prev) stay.
    // Besides the budget, the device itself must hold `need` plus a reserve
    // for the data outside the eigenvector arenas (operators, density
    // matrices, spectral temporaries, workspaces): HBM left free below that
    // also spills.
    static constexpr size_t kHbmReserve = (size_t)8 << 30;
    size_t free_hbm() {
        size_t fr = 0, tot = 0;
        gpu.check(nrg_mem_info(gpu.ctx, &fr, &tot), "nrg_mem_info");
        return fr + gpu.idle_bytes();
    }
    void make_room(size_t need, int keep, size_t reserve = kHbmReserve) {
        while (resident + need > budget || free_hbm() < need + reserve) {
            int victim = -1;
            size_t far = 0;
            for (int N = 0; N < (int)iters.size(); ++N) {
                const Iter &c = iters[N];
                if (!c.base || c.spilled || N == keep || (pass1_N >= 0 && N >= pass1_N - 1)) continue;
                const size_t u = schedule.empty() ? SIZE_MAX - (size_t)N : next_use(N);
                if (victim < 0 || u > far) {
                    victim = N;
                    far = u;
                }
            }
            if (victim < 0) return;  // nothing left to spill: allocate over budget
            spill(iters[victim]);
        }
    }
    // iteration N is about to be used by a sweep or by pass 2
    // `reserve`: HBM the step will allocate besides the eigenvectors.  A
    // shell of pass 2 builds its operators over all states (blocks of the
    // full subspace dimensions, the previous shell's still alive during the
    // recalculation) and the spectral products: about twelve times the
    // shell's eigenvector bytes.
    void use(int N, bool pass2 = false) {
        while (sched_pos < schedule.size() && schedule[sched_pos] != N) ++sched_pos;
        Iter &it = iters[N];
        const size_t reserve = pass2 ? std::max(kHbmReserve, 12 * it.bytes) : kHbmReserve;
        make_room(it.spilled ? it.bytes : 0, N, reserve);
        if (it.spilled) restore(it);
        if (sched_pos < schedule.size()) ++sched_pos;
    }

    // ------------------------------------------------------------ pass 1
    void write_dump_iter(const Iter &it) {
        if (!fdump.is_open()) return;
        int32_t hdr[2] = {it.N, (int32_t)it.subs.size()};
        fdump.write((const char *)hdr, sizeof hdr);
        for (auto &kv : it.subs) {
            int32_t h[4] = {kv.first.first, kv.first.second, kv.second.d, kv.second.kept};
            fdump.write((const char *)h, sizeof h);
            fdump.write((const char *)kv.second.value.data(), sizeof(double) * kv.second.value.size());
        }
    }

    // energies.nrg / energies.dmnrg (dumptofile, src/nrg.cc:1435-1443): all
    // (pre-truncation) energies of an iteration, one line per subspace
    static void dump_energies(std::ofstream &f, int N, const std::map<Invar, std::vector<double>> &vals) {
        if (!f.is_open()) return;
        f << "\n===== Iteration number: " << N << "\n";
        for (auto &kv : vals) {
            f << "Subspace: " << kv.first.first << " " << kv.first.second << "\n";
            for (double x : kv.second) f << x << ' ';
            f << "\n";
        }
    }

    // my_fcmp (src/numerics.h:37-53): equal if both are zero, both below
    // eps, or relatively closer than eps
    static int fcmp(double x, double y, double eps) {
        if (x == 0.0 && y == 0.0) return 0;
        if (std::fabs(x) < eps && std::fabs(y) < eps) return 0;
        if (std::fabs(x - y) / (std::fabs(x) + std::fabs(y)) < eps) return 0;
        return x > y ? 1 : -1;
    }

    // annotated.dat (dump_annotated, src/nrg.cc:2062-2107): the lowest
    // dumpannotated levels of an iteration (extended over a cluster at the
    // cut), grouped by degeneracy with their quantum numbers and total
    // multiplicity; iterations separated by an empty line
    std::ofstream fannotated;
    void dump_annotated(const std::map<Invar, std::vector<double>> &vals, double scale, double totalE) {
        if (!fannotated.is_open()) return;
        std::vector<std::pair<double, Invar>> lst;
        for (auto &kv : vals)
            for (double e : kv.second) lst.push_back({e, kv.first});
        std::sort(lst.begin(), lst.end());
        size_t len = std::min(lst.size(), (size_t)P.dumpannotated);
        for (size_t i = len; i + 1 < lst.size(); ++i) {  // the reference stops before the last level
            if (fcmp(lst[i].first, lst[i - 1].first, P.grouptol) == 0) ++len;
            else break;
        }
        auto en = [&](double e) { return e * (P.dumpscaled ? 1.0 : scale) + (P.dumpabs ? totalE : 0.0); };
        auto qn = [](const Invar &I) { return std::to_string(I.first) + " " + std::to_string(I.second); };
        fannotated << std::setprecision(P.dumpprecision);
        if (P.dumpgroups) {
            for (size_t i = 0; i < len;) {
                fannotated << en(lst[i].first);
                std::vector<std::string> q;
                long deg = 0;
                const size_t i0 = i;
                while (i < len && fcmp(lst[i].first, lst[i0].first, P.grouptol) == 0) {
                    q.push_back(qn(lst[i].second));
                    deg += T.mult(lst[i].second);
                    ++i;
                }
                std::sort(q.begin(), q.end());
                for (auto &x : q) fannotated << " (" << x << ")";
                fannotated << " [" << deg << "]\n";
            }
        } else {
            for (auto &x : lst) fannotated << en(x.first) << " " << qn(x.second) << "\n";
        }
        fannotated << "\n";
    }

    // subspaces.dat (dump_subspace_information, src/nrg.cc:3207-3218)
    void dump_subspaces() {
        std::ofstream o(path("subspaces.dat"));
        for (int N = P.Ninit; N < Nmax; ++N) {
            o << "Iteration " << N << "\n";
            o << "len_dm=" << iters[N].subs.size() << "\n";
            for (auto &kv : iters[N].subs)
                o << "I=" << kv.first.first << " " << kv.first.second << " len=" << kv.second.kept
                  << " kept=" << kv.second.kept << " total=" << kv.second.d << "\n";
            o << "\n";
        }
    }

    // customfdm (nrg_measure_singlet1_fdm, calc_trace_fdm_kept,
    // src/nrg.cc:973-991, 2239-2247, 3241-3245): <O> = sum_I mult(I)
    // sum_{i,j<kept} rho_FDM[I](i,j) O[I](j,i) at iteration fdmexpvn, with
    // the singlet operators' diagonal blocks kept from the first pass
    std::map<std::string, std::map<Invar, std::vector<double>>> fdmexpv_ops;  // kept x kept blocks
    void keep_fdmexpv_ops(const Iter &it, const std::map<std::string, DevOp> &ops) {
        for (auto &name : singlet_names) {
            const DevOp &op = ops.at(name);
            for (auto &kv : it.subs) {
                const int k = kv.second.kept;
                std::vector<double> m((size_t)k * k);
                if (k) {
                    auto b = op.find({kv.first, kv.first});
                    if (b == op.end()) die("fdmexpv: missing block");
                    std::vector<nrg_strided> rows;
                    for (int r = 0; r < k; ++r) rows.push_back({b->second.p + (size_t)r * b->second.ld, k, 1});
                    gpu.gather(rows, m.data());
                }
                fdmexpv_ops[name][kv.first] = std::move(m);
            }
        }
    }
    void write_customfdm() {
        const int N = P.fdmexpvn;
        if (N < P.Ninit || N >= Nmax - 1) die("fdmexpvn must lie in [Ninit, Nmax-1)");
        std::map<Invar, std::vector<double>> rho;
        for (auto &kv : rhofdm[N]) {
            const int k = kv.second.dim;
            std::vector<double> m((size_t)k * k, 0.0);
            if (k) {
                if (kv.second.diag) {
                    std::vector<double> d(k);
                    gpu.gather({{kv.second.p, k, 1}}, d.data());
                    for (int i = 0; i < k; ++i) m[(size_t)i * k + i] = d[i];
                } else {
                    gpu.gather({{kv.second.p, (int64_t)k * k, 1}}, m.data());
                }
            }
            rho[kv.first] = std::move(m);
        }
        std::ofstream f(path("customfdm"));
        auto fo = [&](auto x) { f << std::setw(P.width_custom) << std::setprecision(P.prec_custom) << x << ' '; };
        f << '#';
        fo(1);
        for (size_t k = 1; k <= singlet_names.size(); ++k) fo(1 + k);
        f << '\n' << '#';
        fo("T");
        for (auto &n : singlet_names) fo(n);
        f << '\n' << ' ';
        fo(P.T);
        for (auto &name : singlet_names) {
            double tr = 0.0;
            for (auto &kv : iters[N].subs) {
                const int k = kv.second.kept;
                const std::vector<double> &r = rho.at(kv.first), &o = fdmexpv_ops.at(name).at(kv.first);
                double sum = 0.0;
                for (int i = 0; i < k; ++i)
                    for (int j = 0; j < k; ++j) sum += r[(size_t)i * k + j] * o[(size_t)j * k + i];
                tr += T.mult(kv.first) * sum;
            }
            fo(tr);
        }
        f << '\n';
    }

    // first-pass iteration log for timing (stream drained at each line)
    std::ofstream fiterlog;
    std::chrono::steady_clock::time_point t_start;

    std::function<void()> deferred;  // pass 1: host-side output of the previous iteration
    void pass1() {
        State st;
        fresh_state(st);
        open_td();
        open_custom();
        if (P.dumpenergies) fenergies.open(path("energies.nrg"));
        if (P.dumpannotated > 0) fannotated.open(path("annotated.dat"));
        if (!opt.iter_log.empty()) fiterlog.open(opt.iter_log);
        double totalenergy = D.GS;
        iters.clear();
        iters.resize(Nmax);
        if (P.calc0) {
            // docalc0 (src/nrg.cc:3064-3075): measurements on the initial
            // cluster, labelled iteration Ninit - 1
            std::map<Invar, std::vector<double>> v0(D.energies.begin(), D.energies.end());
            td_row(v0, SCALE(0));
            dump_annotated(v0, SCALE(0), totalenergy);
            custom_row(SCALE(0) / P.betabar, v0, st.ops_s);
            dump_energies(fenergies, P.Ninit - 1, v0);
        }
        for (int N = P.Ninit; N < Nmax; ++N) {
            const bool last = (N + 1 == Nmax);
            if (P.forcestop == N) {  // src/nrg.cc:3113-3114
                if (deferred) {
                    deferred();
                    deferred = nullptr;
                }
                flush_custom();
                throw Stop("*** Stop request at iteration " + std::to_string(N));
            }
            Iter &it = iters[N];
            it.N = N;
            it.scale = SCALE(N + 1);
            it.arena.reset(new Arena(&gpu));
            determine_tasks(st.prev, it);
            if (!opt.quiet && !it.subs.empty()) {
                // sort_task_list's report (src/nrg.cc:2792-2795)
                int mn = 1 << 30, mx = 0;
                for (auto &kv : it.subs) {
                    mn = std::min(mn, kv.second.d);
                    mx = std::max(mx, kv.second.d);
                }
                std::cout << "Stats: nr=" << it.subs.size() << " min=" << mn << " max=" << mx << std::endl;
            }
            const auto t_iter = std::chrono::steady_clock::now();
            size_t tot = 0;
            for (auto &kv : it.subs) tot += (size_t)kv.second.d * kv.second.d + kv.second.d;
            it.arena->chunk = std::max<size_t>(tot + 64 * it.subs.size(), 1 << 16);
            it.bytes = 8 * it.arena->chunk;
            pass1_N = N;
            make_room(it.bytes, N);
            for (auto &kv : it.subs) {
                kv.second.U = it.arena->alloc((size_t)kv.second.d * kv.second.d);
                kv.second.W = it.arena->alloc(kv.second.d);
            }
            if (it.arena->chunks.size() != 1) die("iteration arena: expected one chunk");
            it.base = it.arena->chunks[0];
            resident += it.bytes;
            // ---- assembly + eigensolve
            timed("diag", [&] {
                std::vector<nrg_asm_block> blocks;
                std::vector<nrg_asm_term> terms;
                const double sf = std::sqrt(P.Lambda);
                for (auto &kv : it.subs) {
                    Sub &s = kv.second;
                    nrg_asm_block b{};
                    b.H = s.U;
                    b.d = s.d;
                    b.term_begin = (int)terms.size();
                    for (int i = 0; i < combs; ++i) {
                        if (!s.rmax[i]) continue;
                        nrg_asm_term t{};
                        t.kind = NRG_TERM_DIAGVEC;
                        t.src = st.prev.at(s.anc[i]).E;
                        t.scale = sf;
                        t.r0 = s.off[i];
                        t.rows = s.rmax[i];
                        terms.push_back(t);
                    }
                    for (const TabEntry &e : T.offdiag) {
                        if (!(s.rmax[e.a] && s.rmax[e.b])) continue;
                        const double fac = e.at(kv.first.second);
                        if (!std::isfinite(fac)) die("offdiag factor is not finite");
                        const double fs = fac * D.xi[e.ch][N] / it.scale;
                        if (std::fabs(fs) < 1e-8) continue;
                        auto m = st.f[e.ch].find({s.anc[e.a], s.anc[e.b]});
                        if (m == st.f[e.ch].end()) die("offdiag: <||f||> subspace does not exist");
                        if (m->second.rows != s.rmax[e.a] || m->second.cols != s.rmax[e.b]) die("offdiag: size mismatch");
                        nrg_asm_term t{};
                        t.kind = NRG_TERM_BLOCK;
                        t.src = m->second.p;
                        t.scale = fs;
                        t.r0 = s.off[e.a];
                        t.c0 = s.off[e.b];
                        t.rows = s.rmax[e.a];
                        t.cols = s.rmax[e.b];
                        t.ld = m->second.ld;
                        terms.push_back(t);
                    }
                    for (int i = 0; i < combs; ++i) {
                        if (!s.rmax[i]) continue;
                        for (int ch = 0; ch < nch; ++ch) {
                            const double z = D.zeta[ch][N + 1];
                            if (z == 0.0) continue;
                            nrg_asm_term t{};
                            t.kind = NRG_TERM_SHIFT;
                            t.scale = z * (T.occ[i][ch] - 1.0) / it.scale;
                            t.r0 = s.off[i];
                            t.rows = s.rmax[i];
                            terms.push_back(t);
                        }
                    }
                    b.term_count = (int)terms.size() - b.term_begin;
                    blocks.push_back(b);
                }
                gpu.check(nrg_assemble_batch(gpu.ctx, blocks.data(), (int)blocks.size(), terms.data(), (int)terms.size()),
                          "nrg_assemble_batch");
                // largest first: the reference sorts its task list the same way
                std::vector<nrg_eig_task> tasks;
                for (auto &kv : it.subs) {
                    nrg_eig_task t{};
                    t.A = kv.second.U;
                    t.W = kv.second.W;
                    t.n = kv.second.d;
                    tasks.push_back(t);
                    eigenpairs += kv.second.d;
                }
                std::stable_sort(tasks.begin(), tasks.end(), [](const nrg_eig_task &a, const nrg_eig_task &b) { return a.n > b.n; });
                if (opt.diag) {
                    // shared launch with the other solves of the process: the
                    // assembled matrices must be complete before it reads them
                    gpu.check(nrg_sync(gpu.ctx), "nrg_sync");
                    if (nrg_status st = opt.diag->diagonalise(tasks.data(), (int)tasks.size()))
                        die("shared nrg_diag_batch failed (" + std::to_string(st) + ")");
                } else if (opt.plain) {
                    // the unbatched path: one eigensolver call per block, in turn
                    for (auto &t : tasks) {
                        gpu.check(nrg_diag_batch(gpu.ctx, &t, 1), "nrg_diag_batch");
                        gpu.check(nrg_sync(gpu.ctx), "nrg_sync");
                    }
                } else {
                    gpu.check(nrg_diag_batch(gpu.ctx, tasks.data(), (int)tasks.size()), "nrg_diag_batch");
                }
                // the previous iteration's host-side output while the GPU
                // runs this eigensolve
                if (deferred) {
                    deferred();
                    deferred = nullptr;
                }
                if (!opt.diag && !opt.plain) {
                    // eigenvalues as soon as they are final (in task order),
                    // while the eigenvectors are back-transformed
                    size_t tot = 0;
                    std::map<const double *, size_t> at;  // W -> position among the values
                    for (auto &t : tasks) {
                        at[t.W] = tot;
                        tot += t.n;
                    }
                    std::vector<double> all(tot);
                    gpu.check(nrg_diag_values(gpu.ctx, all.data(), tot), "nrg_diag_values");
                    for (auto &kv : it.subs) {
                        const size_t p = at.at(kv.second.W);
                        kv.second.value.assign(all.begin() + p, all.begin() + p + kv.second.d);
                    }
                } else {
                    std::vector<nrg_strided> items;
                    size_t tot = 0;
                    for (auto &kv : it.subs) {
                        items.push_back({kv.second.W, kv.second.d, 1});
                        tot += kv.second.d;
                    }
                    std::vector<double> all(tot);
                    gpu.gather(items, all.data());
                    size_t pos = 0;
                    for (auto &kv : it.subs) {
                        kv.second.value.assign(all.begin() + pos, all.begin() + pos + kv.second.d);
                        pos += kv.second.d;
                    }
                }
            });
            // ---- ground state, truncation (src/nrg.cc:2881-2929)
            double Egs = 1e300;
            for (auto &kv : it.subs) Egs = std::min(Egs, kv.second.value[0]);
            for (auto &kv : it.subs)
                for (auto &x : kv.second.value) x -= Egs;
            it.Egs = Egs;
            fix_splittings(it);
            truncate_prepare(it, last, false);
            totalenergy += Egs * it.scale;
            // the GPU work of this iteration is queued first (the shifted
            // values become E_prev, then the recalculations); the host-side
            // bookkeeping and output follow while it runs
            {
                std::vector<nrg_strided> items;
                std::vector<double> all;
                for (auto &kv : it.subs) {
                    items.push_back({kv.second.W, kv.second.d, 1});
                    all.insert(all.end(), kv.second.value.begin(), kv.second.value.end());
                }
                gpu.scatter(items, all.data());
            }
            // ---- recalculation with kept states (strategy=kept)
            st.old_op_arena = std::move(st.op_arena);
            st.op_arena.reset(new Arena(&gpu));
            Arena tmp(&gpu);
            if (!last) timed("recalc f", [&] { st.f = recalc_f(it, *st.op_arena); });
            timed("recalc s", [&] {
                for (auto &kv : st.ops_s) kv.second = recalc_general_all(it, kv.second, false, false, *st.op_arena, tmp);
            });
            tmp.release();             // (stream-ordered: later work reuses the memory)
            st.old_op_arena.reset();   // the previous iteration's operators, read by work already queued
            // host-side output of this iteration: run once the next
            // iteration's eigensolve is queued (custom_row then still sees
            // this iteration's singlet operators), or at once when stages
            // are timed
            const double te = totalenergy;
            auto book = [this, &it, &st, N, te, t_iter] {
                std::map<Invar, std::vector<double>> allv, keptv;
                for (auto &kv : it.subs) {
                    Sub &s = kv.second;
                    s.absenergy.resize(s.d);
                    for (int r = 0; r < s.d; ++r) s.absenergy[r] = te + s.value[r] * it.scale;
                    allv[kv.first] = s.value;
                    keptv[kv.first] = std::vector<double>(s.value.begin(), s.value.begin() + s.kept);
                }
                dump_energies(fenergies, N, allv);
                write_dump_iter(it);
                td_row(allv, it.scale);
                dump_annotated(allv, it.scale, te);
                size_t nkept = 0, nall = 0;
                for (auto &kv : it.subs) {
                    nkept += kv.second.kept;
                    nall += kv.second.d;
                }
                if (!opt.quiet)
                    std::cout << "N=" << N << " subspaces=" << it.subs.size() << " Kept: " << nkept << " out of "
                              << nall << " ("
                              << std::chrono::duration<double>(std::chrono::steady_clock::now() - t_iter).count()
                              << " s)" << std::endl;
                custom_row(it.scale / P.betabar, keptv, st.ops_s);
            };
            if (stage_timing() || last) book();
            else deferred = book;
            // ---- E_prev for the next iteration: first `kept` values of W
            std::map<Invar, Prev> np;
            for (auto &kv : it.subs) np[kv.first] = Prev{kv.second.kept, kv.second.W};
            st.prev = std::move(np);
            if (N == P.Ninit) plan_schedule();
            if (P.fdmexpv && N == P.fdmexpvn) keep_fdmexpv_ops(it, st.ops_s);
            if (fiterlog.is_open()) {
                gpu.check(nrg_sync(gpu.ctx), "nrg_sync");
                long long ep = 0;
                for (auto &kv : it.subs) ep += kv.second.d;
                fiterlog << N << ' ' << std::setprecision(9)
                         << std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() << ' '
                         << ep << std::endl;
            }
        }
        pass1_N = -1;
        // no eigensolves after pass 1: its workspace goes back to the device
        gpu.check(nrg_ctx_release_workspace(gpu.ctx), "nrg_ctx_release_workspace");
        GSenergy = totalenergy;
        scale_last = SCALE(Nmax);
        if (!opt.quiet) std::cout << "Total energy: " << std::setprecision(16) << totalenergy << std::endl;
        ftd.close();
        flush_custom();
        fcustom.close();
        fannotated.close();
        if (P.dumpsubspaces) dump_subspaces();  // finalize_nrg (src/nrg.cc:3222-3234)
    }

    // ------------------------------------------------------------ DM sweeps
    // Trace checks of density matrices (src/dmnrg.h:398-412), several at
    // once: one transfer of all diagonals, then the checks in the given order.
    struct TraceCheck {
        const RhoMap *r;
        double expected;
        const char *what;
        int N;
    };
    void rho_trace_check_all(const std::vector<TraceCheck> &checks) {
        std::vector<nrg_strided> items;
        size_t tot = 0;
        for (auto &c : checks)
            for (auto &kv : *c.r) {
                if (!kv.second.dim) continue;
                items.push_back({kv.second.p, kv.second.dim, kv.second.diag ? 1 : kv.second.dim + 1});
                tot += kv.second.dim;
            }
        std::vector<double> dg(tot);
        gpu.gather(items, dg.data());
        size_t pos = 0;
        for (auto &c : checks) {
            double tr = 0;
            for (auto &kv : *c.r) {
                if (!kv.second.dim) continue;
                double s = 0;
                for (int i = 0; i < kv.second.dim; ++i) s += dg[pos + i];
                pos += kv.second.dim;
                tr += T.mult(kv.first) * s;
            }
            const double rel = (tr - c.expected) / c.expected;
            if (!(std::fabs(rel) < 1e-8))
                die(std::string(c.what) + " trace check failed at N=" + std::to_string(c.N) + ": " + std::to_string(tr));
        }
    }
    void rho_trace_check(const RhoMap &r, double expected, const char *what, int N) {
        rho_trace_check_all({TraceCheck{&r, expected, what, N}});
    }

    // calc_densitymatrix_iterN / calc_fulldensitymatrix_iterN with cdmI
    // (src/dmnrg.h:327-429, 527-562): rho_{N-1}[I] = sum_i coef U_i^T rho_N U_i
    RhoMap dm_step(int N, const RhoMap &rhoN, const RhoMap *rhoDD, Arena &out, Arena &tmp) {
        RhoMap prev;
        GemmBatch st1, st2;
        const Iter &itN = iters[N];
        const Iter &itP = iters[N - 1];
        for (auto &kv : itP.subs) {
            const Invar I = kv.first;
            const int dim = kv.second.kept;
            Rho r;
            r.dim = dim;
            r.p = out.alloc(std::max<size_t>(1, (size_t)dim * dim));
            prev[I] = r;
            if (!dim) continue;
            st2.out(r.p, dim, dim, dim);
            for (int i = 0; i < combs; ++i) {
                const Invar I1 = sub(I, T.In[i]);
                auto s = itN.subs.find(I1);
                if (s == itN.subs.end()) continue;
                const Sub &S1 = s->second;
                const double coef = double(T.mult(I1)) / double(T.mult(I));
                for (const RhoMap *src : {&rhoN, rhoDD}) {
                    if (!src) continue;
                    auto rr = src->find(I1);
                    if (rr == src->end()) continue;
                    const Rho &rho = rr->second;
                    const int nrom = rho.dim;
                    if (nrom == 0) continue;
                    const int rmax = S1.rmax[i];
                    if (rmax == 0) continue;
                    if (rmax != dim) die("cdmI: range mismatch");
                    if (nrom > S1.d) die("cdmI: nromega exceeds the number of states");
                    const double *U = S1.U + S1.off[i];
                    if (rho.diag) {
                        st2.term(U, S1.d, true, U, S1.d, false, nrom, coef, rho.p);
                    } else {
                        double *Tm = tmp.alloc((size_t)dim * nrom);
                        st1.out(Tm, dim, nrom, nrom);
                        st1.term(U, S1.d, true, rho.p, nrom, false, nrom, 1.0);
                        st2.term(Tm, nrom, false, U, S1.d, false, nrom, coef);
                    }
                }
            }
        }
        st1.run(gpu, "DM stage 1");
        st2.run(gpu, "DM stage 2");
        return prev;
    }

    void init_rho() {
        // init_rho (src/nrg.cc:1877-1894) on the truncated last shell, dense
        const Iter &it = iters[Nmax - 1];
        double ZZ = 0;
        for (auto &kv : it.subs)
            for (int r = 0; r < kv.second.kept; ++r) ZZ += T.mult(kv.first) * std::exp(-kv.second.value[r] * scale_last / P.T);
        std::vector<nrg_strided> items;
        std::vector<double> h;
        for (auto &kv : it.subs) {
            const int k = kv.second.kept;
            Rho r;
            r.dim = k;
            r.p = rho_arena->alloc(std::max<size_t>(1, (size_t)k * k));
            if (k) gpu.check(nrg_zero(gpu.ctx, r.p, (size_t)k * k), "nrg_zero");
            for (int i = 0; i < k; ++i) h.push_back(std::exp(-kv.second.value[i] * scale_last / P.T) / ZZ);
            if (k) items.push_back({r.p, k, k + 1});
            rho_init[kv.first] = r;
        }
        gpu.scatter(items, h.data());
        rho_trace_check(rho_init, 1.0, "init_rho", Nmax - 1);
    }

    void calc_ZnD() {
        // calc_ZnD_one / calc_ZnD (src/nrg.cc:1897-1913, 2002-2041), long double
        ZnD.assign(Nmax, 0.0);
        wn.assign(Nmax, 0.0);
        wnfactor.assign(Nmax, 0.0);
        for (int N = P.Ninit; N < Nmax; ++N) {
            const bool last = (N + 1 == Nmax);
            long double ZZ = 0;
            for (auto &kv : iters[N].subs) {
                const int mn = last ? 0 : kv.second.kept;
                for (int i = mn; i < kv.second.d; ++i) {
                    const long double Eabs = (long double)kv.second.absenergy[i] - (long double)GSenergy;
                    if (Eabs < 0) die("calc_ZnD: negative excitation energy");
                    ZZ += T.mult(kv.first) * expl_neg(Eabs / (long double)P.T);
                }
            }
            ZnD[N] = (double)ZZ;
        }
        double Z = 0;
        for (int N = P.Ninit; N < Nmax; ++N) Z += std::pow(double(combs), Nmax - N - 1) * ZnD[N];
        double sumwn = 0;
        for (int N = P.Ninit; N < Nmax; ++N) {
            wn[N] = std::pow(double(combs), Nmax - N - 1) * ZnD[N] / Z;
            sumwn += wn[N];
        }
        if (std::fabs(sumwn - 1.0) > 1e-12) die("sum of w_n differs from 1");
        for (int N = P.Ninit; N < Nmax; ++N) {
            const double w = wn[N] / ZnD[N];
            wnfactor[N] = std::isnan(w) ? 0.0 : w;
        }
    }

    // init_rho_FDM (src/dmnrg.h:487-525): diagonal, discarded states only
    RhoMap init_rho_FDM(int N) {
        RhoMap r;
        const bool last = (N + 1 == Nmax);
        std::vector<nrg_strided> items;
        std::vector<double> h;
        for (auto &kv : iters[N].subs) {
            const int mn = last ? 0 : kv.second.kept, mx = kv.second.d;
            for (int i = 0; i < mx; ++i) {
                double hv = 0.0;
                if (i >= mn) {
                    const long double Eabs = (long double)kv.second.absenergy[i] - (long double)GSenergy;
                    long double v = expl_neg(Eabs / (long double)P.T) * ((long double)wn[N] / (long double)ZnD[N]);
                    if (!std::isfinite((double)v)) v = 0;
                    hv = (double)v;
                }
                h.push_back(hv);
            }
            Rho x;
            x.dim = mx;
            x.diag = true;
            x.p = rho_arena->alloc(mx);
            if (mx) items.push_back({x.p, mx, 1});
            r[kv.first] = x;
        }
        gpu.scatter(items, h.data());
        return r;
    }

    void dm_sweeps() {
        rho_arena.reset(new Arena(&gpu));
        rhos.assign(Nmax, RhoMap());
        rhofdm.assign(Nmax, RhoMap());
        if (P.fdm) calc_ZnD();
        init_rho();
        if (P.dmnrg) {
            timed("DM", [&] {
                // the matrices stay resident: their traces are checked after
                // the sweep, so that the sweep never waits for the host
                std::vector<TraceCheck> checks;
                RhoMap cur = rho_init;
                for (int N = Nmax - 1; N > P.Ninit; --N) {
                    if (!opt.quiet) std::cout << "DM sweep N=" << N << std::endl;
                    use(N);
                    Arena tmp(&gpu);
                    RhoMap prev = dm_step(N, cur, nullptr, *rho_arena, tmp);
                    rhos[N - 1] = prev;
                    checks.push_back({&rhos[N - 1], 1.0, "DM", N});
                    cur = prev;
                }
                rho_trace_check_all(checks);
            });
        }
        if (P.fdm) {
            timed("FDM", [&] {
                std::vector<TraceCheck> checks;
                RhoMap cur = init_rho_FDM(Nmax - 1);
                for (int N = Nmax - 1; N > P.Ninit; --N) {
                    if (!opt.quiet) std::cout << "FDM sweep N=" << N << std::endl;
                    use(N);
                    Arena tmp(&gpu);
                    const bool last = (N + 1 == Nmax);
                    RhoMap dd;
                    if (!last) dd = init_rho_FDM(N);
                    RhoMap prev = dm_step(N, cur, last ? nullptr : &dd, *rho_arena, tmp);
                    double expected = 0;
                    for (int n = N; n < Nmax; ++n) expected += wn[n];
                    rhofdm[N - 1] = prev;
                    checks.push_back({&rhofdm[N - 1], expected, "FDM", N});
                    cur = prev;
                }
                rho_trace_check_all(checks);
            });
        }
    }

    // ------------------------------------------------------------ pass 2
    double getfactor() const { return (nch == 1 && !P.NN1) ? P.Lambda : std::sqrt(P.Lambda); }

    void merge_nn2(int N, bool last, Spec &sp, const std::vector<double> &pos, const std::vector<double> &neg,
                   double sc) {
        // SpectrumRealFreq::mergeNN2 / mergeNN2half / windowfunction
        // (src/spectrumrealfreq.cc:116-185)
        if (!P.NN1 && !P.NN2avg) {
            const bool even = (N % 2 == 0);
            if (!(P.NN2even ? even : !even)) return;
        }
        const double fac = getfactor();
        const double Emin = sc * P.goodE, Ex = sc * P.goodE * fac, Emax = sc * P.goodE * fac * fac;
        const bool first = (N == P.Ninit);
        auto fnc = [&](double x) {
            if (P.NNtanh > 0.0) {
                auto f0 = [&](double y) { return std::tanh(P.NNtanh * (y - 0.5)); };
                return (f0(x) - f0(0.0)) / (f0(1.0) - f0(0.0));
            }
            return x;
        };
        auto window = [&](double E) {
            if (E <= Ex && last) return 1.0;
            if (E >= Ex && first) return 1.0;
            if (E <= Emin || E >= Emax) return 0.0;
            if (Emin < E && E <= Ex) return fnc((E - Emin) / (Ex - Emin));
            return 1.0 - fnc((E - Ex) / (Emax - Ex));
        };
        const double factor = P.NN2avg ? 0.5 : 1.0;
        for (int k = 0; k < nbins; ++k) {
            const double E = bin_e[k];
            if (Emin < E && E < Emax) {
                if (pos[k] != 0.0) sp.pos[k] += factor * pos[k] * window(E);
                if (neg[k] != 0.0) sp.neg[k] += factor * neg[k] * window(E);
            }
        }
    }

    void spectra_now(Iter &it, int N, bool last, const std::map<std::string, DevOp> &ops_d, bool all_states,
                     Arena &tmp) {
        for (auto &sp : spectra) {
            const DevOp &op1 = ops_d.at(sp.op1), &op2 = ops_d.at(sp.op2);
            GemmBatch gb;
            std::vector<nrg_spec_pair> pairs;
            const bool fdm = sp.kind == "FDM";
            const RhoMap *rm = nullptr;
            if (fdm) rm = last ? nullptr : &rhofdm[N];
            else rm = last ? &rho_init : &rhos[N];
            // the nonzero blocks of op1, in the order of a double loop over the
            // subspaces (Ii outer, Ij inner), which fixes the order of the pairs
            // (probing every pair of subspaces cost ~0.1 ms per shell)
            std::vector<std::pair<Invar, Invar>> keys;  // (Ii, Ij)
            keys.reserve(op1.size());
            for (auto &blk : op1) keys.push_back({blk.first.second, blk.first.first});
            std::sort(keys.begin(), keys.end());
            for (const auto &key : keys) {
                    const Invar Ii = key.first, Ij = key.second;
                    const auto ki = it.subs.find(Ii), kj = it.subs.find(Ij);
                    if (ki == it.subs.end() || kj == it.subs.end()) continue;
                    const Twoinvar II{Ij, Ii};
                    auto a = op1.find(II), b = op2.find(II);
                    if (b == op2.end()) continue;
                    if (sp.spin != 0 && Ij.second - Ii.second != sp.spin) continue;
                    const DevMat &o1 = a->second, &o2 = b->second;
                    nrg_spec_pair q{};
                    q.op1 = o1.p;
                    q.op2 = o2.p;
                    q.ld1 = o1.ld;
                    q.ld2 = o2.ld;
                    q.sf = T.specdens_factor(Ii, Ij);
                    if (fdm) {
                        q.alli = ki->second.d;
                        q.allj = kj->second.d;
                        if (o1.rows != q.allj || o1.cols != q.alli) die("FDM: operator block has the wrong size");
                        q.reti = last ? 0 : rm->at(Ii).dim;
                        q.retj = last ? 0 : rm->at(Ij).dim;
                        q.Ei = ki->second.Eabs;
                        q.Ej = kj->second.Eabs;
                        if (q.allj > 0 && q.reti > 0) {
                            double *M1 = tmp.alloc((size_t)q.allj * q.reti);
                            gb.out(M1, q.allj, q.reti, q.reti);
                            gb.term(o2.p, o2.ld, false, rm->at(Ii).p, q.reti, false, q.reti, 1.0);
                            q.M1 = M1;
                            q.ldm1 = q.reti;
                        }
                        if (q.retj > 0 && q.alli > 0) {
                            double *M2 = tmp.alloc((size_t)q.retj * q.alli);
                            gb.out(M2, q.retj, q.alli, q.alli);
                            gb.term(rm->at(Ij).p, q.retj, false, o2.p, o2.ld, false, q.retj, 1.0);
                            q.M2 = M2;
                            q.ldm2 = q.alli;
                        }
                    } else {
                        const Rho &rp = rm->at(Ii), &r1 = rm->at(Ij);
                        q.alli = rp.dim;
                        q.allj = r1.dim;
                        q.reti = q.alli;
                        q.retj = q.allj;
                        q.Ei = ki->second.W;
                        q.Ej = kj->second.W;
                        if (q.alli == 0 || q.allj == 0) continue;
                        double *M1 = tmp.alloc((size_t)q.allj * q.alli);
                        double *M2 = tmp.alloc((size_t)q.allj * q.alli);
                        gb.out(M1, q.allj, q.alli, q.alli);
                        gb.term(o2.p, o2.ld, false, rp.p, rp.dim, false, rp.dim, 1.0);
                        gb.out(M2, q.allj, q.alli, q.alli);
                        gb.term(r1.p, r1.dim, false, o1.p, o1.ld, false, r1.dim, 1.0);
                        q.M1 = M1;
                        q.M2 = M2;
                        q.ldm1 = q.ldm2 = q.alli;
                    }
                    pairs.push_back(q);
                }
            gb.run(gpu, "spectral GEMM");
            nrg_spec_params prm{};
            prm.mode = fdm ? NRG_SPEC_FDM : NRG_SPEC_DMNRG;
            prm.nbins = nbins;
            prm.log10emin = log10emin;
            prm.bins_per_decade = P.bins;
            prm.discard = P.discard_immediately;
            prm.T = P.T;
            prm.wnf = fdm ? wnfactor[N] : 0.0;
            prm.escale = fdm ? 1.0 : it.scale;
            const double fac = getfactor();
            prm.emin_w = P.goodE;
            prm.emax_w = P.goodE * fac * fac;
            if (bin_slots.size() >= bin_capacity) die("spectra: more histogram slots than planned");
            double *slot = dbins + bin_slots.size() * 2 * (size_t)nbins;
            gpu.check(nrg_spec_bin(gpu.ctx, &prm, pairs.data(), (int)pairs.size(), slot), "nrg_spec_bin");
            bin_slots.push_back({N, last, (size_t)(&sp - spectra.data()), it.scale});
        }
    }

    // Reads all histograms of pass 2 at once and merges them into the
    // spectra in the order they were filled (as the reference does shell by
    // shell: summed for FDM, merge_nn2 for DM-NRG).
    void merge_bins() {
        if (bin_slots.empty()) return;
        const size_t w = 2 * (size_t)nbins;
        std::vector<double> hb(bin_slots.size() * w);
        gpu.gather({nrg_strided{dbins, (int64_t)hb.size(), 1}}, hb.data());
        for (size_t k = 0; k < bin_slots.size(); ++k) {
            const BinSlot &b = bin_slots[k];
            Spec &sp = spectra[b.spec];
            std::vector<double> pos(hb.begin() + k * w, hb.begin() + k * w + nbins),
                neg(hb.begin() + k * w + nbins, hb.begin() + (k + 1) * w);
            if (sp.kind == "FDM") {
                for (int i = 0; i < nbins; ++i) {
                    sp.pos[i] += pos[i];
                    sp.neg[i] += neg[i];
                }
            } else {
                merge_nn2(b.N, b.last, sp, pos, neg, b.scale);
            }
        }
        bin_slots.clear();
    }

    void pass2() {
        State st;
        fresh_state(st);
        if (!spectra.empty()) {
            bin_capacity = (size_t)std::max(0, Nmax - P.Ninit) * spectra.size();
            dbins = gpu.alloc(std::max<size_t>(1, bin_capacity * 2 * (size_t)nbins));
            gpu.check(nrg_zero(gpu.ctx, dbins, bin_capacity * 2 * (size_t)nbins), "nrg_zero");
        }
        double totalenergy = D.GS;
        const bool recalc_all = P.fdm;  // do_recalc_all (src/nrg.cc:2436-2444)
        // energies.dmnrg: the second run's dump, to be compared with
        // energies.nrg (src/nrg.cc:1805-1810)
        std::ofstream fen2;
        if (P.dumpenergies) fen2.open(path("energies.dmnrg"));
        if (P.calc0) dump_energies(fen2, P.Ninit - 1, std::map<Invar, std::vector<double>>(D.energies.begin(), D.energies.end()));
        for (int N = P.Ninit; N < Nmax; ++N) {
            const bool last = (N + 1 == Nmax);
            if (!opt.quiet) std::cout << "Spectra N=" << N << std::endl;
            use(N, true);
            Iter &it = iters[N];
            if (fen2.is_open()) {
                std::map<Invar, std::vector<double>> allv;
                for (auto &kv : it.subs) allv[kv.first] = kv.second.value;
                dump_energies(fen2, N, allv);
            }
            // the stored values are already shifted by Egs and fixed
            totalenergy += it.Egs * it.scale;
            truncate_prepare(it, last, true);
            Arena eabs(&gpu);
            eabs.chunk = 1 << 16;
            {
                std::vector<nrg_strided> items;
                std::vector<double> e;
                for (auto &kv : it.subs) {
                    Sub &s = kv.second;
                    for (int r = 0; r < s.d; ++r) e.push_back(s.absenergy[r] - GSenergy);
                    s.Eabs = eabs.alloc(s.d);
                    items.push_back({s.Eabs, s.d, 1});
                }
                gpu.scatter(items, e.data());
            }
            st.old_op_arena = std::move(st.op_arena);
            st.op_arena.reset(new Arena(&gpu));
            Arena tmp(&gpu);
            if (recalc_all) {
                timed("recalc d", [&] {
                    for (auto &kv : st.ops_d) kv.second = recalc_general_all(it, kv.second, true, true, *st.op_arena, tmp);
                });
                timed("spec", [&] { spectra_now(it, N, last, st.ops_d, true, tmp); });
                // trim to kept (nrg_trim_matel, src/nrg.cc:2168-2188): views
                for (auto &kv : st.ops_d)
                    for (auto &m : kv.second) {
                        m.second.rows = it.subs.at(m.first.first).kept;
                        m.second.cols = it.subs.at(m.first.second).kept;
                    }
            } else {
                timed("recalc d", [&] {
                    for (auto &kv : st.ops_d) kv.second = recalc_general_all(it, kv.second, true, false, *st.op_arena, tmp);
                });
                timed("spec", [&] { spectra_now(it, N, last, st.ops_d, false, tmp); });
            }
            st.old_op_arena.reset();
        }
        merge_bins();
        totalenergy2 = totalenergy;
        if (std::fabs(totalenergy2 - GSenergy) > 1e-10 * std::max(1.0, std::fabs(GSenergy)))
            die("GS energy differs between the two passes");
    }

    // ------------------------------------------------------------ output
    // Bins::trim (src/bins.h:165-190), savebins + continuous
    // (src/spectrumrealfreq.cc:234-306), BR_NEW (src/spectral.h:56-102)
    void trimmed(const std::vector<double> &w, std::vector<double> &E, std::vector<double> &Wt) const {
        for (int i = 0; i + 1 < nbins; ++i) {
            const double ew = bin_e[i + 1] - bin_e[i];
            if (!(std::fabs(w[i]) < P.discard_trim * ew)) {
                E.push_back(bin_e[i]);
                Wt.push_back(w[i]);
            }
        }
        E.push_back(bin_e[nbins - 1]);
        Wt.push_back(w[nbins - 1]);
    }

    void write_spectra() {
        for (auto &sp : spectra) {
            std::string name = sp.op1 + "-" + sp.op2;
            if (sp.spin == 1) name += "-u";
            if (sp.spin == -1) name += "-d";
            const std::string fn = path("spec_" + sp.kind + "_dens_" + name);
            std::vector<double> ep, wp, en, wq;
            trimmed(sp.pos, ep, wp);
            trimmed(sp.neg, en, wq);
            if (P.savebins) {
                std::ofstream f(fn + ".bin", std::ios::binary);
                for (size_t i = 0; i < ep.size(); ++i) {
                    f.write((const char *)&ep[i], 8);
                    f.write((const char *)&wp[i], 8);
                }
                for (size_t i = 0; i < en.size(); ++i) {
                    const double e = -en[i];
                    f.write((const char *)&e, 8);
                    f.write((const char *)&wq[i], 8);
                }
            }
            double sum = 0;
            for (double x : wp) sum += x;
            for (double x : wq) sum += x;
            if (!opt.quiet) std::cout << "Spectrum: " << name << " " << sp.kind << " sum=" << std::setprecision(12) << sum << std::endl;
            if (!P.broaden) continue;
            const double alpha = P.alpha, gamma = alpha / 4.0;
            const double omega0 = P.omega0 < 0.0 ? P.omega0_ratio * P.T : P.omega0;
            const double sqrtpi = 1.7724538509055160273;
            auto BR_L = [&](double e, double ept) {
                if ((e < 0.0 && ept > 0.0) || (e > 0.0 && ept < 0.0)) return 0.0;
                if (ept == 0.0) return 0.0;
                const double x = std::log(e / ept) / alpha - gamma;
                return std::exp(-x * x) / (alpha * std::fabs(e) * sqrtpi);
            };
            auto BR_G = [&](double e, double ept) {
                const double x = (e - ept) / omega0;
                return std::exp(-x * x) / (omega0 * sqrtpi);
            };
            auto BR_NEW = [&](double e, double ept) {
                const double pl = BR_L(e, ept);
                if (std::fabs(e) > omega0) return pl;
                const double y = std::log(std::fabs(e) / omega0) / alpha;
                const double h = std::exp(-y * y);
                return pl * h + BR_G(e, ept) * (1.0 - h);
            };
            const double bmin = P.broaden_min <= 0.0 ? P.broaden_min_ratio * SCALE(Nmax) : P.broaden_min;
            std::vector<double> mesh;
            for (double E = P.broaden_max; E > bmin; E /= P.broaden_ratio) mesh.push_back(E);
            // every mesh point sums over all bins (a log and two exps per
            // term): the points are independent and spread over host threads,
            // each summed in the same order as a serial loop would
            std::vector<std::pair<double, double>> dpos(mesh.size()), dneg(mesh.size());
            auto point = [&](size_t k) {
                const double E = mesh[k];
                double vp = 0, vn = 0;
                for (size_t i = 0; i < ep.size(); ++i) {
                    vp += wp[i] * BR_NEW(E, ep[i]);
                    vn += wp[i] * BR_NEW(-E, ep[i]);
                }
                for (size_t i = 0; i < en.size(); ++i) {
                    vn += wq[i] * BR_NEW(-E, -en[i]);
                    vp += wq[i] * BR_NEW(E, -en[i]);
                }
                dpos[k] = {E, vp};
                dneg[k] = {-E, vn};
            };
            const size_t nth = std::min<size_t>(mesh.size() / 16 + 1,
                                                std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
            std::vector<std::thread> pool;
            for (size_t t = 1; t < nth; ++t)
                pool.emplace_back([&, t] {
                    for (size_t k = t; k < mesh.size(); k += nth) point(k);
                });
            for (size_t k = 0; k < mesh.size(); k += nth) point(k);
            for (auto &th : pool) th.join();
            std::sort(dpos.begin(), dpos.end());
            std::sort(dneg.begin(), dneg.end());
            std::ofstream f(fn + ".dat");
            f << std::setprecision(P.prec_xy);
            for (auto &x : dneg) f << x.first << " " << x.second << "\n";
            for (auto &x : dpos) f << x.first << " " << x.second << "\n";
        }
    }

    void write_stats(double wall) {
        if (opt.stats.empty()) return;
        std::ofstream f(opt.stats);
        f << std::setprecision(10) << "{\"wall_s\": " << wall << ", \"eigenpairs\": " << eigenpairs
          << ", \"Nmax\": " << Nmax << ", \"spills\": " << spills << ", \"restores\": " << restores
          << ", \"host_spill_peak_bytes\": " << host_peak_bytes << ", \"pinned_exhausted\": " << (pinned_exhausted ? "true" : "false")
          << ", \"stages\": {";
        bool first = true;
        for (auto &kv : stages) {
            if (!first) f << ", ";
            first = false;
            f << "\"" << kv.first << "\": {\"s\": " << kv.second.t << ", \"calls\": " << kv.second.n << "}";
        }
        f << "}, \"kernels\": {";
        nrg_kernel_stat ks[64];
        const int nk = std::min(nrg_timing_get(ks, 64), 64);
        for (int k = 0; k < nk; ++k) {
            if (k) f << ", ";
            f << "\"" << ks[k].name << "\": {\"launches\": " << ks[k].launches << ", \"ms\": " << ks[k].ms
              << ", \"flops\": " << ks[k].flops << ", \"bytes\": " << ks[k].bytes << "}";
        }
        f << "}}\n";
    }

    int run() {
        auto t0 = std::chrono::steady_clock::now();
        // per-kernel accounting is process-wide: not with concurrent solves
        if (!opt.stats.empty() && !opt.diag) {
            nrg_timing_enable(1);
            nrg_timing_reset();
        }
        t_start = t0;
        setup();
        if (!opt.dump_eigen.empty()) fdump.open(opt.dump_eigen, std::ios::binary);
        {
            // the shared eigensolver expects this solve until it leaves,
            // also when the pass ends with an error
            struct Leave {
                DiagService *d;
                ~Leave() {
                    if (d) d->leave();
                }
            } leave{opt.diag};
            if (opt.diag) opt.diag->enter();
            pass1();
        }
        if (P.stopafter == "nrg") throw Stop("*** Stopped after the first sweep.");  // src/nrg.cc:3453
        if (P.dm) {
            dm_sweeps();
            if (P.stopafter == "rho") throw Stop("*** Stopped after the DM calculation.");  // src/nrg.cc:3462
            pass2();
            write_spectra();
            if (P.fdmexpv) write_customfdm();
        }
        for (auto &it : iters) free_host(it);
        iters.clear();
        rhos.clear();
        rhofdm.clear();
        rho_init.clear();
        rho_arena.reset();
        if (dbins) gpu.free(dbins);
        gpu.check(nrg_sync(gpu.ctx), "nrg_sync");
        auto t1 = std::chrono::steady_clock::now();
        const double wall = std::chrono::duration<double>(t1 - t0).count();
        write_stats(wall);
        std::ofstream(path("DONE")) << "DONE\n";
        return 0;
    }
};

int run(const Options &opt) {
    Engine e(opt);
    return e.run();
}

}  // namespace nrg
