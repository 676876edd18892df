// serve.cpp - serving loop: many whole NRG solves in one process, one GPU.
//
// Requests (one solve each: a directory holding param and data, an arrival
// time) are admitted as solve slots and HBM become free.  Every admitted
// solve runs the two-pass engine in its own host thread with its own context
// and stream, so the recalculation, density-matrix and spectral stages of
// different solves overlap on the device.  The eigensolves, which dominate,
// are merged: each iteration's task list goes to one DiagBatcher, which
// launches the pending task lists of all solves that are in their NRG pass
// as one batched nrg_diag_batch.  A solve's results do not depend on what it
// is batched with (every matrix is reduced by its own workgroup or team).
//
// No counterpart in the reference (one solve per process there; z-averaging
// and sweeps are separate runs).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>

#include "nrg_host.h"

namespace nrg {

namespace {

using Clock = std::chrono::steady_clock;

// Merges the per-iteration eigensolver calls of concurrently running solves.
// A launch starts when every solve that is in its NRG pass has submitted, or
// when the oldest submission has waited `window` (solves in other stages do
// not hold the others back).
class DiagBatcher : public DiagService {
  public:
    DiagBatcher(int device, double window_ms) : window_(window_ms) {
        if (nrg_ctx_create(&ctx_, device) != NRG_OK) die("DiagBatcher: nrg_ctx_create failed");
        th_ = std::thread([this] { loop(); });
    }
    ~DiagBatcher() override {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
        nrg_ctx_destroy(ctx_);
    }
    void enter() override {
        std::lock_guard<std::mutex> g(m_);
        ++active_;
    }
    void leave() override {
        {
            std::lock_guard<std::mutex> g(m_);
            --active_;
        }
        cv_.notify_all();
    }
    nrg_status diagonalise(const nrg_eig_task *tasks, int ntasks) override {
        Req r;
        r.tasks.assign(tasks, tasks + ntasks);
        r.t = Clock::now();
        std::unique_lock<std::mutex> lk(m_);
        pending_.push_back(&r);
        cv_.notify_all();
        cv_.wait(lk, [&] { return r.done; });
        return r.st;
    }
    long launches() const { return launches_; }
    long merged_solves() const { return merged_; }

  private:
    struct Req {
        std::vector<nrg_eig_task> tasks;
        Clock::time_point t;
        bool done = false;
        nrg_status st = NRG_OK;
    };
    void loop() {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait_for(lk, std::chrono::microseconds(200), [&] {
                return stop_ || (!pending_.empty() && (int)pending_.size() >= active_);
            });
            if (stop_ && pending_.empty()) return;
            if (pending_.empty()) continue;
            const double waited =
                std::chrono::duration<double, std::milli>(Clock::now() - pending_.front()->t).count();
            if ((int)pending_.size() < active_ && waited < window_ && !stop_) continue;
            std::vector<Req *> batch(pending_.begin(), pending_.end());
            pending_.clear();
            lk.unlock();
            std::vector<nrg_eig_task> all;
            for (Req *r : batch) all.insert(all.end(), r->tasks.begin(), r->tasks.end());
            nrg_status st = nrg_diag_batch(ctx_, all.data(), (int)all.size());
            if (st == NRG_OK) st = nrg_sync(ctx_);
            if (st != NRG_OK) std::cerr << "shared eigensolve: " << nrg_last_error(ctx_) << std::endl;
            lk.lock();
            ++launches_;
            merged_ += (long)batch.size();
            for (Req *r : batch) {
                r->st = st;
                r->done = true;
            }
            cv_.notify_all();
        }
    }
    nrg_ctx *ctx_ = nullptr;
    double window_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Req *> pending_;
    int active_ = 0;
    bool stop_ = false;
    long launches_ = 0, merged_ = 0;
    std::thread th_;
};

std::string json_escape(const std::string &x) {
    std::string o;
    for (char ch : x) {
        if (ch == '"' || ch == '\\') o += '\\';
        if ((unsigned char)ch < 0x20) o += ' ';
        else o += ch;
    }
    return o;
}

struct Request {
    std::string id, dir;
    double arrival = 0.0;
    double footprint = 0.0;  // estimated device bytes
    // measured
    double admitted = -1.0, finished = -1.0;
    long long eigenpairs = 0;
    int rc = -1;
    std::string error;  // message of a failed solve
};

// Device memory a solve keeps until it ends: the eigenvectors of every
// iteration (about sum_I d_I^2 doubles each, with sum_I d_I <= combs * keep)
// plus operators and density matrices; a deliberately generous estimate.
double estimate_footprint(const std::string &dir) {
    Params P;
    P.load(dir + "/param");
    std::ifstream f(dir + "/data");
    std::string line;
    int nch = 1, nmax = 0;
    while (std::getline(f, line)) {
        if (line.rfind("#!", 0) == 0) {
            std::getline(f, line);
            std::istringstream is(line);
            is >> nch >> nmax;
            break;
        }
    }
    const double combs = std::pow(4.0, nch);
    const double sd = combs * P.keep;
    return 8.0 * std::max(1, nmax) * sd * sd / 16.0 + (64ull << 20);
}

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (v.size() - 1);
    const size_t i = (size_t)pos;
    if (i + 1 >= v.size()) return v.back();
    return v[i] + (pos - i) * (v[i + 1] - v[i]);
}

long long read_eigenpairs(const std::string &stats) {
    std::ifstream f(stats);
    std::string s((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const auto p = s.find("\"eigenpairs\":");
    return p == std::string::npos ? 0 : std::atoll(s.c_str() + p + 13);
}

}  // namespace

int serve(const std::string &reqfile, const std::string &report, const std::string &tables, int device, int slots,
          double window_ms, double mem_fraction) {
    std::vector<Request> reqs;
    {
        std::ifstream f(reqfile);
        if (!f) die("cannot open request list " + reqfile);
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream is(line);
            Request r;
            if (!(is >> r.id >> r.arrival >> r.dir)) die("bad request line: " + line);
            r.footprint = estimate_footprint(r.dir);
            reqs.push_back(r);
        }
    }
    std::stable_sort(reqs.begin(), reqs.end(), [](const Request &a, const Request &b) { return a.arrival < b.arrival; });
    nrg_ctx *probe = nullptr;
    if (nrg_ctx_create(&probe, device) != NRG_OK) die("serve: nrg_ctx_create failed");
    size_t fr = 0, tot = 0;
    if (nrg_mem_info(probe, &fr, &tot) != NRG_OK) die("serve: nrg_mem_info failed");
    nrg_ctx_destroy(probe);
    const double budget = mem_fraction * (double)fr;

    DiagBatcher batcher(device, window_ms);
    std::mutex m;
    std::condition_variable cv;
    int running = 0;
    double committed = 0.0;
    std::vector<std::thread> threads;
    const auto t0 = Clock::now();
    auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    size_t next = 0;  // first request not admitted
    std::unique_lock<std::mutex> lk(m);
    while (next < reqs.size()) {
        Request &r = reqs[next];
        const double wait_arrival = r.arrival - now();
        if (wait_arrival > 0) {
            cv.wait_for(lk, std::chrono::duration<double>(std::min(wait_arrival, 0.01)));
            continue;
        }
        // admission: a slot and the solve's estimated footprint must be free
        // (a solve larger than the whole budget runs alone)
        const bool fits = committed + r.footprint <= budget || running == 0;
        if (running >= slots || !fits) {
            cv.wait_for(lk, std::chrono::milliseconds(2));
            continue;
        }
        ++running;
        committed += r.footprint;
        r.admitted = now();
        Request *rp = &r;
        threads.emplace_back([&, rp] {
            Options o;
            o.workdir = rp->dir;
            o.tables_dir = tables;
            o.device = device;
            o.quiet = true;
            o.stats = rp->dir + "/stats.json";
            o.diag = &batcher;
            // a failed solve is reported and the others go on (its engine
            // has left the shared eigensolver and drained its stream)
            int rc = 1;
            std::string err;
            try {
                rc = run(o);
            } catch (const std::exception &e) {
                err = e.what();
            }
            const long long ep = rc == 0 ? read_eigenpairs(o.stats) : 0;
            std::lock_guard<std::mutex> g(m);
            rp->rc = rc;
            rp->error = err;
            rp->eigenpairs = ep;
            rp->finished = now();
            --running;
            committed -= rp->footprint;
            cv.notify_all();
        });
        ++next;
    }
    cv.wait(lk, [&] { return running == 0; });
    lk.unlock();
    for (auto &t : threads) t.join();
    const double makespan = now();

    std::vector<double> lat;
    long long pairs = 0;
    int failed = 0;
    for (auto &r : reqs) {
        lat.push_back(r.finished - r.arrival);
        pairs += r.eigenpairs;
        failed += r.rc != 0;
    }
    std::ofstream f(report);
    f << std::setprecision(10);
    f << "{\"solves\": " << reqs.size() << ", \"failed\": " << failed << ", \"makespan_s\": " << makespan
      << ", \"solves_per_s\": " << reqs.size() / makespan << ", \"eigenpairs\": " << pairs
      << ", \"eigenpairs_per_s\": " << pairs / makespan << ", \"latency_p50_s\": " << pct(lat, 0.5)
      << ", \"latency_p99_s\": " << pct(lat, 0.99) << ", \"slots\": " << slots << ", \"hbm_budget_bytes\": " << budget
      << ", \"shared_launches\": " << batcher.launches() << ", \"merged_iterations\": " << batcher.merged_solves()
      << ", \"requests\": [";
    for (size_t i = 0; i < reqs.size(); ++i) {
        const Request &r = reqs[i];
        f << (i ? ", " : "") << "{\"id\": \"" << r.id << "\", \"dir\": \"" << r.dir << "\", \"arrival\": " << r.arrival
          << ", \"admitted\": " << r.admitted << ", \"finished\": " << r.finished << ", \"latency\": "
          << r.finished - r.arrival << ", \"eigenpairs\": " << r.eigenpairs << ", \"footprint\": " << r.footprint
          << ", \"rc\": " << r.rc << ", \"error\": \"" << json_escape(r.error) << "\"}";
    }
    f << "]}\n";
    return failed ? 1 : 0;
}

}  // namespace nrg

// nrg-serve REQUESTS [--report FILE] [--slots S] [--window-ms W] [--mem-fraction F] [--tables DIR] [--device N]
static int nrg_serve_main_args(int argc, char **argv) {
    std::string req, report = "serve_report.json", tables;
    int device = 0, slots = 8;
    double window = 5.0, memf = 0.85;
    if (const char *t = std::getenv("NRG_TABLES")) tables = t;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) nrg::die("missing value for " + a);
            return argv[++i];
        };
        if (a == "--report") report = next();
        else if (a == "--slots") slots = std::max(1, std::atoi(next().c_str()));
        else if (a == "--window-ms") window = std::atof(next().c_str());
        else if (a == "--mem-fraction") memf = std::atof(next().c_str());
        else if (a == "--tables") tables = next();
        else if (a == "--device") device = std::atoi(next().c_str());
        else if (a == "-h" || a == "--help") {
            std::cout << "usage: nrg-serve REQUESTS [--report FILE] [--slots S] [--window-ms W] [--mem-fraction F]"
                         " [--tables DIR] [--device N]\n"
                         "REQUESTS: lines 'id arrival_seconds directory' (directory holds param and data)\n";
            return 0;
        } else if (req.empty()) req = a;
        else nrg::die("unknown option " + a);
    }
    if (req.empty()) nrg::die("nrg-serve: no request list");
    if (tables.empty()) nrg::die("nrg-serve: --tables or NRG_TABLES required");
    nrg::load_tuning(tables);
    return nrg::serve(req, report, tables, device, slots, window, memf);
}

// Status 0, 1 when a request failed (each is listed with its error in the
// report) or on an error of the loop itself.
extern "C" int nrg_serve_main(int argc, char **argv) {
    try {
        return nrg_serve_main_args(argc, argv);
    } catch (const std::exception &e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
