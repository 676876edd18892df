// cli.cpp - `nrg`: drop-in for the NRG Ljubljana back end.
// Reads ./param and ./data, writes td, custom, spec_*_dens_*.{bin,dat} and
// DONE into the current directory, like the reference binary
// (src/nrg.cc:3527-3582).  The -w workdir option of the reference names its
// scratch directory for unitaryN/rhoN files; this engine keeps them in HBM,
// so the option is accepted and unused.
#include <libgen.h>
#include <unistd.h>

#include <climits>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "nrg_host.h"

// Entry point shared by the `nrg` executable (bin/nrg.cpp) and in-process
// callers (bench.py loads libnrg_engine.so and calls it with argv).

static std::string exe_dir() {
    char buf[PATH_MAX];
    ssize_t n = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
    if (n <= 0) return ".";
    buf[n] = 0;
    return std::string(dirname(buf));
}

static int nrg_main_args(int argc, char **argv) {
    nrg::Options opt;
    opt.tables_dir = exe_dir() + "/../tables";
    if (const char *t = std::getenv("NRG_TABLES")) opt.tables_dir = t;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) nrg::die("missing value for " + a);
            return argv[++i];
        };
        if (a == "-w") next();
        else if (a == "--dump-eigen") opt.dump_eigen = next();
        else if (a == "--stats") opt.stats = next();
        else if (a == "--stats-light") opt.stats_light = true;
        else if (a == "--tables") opt.tables_dir = next();
        else if (a == "--device") opt.device = std::atoi(next().c_str());
        else if (a == "--quiet") opt.quiet = true;
        else if (a == "--plain-diag") opt.plain = true;
        else if (a == "--hbm-budget") opt.hbm_budget_gb = std::atof(next().c_str());
        else if (a == "--spill-all") opt.hbm_budget_gb = 0.0;
        else if (a == "--spill-pageable") opt.spill_pageable = true;
        else if (a == "--iter-log") opt.iter_log = next();
        else if (a == "-h" || a == "--help") {
            std::cout << "usage: nrg [-w workdir] [--dump-eigen file] [--stats file [--stats-light]] [--tables dir] [--device n] [--quiet] [--plain-diag] [--hbm-budget GB] [--spill-all] [--spill-pageable] [--iter-log file]\n";
            return 0;
        } else nrg::die("unknown option " + a);
    }
    nrg::load_tuning(opt.tables_dir);
    return nrg::run(opt);
}

// Status 0, or 1 after an error or a requested stop, whose message goes to
// stderr (the process and the GPU context stay usable for in-process callers).
extern "C" int nrg_main(int argc, char **argv) {
    try {
        return nrg_main_args(argc, argv);
    } catch (const nrg::Stop &e) {
        std::cout << e.what() << std::endl;
        return 1;
    } catch (const std::exception &e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
}
