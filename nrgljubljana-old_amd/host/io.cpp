// io.cpp - input contract of the drop-in: `param`, `data`, coefficient tables.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>

#include "nrg_host.h"

namespace nrg {

void die(const std::string &msg) { throw Error(msg); }

void load_tuning(const std::string &tables_dir) {
    static const std::set<std::string> known = {"NRG_BLK_MIN_N", "NRG_GEMM_MFMA_MIN_TILES"};
    std::ifstream f(tables_dir + "/tune.txt");
    std::string line;
    while (f && std::getline(f, line)) {
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream is(line);
        std::string name, value;
        if (!(is >> name >> value)) continue;
        if (!known.count(name)) die("tune.txt: unknown setting " + name);
        setenv(name.c_str(), value.c_str(), 0);
    }
}

// ---------------------------------------------------------------------------
// param: a [param] block of keyword=value lines; duplicates are fatal,
// unknown keywords are reported (src/misc.h:71-124, src/nrg.cc:681-703).
// ---------------------------------------------------------------------------
static bool to_bool(const std::string &v) {
    return v == "true" || v == "1" || v == "yes" || v == "True" || v == "TRUE";
}

void Params::load(const std::string &path) {
    std::ifstream f(path);
    if (!f) die("Can't load parameters from " + path);
    std::string line;
    bool inblock = false;
    while (std::getline(f, line)) {
        size_t a = line.find_first_not_of(" \t\r");
        if (a == std::string::npos) continue;
        line = line.substr(a);
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        if (line.empty() || line[0] == '#' || line[0] == ';') continue;
        if (line[0] == '[') {
            inblock = (line == "[param]");
            continue;
        }
        if (!inblock) continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string k = line.substr(0, eq), v = line.substr(eq + 1);
        while (!k.empty() && (k.back() == ' ' || k.back() == '\t')) k.pop_back();
        size_t b = v.find_first_not_of(" \t");
        v = (b == std::string::npos) ? "" : v.substr(b);
        if (raw.count(k)) die("Duplicate parameter " + k);
        raw[k] = v;
    }
    std::set<std::string> used;
    auto D = [&](const char *k, double &x) { if (raw.count(k)) { x = std::stod(raw[k]); used.insert(k); } };
    auto I = [&](const char *k, int &x) { if (raw.count(k)) { x = std::stoi(raw[k]); used.insert(k); } };
    auto B = [&](const char *k, bool &x) { if (raw.count(k)) { x = to_bool(raw[k]); used.insert(k); } };
    auto S = [&](const char *k, std::string &x) { if (raw.count(k)) { x = raw[k]; used.insert(k); } };
    D("Lambda", Lambda); D("z", z); D("bandrescale", bandrescale); D("keepenergy", keepenergy);
    D("T", T); D("betabar", betabar); S("discretization", discretization); S("strategy", strategy);
    S("specd", specd); I("keep", keep); I("keepmin", keepmin); I("bins", bins);
    I("safeguardmax", safeguardmax); I("Nmax", Nmax); I("Ninit", Ninit); I("width_td", width_td);
    I("prec_td", prec_td); I("width_custom", width_custom); I("prec_custom", prec_custom);
    I("prec_xy", prec_xy); B("dmnrg", dmnrg); B("fdm", fdm); B("dm", dm); B("savebins", savebins);
    B("broaden", broaden); B("NN2even", NN2even); B("NN2avg", NN2avg); B("NN1", NN1);
    B("lastall", lastall); B("lastalloverride", lastalloverride); B("calc0", calc0);
    B("dumpenergies", dumpenergies); D("goodE", goodE); D("NNtanh", NNtanh);
    B("dumpsubspaces", dumpsubspaces); B("fdmexpv", fdmexpv); I("fdmexpvn", fdmexpvn); I("forcestop", forcestop); S("stopafter", stopafter);
    I("dumpannotated", dumpannotated); B("dumpabs", dumpabs); B("dumpscaled", dumpscaled);
    I("dumpprecision", dumpprecision); B("dumpgroups", dumpgroups); D("grouptol", grouptol);
    D("safeguard", safeguard); D("fixeps", fixeps); D("emin", emin); D("emax", emax);
    D("discard_trim", discard_trim); D("discard_immediately", discard_immediately);
    D("alpha", alpha); D("omega0", omega0); D("omega0_ratio", omega0_ratio);
    D("broaden_max", broaden_max); D("broaden_min", broaden_min);
    D("broaden_min_ratio", broaden_min_ratio); D("broaden_ratio", broaden_ratio);
    for (const char *k : {"diag", "diagth", "checkdiag", "symtype", "ops", "tri", "discretization"}) used.insert(k);
    std::string unused;
    for (auto &kv : raw)
        if (!used.count(kv.first)) unused += " " + kv.first;
    if (!unused.empty()) std::cerr << "Unused settings:" << unused << std::endl;
    discretization = discretization.substr(0, 1);
    // validateparameters (src/read-input.cc:6-36)
    if (keep <= 1) die("keep must be > 1");
    if (Lambda <= 1.0) die("Lambda must be > 1");
    if (discretization != "Y") die("only discretization=Y is supported");
    if (dmnrg || fdm) dm = true;
    if (fdm && !lastalloverride) lastall = true;
    if (strategy != "kept" && strategy != "all") die("unknown strategy " + strategy);
}

// ---------------------------------------------------------------------------
// data
// ---------------------------------------------------------------------------
void Data::load(const std::string &path) {
    std::ifstream fin(path);
    if (!fin) die("Can't load initial data.");
    std::string line, body;
    int version = -1;
    while (std::getline(fin, line)) {
        size_t a = line.find_first_not_of(" \t\r");
        if (a != std::string::npos && line[a] == '#') {
            if (line.compare(a, 2, "#!") == 0) version = std::atoi(line.c_str() + a + 2);
            size_t p = line.find("symtype");
            if (p != std::string::npos) {
                std::istringstream ss(line.substr(p + 7));
                ss >> sym;
            }
            continue;
        }
        body += line;
        body += '\n';
    }
    if (version != 9) die("data file version must be 9");
    std::istringstream in(body);
    size_t nsubs;
    if (!(in >> channels >> Nmax >> nsubs)) die("data: bad header");
    for (size_t s = 0; s < nsubs; ++s) {
        Invar I;
        int n;
        in >> I.first >> I.second >> n;
        std::vector<double> e(n);
        for (auto &x : e) in >> x;
        energies[I] = e;
    }
    auto read_op = [&](HostOp &op) {
        size_t nf;
        in >> nf;
        for (size_t k = 0; k < nf; ++k) {
            Invar I1, I2;
            in >> I1.first >> I1.second >> I2.first >> I2.second;
            if (!energies.count(I1) || !energies.count(I2)) die("data: matrix element between unknown subspaces");
            HostMat m;
            m.rows = (int)energies[I1].size();
            m.cols = (int)energies[I2].size();
            m.v.resize((size_t)m.rows * m.cols);
            for (auto &x : m.v) in >> x;
            op[{I1, I2}] = std::move(m);
        }
    };
    for (int ch = 0; ch < channels; ++ch) {
        std::string tag;
        int i, j;
        in >> tag >> i >> j;
        if (tag != "f" || i != ch || j != 0) die("data: expected 'f ch 0' block");
        f.emplace_back();
        read_op(f.back());
    }
    std::string tag;
    while (in >> tag) {
        if (tag == "e") {
            in >> GS;
        } else if (tag == "s" || tag == "d" || tag == "p" || tag == "g" || tag == "t" || tag == "o" || tag == "q") {
            std::string name;
            in >> name;
            HostOp op;
            read_op(op);
            if (tag == "s") ops_s[name] = std::move(op);
            else if (tag == "d") ops_d[name] = std::move(op);
        } else if (tag == "z") {
            for (auto *tab : {&xi, &zeta})
                for (int ch = 0; ch < channels; ++ch) {
                    int n;
                    in >> n;
                    std::vector<double> v(n + 1);
                    for (auto &x : v) in >> x;
                    tab->push_back(v);
                }
        } else {
            die("Unknown block " + tag + " in data file.");
        }
    }
    if (in.bad()) die("data: read error");
}

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
double TabEntry::at(int ss) const {
    if (vals.size() == 1) return vals[0];
    if (ss < 1 || ss > (int)vals.size()) die("coefficient table exhausted at SS=" + std::to_string(ss));
    return vals[ss - 1];
}

void Tables::load(const std::string &path) {
    std::ifstream f(path);
    if (!f) die("Can't open coefficient table " + path);
    std::string tag;
    f >> tag;
    int ver;
    f >> ver;
    if (tag != "NRGTAB" || ver != 1) die("bad table file " + path);
    auto readvals = [&](TabEntry &e) {
        int nv;
        f >> nv;
        e.vals.resize(nv);
        for (auto &x : e.vals) {
            std::string s;
            f >> s;
            x = (s == "nan") ? NAN : std::stod(s);
        }
    };
    while (f >> tag) {
        if (tag == "sym") f >> sym;
        else if (tag == "channels") f >> channels;
        else if (tag == "combs") f >> combs;
        else if (tag == "ssmax") f >> ssmax;
        else if (tag == "comb") {
            int i;
            Invar in, qn;
            f >> i >> in.first >> in.second >> qn.first >> qn.second;
            std::vector<int> o(channels);
            for (auto &x : o) f >> x;
            In.push_back(in);
            QN.push_back(qn);
            occ.push_back(o);
        } else if (tag == "offdiag") {
            TabEntry e;
            f >> e.a >> e.b >> e.ch;
            readvals(e);
            offdiag.push_back(e);
        } else if (tag == "recalcf") {
            TabEntry e;
            f >> e.ch >> e.kind >> e.a >> e.b;
            readvals(e);
            recalcf.push_back(e);
        } else if (tag == "doublet") {
            TabEntry e;
            f >> e.kind >> e.a >> e.b;
            e.ch = 0;
            readvals(e);
            doublet.push_back(e);
        } else if (tag == "end") {
            break;
        } else {
            die("bad table entry " + tag);
        }
    }
    if ((int)In.size() != combs) die("table: combination count mismatch");
}

// Symmetry::triangle_inequality (src/sym-QS.cc:27-31, src/sym-QSZ.cc:34-38,
// src/symmetry.cc:33-38)
bool Tables::triangle(const Invar &a, const Invar &b, const Invar &c) const {
    if (a.first != b.first + c.first) return false;
    if (sym == "QS") {
        const int x = a.second, y = b.second, zz = c.second;
        return std::abs(x - y) <= zz - 1 && std::abs(y - zz) <= x - 1 && std::abs(zz - x) <= y - 1;
    }
    return a.second == b.second + c.second;
}

// SymmetryQS::specdens_factor (src/sym-QS.cc:95-104); 1 for QSZ.
double Tables::specdens_factor(const Invar &Ip, const Invar &I1) const {
    if (sym != "QS") return 1.0;
    const int ssp = Ip.second, ss1 = I1.second;
    const double Sp = (ssp - 1) / 2.0;
    return ss1 == ssp + 1 ? Sp + 1.0 : Sp;
}

}  // namespace nrg
