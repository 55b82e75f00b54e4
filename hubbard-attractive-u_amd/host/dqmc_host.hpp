// dqmc_host.hpp -- host-side pieces of the drop-in for the reference's sweep path:
// configuration (config.toml keys, defaults and range checks), the square lattice tables used
// by measurements, the six observables with their binning/statistics, and the text formats of
// ising.fields / *.out / *.bins.out / momenta.out / tgrids.out.  Header-only, no GPU code, so
// it builds with any C++17 compiler (the CPU tests compile it with g++).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

namespace dqmc_toml {

// A TOML subset: [a.b] tables, key = value with float / int / bool / string / arrays of those,
// '#' comments.  Values are kept as their source text, typed on lookup.
struct Value {
    std::string text;     // scalar text, or "" for arrays
    std::vector<std::string> items;
    bool is_array = false;
    bool is_string = false;
};

class Table {
   public:
    std::map<std::string, Value> kv;   // "Model.HubbardAttractiveU.u" -> value

    static std::string trim(const std::string& s)
    {
        size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
        return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    }
    static std::string strip_comment(const std::string& s)
    {
        bool in_str = false;
        for (size_t i = 0; i < s.size(); ++i) {
            if (s[i] == '"') in_str = !in_str;
            if (s[i] == '#' && !in_str) return s.substr(0, i);
        }
        return s;
    }
    static Table parse_file(const std::string& path)
    {
        std::ifstream f(path);
        if (!f.is_open()) throw std::runtime_error("toml: cannot open '" + path + "'");
        std::stringstream ss;
        ss << f.rdbuf();
        return parse(ss.str());
    }
    static Table parse(const std::string& src)
    {
        Table t;
        std::string prefix, line, pending;
        std::istringstream in(src);
        int lineno = 0;
        while (std::getline(in, line)) {
            ++lineno;
            std::string l = trim(strip_comment(line));
            if (!pending.empty()) {   // continuation of a multi-line array
                pending += " " + l;
                if (std::count(pending.begin(), pending.end(), '[') > std::count(pending.begin(), pending.end(), ']')) continue;
                l = pending;
                pending.clear();
            }
            if (l.empty()) continue;
            if (l.front() == '[' && l.find('=') == std::string::npos) {
                if (l.back() != ']') throw std::runtime_error("toml: bad table header at line " + std::to_string(lineno));
                prefix = trim(l.substr(1, l.size() - 2));
                continue;
            }
            const size_t eq = l.find('=');
            if (eq == std::string::npos) throw std::runtime_error("toml: expected key = value at line " + std::to_string(lineno));
            const std::string key = trim(l.substr(0, eq));
            std::string val = trim(l.substr(eq + 1));
            if (!val.empty() && val.front() == '[' &&
                std::count(val.begin(), val.end(), '[') > std::count(val.begin(), val.end(), ']')) {
                pending = l;
                continue;
            }
            Value v;
            if (!val.empty() && val.front() == '[') {
                if (val.back() != ']') throw std::runtime_error("toml: bad array at line " + std::to_string(lineno));
                v.is_array = true;
                std::string body = val.substr(1, val.size() - 2), cur;
                bool in_str = false;
                for (char ch : body) {
                    if (ch == '"') in_str = !in_str;
                    if (ch == ',' && !in_str) { if (!trim(cur).empty()) v.items.push_back(trim(cur)); cur.clear(); }
                    else cur += ch;
                }
                if (!trim(cur).empty()) v.items.push_back(trim(cur));
            } else {
                v.text = val;
                v.is_string = !val.empty() && val.front() == '"';
            }
            t.kv[prefix.empty() ? key : prefix + "." + key] = v;
        }
        return t;
    }
    const Value* find(const std::string& k) const
    {
        auto it = kv.find(k);
        return it == kv.end() ? nullptr : &it->second;
    }
    static std::string unquote(const std::string& s)
    {
        if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
        return s;
    }
    double get_double(const std::string& k, double def) const
    {
        const Value* v = find(k);
        if (!v || v->is_array || v->is_string) return def;
        return std::stod(v->text);
    }
    long get_int(const std::string& k, long def) const
    {
        const Value* v = find(k);
        if (!v || v->is_array || v->is_string) return def;
        if (v->text.find_first_of(".eE") != std::string::npos) return def;   // a float is not an int
        return std::stol(v->text);
    }
    bool get_bool(const std::string& k, bool def) const
    {
        const Value* v = find(k);
        if (!v || v->is_array) return def;
        if (v->text == "true") return true;
        if (v->text == "false") return false;
        return def;
    }
};

}  // namespace dqmc_toml

namespace DQMC {

// Mirrors include/dqmc_params.hpp:20-45 of the reference, plus the engine options of the
// [Engine] table (ignored by the reference's parser).
struct Params {
    int nl{}, ns{}, ng{};
    int nt{};
    double dt{}, beta{};
    double nnt{}, mu{}, u{};
    int sweeps_warmup{};
    int stabilization_pace{};
    std::variant<std::string, std::vector<int>> momentum_list{};
    std::set<std::string> observable_list{};
    int bin_num{};
    int bin_capacity{};
    int sweeps_between_bins{};
    bool is_fft{};
    // [Engine]
    int chains{1};
    int precision{0};            // 0 = fp64, 1 = mixed
    int delay{32};
    long seed{-1};               // < 0: time-based like the reference (src/main.cpp:111)
    int use_graphs{1};
};

inline const std::set<std::string>& all_observables()
{
    static const std::set<std::string> s = {"FillingNumber", "DoubleOccupation", "StaticSWavePairingCorrelation",
                                            "DynamicGreenFunctions", "LocalDensityOfStates",
                                            "LocalDynamicTransverseSpinCorrelation"};
    return s;
}
inline bool is_equaltime_observable(const std::string& n)
{
    return n == "FillingNumber" || n == "DoubleOccupation" || n == "StaticSWavePairingCorrelation";
}

namespace Parser {

template <typename T>
void range_check(T val, const std::string& name, T lower)
{
    if (val < lower) throw std::runtime_error("DQMC::Parser::range_check<T>(): '" + name + "' out of range.");
}

// Same keys, defaults, observable keywords and range checks as include/dqmc_parser.hpp:49-133.
// `world_size` splits bin_num like the reference splits it over MPI ranks (:94); here the
// "ranks" are the batched chains.
inline void parse_toml_config(const std::string& path, int world_size, Params& p)
{
    const dqmc_toml::Table c = dqmc_toml::Table::parse_file(path);
    if (world_size <= 0) world_size = (int)c.get_int("Engine.chains", 1);   // chains stand in for ranks
    if (world_size < 1) throw std::runtime_error("DQMC::Parser::range_check<T>(): 'chains' out of range.");
    p.nnt = c.get_double("Model.HubbardAttractiveU.nnt", 1.0);
    p.u = c.get_double("Model.HubbardAttractiveU.u", 4.0);
    p.mu = c.get_double("Model.HubbardAttractiveU.mu", 0.0);
    p.nl = (int)c.get_int("Lattice.SquareLattice.nl", 4);
    p.ns = p.nl * p.nl;
    p.ng = p.ns;
    const dqmc_toml::Value* ml = c.find("Lattice.SquareLattice.momentum_list");
    if (!ml) p.momentum_list = std::string("");
    else if (ml->is_string) p.momentum_list = dqmc_toml::Table::unquote(ml->text);
    else if (ml->is_array) {
        std::vector<int> m;
        for (const auto& it : ml->items) {
            if (it.empty() || it.front() == '"' || it.find('.') != std::string::npos)
                throw std::runtime_error("DQMC::Parser::parse_toml_config(): invalid input momentum list.");
            m.push_back(std::stoi(it));
        }
        p.momentum_list = m;
    } else throw std::runtime_error("DQMC::Parser::parse_toml_config(): invalid input momentum list.");
    p.beta = c.get_double("MonteCarlo.beta", 8.0);
    p.nt = (int)c.get_int("MonteCarlo.nt", 160);
    p.dt = p.beta / p.nt;
    p.stabilization_pace = (int)c.get_int("MonteCarlo.stabilization_pace", 10);
    p.is_fft = c.get_bool("MonteCarlo.is_fft", false);
    p.sweeps_warmup = (int)c.get_int("Measurement.sweeps_warmup", 1000);
    p.bin_num = (int)std::ceil(c.get_int("Measurement.bin_num", 20) / (double)world_size);
    p.bin_capacity = (int)c.get_int("Measurement.bin_capacity", 100);
    p.sweeps_between_bins = (int)c.get_int("Measurement.sweeps_between_bins", 10);
    const dqmc_toml::Value* ob = c.find("Measurement.observables");
    if (!ob || !ob->is_array) throw std::runtime_error("DQMC::Parser::parse_toml_config(): invalid input observables.");
    std::vector<std::string> obs;
    for (const auto& it : ob->items) {
        if (it.empty() || it.front() != '"') throw std::runtime_error("DQMC::Parser::parse_toml_config(): invalid input observables.");
        obs.push_back(dqmc_toml::Table::unquote(it));
    }
    const bool is_all = std::count(obs.begin(), obs.end(), "all") || std::count(obs.begin(), obs.end(), "All");
    const bool is_none = std::count(obs.begin(), obs.end(), "none") || std::count(obs.begin(), obs.end(), "None");
    if (is_all && is_none)
        throw std::runtime_error("DQMC::Parser::parse_toml_config(): recieved conflict observable options 'all/All' and 'none/None'.");
    if (is_all) p.observable_list = all_observables();
    else if (is_none) p.observable_list = {};
    else p.observable_list = std::set<std::string>(obs.begin(), obs.end());
    // [Engine] options
    p.chains = (int)c.get_int("Engine.chains", world_size);
    const dqmc_toml::Value* pr = c.find("Engine.precision");
    p.precision = (pr && dqmc_toml::Table::unquote(pr->text) == "mixed") ? 1 : 0;
    p.delay = (int)c.get_int("Engine.delay", 32);
    p.seed = c.get_int("Engine.seed", -1);
    p.use_graphs = c.get_bool("Engine.graphs", true) ? 1 : 0;

    range_check(p.u, "u", 0.);
    range_check(p.nl, "nl", 1);
    range_check(p.beta, "beta", 0.);
    range_check(p.nt, "nt", 1);
    range_check(p.stabilization_pace, "stabilization_pace", 1);
    range_check(p.sweeps_warmup, "sweeps_warmup", 0);
    range_check(p.bin_num, "bin_num", 0);
    range_check(p.bin_capacity, "bin_capacity", 0);
    range_check(p.sweeps_between_bins, "sweeps_between_bins", 0);
}

}  // namespace Parser
}  // namespace DQMC

namespace Lattice {

// Tables of src/square_lattice.cpp used by measurements: k-stars (:83-95), momentum lists
// (:97-118), displacement (:63-81) and exp(-i k.r) factors (:120-135).
class SquareLattice {
   public:
    int nl = 0, ns = 0, nkstars = 0;
    std::vector<double> kstars;          // [nkstars][2]
    std::vector<int> all_kstars, x2k2x, gamma2m;
    std::vector<int> displacement;       // [i][j] -> site index of rj - ri

    void initialize(const DQMC::Params& p)
    {
        nl = p.nl;
        ns = p.ns;
        const int h = (int)std::floor(nl / 2.0) + 1;
        nkstars = h * h;
        kstars.assign(2 * (size_t)nkstars, 0.0);
        int n = 0;
        for (int i = (int)std::ceil(nl / 2.0); i <= nl; ++i)
            for (int j = (int)std::ceil(nl / 2.0); j <= nl; ++j) {
                kstars[2 * n] = (double)i / nl * 2 * M_PI - M_PI;
                kstars[2 * n + 1] = (double)j / nl * 2 * M_PI - M_PI;
                ++n;
            }
        all_kstars.clear(); x2k2x.clear(); gamma2m.clear();
        for (int i = 0; i < nkstars; ++i) all_kstars.push_back(i);
        for (int i = 0; i < h; ++i) x2k2x.push_back((i + 1) * (int)std::floor(nl / 2.0));
        for (int i = 0; i < h; ++i) gamma2m.push_back(i * ((int)std::floor(nl / 2.0) + 2));
        displacement.assign((size_t)ns * ns, 0);
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ns; ++j) {
                const int dx = ((j % nl) - (i % nl) + nl) % nl, dy = ((j / nl) - (i / nl) + nl) % nl;
                displacement[(size_t)i * ns + j] = dx + dy * nl;
            }
    }
    double kx(int k) const { return kstars[2 * (size_t)k]; }
    double ky(int k) const { return kstars[2 * (size_t)k + 1]; }
    // Re exp(-i k.r) for the displacement site index r
    double fourier_re(int r, int k) const { return std::cos((r % nl) * kx(k) + (r / nl) * ky(k)); }
};

}  // namespace Lattice

namespace Observable {

// Binned observable: data [nbin, shape...], per-bin cache, statistics (include/observable.hpp:160-207).
class ObservableReal {
   public:
    std::string name;
    std::vector<size_t> shape;   // empty = scalar
    size_t nbin = 0, size = 1;
    std::vector<double> data, cache, mean, stddev, stderr_;
    long count = 0;

    void set_shape(size_t nb, std::vector<size_t> sh)
    {
        nbin = nb;
        shape = std::move(sh);
        size = 1;
        for (size_t d : shape) size *= d;
        data.assign(nbin * size, 0.0);
        cache.assign(size, 0.0);
        mean.assign(size, 0.0);
        stddev.assign(size, 0.0);
        stderr_.assign(size, 0.0);
        count = 0;
    }
    void normalize_cache()
    {
        if (count <= 0) throw std::runtime_error("Observable::Observable<_scalar>::normalize_cache(): cache divided by zero.");
        for (auto& v : cache) v /= count;
    }
    void push_cache_to_data(size_t bin)
    {
        if (bin >= nbin) throw std::out_of_range("Observable::Observable<_scalar>::push_cache_to_data(): index of bin out of range.");
        std::copy(cache.begin(), cache.end(), data.begin() + bin * size);
    }
    void clear_cache() { std::fill(cache.begin(), cache.end(), 0.0); count = 0; }
    // mean, stddev with ddof = 0, stderr = stddev / sqrt(nbin)
    void analyse()
    {
        for (size_t e = 0; e < size; ++e) {
            double m = 0.0;
            for (size_t b = 0; b < nbin; ++b) m += data[b * size + e];
            m /= nbin;
            double v = 0.0;
            for (size_t b = 0; b < nbin; ++b) { const double d = data[b * size + e] - m; v += d * d; }
            mean[e] = m;
            stddev[e] = std::sqrt(v / nbin);
            stderr_[e] = stddev[e] / std::sqrt((double)nbin);
        }
    }
};

}  // namespace Observable

namespace DQMC {
namespace IO {

inline std::string fmt(const char* f, double v)
{
    char b[64];
    std::snprintf(b, sizeof(b), f, v);
    return b;
}
inline std::string fmti(long v)
{
    char b[32];
    std::snprintf(b, sizeof(b), "% 15ld", v);
    return b;
}
// A scalar observable reaches the reference's stream as a 0-d array, whose printer drops trailing
// zeros of the 15-decimal fixed form and pads with spaces in their place (example/double_occu.out).
inline std::string fmt_scalar30(double v)
{
    std::string s = fmt("%.15f", v);
    size_t k = 0;
    while (k < 14 && s.size() > 2 && s[s.size() - 1 - k] == '0') ++k;
    s = s.substr(0, s.size() - k) + std::string(k, ' ');
    if (s.size() < 30) s = std::string(30 - s.size(), ' ') + s;
    return s;
}

// print_observable (include/dqmc_io.hpp:130-173)
inline void print_observable(std::ostream& os, const Observable::ObservableReal& o, bool show_header = true)
{
    if (o.shape.empty()) {
        if (show_header) os << fmti(1) << "\n";
        os << fmti(0) << fmt_scalar30(o.mean[0]) << fmt_scalar30(o.stddev[0]) << fmt_scalar30(o.stderr_[0]) << "\n";
        return;
    }
    if (show_header) {
        for (size_t d : o.shape) os << fmti((long)d);
        os << "\n";
    }
    std::vector<size_t> acc(o.shape.size(), 1);
    for (size_t d = o.shape.size(); d-- > 1;) acc[d - 1] = acc[d] * o.shape[d];
    for (size_t i = 0; i < o.size; ++i) {
        for (size_t d = 0; d < o.shape.size(); ++d) os << fmti((long)((i / acc[d]) % o.shape[d]));
        os << fmt("% 30.15f", o.mean[i]) << fmt("% 30.15f", o.stddev[i]) << fmt("% 30.15f", o.stderr_[i]) << "\n";
    }
}

// print_observable_data (include/dqmc_io.hpp:175-207): data shape [nbin, shape...]
inline void print_observable_data(std::ostream& os, const Observable::ObservableReal& o, bool show_header = true)
{
    std::vector<size_t> shape = {o.nbin};
    shape.insert(shape.end(), o.shape.begin(), o.shape.end());
    if (show_header) {
        for (size_t d : shape) os << fmti((long)d);
        os << "\n";
    }
    std::vector<size_t> acc(shape.size(), 1);
    for (size_t d = shape.size(); d-- > 1;) acc[d - 1] = acc[d] * shape[d];
    const size_t total = o.nbin * o.size;
    for (size_t i = 0; i < total; ++i) {
        for (size_t d = 0; d < shape.size(); ++d) os << fmti((long)((i / acc[d]) % shape[d]));
        os << fmt("% 30.15f", o.data[i]) << "\n";
    }
}

// save_ising_fields_configs (include/dqmc_io.hpp:234-255); s is [t][i]
inline void save_ising_fields(const std::string& file, int nt, int ns, const std::vector<int>& s)
{
    std::ofstream f(file, std::ios::out | std::ios::trunc);
    if (!f.is_open()) throw std::runtime_error("DQMC::IO::save_ising_fields_configs(): fail to open '" + file + "'.");
    f << fmti(nt) << fmti(ns) << "\n";
    for (int t = 0; t < nt; ++t)
        for (int i = 0; i < ns; ++i) f << fmti(t) << fmti(i) << fmti(s[(size_t)t * ns + i]) << "\n";
}

// load_ising_fields_configs (include/dqmc_io.hpp:257-290): dims must match, rows in any order
inline void load_ising_fields(const std::string& file, int nt, int ns, std::vector<int>& s)
{
    std::ifstream f(file);
    if (!f.is_open()) throw std::runtime_error("DQMC::IO::load_ising_fields_configs(): fail to open '" + file + "'.");
    long fnt = -1, fns = -1;
    std::string line;
    if (!std::getline(f, line)) throw std::runtime_error("DQMC::IO::load_ising_fields_configs(): empty file.");
    std::istringstream h(line);
    h >> fnt >> fns;
    if (fnt != nt || fns != ns)
        throw std::runtime_error("DQMC::IO::load_ising_fields_configs(): contradictory dimensions of the ising fields.");
    s.assign((size_t)nt * ns, 0);
    while (std::getline(f, line)) {
        std::istringstream r(line);
        long t, i, v;
        if (!(r >> t >> i >> v)) continue;
        if (t < 0 || t >= nt || i < 0 || i >= ns)
            throw std::runtime_error("DQMC::IO::load_ising_fields_configs(): row index out of range.");
        s[(size_t)t * ns + i] = (int)v;
    }
}

// print_momentum_list (include/dqmc_io.hpp:297-313)
inline void print_momentum_list(std::ostream& os, const std::vector<int>& mlist, const Lattice::SquareLattice& lat)
{
    os << fmti((long)mlist.size()) << fmti((long)lat.all_kstars.size()) << "\n";
    for (size_t k = 0; k < mlist.size(); ++k)
        os << fmti((long)k) << fmti(mlist[k]) << fmt("% 30.15f", lat.kx(mlist[k])) << fmt("% 30.15f", lat.ky(mlist[k])) << "\n";
}

// print_imaginary_time_grids (include/dqmc_io.hpp:315-328)
inline void print_imaginary_time_grids(std::ostream& os, const Params& p)
{
    os << fmti(p.nt) << fmt("% 30.8f", p.beta) << fmt("% 30.8f", p.dt) << "\n";
    for (int t = 0; t < p.nt; ++t) os << fmti(t) << fmt("% 30.8f", (double)(t * p.dt)) << "\n";
}

}  // namespace IO
}  // namespace DQMC

namespace Measurement {

// Observable registry, flags and bins of one chain (src/measure_handle.cpp).  The estimators of
// include/observable_methods.hpp run on the device (dqmc_measure_*); the bins land in `obs`.
class Handle {
   public:
    int sweeps_warmup = 0, bin_num = 0, bin_capacity = 0, sweeps_between_bins = 0;
    bool warmup = false, equaltime = false, dynamic = false;
    std::vector<int> momentum_list;
    std::map<std::string, Observable::ObservableReal> obs;
    std::vector<std::string> eq_list, dyn_list;

    // One handle holds one chain's bins (bin_num of them), like one MPI rank of the reference.
    void initialize(const DQMC::Params& p, const Lattice::SquareLattice& lat)
    {
        sweeps_warmup = p.sweeps_warmup;
        bin_num = p.bin_num;
        bin_capacity = p.bin_capacity;
        sweeps_between_bins = p.sweeps_between_bins;
        if (auto s = std::get_if<std::string>(&p.momentum_list)) {
            if (*s == "AllKstars") momentum_list = lat.all_kstars;
            else if (*s == "X2K2XLine") momentum_list = lat.x2k2x;
            else if (*s == "Gamma2MLine") momentum_list = lat.gamma2m;
            else throw std::runtime_error("Measurement::Handle::initialize(): invalid momentum list.");
        } else {
            momentum_list = std::get<std::vector<int>>(p.momentum_list);
            for (int k : momentum_list)
                if (k >= lat.nkstars) throw std::runtime_error("Measurement::Handle::initialize(): index of momentum out of range.");
        }
        for (const auto& n : p.observable_list) {
            if (!DQMC::all_observables().count(n))
                throw std::runtime_error("Observable::Handle::initialize(): undefined observable '" + n + "'.");
            (DQMC::is_equaltime_observable(n) ? eq_list : dyn_list).push_back(n);
        }
        warmup = sweeps_warmup > 0;
        equaltime = !eq_list.empty() && bin_num > 0 && bin_capacity > 0;
        dynamic = !dyn_list.empty() && bin_num > 0 && bin_capacity > 0;
        const size_t nb = (size_t)bin_num, nt = p.nt, nk = momentum_list.size();
        for (const auto& n : eq_list) { obs[n].name = n; obs[n].set_shape(nb, {}); }
        for (const auto& n : dyn_list) {
            obs[n].name = n;
            if (n == "DynamicGreenFunctions") obs[n].set_shape(nb, {nk, nt});
            else obs[n].set_shape(nb, {nt});
        }
    }

    void normalize_cache() { for (auto& kv : obs) kv.second.normalize_cache(); }
    void push_cache_to_data(size_t bin) { for (auto& kv : obs) kv.second.push_cache_to_data(bin); }
    void clear_cache() { for (auto& kv : obs) kv.second.clear_cache(); }
    void analyse() { for (auto& kv : obs) kv.second.analyse(); }
    bool is_found(const std::string& n) const { return obs.count(n) > 0; }

    // Concatenate the chains' bins in chain order, as the reference's post-run gather
    // concatenates the ranks' bins along the bin axis (include/utils/observable_gather.hpp:30-69).
    static Handle gather(const std::vector<Handle>& per_chain)
    {
        Handle h = per_chain.at(0);
        for (auto& kv : h.obs) {
            auto& o = kv.second;
            const size_t nb = o.nbin * per_chain.size();
            std::vector<double> data;
            data.reserve(nb * o.size);
            for (const auto& c : per_chain) {
                const auto& oc = c.obs.at(kv.first);
                data.insert(data.end(), oc.data.begin(), oc.data.end());
            }
            const auto shape = o.shape;
            o.set_shape(nb, shape);
            o.data = data;
        }
        h.bin_num = (int)(h.bin_num * per_chain.size());
        return h;
    }
};

}  // namespace Measurement
