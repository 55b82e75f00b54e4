// dqmc_facade.hpp -- C++ facades with the reference's class and method names
// (Model::HubbardAttractiveU, DQMC::Core, DQMC::Handle; include/hubbard.h:31-111,
// include/dqmc_core.h:35-112, include/dqmc_handle.h) implemented over the C ABI of the MI355X
// engine (include/dqmc_amd.h).  One engine context holds all chains; each chain plays the role
// of one MPI rank of the reference.
#pragma once

#include <chrono>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dqmc_amd.h"
#include "dqmc_host.hpp"

namespace DQMC {

inline void check(dqmc_ctx* c, int rc, const char* what)
{
    if (rc != DQMC_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + dqmc_last_error(c));
}

// Column-major N x N host matrix: the reference's Eigen::MatrixXd for G accessors.
struct Matrix {
    int n = 0;
    std::vector<double> a;
    Matrix() = default;
    explicit Matrix(int n_) : n(n_), a((size_t)n_ * n_, 0.0) {}
    double& operator()(int i, int j) { return a[i + (size_t)j * n]; }
    double operator()(int i, int j) const { return a[i + (size_t)j * n]; }
    double* data() { return a.data(); }
    const double* data() const { return a.data(); }
    int rows() const { return n; }
    int cols() const { return n; }
    double trace() const { double s = 0; for (int i = 0; i < n; ++i) s += (*this)(i, i); return s; }
};

}  // namespace DQMC

namespace Model {

class HubbardAttractiveU {
   public:
    dqmc_ctx* ctx = nullptr;
    int nl = 0, ns = 0, nt = 0, chains = 1;
    double dt = 0, nnt = 0, mu = 0, u = 0, alpha = 0;

    ~HubbardAttractiveU() { if (ctx) dqmc_destroy(ctx); }

    // HubbardAttractiveU::initialize (src/hubbard.cpp:25-48): builds the engine context.
    void initialize(const DQMC::Params& p, const Lattice::SquareLattice&)
    {
        nl = p.nl; ns = p.ns; nt = p.nt; dt = p.dt; nnt = p.nnt; mu = p.mu; u = p.u; chains = p.chains;
        alpha = std::acosh(std::exp(0.5 * dt * u));
        dqmc_params_c c{p.nl, p.nt, p.beta, p.nnt, p.u, p.mu, p.stabilization_pace, p.is_fft ? 1 : 0, 0, 0};
        DQMC::check(nullptr, dqmc_create(&c, chains, p.precision, std::max(1, std::min(32, p.delay)), &ctx), "dqmc_create");
        dqmc_use_graphs(ctx, p.use_graphs);
    }
    void set_seed(int chain, uint64_t seed) { DQMC::check(ctx, dqmc_set_seed(ctx, chain, seed), "dqmc_set_seed"); }
    // set_ising_fields_to_random (src/hubbard.cpp:102-110), every chain from its own stream
    void set_ising_fields_to_random()
    {
        for (int c = 0; c < chains; ++c) DQMC::check(ctx, dqmc_fields_random(ctx, c), "dqmc_fields_random");
    }
    std::vector<int> fields(int chain) const
    {
        std::vector<int8_t> s((size_t)nt * ns);
        DQMC::check(ctx, dqmc_get_fields(ctx, chain, s.data()), "dqmc_get_fields");
        return std::vector<int>(s.begin(), s.end());
    }
    void set_fields(int chain, const std::vector<int>& s)
    {
        std::vector<int8_t> b(s.begin(), s.end());
        DQMC::check(ctx, dqmc_set_fields(ctx, chain, b.data()), "dqmc_set_fields");
    }
    // multiply_B_from_left ... multiply_adjB_from_left (include/hubbard.h:58-62), in place
    void multiply_B_from_left(DQMC::Matrix& g, int t, int chain = 0) const { bop(g, 0, t, chain); }
    void multiply_B_from_right(DQMC::Matrix& g, int t, int chain = 0) const { bop(g, 1, t, chain); }
    void multiply_invB_from_left(DQMC::Matrix& g, int t, int chain = 0) const { bop(g, 2, t, chain); }
    void multiply_invB_from_right(DQMC::Matrix& g, int t, int chain = 0) const { bop(g, 3, t, chain); }
    void multiply_adjB_from_left(DQMC::Matrix& g, int t, int chain = 0) const { bop(g, 4, t, chain); }
    // locally_update_ising_field / get_acceptance_ratio / update_green_function (include/hubbard.h:71-73)
    // on the chain's current G; the Core argument of the reference is the engine context itself
    void locally_update_ising_field(int t, int i, int chain = 0)
    {
        DQMC::check(ctx, dqmc_flip_field(ctx, chain, t, i), "dqmc_flip_field");
    }
    double get_acceptance_ratio(int t, int i, int chain = 0) const
    {
        double r = 0.0;
        DQMC::check(ctx, dqmc_acceptance_ratio(ctx, chain, t, i, &r), "dqmc_acceptance_ratio");
        return r;
    }
    void update_green_function(int t, int i, int chain = 0)
    {
        DQMC::check(ctx, dqmc_update_green_function(ctx, chain, t, i), "dqmc_update_green_function");
    }

   private:
    void bop(DQMC::Matrix& g, int op, int t, int chain) const
    {
        DQMC::check(ctx, dqmc_apply_bop(ctx, chain, op, t, g.data()), "dqmc_apply_bop");
    }
};

}  // namespace Model

namespace DQMC {

class Core {
   public:
    Model::HubbardAttractiveU* model = nullptr;
    int ns = 0, nt = 0, chains = 1;
    bool is_equaltime = false, is_dynamic = false;

    // Core::initialize (src/dqmc_core.cpp:20-30)
    void initialize(const Params& p, const Measurement::Handle& m)
    {
        ns = p.ns; nt = p.nt; chains = p.chains;
        is_equaltime = m.equaltime;
        is_dynamic = m.dynamic;
    }
    void initialize_svdstacks(Model::HubbardAttractiveU& mdl)
    {
        model = &mdl;
        check(model->ctx, dqmc_set_collect(model->ctx, is_equaltime, is_dynamic), "dqmc_set_collect");
    }
    // initialize_svdstacks + initialize_green_functions run as one engine call
    void initialize_green_functions() { check(model->ctx, dqmc_init(model->ctx), "dqmc_init"); }
    void set_thermalization(bool th) { check(model->ctx, dqmc_set_thermalization(model->ctx, th), "set_thermalization"); }
    void sweep_from_0_to_beta(Model::HubbardAttractiveU&) { check(model->ctx, dqmc_sweep_up(model->ctx), "sweep_up"); }
    void sweep_from_beta_to_0(Model::HubbardAttractiveU&) { check(model->ctx, dqmc_sweep_down(model->ctx), "sweep_down"); }
    void sweep_for_dynamic_green_functions(Model::HubbardAttractiveU&)
    {
        check(model->ctx, dqmc_sweep_dynamic(model->ctx), "sweep_dynamic");
    }
    // accessors (include/dqmc_core.h:76-81), per chain
    Matrix current_gftt(int chain = 0) const { return get(DQMC_GF_CURRENT, 0, chain); }
    Matrix gftt(int t, int chain = 0) const { return get(DQMC_GF_TT, t, chain); }
    Matrix gft0(int t, int chain = 0) const { return get(DQMC_GF_T0, t, chain); }
    Matrix gf0t(int t, int chain = 0) const { return get(DQMC_GF_0T, t, chain); }
    double equaltimeWrapError() const { return stat(0); }
    double dynamicWrapError() const { return stat(1); }

    // Observables on the device (Measurement::Handle::initialize and the Observable::Methods
    // estimators, src/measure_handle.cpp:18-104): only the bins ever reach the host.
    void setup_measurement(const Measurement::Handle& m, const Lattice::SquareLattice& lat)
    {
        unsigned mask = 0;
        for (const auto& n : m.eq_list) mask |= obs_bit(n);
        for (const auto& n : m.dyn_list) mask |= obs_bit(n);
        if (!mask || (!m.equaltime && !m.dynamic)) return;
        std::vector<double> kxy;
        for (int k : m.momentum_list) { kxy.push_back(lat.kx(k)); kxy.push_back(lat.ky(k)); }
        check(model->ctx, dqmc_measure_setup(model->ctx, mask, m.bin_num, (int)m.momentum_list.size(), kxy.data()),
              "dqmc_measure_setup");
    }
    void measure_equaltime() { check(model->ctx, dqmc_measure_equaltime(model->ctx), "dqmc_measure_equaltime"); }
    void measure_dynamic() { check(model->ctx, dqmc_measure_dynamic(model->ctx), "dqmc_measure_dynamic"); }
    void push_bin(int bin) { check(model->ctx, dqmc_measure_push_bin(model->ctx, bin), "dqmc_measure_push_bin"); }
    // every chain's bins into its handle's observables (one handle per chain, like one per rank)
    void fetch_bins(std::vector<Measurement::Handle>& per_chain) const
    {
        for (int c = 0; c < chains; ++c)
            for (auto& kv : per_chain.at(c).obs)
                check(model->ctx, dqmc_get_bins(model->ctx, c, obs_bit(kv.first), kv.second.data.data()), "dqmc_get_bins");
    }
    static unsigned obs_bit(const std::string& n)
    {
        if (n == "FillingNumber") return DQMC_OBS_FILLING;
        if (n == "DoubleOccupation") return DQMC_OBS_DOUBLE_OCCU;
        if (n == "StaticSWavePairingCorrelation") return DQMC_OBS_SWAVE;
        if (n == "DynamicGreenFunctions") return DQMC_OBS_GREEN_KT;
        if (n == "LocalDensityOfStates") return DQMC_OBS_LDOS;
        return DQMC_OBS_SPIN;
    }

   private:
    Matrix get(int kind, int t, int chain) const
    {
        Matrix m(ns);
        check(model->ctx, dqmc_get_gf(model->ctx, chain, kind, t, m.data()), "dqmc_get_gf");
        return m;
    }
    double stat(int which) const
    {
        double best = 0.0;
        for (int c = 0; c < chains; ++c) {
            double we = 0, wd = 0, dg = 0;
            long long a = 0, v = 0;
            check(model->ctx, dqmc_get_stats(model->ctx, c, &we, &wd, &dg, &a, &v), "dqmc_get_stats");
            best = std::max(best, which == 0 ? we : wd);
        }
        return best;
    }
};

// Sweep schedule of DQMC::Handle (src/dqmc_handle.cpp:56-168) for all chains at once.
class Handle {
   public:
    std::chrono::steady_clock::time_point begin, end;
    void timer_begin() { begin = std::chrono::steady_clock::now(); }
    void timer_end() { end = std::chrono::steady_clock::now(); }
    double timer_ms() const { return (double)std::chrono::duration_cast<std::chrono::milliseconds>(end - begin).count(); }

    void thermalize(Core& core, const Measurement::Handle& m, Model::HubbardAttractiveU& model)
    {
        if (!m.warmup) return;
        core.set_thermalization(true);
        for (int s = 1; s <= (int)std::ceil(m.sweeps_warmup / 2.0); ++s) {
            core.sweep_from_0_to_beta(model);
            core.sweep_from_beta_to_0(model);
        }
    }

    void measure(Core& core, std::vector<Measurement::Handle>& per_chain, Model::HubbardAttractiveU& model,
                 const Lattice::SquareLattice& lat)
    {
        const Measurement::Handle& m = per_chain.at(0);
        if (!m.equaltime && !m.dynamic) return;
        const int binsize = m.dynamic ? m.bin_capacity : (int)std::ceil(m.bin_capacity / 2.0);
        core.set_thermalization(false);
        for (int bin = 0; bin < m.bin_num; ++bin) {
            for (int s = 1; s <= (int)std::ceil(m.sweeps_between_bins / 2.0); ++s) {
                core.sweep_from_0_to_beta(model);
                core.sweep_from_beta_to_0(model);
            }
            for (int s = 1; s <= binsize; ++s) {
                if (m.dynamic) {
                    core.sweep_for_dynamic_green_functions(model);
                    core.measure_dynamic();
                } else {
                    core.sweep_from_0_to_beta(model);
                    core.measure_equaltime();
                }
                core.sweep_from_beta_to_0(model);
                if (m.equaltime) core.measure_equaltime();
            }
            core.push_bin(bin);   // normalize_cache + push_cache_to_data + clear_cache, every chain
        }
        core.fetch_bins(per_chain);
        (void)lat;
    }
};

}  // namespace DQMC
