// main.cpp -- drop-in application for the reference's executable (src/main.cpp): same command
// line (-c config.toml, -o output dir, -f ising.fields, -h), same configuration keys, same
// output files.  The MPI ranks of the reference become chains batched on one GPU; an optional
// [Engine] table selects chains, precision ("fp64" / "mixed"), delay, seed and graphs.
#include <sys/stat.h>

#include <cstdio>
#include <cstring>
#include <ctime>
#include <iostream>

#include "dqmc_facade.hpp"

namespace {

void usage(const char* argv0)
{
    std::cerr << argv0 << "\nProgram options:\n"
              << "  -h [ --help ]                 display this information.\n"
              << "  -c [ --config ] arg (=./example/config.toml)\n"
              << "                                path of the toml configuration file.\n"
              << "  -o [ --output ] arg (=./example)\n"
              << "                                path of the folder where the program output are saved.\n"
              << "  -f [ --fields ] arg           path of the file with initial configurations of the\n"
              << "                                auxiliary Ising fields (random if not assigned).\n";
}

std::string fmt_line_str(const char* name, const std::string& v)
{
    char b[128];
    std::snprintf(b, sizeof(b), "%30s%7s%20s\n", name, "->", v.c_str());
    return b;
}
std::string fmt_line_int(const char* name, long v)
{
    char b[128];
    std::snprintf(b, sizeof(b), "%30s%7s%20ld\n", name, "->", v);
    return b;
}
std::string fmt_line_dbl(const char* name, double v)
{
    char b[128];
    std::snprintf(b, sizeof(b), "%30s%7s%20.3f\n", name, "->", v);
    return b;
}

// print_initialization_info (include/dqmc_io.hpp:41-90)
void print_initialization_info(const DQMC::Params& p, const Measurement::Handle& m, int nproc)
{
    auto b2s = [](bool b) { return std::string(b ? "True" : "False"); };
    std::cout << ">> Model: HubbardAttractiveU\n\n"
              << fmt_line_dbl("Nearest hopping 'nnt'", p.nnt) << fmt_line_dbl("Chemical potential 'mu'", p.mu)
              << fmt_line_dbl("On-site interaction 'u'", p.u) << "\n";
    std::cout << ">> Lattice: SquareLattice\n\n" << fmt_line_int("Linear size 'nl'", p.nl);
    if (auto s = std::get_if<std::string>(&p.momentum_list)) std::cout << fmt_line_str("Momentum list", *s) << "\n";
    else std::cout << fmt_line_str("Momentum list", "Custom") << "\n";
    std::cout << ">> MonteCarlo Params:\n\n" << fmt_line_dbl("Inverse temperature", p.beta)
              << fmt_line_int("Number of imag-time slice ", p.nt) << fmt_line_dbl("Imag-time spacing", p.dt)
              << fmt_line_int("Stabilization pace", p.stabilization_pace)
              << fmt_line_str("FFT-implemented expK mult.", b2s(p.is_fft)) << "\n";
    std::cout << ">> Measurement Params:\n\n" << fmt_line_str("Warm up", b2s(m.warmup))
              << fmt_line_str("Equal-time measure", b2s(m.equaltime)) << fmt_line_str("Dynamic measure", b2s(m.dynamic))
              << fmt_line_int("Sweeps for warmup", (long)(2 * std::ceil(m.sweeps_warmup / 2.0)))
              << fmt_line_int("Number of bins", (long)m.bin_num * nproc)
              << fmt_line_int("Sweeps per bin", (!m.dynamic && m.equaltime) ? (long)(2 * std::ceil(m.bin_capacity / 2.0)) : m.bin_capacity)
              << fmt_line_int("Sweeps between bins", (long)(2 * std::ceil(m.sweeps_between_bins / 2.0))) << std::endl;
}

// print_dqmc_summary (include/dqmc_io.hpp:92-123)
void print_dqmc_summary(const DQMC::Core& core, const Measurement::Handle& m, double ms)
{
    const int day = (int)std::floor(ms / 86400000), hour = (int)std::floor((ms / 1000 - day * 86400) / 3600);
    const int minute = (int)std::floor((ms / 1000 - day * 86400 - hour * 3600) / 60);
    const double sec = ms / 1000 - 86400 * day - 3600 * hour - 60 * minute;
    char b[256];
    if (day) std::snprintf(b, sizeof(b), "\n>> The simulation finished in %d d %d h %d m %.2f s.\n", day, hour, minute, sec);
    else if (hour) std::snprintf(b, sizeof(b), "\n>> The simulation finished in %d h %d m %.2f s.\n", hour, minute, sec);
    else if (minute) std::snprintf(b, sizeof(b), "\n>> The simulation finished in %d m %.2f s.\n", minute, sec);
    else std::snprintf(b, sizeof(b), "\n>> The simulation finished in %.2f s.\n", sec);
    std::cout << b << std::endl;
    if ((m.warmup || m.equaltime) && m.dynamic)
        std::snprintf(b, sizeof(b), ">> Maximum of the wrapping error: %.5e (equal-time) and %.5e (dynamic).\n",
                      core.equaltimeWrapError(), core.dynamicWrapError());
    else if (m.warmup || m.equaltime)
        std::snprintf(b, sizeof(b), ">> Maximum of the wrapping error: %.5e (equal-time).\n", core.equaltimeWrapError());
    else if (m.dynamic)
        std::snprintf(b, sizeof(b), ">> Maximum of the wrapping error: %.5e (dynamic).\n", core.dynamicWrapError());
    else b[0] = 0;
    std::cout << b << std::endl;
}

void write_obs(const Measurement::Handle& m, const std::string& name, const std::string& dir, const std::string& stem)
{
    if (!m.is_found(name)) return;
    const auto& o = m.obs.at(name);
    std::ofstream f(dir + "/" + stem + ".out", std::ios::out | std::ios::trunc);
    DQMC::IO::print_observable(f, o, true);
    std::ofstream g(dir + "/" + stem + ".bins.out", std::ios::out | std::ios::trunc);
    DQMC::IO::print_observable_data(g, o, true);
}

}  // namespace

int main(int argc, char* argv[])
{
    std::string config = "./example/config.toml", output = "./example", fields;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](std::string& dst) {
            if (i + 1 >= argc) { usage(argv[0]); std::exit(2); }
            dst = argv[++i];
        };
        if (a == "-h" || a == "--help") { usage(argv[0]); return 0; }
        else if (a == "-c" || a == "--config") need(config);
        else if (a == "-o" || a == "--output") need(output);
        else if (a == "-f" || a == "--fields") need(fields);
        else { std::cerr << "main(): invalid program options got from the command line." << std::endl; usage(argv[0]); return 2; }
    }
    try {
        mkdir(output.c_str(), 0755);
        const std::time_t now = std::time(nullptr);
        std::cout << ">> Current time: " << std::asctime(std::localtime(&now)) << std::endl;

        DQMC::Params params;
        DQMC::Parser::parse_toml_config(config, 0, params);
        std::cout << ">> Distribute tasks to " << params.chains << " chains on one GPU.\n" << std::endl;

        Lattice::SquareLattice lattice;
        Model::HubbardAttractiveU model;
        DQMC::Core core;
        lattice.initialize(params);
        model.initialize(params, lattice);
        std::vector<Measurement::Handle> meas(params.chains);
        for (auto& m : meas) m.initialize(params, lattice);
        core.initialize(params, meas[0]);

        // seeds: base + chain (the reference uses time(nullptr) + rank, src/main.cpp:111)
        const uint64_t base = params.seed >= 0 ? (uint64_t)params.seed : (uint64_t)std::time(nullptr);
        for (int c = 0; c < params.chains; ++c) model.set_seed(c, base + c);
        if (!fields.empty()) {
            std::vector<int> s;
            DQMC::IO::load_ising_fields(fields, params.nt, params.ns, s);
            for (int c = 0; c < params.chains; ++c) model.set_fields(c, s);
            std::cout << ">> Configurations of ising fields loaded from file.\n" << std::endl;
        } else {
            model.set_ising_fields_to_random();
            std::cout << ">> Configurations of ising fields initialized randomly.\n" << std::endl;
        }
        core.initialize_svdstacks(model);
        core.setup_measurement(meas[0], lattice);
        core.initialize_green_functions();
        std::cout << ">> Initialization finished.\n\n>> The simulation is going to get started with parameters below:\n"
                  << std::endl;
        print_initialization_info(params, meas[0], params.chains);

        DQMC::Handle handle;
        handle.timer_begin();
        handle.thermalize(core, meas[0], model);
        handle.measure(core, meas, model, lattice);
        Measurement::Handle all = Measurement::Handle::gather(meas);
        all.analyse();
        handle.timer_end();
        print_dqmc_summary(core, meas[0], handle.timer_ms());

        write_obs(all, "FillingNumber", output, "filling");
        write_obs(all, "DoubleOccupation", output, "double_occu");
        write_obs(all, "DynamicGreenFunctions", output, "gf");
        write_obs(all, "LocalDensityOfStates", output, "ldos");
        write_obs(all, "StaticSWavePairingCorrelation", output, "swave_corr");
        write_obs(all, "LocalDynamicTransverseSpinCorrelation", output, "spin_corr");
        const std::string fout = fields.empty() ? output + "/ising.fields" : fields;
        DQMC::IO::save_ising_fields(fout, params.nt, params.ns, model.fields(0));
        for (int c = 1; c < params.chains; ++c)
            DQMC::IO::save_ising_fields(output + "/ising.fields." + std::to_string(c), params.nt, params.ns, model.fields(c));
        {
            std::ofstream f(output + "/momenta.out", std::ios::out | std::ios::trunc);
            DQMC::IO::print_momentum_list(f, all.momentum_list, lattice);
            std::ofstream g(output + "/tgrids.out", std::ios::out | std::ios::trunc);
            DQMC::IO::print_imaginary_time_grids(g, params);
        }
        std::cout << ">> See the results under folder '" << output << "'." << std::endl;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
