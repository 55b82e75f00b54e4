// CPU checks of the host-side drop-in pieces (hubbard-attractive-u_amd/host/dqmc_host.hpp):
//   host_check parse <config.toml>                      -> key=value lines of the parsed Params
//   host_check fields <in> <out> <nt> <ns>              -> load then save an ising.fields file
//   host_check stats <bins.out> <out> <ndim> <d1> ...   -> *.out recomputed from a *.bins.out
//   host_check momenta <nl> <out>                       -> momenta.out for AllKstars
//   host_check tgrids <nt> <beta> <out>                 -> tgrids.out
#include <cstdlib>
#include <iostream>

#include "../../hubbard-attractive-u_amd/host/dqmc_host.hpp"

int main(int argc, char** argv)
{
    const std::string cmd = argc > 1 ? argv[1] : "";
    try {
        if (cmd == "parse") {
            DQMC::Params p;
            DQMC::Parser::parse_toml_config(argv[2], argc > 3 ? std::atoi(argv[3]) : 1, p);
            std::printf("nl=%d ns=%d nt=%d beta=%.17g dt=%.17g nnt=%.17g u=%.17g mu=%.17g pace=%d warmup=%d bins=%d cap=%d between=%d fft=%d nobs=%zu\n",
                        p.nl, p.ns, p.nt, p.beta, p.dt, p.nnt, p.u, p.mu, p.stabilization_pace, p.sweeps_warmup,
                        p.bin_num, p.bin_capacity, p.sweeps_between_bins, (int)p.is_fft, p.observable_list.size());
            if (auto s = std::get_if<std::string>(&p.momentum_list)) std::printf("momentum=%s\n", s->c_str());
            return 0;
        }
        if (cmd == "fields") {
            const int nt = std::atoi(argv[4]), ns = std::atoi(argv[5]);
            std::vector<int> s;
            DQMC::IO::load_ising_fields(argv[2], nt, ns, s);
            DQMC::IO::save_ising_fields(argv[3], nt, ns, s);
            return 0;
        }
        if (cmd == "stats") {
            const int ndim = std::atoi(argv[4]);
            std::vector<size_t> shape;
            for (int d = 0; d < ndim; ++d) shape.push_back((size_t)std::atol(argv[5 + d]));
            std::ifstream in(argv[2]);
            std::vector<double> vals;
            std::string line;
            while (std::getline(in, line)) {
                std::istringstream r(line);
                std::vector<double> cols;
                double x;
                while (r >> x) cols.push_back(x);
                if ((int)cols.size() == ndim + 2) vals.push_back(cols.back());
            }
            size_t sz = 1;
            for (size_t d : shape) sz *= d;
            Observable::ObservableReal o;
            o.set_shape(vals.size() / sz, shape);
            o.data = vals;
            o.analyse();
            std::ofstream out(argv[3]);
            DQMC::IO::print_observable(out, o, false);
            return 0;
        }
        if (cmd == "momenta") {
            DQMC::Params p;
            p.nl = std::atoi(argv[2]);
            p.ns = p.nl * p.nl;
            Lattice::SquareLattice lat;
            lat.initialize(p);
            std::ofstream out(argv[3]);
            DQMC::IO::print_momentum_list(out, lat.all_kstars, lat);
            return 0;
        }
        if (cmd == "tgrids") {
            DQMC::Params p;
            p.nt = std::atoi(argv[2]);
            p.beta = std::atof(argv[3]);
            p.dt = p.beta / p.nt;
            std::ofstream out(argv[4]);
            DQMC::IO::print_imaginary_time_grids(out, p);
            return 0;
        }
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 3;
    }
    std::cerr << "unknown command" << std::endl;
    return 2;
}
