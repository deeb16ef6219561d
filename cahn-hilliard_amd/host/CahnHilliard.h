// CahnHilliard.h -- the simulation object of the drop-in (host side).
//
// Follows src/CahnHilliard.cpp:34-361: N must be a power of 2; the initial
// field comes from `load_from` (native text, ~6 significant digits) or from
// drand48 (initial_density, initial_A, initial_N_peaks; one draw per bin,
// x fastest, the generator seeded by the Manager with srand48(seed)); unit
// scaling; mobility before integrator; the integrator factory keyed on
// `integrator`.  The factory builds the HIP integrators whatever `use_CUDA`
// says: the HIP drivers have no CPU integrators.  average_mass / average_free_energy
// are the library's FP64 device reductions; a NaN is fatal as in the reference.
#pragma once

#include <cstdlib>
#include <fstream>
#include <memory>
#include <string>

#include "integrators.h"

namespace ch {

inline void trim(std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    s = (a == std::string::npos) ? std::string() : s.substr(a, b - a + 1);
}

inline std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    size_t i = 0;
    while(i < s.size()) {
        while(i < s.size() && (s[i] == ' ' || s[i] == '\t')) i++;
        size_t j = i;
        while(j < s.size() && s[j] != ' ' && s[j] != '\t') j++;
        if(j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    if(out.empty()) out.push_back("");
    return out;
}

template <int dims>
class CahnHilliard : public Object {
public:
    int N = 0, N_minus_one = 0, bits = 0, grid_size = 0;
    double dt = 0.0, dx = 0.0, V_bin = 0.0;
    std::vector<double> k_laplacian;
    FreeEnergyModel *model = nullptr;
    std::unique_ptr<HIPIntegrator<dims>> integrator;
    std::unique_ptr<IMobility<dims>> mobility;

    CahnHilliard(SimulationState<dims> &sim_state, FreeEnergyModel *m, const chtoml::Table &config)
        : model(m), _sim_state(sim_state) {
        N = _config_value<int>(config, "N");
        k_laplacian = _config_array_values<double>(config, "k", model->N_species());
        dt = _config_value<double>(config, "dt");
        dx = _config_optional_value<double>(config, "dx", 1.0);
        _internal_to_user = _config_optional_value<double>(config, "distance_scaling_factor", 1.0);
        _user_to_internal = 1.0 / _internal_to_user;
        std::string user_integrator = _config_optional_value<std::string>(config, "integrator", "euler");
        sim_state.use_CUDA = _config_optional_value<bool>(config, "use_CUDA", false);
        info("Running a simulation with N = {}, dt = {}, dx = {}, scaling factor = {}", N, dt, dx, _internal_to_user);

        double log2N = std::log2((double)N);
        if(N <= 0 || std::ceil(log2N) != std::floor(log2N)) critical("N should be a power of 2");
        N_minus_one = N - 1;
        bits = (int)log2N;
        grid_size = N;
        for(int i = 1; i < dims; i++) grid_size *= N;

        _sim_state.rho = MultiField<double>(grid_size, model->N_species());
        _sim_state.mobility = MultiField<double>(grid_size, model->N_species());

        if(!chtoml::find(config, "initial_density") && !chtoml::find(config, "load_from"))
            critical("Either 'initial_density' or 'load_from' should be specified");
        if(chtoml::find(config, "load_from")) _load(config);
        else _init_random(config);

        dx *= _user_to_internal;
        for(auto &k : k_laplacian) k *= std::pow(_user_to_internal, 5);
        for(int idx = 0; idx < _sim_state.rho.bins(); idx++)
            for(int s = 0; s < model->N_species(); s++) _sim_state.rho(idx, s) /= CUB(_user_to_internal);
        V_bin = CUB(dx);

        mobility.reset(build_mobility<dims>(config, _sim_state));
        if(!sim_state.use_CUDA) info("use_CUDA = false: the HIP drivers always integrate on the GPU");
        if(user_integrator == "euler") integrator = std::make_unique<EulerHIP<dims>>(_sim_state, m, config);
        else if(user_integrator == "euler_mobility")
            integrator = std::make_unique<EulerMobilityHIP<dims>>(_sim_state, m, config);
        else if(user_integrator == "pseudospectral")
            integrator = std::make_unique<PseudospectralHIP<dims>>(_sim_state, m, config);
        else if(user_integrator == "pseudospectral_mobility")
            integrator = std::make_unique<PseudospectralMobilityHIP<dims>>(_sim_state, m, config);
        else if(user_integrator == "bailo")
            critical("Unsupported CUDA integrator {}", user_integrator);
        else critical("Unsupported integrator {}", user_integrator);
        integrator->validate();
        integrator->setup(*mobility);
        _sim_state.integrator = integrator.get();
    }

    int cell_idx(const int coords[dims]) const {
        int idx = 0, mul = 1;
        for(int d = 0; d < dims; d++) {
            idx += coords[d] * mul;
            mul <<= bits;
        }
        return idx;
    }

    void evolve() { evolve(1); }
    // n x { update_mobility(); integrator->evolve(); } (CahnHilliard.cpp:279-282)
    void evolve(long long n) {
        mobility->update_mobility();
        integrator->evolve(n);
    }

    double average_mass() {
        double m, f;
        if(!integrator->observables(m, f)) critical("Encountered a nan while computing the total mass");
        return m;
    }
    double average_free_energy() {
        double m, f;
        if(!integrator->observables(m, f)) critical("Encountered a nan while computing the total free energy");
        return f;
    }
    // average_free_energy and average_mass of the current state, read later
    // with fetch_observables (first in, first out)
    void queue_observables() { integrator->queue_observables(); }
    // evolve(n), then queue_observables() (one captured graph for a recurring n)
    void evolve_observe(long long n) {
        mobility->update_mobility();
        integrator->evolve_observe(n);
    }
    void fetch_observables(double &F, double &m) {
        if(!integrator->fetch_observables(m, F)) critical("Encountered a nan while computing the total free energy");
    }
    // src/CahnHilliard.cpp:322-361 (host side, output only)
    MultiField<double> pressure() {
        integrator->sync();
        MultiField<double> out(grid_size, model->N_species());
        for(int idx = 0; idx < grid_size; idx++) {
            double r = _sim_state.rho(idx, 0);
            double v = model->pressure(r);
            int c[3] = {0, 0, 0}, cp[3];
            for(int d = 0, t = idx; d < dims; d++, t >>= bits) c[d] = t & N_minus_one;
            for(int d = 0; d < dims; d++) {
                for(int e = 0; e < dims; e++) cp[e] = c[e];
                cp[d] = (c[d] + 1) & N_minus_one;
                double g = (_sim_state.rho(cell_idx(cp), 0) - r) / dx;
                v -= 0.5 * k_laplacian[0] * g * g;
            }
            out(idx, 0) = v;
        }
        return out;
    }
    double average_pressure() {
        auto p = pressure();
        double s = 0;
        for(int i = 0; i < grid_size; i++) s += p(i, 0);
        return s / grid_size;
    }
    std::string_view name() const override { return "Simulation manager"; }

private:
    SimulationState<dims> &_sim_state;
    double _user_to_internal = 1.0, _internal_to_user = 1.0;

    // src/CahnHilliard.cpp:133-151
    void _init_random(const chtoml::Table &config) {
        std::vector<double> densities = _config_array_values<double>(config, "initial_density", model->N_species());
        double A = _config_optional_value<double>(config, "initial_A", 1e-2);
        int N_peaks = _config_optional_value<int>(config, "initial_N_peaks", 0);
        double initial_k = 2 * M_PI * N_peaks / (double)N;
        for(int bin = 0; bin < grid_size; bin++) {
            double modulation = A * std::cos(initial_k * bin);
            for(int i = 0; i < model->N_species(); i++) {
                double r = (N_peaks == 0) ? (drand48() - 0.5) : 1.0 + 0.02 * (drand48() - 0.5);
                double avg = densities[i];
                _sim_state.rho(bin, i) = (avg != 0.0) ? avg * (1.0 + 2.0 * modulation * r) : 2.0 * modulation * r;
            }
        }
    }

    // src/CahnHilliard.cpp:72-132
    void _load(const chtoml::Table &config) {
        std::string filename = _config_value<std::string>(config, "load_from");
        std::ifstream in(filename.c_str());
        if(!in.good()) critical("The initial conditions file '{}' is not readable", filename);
        info("Using filename '{}' for the initial conditions", filename);
        std::string line;
        int lines = 0;
        // N^(dims-1) data lines per species (y-major, then z); the reference's
        // loop stops after N lines, which in 3D loads only the z = 0 plane
        const int want = dims == 3 ? N * N : N;
        for(int s = 0; s < model->N_species(); s++) {
            int ok = 0;
            while(ok < want && in.good()) {
                std::getline(in, line);
                trim(line);
                auto spl = split_ws(line);
                if(line.size() > 0 && line[0] != '#') {
                    if(dims == 1) {
                        if(spl.size() != 1) critical("Line n. {} in the initial configuration file contains {} fields, should be 1", lines, spl.size());
                        _sim_state.rho(ok, s) = std::stod(line);
                    }
                    else {
                        if((int)spl.size() != N)
                            critical("Line n. {} in the initial configuration file contains only {} fields, should be {}", lines, spl.size(), N);
                        int c[3] = {0, dims == 2 ? ok : ok % N, ok / N};
                        for(c[0] = 0; c[0] < N; c[0]++) _sim_state.rho(cell_idx(c), s) = std::stod(spl[c[0]]);
                    }
                    ok++;
                }
                lines++;
            }
            if(ok != want)
                critical("The initial configuration file contains only {} valid lines for species {}, should be {}", ok, s, N);
        }
        info("Initial configuration parsed: found {} lines", lines);
    }
};

}  // namespace ch
