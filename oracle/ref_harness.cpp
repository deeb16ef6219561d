// ref_harness.cpp -- drives the reference's own CPU integrators (compiled from
// the unmodified sources under /root/reference/src) to write FP64 golden vectors.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by oracle/Makefile, run in
// the development container by tests/golden/make_golden.py.  Never shipped.
//
// Usage: ref_<D>D config.toml in.bin steps out.bin
//   config.toml  reference TOML keys (integrator, free_energy, N, dt, dx, k,
//                mobility.*, landau.*, wertheim.*)
//   in.bin       N^D raw little-endian doubles (the initial field)
//   out.bin      N^D raw doubles after `steps` calls of CahnHilliard::evolve's
//                body: mobility->update_mobility(); integrator->evolve();
//                (src/CahnHilliard.cpp:279-282)
#include "integrators/EulerCPU.h"
#include "integrators/EulerMobilityCPU.h"
#include "mobility/IMobility.h"
#include "models/Landau.h"
#include "models/SimpleWertheim.h"

#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>

// mobility.type = "linear" (M rho / rho_ref) has no class in the reference: the
// harness supplies it through the reference's own IMobility interface
// (src/mobility/IMobility.h), which EulerMobilityCPU reads through
// SimulationState::mobility like every other rule.
template <int dims>
class HarnessLinearMobility : public ch::IMobility<dims> {
public:
    HarnessLinearMobility(ch::SimulationState<dims> &state, double M, double rho_ref)
        : ch::IMobility<dims>(state), _M(M), _rho_ref(rho_ref) {}
    void update_mobility() override {
        for(unsigned int idx = 0; idx < this->_sim_state.rho.bins(); idx++)
            for(int s = 0; s < this->_sim_state.rho.species(); s++)
                this->_sim_state.mobility(idx, s) = _M * this->_sim_state.rho(idx, s) / _rho_ref;
    }

private:
    double _M, _rho_ref;
};

int main(int argc, char **argv) {
    if(argc < 5) {
        std::fprintf(stderr, "usage: %s config.toml in.bin steps out.bin\n", argv[0]);
        return 2;
    }
    toml::parse_result res = toml::parse_file(argv[1]);
    if(!res) {
        std::fprintf(stderr, "toml parse failed\n");
        return 2;
    }
    toml::table config = std::move(res).table();
    long long steps = std::atoll(argv[3]);

    ch::SimulationState<DIM> state;
    state.internal_to_user = 1.0;
    state.user_to_internal = 1.0;
    state.use_CUDA = false;

    std::string model_name = config["free_energy"].value_or(std::string("landau"));
    if(model_name == "landau") state.model = std::make_unique<ch::Landau>(config);
    else if(model_name == "simple_wertheim") state.model = std::make_unique<ch::SimpleWertheim>(config);
    else { std::fprintf(stderr, "unsupported model\n"); return 2; }

    int N = config["N"].value_or(0);
    int bins = 1;
    for(int d = 0; d < DIM; d++) bins *= N;
    state.rho = ch::MultiField<double>(bins, 1);
    state.mobility = ch::MultiField<double>(bins, 1);

    std::ifstream in(argv[2], std::ios::binary);
    in.read(reinterpret_cast<char *>(state.rho.data()), sizeof(double) * bins);
    if(!in) { std::fprintf(stderr, "short input\n"); return 2; }

    std::unique_ptr<ch::IMobility<DIM>> mobility;
    if(config["mobility"]["type"].value_or(std::string("constant")) == "linear")
        mobility = std::make_unique<HarnessLinearMobility<DIM>>(state, config["mobility"]["M"].value_or(1.0),
                                                                config["mobility"]["rho_ref"].value_or(1.0));
    else
        mobility.reset(ch::build_mobility(config, state));
    std::string integ = config["integrator"].value_or(std::string("euler"));
    std::unique_ptr<ch::Integrator<DIM>> integrator;
    if(integ == "euler") integrator = std::make_unique<ch::EulerCPU<DIM>>(state, state.model.get(), config);
    else if(integ == "euler_mobility") integrator = std::make_unique<ch::EulerMobilityCPU<DIM>>(state, state.model.get(), config);
    else { std::fprintf(stderr, "unsupported integrator\n"); return 2; }
    integrator->validate();
    state.integrator = integrator.get();

    for(long long t = 0; t < steps; t++) {
        mobility->update_mobility();
        integrator->evolve();
    }

    std::ofstream out(argv[4], std::ios::binary);
    out.write(reinterpret_cast<const char *>(state.rho.data()), sizeof(double) * bins);
    return out ? 0 : 1;
}
