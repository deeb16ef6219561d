// main.cpp -- ch_{1,2,3}D: the drop-in driver (Manager) of the HIP build.
//
// Same input file, run loop and outputs as the reference driver
// (src/main.cpp:20-200): srand48(seed); model factory on `free_energy`;
// init_* snapshots; for t in [t0, t0 + steps): last_* every print_last_every,
// trajectory and pressure frames, the energy line "t*dt F mass t" every print_every
// (printed *before* the step, as the reference does), then evolve; last_*
// at the end.  Steps between two output events are queued as one batch
// (CahnHilliard::evolve(n)), which the library replays as captured graphs.
#include <cstdio>
#include <ctime>
#include <deque>
#include <fstream>
#include <iostream>

#include "CahnHilliard.h"
#include "Printer.h"

#ifndef DIM
#define DIM 2
#endif

namespace ch {

class Manager final : public Object {
public:
    explicit Manager(const chtoml::Table &config) {
        _steps = _config_value<long long>(config, "steps");
        if(_steps < 0) critical("steps should be a number larger than 0");
        _state.internal_to_user = _config_optional_value<double>(config, "distance_scaling_factor", 1.0);
        _state.user_to_internal = 1.0 / _state.internal_to_user;
        _print_mass_every = _config_optional_value<long long>(config, "print_every", 0);
        _print_pressure = _config_optional_value<bool>(config, "output.print_pressure", false);
        srand48(_config_optional_value<long long>(config, "seed", (long long)std::time(NULL)));
        std::string outp = _config_optional_value<std::string>(config, "output_path", ".");
        _output_path = std::filesystem::path(outp);
        if(!std::filesystem::is_directory(_output_path))
            critical("Output path '{}' does not exist or is not a directory", outp);

        std::string model_name = _config_value<std::string>(config, "free_energy");
        if(model_name == "landau") _state.model = std::make_unique<Landau>(config);
        else if(model_name == "simple_wertheim") _state.model = std::make_unique<SimpleWertheim>(config);
        else if(model_name == "saleh") _state.model = std::make_unique<SalehWertheim>(config);
        else if(model_name == "generic_wertheim") _state.model = std::make_unique<GenericWertheim>(config);
        else if(model_name == "ricci") _state.model = std::make_unique<RicciWertheim>(config);
        else critical("Unsupported free energy model '{}'", model_name);

        _system = std::make_unique<CahnHilliard<DIM>>(_state, _state.model.get(), config);
        _printer = std::make_unique<Printer<DIM>>(_state, config);

        if(chtoml::find(config, "load_from")) {
            _openmode = std::ios_base::app;
            std::ifstream in(_config_value<std::string>(config, "load_from"));
            std::string line;
            std::getline(in, line);
            trim(line);
            auto spl = split_ws(line);
            if(spl.size() > 3 && spl[0] == "#" && spl[1] == "step" && spl[2] == "=") {
                _initial_t = std::stoll(spl[3]);
                info("Initial time step (as parsed from the 'load_from' file): {}", _initial_t);
            }
            else warning("The header line of the 'load_from' file does not contain information about the time step");
        }
    }

    void run() {
        _print_current_state("init_", 0);
        const double rho_factor = _state.density_to_user(1.0);
        std::ofstream mass_output((_output_path / "energy.dat").string(), _openmode);
        const long long t_end = _initial_t + _steps;
        long long t = _initial_t;
        // energy lines: the sums of step t are queued before the next steps are
        // issued and the line is written once those steps are on their way, so
        // the device does not wait for the host at every print; the lines and
        // their order are the reference's
        std::deque<long long> pending;
        auto energy_line = [&](long long ts, double F, double m) {
            std::string out = fmt_general(ts * _system->dt, 5) + " " + fmt_general(F, 8) + " " + fmt_general(m, 5) +
                              " " + std::to_string(ts);
            if(_print_pressure) out += " " + fmt_general(_system->average_pressure(), 5);
            mass_output << out << std::endl;
            std::cout << out << std::endl;
        };
        auto write_pending = [&](size_t keep) {
            while(pending.size() > keep) {
                double F, m;
                _system->fetch_observables(F, m);
                energy_line(pending.front(), F, m);
                pending.pop_front();
            }
        };
        bool queued = false;  // the observables of t were queued with the steps that led to t
        while(t < t_end) {
            if(_printer->should_print_last(t)) _print_current_state("last_", t);
            if(_printer->should_print_traj(t)) {
                _state.integrator->sync();
                for(int i = 0; i < _state.model->N_species(); i++)
                    _printer->add_to_trajectory("traj", "density", _state.rho, rho_factor, i, t);
                _printer->trajectory_frame_done();
            }
            // pressure frames, one file per species (src/main.cpp:114-119)
            if(_printer->should_print_pressure(t)) {
                auto p = _system->pressure();
                for(int i = 0; i < _state.model->N_species(); i++)
                    _printer->add_to_trajectory("pressure", "pressure", p, rho_factor, i, t);
                _printer->pressure_frame_done();
            }
            if(_print_mass_every > 0 && t % _print_mass_every == 0) {
                if(_print_pressure) {
                    // the pressure needs the host copy of this state: no queueing
                    write_pending(0);
                    double F = _system->average_free_energy(), m = _system->average_mass();
                    energy_line(t, F, m);
                }
                else if(!queued) {
                    write_pending(CH_OBS_QUEUE - 1);
                    _system->queue_observables();
                    pending.push_back(t);
                }
            }
            // steps until the next output event
            long long next = _printer->next_event(t, t_end);
            if(_print_mass_every > 0) next = std::min(next, (t / _print_mass_every + 1) * _print_mass_every);
            next = std::max(next, t + 1);
            // an energy line due at `next`: its reductions go with the steps
            queued = _print_mass_every > 0 && !_print_pressure && next < t_end && next % _print_mass_every == 0;
            if(queued) {
                write_pending(CH_OBS_QUEUE - 1);
                _system->evolve_observe(next - t);
                pending.push_back(next);
            }
            else {
                _system->evolve(next - t);
            }
            t = next;
            write_pending(1);
        }
        write_pending(0);
        mass_output.close();
        _print_current_state("last_", _steps);
    }
    std::string_view name() const override { return "Manager"; }

private:
    void _print_current_state(const std::string &prefix, long long t) {
        _state.integrator->sync();
        _printer->print_current_state(prefix, t);
        if(_print_pressure) {
            auto p = _system->pressure();
            _printer->write_native((_output_path / (prefix + "pressure.dat")).string(), "pressure", p,
                                   _state.density_to_user(1.0), -1, t);
        }
    }

    SimulationState<DIM> _state;
    bool _print_pressure = false;
    long long _initial_t = 0, _steps = 0, _print_mass_every = 0;
    std::ios_base::openmode _openmode = std::ios_base::out;
    std::unique_ptr<CahnHilliard<DIM>> _system;
    std::unique_ptr<Printer<DIM>> _printer;
    std::filesystem::path _output_path = ".";
};

}  // namespace ch

int main(int argc, char *argv[]) {
    if(argc < 2) {
        std::cerr << "Usage is " << argv[0] << " configuration_file" << std::endl;
        return 0;
    }
    chtoml::Table config;
    std::string err;
    if(!chtoml::parse_file(argv[1], config, err)) {
        std::cerr << "Parsing failed with error '" << err << "'" << std::endl;
        return 1;
    }
    ch::Manager manager(config);
    manager.run();
    return 0;
}
