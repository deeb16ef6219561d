// Printer.h -- native / VTK output of the drop-in (host side).
//
// Follows src/utils/Printer.cpp:13-262: trajectory cadence "linear"
// (print_trajectory_every, print_last_every defaulting to it) or "log"
// (output.log_n0, output.log_fact); with output.print_pressure, pressure
// frames every output.print_pressure_every steps or on the log schedule
// (output.print_pressure_strategy); native files start with
//   # step = S, t = T, size = NXxNYxNZ | obs
// followed by the values at the default ostream precision (6 significant
// digits), a newline every N values; VTK as STRUCTURED_POINTS ASCII.
// output.lossless = true (no reference counterpart) writes the native values
// with 17 significant digits instead, so a load_from restart continues from
// the exact FP64 state (with distance_scaling_factor = 1, where the unit
// factor is 1).
#pragma once

#include <filesystem>
#include <fstream>
#include <iomanip>
#include <set>
#include <string>

#include "Object.h"
#include "SimulationState.h"
#include "fmt_general.h"

namespace ch {

template <int dims>
class Printer : public Object {
public:
    Printer(SimulationState<dims> &state, const chtoml::Table &config) : _state(state) {
        N = _config_value<int>(config, "N");
        dt = _config_value<double>(config, "dt");
        dx = _config_optional_value<double>(config, "dx", 1.0);
        _strategy = _config_optional_value<std::string>(config, "print_trajectory_strategy", "linear");
        if(_strategy == "linear") {
            _traj_every = _config_optional_value<long long>(config, "print_trajectory_every", 0);
            _last_every = _config_optional_value<long long>(config, "print_last_every", _traj_every);
        }
        else if(_strategy == "log") {
            _log_n0 = _config_value<int>(config, "output.log_n0");
            _log_fact = _config_value<double>(config, "output.log_fact");
            _last_every = _config_value<long long>(config, "print_last_every");
        }
        else critical("Unsupported trajectory printing strategy '{}'", _strategy);

        // pressure frames (output.print_pressure): every print_pressure_every
        // steps, or on the log schedule round(log_n0 log_fact^n) (Printer.cpp:33-46)
        _print_pressure = _config_optional_value<bool>(config, "output.print_pressure", false);
        if(_print_pressure) {
            _pressure_strategy = _config_optional_value<std::string>(config, "output.print_pressure_strategy", "linear");
            if(_pressure_strategy == "linear") {
                _pressure_every = _config_value<long long>(config, "output.print_pressure_every");
            }
            else if(_pressure_strategy == "log") {
                _log_n0 = _config_value<int>(config, "output.log_n0");
                _log_fact = _config_value<double>(config, "output.log_fact");
            }
            else critical("Unsupported pressure printing strategy '{}'", _pressure_strategy);
        }

        std::string outp = _config_optional_value<std::string>(config, "output.path", ".");
        _out = std::filesystem::path(outp);
        if(!std::filesystem::exists(_out)) critical("Output path '{}' does not exist", outp);
        if(!std::filesystem::is_directory(_out)) critical("Output path '{}' is not a directory", outp);
        _print_vtk = _config_optional_value<bool>(config, "output.print_vtk", false);
        _lossless = _config_optional_value<bool>(config, "output.lossless", false);
        std::string trajp = _print_vtk ? _config_value<std::string>(config, "output.trajectory_path")
                                       : _config_optional_value<std::string>(config, "output.trajectory_path", outp);
        _traj = std::filesystem::path(trajp);
        if(!std::filesystem::is_directory(_traj)) critical("Trajectory output path '{}' does not exist", trajp);
        if(!_print_vtk && !chtoml::find(config, "load_from")) {
            for(int i = 0; i < _state.model->N_species(); i++)
                for(const char *prefix : {"traj", "pressure", "mu"}) {
                    std::ofstream o((_traj / (std::string(prefix) + "_" + std::to_string(i) + ".dat")).string(),
                                    std::ios_base::out);
                }
        }
    }

    bool should_print_last(long long t) const { return _last_every > 0 && t % _last_every == 0; }
    // Printer.cpp:157-170; the log schedule counts pressure frames on their own
    // (pressure_frame_done), so it is the trajectory's schedule
    bool should_print_pressure(long long t) const {
        if(!_print_pressure) return false;
        if(_pressure_strategy == "linear") return _pressure_every > 0 && t % _pressure_every == 0;
        return (long long)std::round(_log_n0 * std::pow(_log_fact, _pressure_printed)) == t;
    }
    void pressure_frame_done() { _pressure_printed++; }
    bool should_print_traj(long long t) const {
        if(_strategy == "linear") return _traj_every > 0 && t % _traj_every == 0;
        long long next_t = (long long)std::round(_log_n0 * std::pow(_log_fact, _traj_printed));
        return next_t == t;
    }
    // earliest t' in (t, t_end] at which should_print_traj/last fires (t_end when none)
    long long next_event(long long t, long long t_end) const {
        long long best = t_end;
        auto upd = [&](long long every) {
            if(every > 0) best = std::min(best, (t / every + 1) * every);
        };
        auto upd_log = [&](int printed) {
            long long nt = (long long)std::round(_log_n0 * std::pow(_log_fact, printed));
            if(nt > t) best = std::min(best, nt);
            else best = std::min(best, t + 1);
        };
        upd(_last_every);
        if(_strategy == "linear") upd(_traj_every);
        else upd_log(_traj_printed);
        if(_print_pressure) {
            if(_pressure_strategy == "linear") upd(_pressure_every);
            else upd_log(_pressure_printed);
        }
        return best;
    }

    // Printer.cpp:97-114: one file per species and the total density
    void print_current_state(const std::string &prefix, long long t) {
        double factor = _state.density_to_user(1.0);
        for(int i = 0; i < _state.model->N_species(); i++) {
            const std::string s = std::to_string(i);
            write_native((_out / (prefix + s + ".dat")).string(), "density", _state.rho, factor, i, t);
            write_vtk((_out / (prefix + s + ".vtk")).string(), "density", _state.rho, factor, i, t);
        }
        write_native((_out / (prefix + "density.dat")).string(), "density", _state.rho, factor, -1, t);
        write_vtk((_out / (prefix + "density.vtk")).string(), "density", _state.rho, factor, -1, t);
    }

    // Printer.cpp:116-138, one species; trajectory_frame_done() after the
    // species of a frame (advances the log cadence once per frame)
    void add_to_trajectory(const std::string &prefix, const std::string &obs, const MultiField<double> &field,
                           double factor, int species, long long t) {
        const std::string s = std::to_string(species);
        if(_print_vtk) {
            write_vtk((_traj / (prefix + "_" + s + "_" + std::to_string(t) + ".vtk")).string(), obs, field, factor,
                      species, t);
        }
        else {
            std::ofstream o((_traj / (prefix + "_" + s + ".dat")).string(), std::ios_base::app);
            write_native(o, obs, field, factor, species, t);
        }
    }
    void trajectory_frame_done() { _traj_printed++; }

    void write_native(std::ostream &o, const std::string &obs, const MultiField<double> &field, double factor,
                      int species, long long t) {
        const int nx = N, ny = dims >= 2 ? N : 1, nz = dims >= 3 ? N : 1;
        const int grid = nx * ny * nz;
        o << "# step = " << t << ", t = " << fmt_general(t * dt, 5) << ", size = " << nx << "x" << ny << "x" << nz
          << " | " << obs << std::endl;
        if(_lossless) o << std::setprecision(17);
        for(int idx = 0; idx < grid; idx++) {
            if(idx > 0 && idx % ny == 0) o << std::endl;
            double v = species == -1 ? field.field_sum(idx) : field(idx, species);
            o << v * factor << " ";
        }
        o << std::endl;
    }
    void write_native(const std::string &file, const std::string &obs, const MultiField<double> &field, double factor,
                      int species, long long t) {
        std::ofstream o(file);
        write_native(o, obs, field, factor, species, t);
    }
    void write_vtk(const std::string &file, const std::string &obs, const MultiField<double> &field, double factor,
                   int species, long long t) {
        if(!_print_vtk) return;
        if(dims == 1) critical("VTK output is not supported for 1D simulations");
        std::ofstream o(file);
        const int nx = N, ny = dims >= 2 ? N : 1, nz = dims >= 3 ? N : 1;
        const double s = _state.length_to_user(dx);
        o << "# vtk DataFile Version 3.0\n";
        o << "step = " << t << ", t = " << fmt_general(t * dt, 5) << ", size = " << nx << "x" << ny << "x" << nz << " | "
          << obs << std::endl;
        o << "ASCII\nDATASET STRUCTURED_POINTS\n";
        o << "DIMENSIONS " << nx << " " << ny << " " << nz << "\nORIGIN 0 0 0\n";
        o << "SPACING " << s << " " << s << " " << s << "\n";
        o << "POINT_DATA " << (size_t)nx * ny * nz << "\nSCALARS species_" << species << " double 1\nLOOKUP_TABLE default\n";
        o << std::scientific;
        for(int idx = 0; idx < nx * ny * nz; idx++) {
            double v = species == -1 ? field.field_sum(idx) : field(idx, species);
            o << v * factor << "\n";
        }
    }
    std::string_view name() const override { return "Printer"; }

    int N = 0;
    double dt = 0, dx = 1;

private:
    SimulationState<dims> &_state;
    std::string _strategy, _pressure_strategy;
    long long _traj_every = 0, _last_every = 0, _pressure_every = 0;
    int _log_n0 = 0, _traj_printed = 0, _pressure_printed = 0;
    bool _print_pressure = false;
    double _log_fact = 1.0;
    bool _print_vtk = false;
    bool _lossless = false;
    std::filesystem::path _out, _traj;
};

}  // namespace ch
