// models.h -- FreeEnergyModel plugins of the drop-in (host side).
//
// Same keys, unit conversions and host formulas as the reference's
// FreeEnergyModel / Landau / SimpleWertheim (src/models/FreeEnergyModel.h:16-59,
// Landau.cpp:16-71, SimpleWertheim.cpp:16-111, src/utils/Delta.cpp:15-42).
// The device evaluation happens inside the fused HIP kernels: describe()
// writes the model's parameters into the ch_desc the integrators hand to the
// C ABI, and der_bulk_free_energy on device buffers calls
// ch_der_bulk_free_energy (the reference's FreeEnergyModel.h:29 entry).
#pragma once

#include <cmath>
#include <string>

#include "../../include/ch_hip.h"
#include "Object.h"
#include "SimulationState.h"

namespace ch {

class FreeEnergyModel : public Object {
public:
    explicit FreeEnergyModel(const chtoml::Table &config) {
        _use_CUDA = _config_optional_value<bool>(config, "use_CUDA", false);
        _user_to_internal = 1.0 / _config_optional_value<double>(config, "distance_scaling_factor", 1.0);
        _density_conversion_factor = CUB(_user_to_internal);
    }
    virtual int N_species() = 0;
    virtual double bulk_free_energy(double rho) = 0;
    virtual double der_bulk_free_energy(double rho) = 0;
    virtual double pressure(double rho) { return rho * der_bulk_free_energy(rho) - bulk_free_energy(rho); }
    // model part of the C-ABI descriptor
    virtual void describe(ch_desc &d) const = 0;
    // FreeEnergyModel::der_bulk_free_energy(field_type *, float *, int) on device buffers
    ch_status der_bulk_free_energy(ch_ctx *ctx, const ch_desc &d, const void *rho_dev, void *out_dev, int64_t n) const {
        return ch_der_bulk_free_energy(ctx, &d, rho_dev, out_dev, n);
    }

protected:
    double _user_to_internal = 1.0, _density_conversion_factor = 1.0;
    bool _use_CUDA = false;
};

class Landau final : public FreeEnergyModel {
public:
    explicit Landau(const chtoml::Table &config) : FreeEnergyModel(config) {
        if(chtoml::find(config, "landau.epsilon")) {
            _epsilon = _config_value<double>(config, "landau.epsilon");
        }
        else if(chtoml::find(config, "landau.T")) {
            _epsilon = 1 - _config_value<double>(config, "landau.T");
        }
        else {
            critical("The Landau free energy requires either 'epsilon' or 'T' to be set");
        }
        if(_user_to_internal != 1.0)
            critical("The Landau free energy does not support distance_scaling_factor values different from 1.0");
        info("epsilon = {}", _epsilon);
    }
    int N_species() override { return 1; }
    double bulk_free_energy(double op) override { return -0.5 * _epsilon * SQR(op) + 0.25 * SQR(SQR(op)); }
    double der_bulk_free_energy(double op) override { return -_epsilon * op + op * op * op; }
    void describe(ch_desc &d) const override {
        d.model = CH_LANDAU;
        d.epsilon = _epsilon;
    }
    std::string_view name() const override { return "Landau free energy model"; }

private:
    double _epsilon = 0.0;
};

class SimpleWertheim final : public FreeEnergyModel {
public:
    explicit SimpleWertheim(const chtoml::Table &config) : FreeEnergyModel(config) {
        _valence = _config_value<int>(config, "wertheim.valence");
        _B2 = _config_value<double>(config, "wertheim.B2");
        _regularisation_delta = _config_optional_value<double>(config, "wertheim.regularisation_delta", 0.0);
        _delta = parse_delta(config, "wertheim.delta");
        info("valence = {}, delta = {}", _valence, _delta);
        _B2 *= CUB(_user_to_internal);
        _delta *= CUB(_user_to_internal);
        _regularisation_delta /= CUB(_user_to_internal);
        _log_delta = std::log(_regularisation_delta);
        _two_valence_delta = 2.0 * _valence * _delta;
    }
    int N_species() override { return 1; }
    double X(double rho) const { return (-1.0 + std::sqrt(1.0 + 2.0 * _two_valence_delta * rho)) / (_two_valence_delta * rho); }
    double bulk_free_energy(double rho) override {
        double f_ref = _B2 * SQR(rho);
        if(rho < _regularisation_delta) {
            double a = 1 / (2.0 * _regularisation_delta), b = _log_delta - 1.0;
            double c = _regularisation_delta * (_log_delta - 1.0) - a * SQR(_regularisation_delta) - b * _regularisation_delta;
            f_ref += a * SQR(rho) + b * rho + c;
        }
        else {
            f_ref += rho * (std::log(rho * _density_conversion_factor) - 1.0);
        }
        double f_bond = (rho > 0.) ? _valence * rho * (std::log(X(rho)) + 0.5 * (1. - X(rho))) : 0.0;
        return f_ref + f_bond;
    }
    double der_bulk_free_energy(double rho) override {
        double d = 2 * _B2 * rho;
        d += (rho < _regularisation_delta) ? _log_delta + (rho - _regularisation_delta) / _regularisation_delta : std::log(rho);
        return d + ((rho > 0.) ? _valence * std::log(X(rho)) : 0.0);
    }
    void describe(ch_desc &d) const override {
        d.model = CH_SIMPLE_WERTHEIM;
        d.valence = _valence;
        d.B2 = _B2;
        d.delta = _delta;
        d.regularisation_delta = _regularisation_delta;
    }
    std::string_view name() const override { return "Simple Wertheim free energy model"; }

    // src/utils/Delta.cpp:15-42: a number, or a table {value} / {T, deltaH, deltaS, salt, sticky_size}
    double parse_delta(const chtoml::Table &config, const std::string &path) const {
        const chtoml::Node *n = _config_node(config, path, true);
        if(n->is_number()) return n->as_double();
        if(n->kind != chtoml::Node::TABLE) critical("A delta must be specified either as a single number or as a table of values");
        const chtoml::Table &t = *n->tab;
        double delta = _config_optional_value<double>(t, "value", 0.0);
        if(delta == 0.0) {
            double T = _config_value<double>(t, "T");
            double salt = _config_optional_value<double>(t, "salt", 1.0);
            int L_DNA = _config_optional_value<int>(t, "sticky_size", 6);
            double dH = _config_value<double>(t, "deltaH"), dS = _config_value<double>(t, "deltaS");
            double dS_salt = 0.368 * (L_DNA - 1.0) * std::log(salt);
            double dG = dH - T * (dS + dS_salt);
            const double k_B = 1.9872036;
            delta = 1.6606 * std::exp(-dG / (k_B * T));
        }
        return delta;
    }

private:
    double _B2 = 0, _delta = 0, _two_valence_delta = 0, _regularisation_delta = 0, _log_delta = 0;
    int _valence = 0;
};

}  // namespace ch
