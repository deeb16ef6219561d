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

#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

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
    // multi-species models: the ch_species_model of ch_sys_create_species
    virtual bool multi_species() const { return false; }
    virtual void describe_species(ch_species_model &) const {}
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
        d.density_conversion = _density_conversion_factor;
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

// src/utils/Delta.cpp:15-42 on a table node: {value} or {T, deltaH, deltaS, salt, sticky_size}
inline double delta_of_table(const Object &o, const chtoml::Table &t) {
    double delta = o._config_optional_value<double>(t, "value", 0.0);
    if(delta == 0.0) {
        double T = o._config_value<double>(t, "T");
        double salt = o._config_optional_value<double>(t, "salt", 1.0);
        int L_DNA = o._config_optional_value<int>(t, "sticky_size", 6);
        double dH = o._config_value<double>(t, "deltaH"), dS = o._config_value<double>(t, "deltaS");
        double dS_salt = 0.368 * (L_DNA - 1.0) * std::log(salt);
        double dG = dH - T * (dS + dS_salt);
        delta = 1.6606 * std::exp(-dG / (1.9872036 * T));
    }
    return delta;
}

// Multi-species models: the device evaluates f', X and the observables
// (species.hip); the host keeps the keys and unit scaling of the reference and
// describes the model to the C ABI.  Their scalar interface is not used.
class MultiSpeciesModel : public FreeEnergyModel {
public:
    using FreeEnergyModel::FreeEnergyModel;
    bool multi_species() const override { return true; }
    double bulk_free_energy(double) override {
        critical("{} has several species: its free energy is evaluated on the device", name());
        return 0.0;
    }
    double der_bulk_free_energy(double) override { return bulk_free_energy(0.0); }
    double pressure(double) override {
        critical("this model does not support the pressure coupling");  // FreeEnergyModel.h:43-46
        return 0.0;
    }
    double delta_at(const chtoml::Table &config, const std::string &path) const {
        const chtoml::Node *n = _config_node(config, path, true);
        if(n->is_number()) return n->as_double();
        if(n->kind != chtoml::Node::TABLE) critical("A delta must be specified either as a single number or as a table of values");
        return delta_of_table(*this, *n->tab);
    }
};

// src/models/SalehWertheim.cpp:18-37
class SalehWertheim final : public MultiSpeciesModel {
public:
    explicit SalehWertheim(const chtoml::Table &config) : MultiSpeciesModel(config) {
        _B2 = _config_value<double>(config, "saleh.B2");
        _B3 = _config_optional_value<double>(config, "saleh.B3", 0.);
        _delta_AA = delta_at(config, "saleh.delta_AA");
        _delta_BB = delta_at(config, "saleh.delta_BB");
        _valence = _config_array_values<int>(config, "saleh.valence", 3);
        info("valences = ({}, {}, {}), B2 = {}, delta_AA = {}, delta_BB = {}", _valence[0], _valence[1], _valence[2], _B2,
             _delta_AA, _delta_BB);
        _B2 *= CUB(_user_to_internal);
        _B3 *= SQR(CUB(_user_to_internal));
        _delta_AA *= CUB(_user_to_internal);
        _delta_BB *= CUB(_user_to_internal);
    }
    int N_species() override { return 3; }
    void describe(ch_desc &d) const override {
        d.model = CH_SALEH_WERTHEIM;
        d.density_conversion = _density_conversion_factor;
    }
    void describe_species(ch_species_model &m) const override {
        m.model = CH_SALEH_WERTHEIM;
        m.n_species = 3;
        for(int i = 0; i < 3; i++) m.saleh_valence[i] = _valence[i];
        m.saleh_B2 = _B2;
        m.saleh_B3 = _B3;
        m.saleh_delta_AA = _delta_AA;
        m.saleh_delta_BB = _delta_BB;
    }
    std::string_view name() const override { return "Saleh's system Wertheim free energy"; }

private:
    std::vector<int> _valence;
    double _B2 = 0, _B3 = 0, _delta_AA = 0, _delta_BB = 0;
};

// src/models/RicciWertheim.cpp:19-36
class RicciWertheim final : public MultiSpeciesModel {
public:
    explicit RicciWertheim(const chtoml::Table &config) : MultiSpeciesModel(config) {
        _B2 = _config_value<double>(config, "ricci.B2");
        _delta_00 = delta_at(config, "ricci.delta_00");
        _delta_12 = delta_at(config, "ricci.delta_12");
        info("B2 = {}, delta_00 = {}, delta_12 = {}", _B2, _delta_00, _delta_12);
        _B2 *= CUB(_user_to_internal);
        _delta_00 *= CUB(_user_to_internal);
        _delta_12 *= CUB(_user_to_internal);
    }
    int N_species() override { return 2; }
    void describe(ch_desc &d) const override {
        d.model = CH_RICCI_WERTHEIM;
        d.density_conversion = _density_conversion_factor;
    }
    void describe_species(ch_species_model &m) const override {
        m.model = CH_RICCI_WERTHEIM;
        m.n_species = 2;
        m.ricci_B2 = _B2;
        m.ricci_delta_00 = _delta_00;
        m.ricci_delta_12 = _delta_12;
    }
    std::string_view name() const override { return "Ricci's system Wertheim free energy"; }

private:
    double _B2 = 0, _delta_00 = 0, _delta_12 = 0;
};

// src/models/GenericWertheim.cpp:22-121: [[generic_wertheim.species]] patches,
// [[generic_wertheim.deltas]] "a-b" + Delta, [[generic_wertheim.B2s]] "a-b" + value
class GenericWertheim final : public MultiSpeciesModel {
public:
    explicit GenericWertheim(const chtoml::Table &config) : MultiSpeciesModel(config) {
        const chtoml::Node *sp = chtoml::find(config, "generic_wertheim.species");
        if(!sp || sp->kind != chtoml::Node::ARRAY) critical("Missing 'generic_wertheim.species' array");
        for(const auto &e : sp->arr) {
            if(e.kind != chtoml::Node::TABLE) critical("generic_wertheim.species must hold tables");
            _patches.push_back(_config_array_values<int>(*e.tab, "patches"));
        }
        if((int)_patches.size() > CH_MAX_SPECIES) critical("at most {} species are supported", CH_MAX_SPECIES);
        for(auto &pl : _patches)
            for(int q : pl) {
                if(q < 0 || q >= CH_MAX_PATCHES) critical("patch types must lie in [0, {})", CH_MAX_PATCHES);
                _n_patches = std::max(_n_patches, q + 1);
            }
        const int S = (int)_patches.size();
        _delta.assign(CH_MAX_PATCHES * CH_MAX_PATCHES, 0.0);
        const chtoml::Node *ds = chtoml::find(config, "generic_wertheim.deltas");
        if(!ds || ds->kind != chtoml::Node::ARRAY) critical("Missing 'generic_wertheim.deltas' array");
        for(const auto &e : ds->arr) {
            auto [a, b] = _pair(_config_value<std::string>(*e.tab, "interaction"), "delta", _n_patches);
            double d = delta_of_table(*this, *e.tab);
            info("Adding {}-{} interaction with delta = {}", a, b, d);
            _delta[a * CH_MAX_PATCHES + b] = _delta[b * CH_MAX_PATCHES + a] = d * CUB(_user_to_internal);
        }
        _B2.assign(CH_MAX_SPECIES * CH_MAX_SPECIES, 0.0);
        const chtoml::Node *bs = chtoml::find(config, "generic_wertheim.B2s");
        if(!bs || bs->kind != chtoml::Node::ARRAY) critical("Missing 'generic_wertheim.B2s' array");
        if((int)bs->arr.size() != S * (S + 1) / 2) {
            bool allow = _config_optional_value<bool>(config, "generic_wertheim.allow_unspecified_B2s", false);
            if(!allow)
                critical("The number of B2 coefficients specified in the input ({}) is different from the one expected ({}), "
                         "calculated as N * (N + 1) / 2. Set 'generic_wertheim.allow_unspecified_B2s = true' to run "
                         "the simulation regardless", bs->arr.size(), S * (S + 1) / 2);
            warning("The number of B2 coefficients differs from N * (N + 1) / 2; running regardless");
        }
        for(const auto &e : bs->arr) {
            auto [a, b] = _pair(_config_value<std::string>(*e.tab, "interaction"), "B2", S);
            double v = _config_value<double>(*e.tab, "value");
            info("Adding {}-{} virial coefficient, B2 = {}", a, b, v);
            _B2[a * CH_MAX_SPECIES + b] = _B2[b * CH_MAX_SPECIES + a] = v * CUB(_user_to_internal);
        }
    }
    int N_species() override { return (int)_patches.size(); }
    void describe(ch_desc &d) const override {
        d.model = CH_GENERIC_WERTHEIM;
        d.density_conversion = _density_conversion_factor;
    }
    void describe_species(ch_species_model &m) const override {
        m.model = CH_GENERIC_WERTHEIM;
        m.n_species = (int)_patches.size();
        m.n_patches = _n_patches;
        for(size_t s = 0; s < _patches.size(); s++)
            for(int q : _patches[s]) m.patch_count[s][q]++;
        for(int a = 0; a < CH_MAX_PATCHES; a++)
            for(int b = 0; b < CH_MAX_PATCHES; b++) m.delta[a][b] = _delta[a * CH_MAX_PATCHES + b];
        for(int a = 0; a < CH_MAX_SPECIES; a++)
            for(int b = 0; b < CH_MAX_SPECIES; b++) m.B2[a][b] = _B2[a * CH_MAX_SPECIES + b];
    }
    std::string_view name() const override { return "Generic Wertheim free energy"; }

private:
    std::vector<std::vector<int>> _patches;
    std::vector<double> _delta, _B2;
    int _n_patches = 0;

    // GenericWertheim.cpp:160-171: "a-b"
    std::pair<int, int> _pair(const std::string &text, const char *context, int limit) const {
        const size_t dash = text.find('-');
        if(dash == std::string::npos || text.find('-', dash + 1) != std::string::npos)
            critical("The following {} interaction specifier is malformed: {}", context, text);
        const int a = std::atoi(text.substr(0, dash).c_str()), b = std::atoi(text.substr(dash + 1).c_str());
        if(a < 0 || b < 0 || a >= limit || b >= limit) critical("{} interaction {} is out of range", context, text);
        return {a, b};
    }
};

}  // namespace ch
