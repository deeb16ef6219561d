// integrators.h -- Integrator<dims> plugins backed by the HIP library.
//
// Integrator<dims> reads the same keys with the same unit scaling as
// src/integrators/Integrator.cpp:6-46 (mobility.type, k per species, dt, dx,
// N) and keeps validate() / evolve() / sync().  HIPIntegrator<dims> plays the
// role of CUDAIntegrator (src/CUDA/integrators/CUDAIntegrator.cu:13-70): it
// owns the device state through a ch_ctx / ch_sys pair, uploads rho at
// construction, and sync() copies rho (and the mobility field) back lazily.
// EulerHIP, EulerMobilityHIP and PseudospectralHIP replace EulerCUDA,
// EulerMobilityCUDA and PseudospectralCUDA.  New optional keys:
//   field_precision = "double" (default) | "float"   (the reference's
//                     compile-time CUDA_FIELD_FLOAT, src/defs_CUDA.h:74-86)
//   device          = GPU index (default 0)
#pragma once

#include <cmath>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/ch_hip.h"
#include "mobility.h"
#include "models.h"

namespace ch {

template <int dims>
class Integrator : public Object {
public:
    Integrator(SimulationState<dims> &sim_state, FreeEnergyModel *model, const chtoml::Table &config)
        : _sim_state(sim_state), _rho(sim_state.rho), _model(model) {
        _mobility_type = _config_optional_value<std::string>(config, "mobility.type", "constant");
        _k_laplacian = _config_array_values<double>(config, "k", model->N_species());
        _dt = _config_value<double>(config, "dt");
        _dx = _config_optional_value<double>(config, "dx", 1.0);
        _dx *= sim_state.user_to_internal;
        for(auto &k : _k_laplacian) k *= std::pow(sim_state.user_to_internal, 5);
        _N_per_dim = _config_value<int>(config, "N");
        _N_bins = _N_per_dim;
        for(int i = 1; i < dims; i++) _N_bins *= _N_per_dim;
        _N_species = model->N_species();
        info("Integrator initialized with dt = {}, dx = {}, k = {}", _dt, _dx, _k_laplacian[0]);
    }
    virtual void validate() {
        if(!_supports_nonconstant_mobility() && _mobility_type != "constant")
            this->critical("The selected integrator only supports constant mobility");
    }
    virtual void evolve() = 0;
    // n consecutive evolve() calls (one queued batch on the device)
    virtual void evolve(long long n) {
        for(long long i = 0; i < n; i++) evolve();
    }
    virtual void sync() {}
    std::string_view name() const override { return "Integrator"; }

protected:
    SimulationState<dims> &_sim_state;
    MultiField<double> &_rho;
    FreeEnergyModel *_model;
    int _N_per_dim = 0, _N_bins = 0, _N_species = 0;
    double _dt = 0.0, _dx = 0.0;
    std::vector<double> _k_laplacian;
    std::string _mobility_type;
    virtual bool _supports_nonconstant_mobility() const { return false; }
};

template <int dims>
class HIPIntegrator : public Integrator<dims> {
public:
    HIPIntegrator(SimulationState<dims> &s, FreeEnergyModel *model, const chtoml::Table &config, int integrator_id)
        : Integrator<dims>(s, model, config) {
        if(model->N_species() != 1 && !model->multi_species())
            this->critical("The HIP integrators support single-species models only");
        _desc = ch_desc{};
        _desc.dims = dims;
        _desc.N = this->_N_per_dim;
        _desc.integrator = integrator_id;
        std::string prec = this->template _config_optional_value<std::string>(config, "field_precision", "double");
        if(prec != "double" && prec != "float") this->critical("field_precision must be 'double' or 'float'");
        _desc.precision = prec == "double" ? CH_FP64 : CH_FP32;
        _desc.dt = this->_dt;
        _desc.dx = this->_dx;
        _desc.k = this->_k_laplacian[0];
        model->describe(_desc);
        if(model->multi_species()) {
            _multi = true;
            _species = ch_species_model{};
            model->describe_species(_species);
            for(int s = 0; s < model->N_species(); s++) _species.k[s] = this->_k_laplacian[s];
        }
        _device = this->template _config_optional_value<int>(config, "device", 0);
    }
    ~HIPIntegrator() override {
        if(_sys) ch_sys_destroy(_sys);
        if(_ctx) ch_ctx_destroy(_ctx);
    }
    // creates the device system; called once the mobility is attached
    void setup(const IMobility<dims> &mobility) {
        mobility.describe(_desc);
        _check(ch_ctx_create(_device, &_ctx), "ch_ctx_create");
        if(_multi) _check(ch_sys_create_species(_ctx, &_desc, &_species, &_sys), "ch_sys_create_species");
        else _check(ch_sys_create(_ctx, &_desc, &_sys), "ch_sys_create");
        _check(ch_sys_upload(_sys, this->_rho.data()), "ch_sys_upload");
        _output_ready = true;
    }
    void evolve() override { evolve(1); }
    void evolve(long long n) override {
        _check(ch_sys_evolve(_sys, n), "ch_sys_evolve");
        _output_ready = false;
    }
    void sync() override {
        if(_output_ready) return;
        _check(ch_sys_download(_sys, this->_rho.data()), "ch_sys_download");
        _check(ch_sys_download_mobility(_sys, this->_sim_state.mobility.data()), "ch_sys_download_mobility");
        _output_ready = true;
    }
    // FP64 device reductions (src/CahnHilliard.cpp:285-319); false on NaN
    bool observables(double &mass, double &fe) {
        ch_status st = ch_sys_observables(_sys, &mass, &fe);
        if(st == CH_ENAN) return false;
        _check(st, "ch_sys_observables");
        return true;
    }
    // the same pair, queued behind the issued steps and read later (FIFO)
    void queue_observables() { _check(ch_sys_observables_queue(_sys), "ch_sys_observables_queue"); }
    // evolve(n), then queue_observables(): one captured graph for a recurring n
    void evolve_observe(long long n) {
        _check(ch_sys_evolve_observe(_sys, n), "ch_sys_evolve_observe");
        _output_ready = false;
    }
    bool fetch_observables(double &mass, double &fe) {
        ch_status st = ch_sys_observables_fetch(_sys, &mass, &fe);
        if(st == CH_ENAN) return false;
        _check(st, "ch_sys_observables_fetch");
        return true;
    }
    const ch_desc &desc() const { return _desc; }
    ch_ctx *context() { return _ctx; }
    ch_sys *system() { return _sys; }

protected:
    ch_desc _desc{};
    ch_species_model _species{};
    bool _multi = false;
    int _device = 0;
    ch_ctx *_ctx = nullptr;
    ch_sys *_sys = nullptr;
    bool _output_ready = false;

    void _check(ch_status st, const char *what) {
        if(st != CH_OK)
            this->critical("{} failed: {} ({})", what, ch_status_string(st), _ctx ? ch_ctx_last_error(_ctx) : "no context");
    }
};

template <int dims>
class EulerHIP : public HIPIntegrator<dims> {
public:
    EulerHIP(SimulationState<dims> &s, FreeEnergyModel *m, const chtoml::Table &c) : HIPIntegrator<dims>(s, m, c, CH_EULER) {}
    std::string_view name() const override { return "EulerHIP"; }
};

template <int dims>
class EulerMobilityHIP : public HIPIntegrator<dims> {
public:
    EulerMobilityHIP(SimulationState<dims> &s, FreeEnergyModel *m, const chtoml::Table &c)
        : HIPIntegrator<dims>(s, m, c, CH_EULER_MOBILITY) {
        // src/integrators/EulerMobilityCPU.cpp:10-22 (keys, defaults, noise factor)
        if(this->template _config_optional_value<bool>(c, "mobility.with_noise", false)) {
            this->_desc.with_noise = 1;
            this->_desc.noise_rescale_factor =
                this->template _config_optional_value<double>(c, "mobility.noise_rescale_factor", 1.0);
            this->_desc.seed =
                (uint64_t)this->template _config_optional_value<long long>(c, "seed", (long long)std::time(NULL));
            double dxd = 1.0;
            for(int d = 0; d < dims; d++) dxd *= this->_dx;
            this->info("Integrating the Cahn-Hilliard equation with non-constant mobility and noise (noise_factor = {})",
                       std::sqrt(2.0 / (dxd * this->_dt)) * this->_desc.noise_rescale_factor);
        }
        else {
            this->info("Integrating the Cahn-Hilliard equation with non-constant mobility and without noise");
        }
    }
    std::string_view name() const override { return "EulerMobilityHIP"; }

protected:
    bool _supports_nonconstant_mobility() const override { return true; }
};

template <int dims>
class PseudospectralHIP : public HIPIntegrator<dims> {
public:
    PseudospectralHIP(SimulationState<dims> &s, FreeEnergyModel *m, const chtoml::Table &c)
        : HIPIntegrator<dims>(s, m, c, CH_PSEUDOSPECTRAL) {
        // src/integrators/PseudospectralCPU.cpp:23-25
        double S = this->template _config_optional_value<double>(c, "pseudospectral.S", 0.0);
        this->_desc.S = S * std::pow(s.user_to_internal, 3);
        this->_desc.use_dealias = this->template _config_optional_value<bool>(c, "pseudospectral.use_dealias", false);
        this->info("S = {}, use_dealias = {}", this->_desc.S, this->_desc.use_dealias);
    }
    std::string_view name() const override { return "PseudospectralHIP"; }
};

template <int dims>
class PseudospectralMobilityHIP : public HIPIntegrator<dims> {
public:
    PseudospectralMobilityHIP(SimulationState<dims> &s, FreeEnergyModel *m, const chtoml::Table &c)
        : HIPIntegrator<dims>(s, m, c, CH_PSEUDOSPECTRAL_MOBILITY) {
        // src/integrators/PseudospectralMobilityCPU.cpp:22-23, 203-208 (S is not rescaled here)
        this->_desc.S = this->template _config_optional_value<double>(c, "pseudospectral.S", 0.0);
        this->_desc.use_dealias = this->template _config_optional_value<bool>(c, "pseudospectral.use_dealias", false);
        this->_desc.gmres_restart = this->template _config_optional_value<int>(c, "pseudospectral.gmres_restart", 30);
        this->_desc.gmres_max_iter = this->template _config_optional_value<int>(c, "pseudospectral.gmres_max_iter", 200);
        this->_desc.gmres_tol = this->template _config_optional_value<double>(c, "pseudospectral.gmres_tol", 1e-10);
        this->_desc.use_gmres = this->template _config_optional_value<bool>(c, "pseudospectral.use_gmres", false);
        this->info("S = {}, use_dealias = {}, use_gmres = {}", this->_desc.S, this->_desc.use_dealias,
                   this->_desc.use_gmres);
    }
    std::string_view name() const override { return "PseudospectralMobilityHIP"; }

protected:
    bool _supports_nonconstant_mobility() const override { return true; }
};

}  // namespace ch
