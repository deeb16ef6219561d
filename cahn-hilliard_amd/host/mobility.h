// mobility.h -- IMobility plugins of the drop-in (host side).
//
// build_mobility follows src/mobility/IMobility.cpp:18-57 (keys mobility.type,
// mobility.M default 1.0, mobility.rho_min; unit scaling M ~ m^-1 as in
// ConstantMobility.h:20, FreeEnergyMobility.h:21, RegularisedMobility.h:20).
// "linear" (M rho / mobility.rho_ref) is new.  The HIP integrators evaluate
// the mobility inside the fused step kernel from the same rho the reference's
// update_mobility() would see (CahnHilliard.cpp:279-282), so update_mobility()
// has no device work left to do; describe() hands the rule to the C ABI.
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/ch_hip.h"
#include "Object.h"
#include "SimulationState.h"

namespace ch {

template <int dims>
class IMobility : public Object {
public:
    IMobility(SimulationState<dims> &state, int type, double M, double rho_min = 0.0, double rho_ref = 1.0)
        : _sim_state(state), _type(type), _M(M), _rho_min(rho_min), _rho_ref(rho_ref) {
        if(type == CH_MOB_CONSTANT) state.mobility.fill(M);  // ConstantMobility.h:19-23
    }
    virtual void update_mobility() {}
    void describe(ch_desc &d) const {
        d.mobility = _type;
        d.M = _M;
        d.rho_min = _rho_min;
        d.rho_ref = _rho_ref;
    }
    int type() const { return _type; }
    std::string_view name() const override { return "IMobility"; }

protected:
    SimulationState<dims> &_sim_state;
    int _type;
    double _M, _rho_min, _rho_ref;
};

template <int dims>
IMobility<dims> *build_mobility(const chtoml::Table &config, SimulationState<dims> &state) {
    struct Helper : Object {
        std::string_view name() const override { return "build_mobility"; }
    } h;
    std::string type = h._config_optional_value<std::string>(config, "mobility.type", "constant");
    double M = h._config_optional_value<double>(config, "mobility.M", 1.0) / state.user_to_internal;
    if(type == "constant") return new IMobility<dims>(state, CH_MOB_CONSTANT, M);
    if(type == "free_energy") return new IMobility<dims>(state, CH_MOB_FREE_ENERGY, M);
    if(type == "regularised") {
        if(!chtoml::find(config, "mobility.rho_min")) h.critical("For 'regularised' mobility, 'rho_min' must be specified");
        double rho_min = h._config_value<double>(config, "mobility.rho_min") / state.user_to_internal;
        return new IMobility<dims>(state, CH_MOB_REGULARISED, M, rho_min);
    }
    if(type == "linear") {
        if(!chtoml::find(config, "mobility.rho_ref")) h.critical("For 'linear' mobility, 'rho_ref' must be specified");
        // a density, converted like rho itself (CahnHilliard.cpp:158-161)
        double rho_ref = h._config_value<double>(config, "mobility.rho_ref") / CUB(state.user_to_internal);
        return new IMobility<dims>(state, CH_MOB_LINEAR, M, 0.0, rho_ref);
    }
    if(type == "gel") h.critical("The 'gel' mobility is not supported by the HIP library");
    throw std::runtime_error("Unknown mobility type: " + type);
}

}  // namespace ch
