// SimulationState.h -- host fields and shared state of the drop-in.
//
// MultiField keeps the reference's species-major layout data[s * bins + idx]
// (src/utils/MultiField.h:49-108); SimulationState carries the same members as
// src/SimulationState.h:18-51 (rho, mobility, model, integrator, unit factors).
#pragma once

#include <memory>
#include <vector>

namespace ch {

template <class T>
class MultiField {
    std::vector<T> _data;
    int _bins = 0, _species = 0;

public:
    MultiField() = default;
    MultiField(int bins, int species) : _data((size_t)bins * species), _bins(bins), _species(species) {}
    T &operator()(int idx, int s) { return _data[(size_t)s * _bins + idx]; }
    T operator()(int idx, int s) const { return _data[(size_t)s * _bins + idx]; }
    void fill(const T &v) { std::fill(_data.begin(), _data.end(), v); }
    T *data() { return _data.data(); }
    const T *data() const { return _data.data(); }
    T field_sum(int idx) const {
        T t{};
        for(int s = 0; s < _species; s++) t += (*this)(idx, s);
        return t;
    }
    int bins() const { return _bins; }
    int species() const { return _species; }
    int size() const { return (int)_data.size(); }
};

class FreeEnergyModel;
template <int dims> class Integrator;

template <int dims>
struct SimulationState {
    MultiField<double> rho;
    MultiField<double> mobility;
    std::unique_ptr<FreeEnergyModel> model;
    Integrator<dims> *integrator = nullptr;
    double user_to_internal = 1.0;
    double internal_to_user = 1.0;
    bool use_CUDA = true;

    double density_to_user(double v) const { return v / (internal_to_user * internal_to_user * internal_to_user); }
    double length_to_user(double v) const { return v * internal_to_user; }
};

}  // namespace ch
