#pragma once
// Neuron models given as source and compiled at run time (naer_model_create
// in naer.h): the reference's NeuronType, which takes a neuron's struct and
// dynamics as source strings (include/naer/NeuronType.h:12-50 of the
// reference).  A model population is added with
// Network::addPopulation(model, size, ModelParams).
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Population.h"
#include "naer.h"

namespace Naer {

class NeuronModel {
public:
    // `source` defines naer_model_step (see naer_model_create); n_state floats
    // of state per neuron
    NeuronModel(const std::string& source, uint32_t n_state)
        : m_nState(n_state)
    {
        if (naer_model_create(source.c_str(), n_state, &m_model) != NAER_OK)
            throw std::runtime_error(std::string("Couldn't compile the neuron model: ") + naer_last_error());
    }

    // a model the library ships: "lif" (2 state floats) or "cuba_lif" (3)
    static std::shared_ptr<NeuronModel> builtin(const std::string& name)
    {
        const char* src = nullptr;
        if (naer_model_source(name.c_str(), &src) != NAER_OK)
            throw std::runtime_error(std::string("Couldn't find the neuron model: ") + naer_last_error());
        return std::make_shared<NeuronModel>(src, name == "cuba_lif" ? 3u : 2u);
    }

    ~NeuronModel()
    {
        if (m_model)
            naer_model_destroy(m_model);
    }

    NeuronModel(const NeuronModel&) = delete;
    NeuronModel& operator=(const NeuronModel&) = delete;

    const naer_model* handle() const { return m_model; }
    uint32_t stateSize() const { return m_nState; }

private:
    naer_model* m_model = nullptr;
    uint32_t m_nState = 0;
};

// A model population's parameters (up to NAER_MODEL_PARAMS floats) and the
// LIF fields that keep their meaning: initial potential range and drive.
struct ModelParams {
    std::vector<float> params;
    float v0_min = 0.0f;
    float v0_max = 0.0f;
    float ext_rate_hz = 0.0f;
    float ext_weight = 0.0f;
};

// Parameter vectors of the shipped models from physical constants, computed
// as the runtime computes its LIF constants (double arithmetic, rounded to
// float).
inline std::vector<float> lifModelParams(const LifParams& q, float dt_ms)
{
    const double P = std::exp(-(double)dt_ms / (double)q.tau_m_ms);
    const double n_ref = (double)std::llround((double)q.t_ref_ms / (double)dt_ms);
    return { (float)P, q.v_th, q.v_reset, q.v_rest, (float)n_ref };
}

inline std::vector<float> cubaLifParams(const LifParams& q, float tau_s_ms, float dt_ms)
{
    if (tau_s_ms == q.tau_m_ms)
        throw std::runtime_error("Couldn't set up cuba_lif: tau_s must differ from tau_m");
    std::vector<float> p = lifModelParams(q, dt_ms);
    const double p22 = std::exp(-(double)dt_ms / (double)q.tau_m_ms);
    const double p11 = std::exp(-(double)dt_ms / (double)tau_s_ms);
    const double p21 = (double)q.tau_m_ms * (p11 - p22) / ((double)tau_s_ms - (double)q.tau_m_ms);
    p.push_back((float)p11);
    p.push_back((float)p21);
    return p;
}

}  // namespace Naer
