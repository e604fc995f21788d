#pragma once
// Population descriptor (reference include/naer/Population.h:9-17), plus the
// LIF parameters the clock-driven step needs.
#include <cstdint>
#include <memory>
#include <vector>

#include "NeuronType.h"

namespace Naer {

class NeuronModel;

// LIF parameters in the units of the model; the defaults are the reference's
// constants (lif.cl:3-6): tau 20 ms, threshold 0.020 V, reset 0 V, t_ref 2 ms.
struct LifParams {
    float tau_m_ms = 20.0f;
    float v_th = 0.020f;
    float v_reset = 0.0f;
    float v_rest = 0.0f;
    float t_ref_ms = 2.0f;
    float v0_min = 0.0f;
    float v0_max = 0.0f;
    float ext_rate_hz = 0.0f;  // independent Poisson drive per neuron
    float ext_weight = 0.0f;
    float src_rate_hz = 20.0f;  // POISSON sources (input_neuron_group.cl:4)
};

struct Population {
    NeuronType type;
    uint32_t baseAddress;  // address within its type (or model) until compile(), absolute after
    uint32_t size;
    LifParams params;
    // NeuronType::MODEL: the model and its parameters
    std::shared_ptr<const NeuronModel> model;
    std::vector<float> modelParams;
    Population(NeuronType _type, uint32_t _baseAddress, uint32_t _size, const LifParams& _params = {})
        : type(_type)
        , baseAddress(_baseAddress)
        , size(_size)
        , params(_params)
    {
    }
};

}  // namespace Naer
