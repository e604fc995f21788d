#pragma once
// Naer::Network — the reference's model-description and simulation API
// (include/naer/Network.h:15-56 of the reference), implemented over the C ABI
// in naer.h.  Every reference signature is kept, so the reference's driver
// models (src/main.cpp:12-50) compile against this header unchanged; errors
// are thrown as std::runtime_error like the reference's.
//
// step() advances one reference epoch (1500 us, src/Network.cpp:125), i.e.
// 1500 us / dt clock-driven steps; run(n) advances n steps of dt.
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Connection.h"
#include "NeuronModel.h"
#include "naer.h"

namespace Naer {

struct NetworkOptions {
    int device = 0;
    float dt_ms = 0.1f;
    uint32_t mode = NAER_MODE_FAST;
    uint64_t seed = 0xbf785e3215ac7ebdull;    // src/Network.cpp:91
    uint32_t poisson_seed = 0x15ac7ebdu;        // (uint32_t) of the same, src/Network.cpp:124
    uint32_t record_capacity = 0;
    float default_delay_ms = 1.5f;              // NAER_AXONAL_DELAY, naer_types.h:4
    uint32_t epoch_us = 1500;
    naer_stdp_desc stdp{ 20.0f, 20.0f, 0.0f, 0.0f, 0.0f, 0.0f };
};

class Network {
public:
    Network();
    explicit Network(const NetworkOptions& options);

    ~Network();

    Network(const Network&) = delete;
    Network& operator=(const Network&) = delete;

    std::shared_ptr<const Naer::Population> addPopulation(Naer::NeuronType type, uint32_t size);
    std::shared_ptr<const Naer::Population> addPopulation(Naer::NeuronType type, uint32_t size,
                                                          const LifParams& params);
    // a population of a source-defined model (the reference's NeuronType
    // source-string path); the model must live until compile() returns
    std::shared_ptr<const Naer::Population> addPopulation(std::shared_ptr<const NeuronModel> model, uint32_t size,
                                                          const ModelParams& params);

    void addConnection(std::shared_ptr<const Naer::Population> source,
                       std::shared_ptr<const Naer::Population> destination, float defaultWeights,
                       double probability);

    void addConnection(std::shared_ptr<const Naer::Population> source,
                       std::shared_ptr<const Naer::Population> destination, float defaultWeights,
                       float minRandomWeights, float maxRandomWeights, double probability, uint64_t seed);

    // Every destination neuron receives `indegree` synapses from sources drawn
    // uniformly from `source`; weight == 0 draws weights in [wmin, wmax).
    void addConnectionFixedIndegree(std::shared_ptr<const Naer::Population> source,
                                    std::shared_ptr<const Naer::Population> destination, uint32_t indegree,
                                    float weight, float wmin, float wmax, uint64_t seed,
                                    DelaySteps delay = {});

    void compile();

    void step();

    void finish();

    uint32_t getTime();

    uint32_t getSynapses();

    uint32_t getNeurons();

    // the reference's public timers (seconds); the update kernel's time is
    // reported in nop_time, the scatter's in scatter_time, potentiation in
    // stdp_time
    double nop_time = 0;
    double stdp_time = 0;
    double poisson_time = 0;
    double scan_time = 0;
    double scatter_time = 0;
    double offset_time = 0;

    // additions
    void run(uint32_t steps);
    void setProfiling(bool on);
    naer_stats stats();
    std::vector<naer_spike> spikes();
    const NetworkOptions& options() const { return m_options; }

private:
    NetworkOptions m_options;
    std::vector<std::shared_ptr<Population>> m_populations;
    std::vector<std::shared_ptr<Connection>> m_connections;
    uint32_t m_typeCounts[3] = { 0, 0, 0 };
    uint32_t m_modelCount = 0;  // neurons of MODEL populations
    naer_ctx* m_ctx = nullptr;
    naer_net* m_net = nullptr;
};

}  // namespace Naer
