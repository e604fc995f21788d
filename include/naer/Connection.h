#pragma once
// Connection descriptor (reference include/naer/Connection.h:8-30), plus the
// rule, in-degree and delay range of the clock-driven step.
#include <cstdint>
#include <memory>

#include "Population.h"

namespace Naer {

enum class ConnectionRule { Bernoulli = 0, FixedIndegree = 1 };

// Delays in steps: min + m * quantum, m uniform; {0,0,0} = network default.
struct DelaySteps {
    uint32_t min = 0, max = 0, quantum = 1;
};

struct Connection {
    std::shared_ptr<const Population> source;
    std::shared_ptr<const Population> destination;
    float defaultWeights;
    float minRandomWeights;
    float maxRandomWeights;
    double probability;
    uint64_t seed;
    ConnectionRule rule = ConnectionRule::Bernoulli;
    uint32_t indegree = 0;
    DelaySteps delay{};
    Connection(std::shared_ptr<const Population> _source, std::shared_ptr<const Population> _destination,
               float _defaultWeights, float _minRandomWeights, float _maxRandomWeights, double _probability,
               uint64_t _seed)
        : source(_source)
        , destination(_destination)
        , defaultWeights(_defaultWeights)
        , minRandomWeights(_minRandomWeights)
        , maxRandomWeights(_maxRandomWeights)
        , probability(_probability)
        , seed(_seed)
    {
    }
};

}  // namespace Naer
