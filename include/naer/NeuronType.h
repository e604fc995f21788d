#pragma once
// Population types of the reference (include/naer/Network.h:82-84 uses
// NeuronType::STDP / NOP / POISSON as an enumeration; the reference's
// NeuronType.h:12-50 declares a class that does not compile).  NOP and STDP
// are leaky integrate-and-fire populations; synapses from an STDP population
// onto an STDP population are plastic (stdp.cl:100).  POISSON is a spike
// source with the reference's exponential-ISI trains.
namespace Naer {

// MODEL: a population of a source-defined neuron model (NeuronModel.h); its
// neurons follow the NOP populations in the address order.
enum class NeuronType { STDP = 0, NOP = 1, POISSON = 2, MODEL = 3 };

}  // namespace Naer
