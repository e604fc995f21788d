// naer — command-line driver with the reference's interface (src/main.cpp):
//   naer [--synapses S] [--model brunel|brunel+|vogels]
// builds the model, runs 10 s of simulated time with `while (getTime() <
// 10'000'000) step();` and prints (N, S), Took and the kernel timers.
// Additions: --workload balanced_10k|large_1m (fixed in-degree workloads,
// 1 s), --mode fast|det|event (event: the reference's event-driven step),
// --seconds T.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>

#include "naer/Network.h"

namespace {

// src/main.cpp:12-30
std::unique_ptr<Naer::Network> brunel_plus(int num_synapses, const Naer::NetworkOptions& o)
{
    double N = std::floor(std::sqrt(num_synapses / 0.05));
    auto opts = o;
    opts.stdp = naer_stdp_desc{ 20.0f, 20.0f, 0.01f * 0.0003f, 0.0105f * 0.0003f, 0.0f, 0.0003f };
    auto net = std::make_unique<Naer::Network>(opts);
    auto P = net->addPopulation(Naer::NeuronType::POISSON, N / 2);
    auto E = net->addPopulation(Naer::NeuronType::STDP, N * 4 / 10);
    auto I = net->addPopulation(Naer::NeuronType::NOP, N / 10);
    net->addConnection(P, E, 2.0 / N, 0.1);
    net->addConnection(P, I, 2.0 / N, 0.1);
    net->addConnection(E, E, 2.0 / N, 0.1);
    net->addConnection(E, I, 2.0 / N, 0.1);
    net->addConnection(I, E, -10.0 / N, 0.1);
    net->addConnection(I, I, -10.0 / N, 0.1);
    return net;
}

// src/main.cpp:32-50
std::unique_ptr<Naer::Network> brunel(int num_synapses, const Naer::NetworkOptions& o)
{
    double N = std::floor(std::sqrt(num_synapses / 0.05));
    auto net = std::make_unique<Naer::Network>(o);
    auto P = net->addPopulation(Naer::NeuronType::POISSON, N / 2);
    auto E = net->addPopulation(Naer::NeuronType::NOP, N * 4 / 10);
    auto I = net->addPopulation(Naer::NeuronType::NOP, N / 10);
    net->addConnection(P, E, 2.0 / N, 0.1);
    net->addConnection(P, I, 2.0 / N, 0.1);
    net->addConnection(E, E, 2.0 / N, 0.1);
    net->addConnection(E, I, 2.0 / N, 0.1);
    net->addConnection(I, E, -10.0 / N, 0.1);
    net->addConnection(I, I, -10.0 / N, 0.1);
    return net;
}

// src/main.cpp:52-69, commented out there: its CubaLIF type is not defined.
// Here E and I are the library's cuba_lif model (current-based LIF) with the
// reference's LIF constants (lif.cl:3-6) and tau_s = 5 ms.
std::unique_ptr<Naer::Network> vogels(int num_synapses, const Naer::NetworkOptions& o)
{
    double N = std::floor(std::sqrt(num_synapses / 0.05));
    auto net = std::make_unique<Naer::Network>(o);
    auto cuba = Naer::NeuronModel::builtin("cuba_lif");
    Naer::ModelParams mp;
    mp.params = Naer::cubaLifParams(Naer::LifParams{}, 5.0f, o.dt_ms);
    auto P = net->addPopulation(Naer::NeuronType::POISSON, N / 2);
    auto E = net->addPopulation(cuba, N * 4 / 10, mp);
    auto I = net->addPopulation(cuba, N / 10, mp);
    net->addConnection(P, E, 2.0 / N, 0.1);
    net->addConnection(P, I, 2.0 / N, 0.1);
    net->addConnection(E, E, 2.0 / N, 0.1);
    net->addConnection(E, I, 2.0 / N, 0.1);
    net->addConnection(I, E, -10.0 / N, 0.1);
    net->addConnection(I, I, -10.0 / N, 0.1);
    return net;
}

// Balanced fixed in-degree network in mV (tau 20 ms, theta 20, reset 10,
// t_ref 2 ms, V0 ~ U[0,20)).
std::unique_ptr<Naer::Network> balanced(uint32_t n, uint32_t k, float je, float g, float drive_hz, float drive_w,
                                        Naer::DelaySteps delay, const Naer::NetworkOptions& o)
{
    auto net = std::make_unique<Naer::Network>(o);
    Naer::LifParams lif;
    lif.tau_m_ms = 20.0f;
    lif.v_th = 20.0f;
    lif.v_reset = 10.0f;
    lif.t_ref_ms = 2.0f;
    lif.v0_min = 0.0f;
    lif.v0_max = 20.0f;
    lif.ext_rate_hz = drive_hz;
    lif.ext_weight = drive_w;
    const uint32_t ne = (uint32_t)std::lround(n * 0.8), ni = n - ne;
    const uint32_t ke = (uint32_t)std::lround(k * 0.8), ki = k - ke;
    auto E = net->addPopulation(Naer::NeuronType::NOP, ne, lif);
    auto I = net->addPopulation(Naer::NeuronType::NOP, ni, lif);
    const uint64_t s = o.seed;
    net->addConnectionFixedIndegree(E, E, ke, je, 0, 0, s + 1, delay);
    net->addConnectionFixedIndegree(E, I, ke, je, 0, 0, s + 2, delay);
    net->addConnectionFixedIndegree(I, E, ki, -g * je, 0, 0, s + 3, delay);
    net->addConnectionFixedIndegree(I, I, ki, -g * je, 0, 0, s + 4, delay);
    return net;
}

}  // namespace

int main(int argc, char* argv[])
{
    double S = 500'000'000.0;
    std::string model = "brunel";
    std::string workload;
    double seconds = 10.0;
    Naer::NetworkOptions opts;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto need = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::cout << "you need to specify " << what << "!" << std::endl;
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--synapses") {
            S = std::floor(std::atof(need("a synapse count").c_str()));
            std::cout << "Setting synapse count: " << S << std::endl;
        } else if (a == "--model") {
            model = need("a model from: [brunel, brunel+, vogels]");
            std::cout << "Setting model: " << model << std::endl;
        } else if (a == "--workload") {
            workload = need("a workload from: [balanced_10k, large_1m]");
            seconds = 1.0;
        } else if (a == "--mode") {
            const std::string m = need("a mode from: [fast, det, event]");
            opts.mode = m == "det" ? NAER_MODE_DETERMINISTIC : m == "event" ? NAER_MODE_EVENT : NAER_MODE_FAST;
        } else if (a == "--seconds") {
            seconds = std::atof(need("a duration").c_str());
        }
    }

    std::unique_ptr<Naer::Network> net;
    try {
        if (!workload.empty()) {
            if (workload == "balanced_10k")
                net = balanced(10000, 1000, 0.1f, 5.0f, 20000.0f, 0.1f, { 15, 15, 1 }, opts);
            else if (workload == "large_1m")
                net = balanced(1000000, 2000, 0.05f, 5.0f, 19700.0f, 0.05f, { 10, 80, 10 }, opts);
            else {
                std::cout << workload << " is not a valid workload name" << std::endl;
                return 1;
            }
        } else if (model == "brunel") {
            net = brunel(S, opts);
        } else if (model == "brunel+") {
            net = brunel_plus(S, opts);
        } else if (model == "vogels") {
            net = vogels(S, opts);
        } else {
            std::cout << model << " is not a valid model name" << std::endl;
            std::cout << "you can choose from: [brunel, brunel+, vogels]" << std::endl;
            return 1;
        }
        net->compile();
    } catch (const std::exception& exc) {
        // the reference prints and continues (src/main.cpp:106-110); without a
        // compiled network there is nothing to run, so stop here
        std::cerr << exc.what() << std::endl;
        return 1;
    }

    std::cout << "(N, S): "
              << "(" << net->getNeurons() << ", " << net->getSynapses() << ")" << std::endl;
    const uint64_t end_us = (uint64_t)std::llround(seconds * 1e6);
    auto start = std::chrono::system_clock::now();
    uint64_t t_us = 0;
    const uint64_t epoch = net->options().epoch_us;
    try {
        while (t_us < end_us) {
            net->step();
            t_us += epoch;
        }
        net->finish();
    } catch (const std::exception& exc) {
        std::cerr << exc.what() << std::endl;
        return 1;
    }
    auto end = std::chrono::system_clock::now();
    std::chrono::duration<double> diff = end - start;
    std::cout << "Took: " << diff.count() << std::endl;
    naer_stats s = net->stats();
    std::cout << "Simulated: " << s.time_us << " us, spikes " << s.spikes << ", synaptic events " << s.events
              << std::endl;
    std::cout << "NOP: " << net->nop_time << std::endl;
    std::cout << "STDP: " << net->stdp_time << std::endl;
    std::cout << "Offset: " << net->offset_time << std::endl;
    std::cout << "Scatter: " << net->scatter_time << std::endl;
    std::cout << "Scan: " << net->scan_time << std::endl;
    std::cout << "Poisson: " << net->poisson_time << std::endl;
    return 0;
}
