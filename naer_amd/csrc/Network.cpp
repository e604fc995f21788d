// Naer::Network over the C ABI (reference: src/Network.cpp).  The address
// layout, defaults and error style follow the reference; device work goes
// through naer.h.
#include "naer/Network.h"

#include <algorithm>
#include <cmath>

namespace {

void check(int rc, const char* what)
{
    if (rc != NAER_OK)
        throw std::runtime_error(std::string("Couldn't ") + what + ": " + naer_last_error());
}

uint32_t steps_of(float ms, float dt_ms) { return (uint32_t)std::lround((double)ms / (double)dt_ms); }

}  // namespace

Naer::Network::Network()
    : Network(NetworkOptions{})
{
}

Naer::Network::Network(const NetworkOptions& options)
    : m_options(options)
{
    check(naer_ctx_create(options.device, &m_ctx), "create the device context");
}

Naer::Network::~Network()
{
    if (m_net)
        naer_net_destroy(m_net);
    if (m_ctx)
        naer_ctx_destroy(m_ctx);
}

std::shared_ptr<const Naer::Population> Naer::Network::addPopulation(Naer::NeuronType type, uint32_t size)
{
    return addPopulation(type, size, LifParams{});
}

std::shared_ptr<const Naer::Population> Naer::Network::addPopulation(Naer::NeuronType type, uint32_t size,
                                                                     const LifParams& params)
{
    if (m_net)
        throw std::runtime_error("Couldn't add a population: network already compiled");
    if (type == NeuronType::MODEL)
        throw std::runtime_error("Couldn't add a population: MODEL populations are added with their model");
    auto pop = std::make_shared<Population>(type, m_typeCounts[(int)type], size, params);
    m_populations.push_back(pop);
    m_typeCounts[(int)type] += size;
    return pop;
}

std::shared_ptr<const Naer::Population> Naer::Network::addPopulation(std::shared_ptr<const NeuronModel> model,
                                                                     uint32_t size, const ModelParams& params)
{
    if (m_net)
        throw std::runtime_error("Couldn't add a population: network already compiled");
    if (!model)
        throw std::runtime_error("Couldn't add a population: no model");
    if (params.params.size() > NAER_MODEL_PARAMS)
        throw std::runtime_error("Couldn't add a population: too many model parameters");
    LifParams lif;
    lif.v0_min = params.v0_min;
    lif.v0_max = params.v0_max;
    lif.ext_rate_hz = params.ext_rate_hz;
    lif.ext_weight = params.ext_weight;
    auto pop = std::make_shared<Population>(NeuronType::MODEL, 0, size, lif);
    pop->model = model;
    pop->modelParams = params.params;
    m_populations.push_back(pop);
    m_modelCount += size;
    return pop;
}

void Naer::Network::addConnection(std::shared_ptr<const Naer::Population> source,
                                  std::shared_ptr<const Naer::Population> destination, float defaultWeights,
                                  double probability)
{
    // defaults of src/Network.cpp:84-91
    addConnection(source, destination, defaultWeights, 0.f, 1.f, probability, 0xbf785e3215ac7ebd);
}

void Naer::Network::addConnection(std::shared_ptr<const Naer::Population> source,
                                  std::shared_ptr<const Naer::Population> destination, float defaultWeights,
                                  float minRandomWeights, float maxRandomWeights, double probability,
                                  uint64_t seed)
{
    m_connections.push_back(std::make_shared<Connection>(source, destination, defaultWeights, minRandomWeights,
                                                         maxRandomWeights, probability, seed));
}

void Naer::Network::addConnectionFixedIndegree(std::shared_ptr<const Naer::Population> source,
                                               std::shared_ptr<const Naer::Population> destination,
                                               uint32_t indegree, float weight, float wmin, float wmax,
                                               uint64_t seed, DelaySteps delay)
{
    auto c = std::make_shared<Connection>(source, destination, weight, wmin, wmax, 0.0, seed);
    c->rule = ConnectionRule::FixedIndegree;
    c->indegree = indegree;
    c->delay = delay;
    m_connections.push_back(c);
}

void Naer::Network::compile()
{
    if (m_net)
        throw std::runtime_error("Couldn't compile: network already compiled");
    // address layout by type [STDP | NOP | POISSON] (src/Network.cpp:269-285),
    // with the MODEL populations after NOP, grouped by model in order of first
    // use
    std::vector<std::shared_ptr<Population>> order;
    for (int t : { 0, 1 })
        for (auto& p : m_populations)
            if ((int)p->type == t)
                order.push_back(p);
    std::vector<const NeuronModel*> models;
    for (auto& p : m_populations)
        if (p->type == NeuronType::MODEL &&
            std::find(models.begin(), models.end(), p->model.get()) == models.end())
            models.push_back(p->model.get());
    for (const NeuronModel* m : models)
        for (auto& p : m_populations)
            if (p->type == NeuronType::MODEL && p->model.get() == m)
                order.push_back(p);
    for (auto& p : m_populations)
        if (p->type == NeuronType::POISSON)
            order.push_back(p);
    std::vector<naer_pop_desc> pops;
    uint32_t address = 0;
    for (auto& p : order) {
        p->baseAddress = address;
        address += p->size;
        const LifParams& q = p->params;
        const uint32_t kind = p->type == NeuronType::POISSON ? NAER_POP_POISSON
                            : p->type == NeuronType::MODEL   ? NAER_POP_MODEL
                                                             : NAER_POP_LIF;
        naer_pop_desc d{ kind, p->size, q.tau_m_ms, q.v_th, q.v_reset, q.v_rest, q.t_ref_ms, q.v0_min, q.v0_max,
                         q.ext_rate_hz, q.ext_weight, q.src_rate_hz };
        if (p->type == NeuronType::MODEL) {
            d.model = p->model->handle();
            for (size_t k = 0; k < p->modelParams.size(); ++k)
                d.params[k] = p->modelParams[k];
        }
        pops.push_back(d);
    }
    auto index_of = [&](const std::shared_ptr<const Population>& p) -> uint32_t {
        for (size_t i = 0; i < order.size(); ++i)
            if (order[i].get() == p.get())
                return (uint32_t)i;
        throw std::runtime_error("Couldn't compile: connection refers to a population of another network");
    };
    const uint32_t d = steps_of(m_options.default_delay_ms, m_options.dt_ms);
    std::vector<naer_conn_desc> conns;
    for (auto& c : m_connections) {
        naer_conn_desc k{};
        k.src_pop = index_of(c->source);
        k.dst_pop = index_of(c->destination);
        k.rule = c->rule == ConnectionRule::FixedIndegree ? NAER_RULE_FIXED_INDEGREE : NAER_RULE_BERNOULLI;
        k.indegree = c->indegree;
        k.probability = c->probability;
        k.weight = c->defaultWeights;
        k.wmin = c->minRandomWeights;
        k.wmax = c->maxRandomWeights;
        if (c->delay.min == 0) {
            k.delay_min = k.delay_max = d;
            k.delay_quantum = 1;
        } else {
            k.delay_min = c->delay.min;
            k.delay_max = c->delay.max;
            k.delay_quantum = c->delay.quantum ? c->delay.quantum : 1;
        }
        k.plastic = (c->source->type == NeuronType::STDP && c->destination->type == NeuronType::STDP) ? 1u : 0u;
        k.seed = c->seed;
        conns.push_back(k);
    }
    naer_net_desc desc{};
    desc.pops = pops.data();
    desc.n_pops = (uint32_t)pops.size();
    desc.conns = conns.data();
    desc.n_conns = (uint32_t)conns.size();
    desc.stdp = m_options.stdp;
    desc.dt_ms = m_options.dt_ms;
    desc.mode = m_options.mode;
    desc.seed = m_options.seed;
    desc.poisson_seed = m_options.poisson_seed;
    desc.record_capacity = m_options.record_capacity;
    check(naer_net_create(m_ctx, &desc, &m_net), "create the network");
    int rc = naer_net_build(m_net);
    if (rc != NAER_OK) {
        std::string msg = naer_last_error();
        naer_net_destroy(m_net);
        m_net = nullptr;
        throw std::runtime_error("Couldn't build the network: " + msg);
    }
}

void Naer::Network::step()
{
    if (!m_net)
        throw std::runtime_error("Couldn't step: network not compiled");
    // the event-driven mode advances whole 1500 us windows
    const uint32_t n = m_options.mode == NAER_MODE_EVENT
                           ? 1u
                           : (uint32_t)std::lround((double)m_options.epoch_us / (1000.0 * m_options.dt_ms));
    check(naer_net_step(m_net, n), "enqueue the step kernels");
}

void Naer::Network::run(uint32_t steps)
{
    if (!m_net)
        throw std::runtime_error("Couldn't run: network not compiled");
    check(naer_net_step(m_net, steps), "enqueue the step kernels");
}

void Naer::Network::finish()
{
    if (!m_net)
        return;
    check(naer_net_sync(m_net), "finish the queued kernels");
    naer_stats s = stats();
    nop_time = s.kernel_ms[NAER_K_UPDATE] * 1e-3;
    scatter_time = s.kernel_ms[NAER_K_SCATTER] * 1e-3;
    stdp_time = s.kernel_ms[NAER_K_POTENTIATE] * 1e-3;
}

void Naer::Network::setProfiling(bool on)
{
    if (m_net)
        check(naer_net_set_profiling(m_net, on ? 1 : 0), "set profiling");
}

naer_stats Naer::Network::stats()
{
    naer_stats s{};
    if (m_net)
        check(naer_net_query(m_net, &s), "query the network");
    return s;
}

std::vector<naer_spike> Naer::Network::spikes()
{
    std::vector<naer_spike> out(m_options.record_capacity);
    uint64_t n = 0;
    if (m_net && !out.empty())
        check(naer_net_read_spikes(m_net, out.data(), out.size(), &n), "read the spikes");
    out.resize(n);
    return out;
}

uint32_t Naer::Network::getTime() { return (uint32_t)stats().time_us; }

uint32_t Naer::Network::getSynapses() { return (uint32_t)stats().n_synapses; }

uint32_t Naer::Network::getNeurons()
{
    return m_typeCounts[0] + m_typeCounts[1] + m_typeCounts[2] + m_modelCount;
}
