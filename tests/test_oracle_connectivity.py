"""The oracle's fixed in-degree rule, pinned without the reference (no GPU).

The configs' networks are built by a fixed in-degree rule that the reference
does not have (its only rule is pairwise Bernoulli, network.cl:14-107), so
the oracle's restatement is pinned by what the rule's definition fixes:

  * every source of a connection lies in its source population and every
    target takes exactly `indegree` synapses from each connection;
  * sources are uniform over the source population (chi-square), delay
    classes are uniform (chi-square), random weights are uniform on
    [wmin, wmax) (Kolmogorov-Smirnov), constant weights are the constant;
  * source, delay class and weight are independent (contingency chi-square);
  * the documented stream layout (include/naer/naer.h: draw k of target j of
    connection c is Philox block {j, k, NAER_STREAM_CONN, c} under the
    connection seed, word 0 -> source, 1 -> delay, 2 -> weight), evaluated by
    an independent numpy Philox4x32-10 that is itself checked against the
    Random123 known answers and ATen's engine.

Each planted mistake (oracle.set_conn_mutation) must fail at least one check.
"""
import json
import os

import numpy as np
import pytest
from scipy import stats

import oracle as O
from conftest import GOLDEN
from naer_amd import models
from naer_amd.network import Network, NeuronType

P_MIN = 1e-6  # a correct rule passes each test at this level; the seeds are fixed
MASK32 = np.uint64(0xFFFFFFFF)
STREAM_CONN = 3


def philox_np(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 over uint64 arrays holding 32-bit words (Salmon et al.,
    SC'11): round (x0,x1,x2,x3) -> (hi(M1 x2)^x1^k0, lo(M1 x2), hi(M0 x0)^x3^k1,
    lo(M0 x0)), Weyl key schedule."""
    c = [np.asarray(x, np.uint64) & MASK32 for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0), np.uint64(k1)
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & MASK32, p1 & MASK32,
             ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & MASK32, p0 & MASK32]
        k0 = (k0 + np.uint64(0x9E3779B9)) & MASK32
        k1 = (k1 + np.uint64(0xBB67AE85)) & MASK32
    return c


def test_numpy_philox_known_answers(golden):
    h = lambda x: int(x, 16)
    for v in golden["philox_kat"]:
        ctr = [h(x) for x in v["ctr"]]
        key = [h(x) for x in v["key"]]
        out = philox_np(*[np.array([x]) for x in ctr], *key)
        assert [int(o[0]) for o in out] == [h(x) for x in v["out"]]
    d = json.load(open(os.path.join(GOLDEN, "philox_vectors.json")))["vectors"][:64]
    for v in d:
        out = philox_np(*[np.array([x]) for x in v["ctr"]], *v["key"])
        assert [int(o[0]) for o in out] == v["out"]


# ---------------------------------------------------------------------------
# The test network: populations of unequal size at unequal bases, one
# connection of random weights with wmin > 0, delays on three grids.
N_E, N_I = 1500, 400
K_E, K_I = 120, 30
W_LO, W_HI = 0.05, 0.3


def test_net():
    net = Network(stdp=models.PAIR_STDP, seed=0x1234_5678_9ABC_DEF0)
    E = net.addPopulation(NeuronType.STDP, N_E)
    I = net.addPopulation(NeuronType.NOP, N_I)
    net.addConnectionFixedIndegree(E, E, K_E, 0.0, W_LO, W_HI, seed=0xA5A5_0000_0000_0011, delay_ms=(1.0, 5.0, 1.0))
    net.addConnectionFixedIndegree(E, I, K_E, 0.1, seed=0x0000_0001_0000_0022, delay_ms=(1.0, 5.0, 1.0))
    net.addConnectionFixedIndegree(I, E, K_I, -0.5, seed=0xFFFF_FFFF_FFFF_FF33, delay_ms=(0.5, 3.0, 0.5))
    net.addConnectionFixedIndegree(I, I, K_I, -0.5, seed=77, delay_ms=1.5)
    return net


test_net.__test__ = False  # a builder, not a test


def build_csr(mutation=O.MUT_NONE):
    net = test_net()
    desc = net.description()
    O.set_conn_mutation(mutation)
    try:
        o = O.OracleNet(desc, O.FIXED, threads=4)
    finally:
        O.set_conn_mutation(O.MUT_NONE)
    rp, col, w = o.csr()
    o.close()
    conns = [desc.conns[i] for i in range(desc.n_conns)]
    bases = [0, N_E]
    sizes = [N_E, N_I]
    return rp, col, w, conns, bases, sizes


def independence_p(tab):
    """p-value of the contingency chi-square, empty rows and columns dropped
    (a planted mistake can leave some bins without synapses)."""
    tab = tab[tab.sum(axis=1) > 0][:, tab.sum(axis=0) > 0]
    if tab.shape[0] < 2 or tab.shape[1] < 2:
        return 0.0
    return stats.chi2_contingency(tab).pvalue


def synapse_table(rp, col, w):
    src = np.repeat(np.arange(len(rp) - 1, dtype=np.int64), np.diff(rp).astype(np.int64))
    tgt = (col & 0x00FFFFFF).astype(np.int64)
    dly = ((col >> 24) & 0x7F).astype(np.int64)
    return src, tgt, dly, w


def check_fixed_indegree(rp, col, w, conns, bases, sizes, n_spec=25):
    """All checks of the rule; returns the list of failures (empty: pinned)."""
    fail = []
    src, tgt, dly, wt = synapse_table(rp, col, w)
    tgt_pop = (tgt >= bases[1]).astype(np.int64)
    src_pop = (src >= bases[1]).astype(np.int64)
    # every target: exactly the in-degree of each connection into its population
    for c in conns:
        in_c = (tgt_pop == c.dst_pop) & (src_pop == c.src_pop)
        per_t = np.bincount(tgt[in_c] - bases[c.dst_pop], minlength=sizes[c.dst_pop])
        if len(per_t) != sizes[c.dst_pop] or not np.all(per_t == c.indegree):
            fail.append(f"conn {c.src_pop}->{c.dst_pop}: in-degrees {per_t.min()}..{per_t.max()}, want {c.indegree}")
    if np.any(src >= bases[1] + sizes[1]) or np.any(tgt >= bases[1] + sizes[1]):
        fail.append("a source or target outside every population")
    for ci, c in enumerate(conns):
        sel = (tgt_pop == c.dst_pop) & (src_pop == c.src_pop)
        s = src[sel] - bases[c.src_pop]
        d = dly[sel]
        x = wt[sel]
        if len(s) == 0:
            fail.append(f"conn {ci}: no synapses")
            continue
        # sources uniform over the source population
        cnt = np.bincount(s, minlength=sizes[c.src_pop])
        p = stats.chisquare(cnt).pvalue
        if p < P_MIN:
            fail.append(f"conn {ci}: sources not uniform (p={p:.2e})")
        # delay classes uniform on the grid
        q = c.delay_quantum or 1
        choices = (c.delay_max - c.delay_min) // q + 1
        on_grid = (d >= c.delay_min) & (d <= c.delay_max) & ((d - c.delay_min) % q == 0)
        if not np.all(on_grid):
            fail.append(f"conn {ci}: delay off the grid")
            s, d, x = s[on_grid], d[on_grid], x[on_grid]
        cls = (d - c.delay_min) // q
        if choices > 1:
            p = stats.chisquare(np.bincount(cls, minlength=choices)).pvalue
            if p < P_MIN:
                fail.append(f"conn {ci}: delay classes not uniform (p={p:.2e})")
            # source and delay class independent
            sb = (s * 10) // sizes[c.src_pop]
            tab = np.zeros((10, choices))
            np.add.at(tab, (sb, cls), 1)
            p = independence_p(tab)
            if p < P_MIN:
                fail.append(f"conn {ci}: delay depends on the source (p={p:.2e})")
        # weights
        if c.weight != 0.0:
            if not np.all(x == np.float32(c.weight)):
                fail.append(f"conn {ci}: weights differ from the constant")
        else:
            if np.any((x < np.float32(c.wmin)) | (x >= np.float32(c.wmax))):
                fail.append(f"conn {ci}: weight outside [wmin, wmax)")
            p = stats.kstest(x.astype(np.float64), "uniform", args=(c.wmin, c.wmax - c.wmin)).pvalue
            if p < P_MIN:
                fail.append(f"conn {ci}: weights not uniform on [wmin, wmax) (p={p:.2e})")
            wb = np.clip(((x - c.wmin) / (c.wmax - c.wmin) * 10).astype(np.int64), 0, 9)
            sb = (s * 10) // sizes[c.src_pop]
            tab = np.zeros((10, 10))
            np.add.at(tab, (sb, wb), 1)
            p = independence_p(tab)
            if p < P_MIN:
                fail.append(f"conn {ci}: weight depends on the source (p={p:.2e})")
        # the documented stream layout on a sample of targets
        k0, k1 = c.seed & 0xFFFFFFFF, c.seed >> 32
        for j in np.linspace(0, sizes[c.dst_pop] - 1, n_spec).astype(np.int64):
            kk = np.arange(c.indegree, dtype=np.uint64)
            r = philox_np(np.full_like(kk, j), kk, np.full_like(kk, STREAM_CONN), np.full_like(kk, ci), k0, k1)
            want_s = bases[c.src_pop] + ((r[0] * np.uint64(sizes[c.src_pop])) >> np.uint64(32)).astype(np.int64)
            want_d = c.delay_min + q * ((r[1] * np.uint64(choices)) >> np.uint64(32)).astype(np.int64)
            if c.weight != 0.0:
                want_w = np.full(c.indegree, np.float32(c.weight))
            else:
                u = (r[2] >> np.uint64(8)).astype(np.float32) * np.float32(2.0 ** -24)
                want_w = np.float32(c.wmin) + np.float32(c.wmax - c.wmin) * u
            want = sorted(zip(want_s.tolist(), want_d.tolist(), want_w.view(np.uint32).tolist()))
            at = sel & (tgt == bases[c.dst_pop] + j)
            got = sorted(zip(src[at].tolist(), dly[at].tolist(), wt[at].view(np.uint32).tolist()))
            if got != want:
                fail.append(f"conn {ci}: target {j} differs from the documented stream layout")
                break
    return fail


def test_fixed_indegree_rule_pinned():
    rp, col, w, conns, bases, sizes = build_csr()
    assert check_fixed_indegree(rp, col, w, conns, bases, sizes) == []


@pytest.mark.parametrize("mutation", [O.MUT_SRC_RANGE, O.MUT_SRC_BASE, O.MUT_WORDS_SWAPPED, O.MUT_DELAY_REUSE,
                                      O.MUT_WEIGHT_REUSE, O.MUT_WEIGHT_SPAN])
def test_planted_mistake_is_caught(mutation):
    rp, col, w, conns, bases, sizes = build_csr(mutation)
    assert check_fixed_indegree(rp, col, w, conns, bases, sizes) != []


def test_planted_mistakes_caught_by_statistics_alone():
    """Every mistake except the word swap (whose marginals and joint law are
    those of the correct rule) fails a check that does not use the stream
    layout."""
    for m in (O.MUT_SRC_RANGE, O.MUT_SRC_BASE, O.MUT_DELAY_REUSE, O.MUT_WEIGHT_REUSE, O.MUT_WEIGHT_SPAN):
        rp, col, w, conns, bases, sizes = build_csr(m)
        assert check_fixed_indegree(rp, col, w, conns, bases, sizes, n_spec=0) != [], m


def test_configs_networks_use_the_pinned_rule():
    """The benchmark builders describe their networks with this rule: the
    in-degree split and weights of balanced_10k, plastic_100k and large_1m."""
    for net, k, je in ((models.balanced_10k(n=1000), 1000, 0.1), (models.plastic_100k(n=1000), 1000, 0.1),
                       (models.large_1m(n=1000), 2000, 0.05)):
        d = net.description()
        cs = [d.conns[i] for i in range(d.n_conns)]
        assert all(c.rule == 1 for c in cs)
        assert [c.indegree for c in cs] == [int(round(0.8 * k))] * 2 + [k - int(round(0.8 * k))] * 2
        assert np.float32(cs[1].weight) == np.float32(je)
        assert np.float32(cs[2].weight) == np.float32(-5.0 * je)


def test_canonical_row_order():
    """Rows are sorted by (target group within the net, word, weight bits),
    group = target >> 14: a net of 40,000 neurons spans three groups."""
    net = models.large_1m(n=40_000, indegree=20)
    o = O.OracleNet(net.description(), O.FIXED, threads=4)
    rp, col, w = o.csr()
    o.close()
    assert int(np.max(col & 0xFFFFFF)) >= 2 << 14
    key = (((col & 0xFFFFFF).astype(np.uint64) >> np.uint64(14)) << np.uint64(33)) | \
        (col.astype(np.uint64) << np.uint64(1))
    src = np.repeat(np.arange(len(rp) - 1), np.diff(rp).astype(np.int64))
    same_row = src[1:] == src[:-1]
    k0, k1 = key[:-1], key[1:]
    w0, w1 = w[:-1].view(np.uint32), w[1:].view(np.uint32)
    ordered = (k0 < k1) | ((k0 == k1) & (w0 <= w1))
    assert np.all(ordered[same_row])
