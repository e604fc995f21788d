"""The oracle against the reference's golden values (no GPU).

Pins: Philox4x32-10 (Random123 KATs + 256 ATen vectors), the reference's
uniform_rand map (random.cl:6-24), its exponential-ISI Poisson train
(input_neuron_group.cl:29) and its pairwise-Bernoulli connectivity
(network.cl:14-107) on the brunel model of src/main.cpp:32-50.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN


def h(x):
    return int(x, 16) if isinstance(x, str) else int(x)


def test_philox_known_answers(golden):
    for v in golden["philox_kat"]:
        out = O.philox([h(x) for x in v["ctr"]], [h(x) for x in v["key"]])
        assert [int(x) for x in out[0]] == [h(x) for x in v["out"]]


def test_philox_matches_aten_engine():
    d = json.load(open(os.path.join(GOLDEN, "philox_vectors.json")))["vectors"]
    ctr = np.array([v["ctr"] for v in d], np.uint32)
    key = np.array([v["key"] for v in d], np.uint32)
    ref = np.array([v["out"] for v in d], np.uint32)
    np.testing.assert_array_equal(O.philox(ctr, key), ref)


def test_philox_planted_mistake_is_caught(golden):
    # swapping the two multipliers (a plausible transcription slip) must fail the KAT
    v = golden["philox_kat"][2]
    ctr = [h(x) for x in v["ctr"]]
    key = [h(x) for x in v["key"]]

    def philox_py(c, k, m0, m1):
        c = list(c)
        k = list(k)
        for _ in range(10):
            p0, p1 = m0 * c[0], m1 * c[2]
            c = [((p1 >> 32) ^ c[1] ^ k[0]) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
                 ((p0 >> 32) ^ c[3] ^ k[1]) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
            k = [(k[0] + 0x9E3779B9) & 0xFFFFFFFF, (k[1] + 0xBB67AE85) & 0xFFFFFFFF]
        return c

    want = [h(x) for x in v["out"]]
    assert philox_py(ctr, key, 0xD2511F53, 0xCD9E8D57) == want
    assert philox_py(ctr, key, 0xCD9E8D57, 0xD2511F53) != want


def test_reference_uniform(golden):
    g = golden["uniform_rand"]
    u = O.ref_uniform(g["gid"], h(g["seed"]), g["iteration"])
    assert u == h(g["z_word"]) / 4294967295.0
    assert abs(u - g["u"]) < 1e-17


def test_reference_poisson_first_spikes(golden):
    g = golden["poisson_first_isi"]
    for idx, want in enumerate(g["times"]):
        tr = O.ref_poisson_epoch(idx, 22360, h(g["seed"]), g["rate_hz"], 0, 10**9)
        assert int(tr[0]) == want


def _brunel_sizes(s):
    N = math.floor(math.sqrt(s / 0.05))
    return int(N / 2), int(N * 4 / 10), int(N / 10)


@pytest.mark.parametrize("idx", [0, 1])
def test_reference_bernoulli_synapse_counts(golden, idx):
    g = golden["brunel_counts"][idx]
    P, E, I = _brunel_sizes(g["synapses_arg"])
    assert P + E + I == g["N"]
    tot = 0
    for a, b in [(P, E), (P, I), (E, E), (E, I), (I, E), (I, I)]:
        for s in range(a):
            tot += len(O.ref_bernoulli_row(s, b, 0.1, 0x15AC7EBD))
    assert tot == g["S"]


def test_reference_bernoulli_rows(golden):
    g = golden["bernoulli_rows"]
    for src, key in ((0, "src0_below200"), (1, "src1_below200")):
        row = O.ref_bernoulli_row(src, g["n_dst"], g["p"], h(g["seed"]))
        assert [int(x) for x in row if x < 200] == g[key]


def test_published_population_sizes(golden):
    for g in golden["published_N"]:
        assert sum(_brunel_sizes(g["synapses_arg"])) == g["N"]


def test_reference_poisson_gid_map():
    # 128 work-items per 2048-neuron group (Network.cpp:181-184)
    assert O.lib().orc_ref_poisson_gid(0) == 0
    assert O.lib().orc_ref_poisson_gid(127) == 127
    assert O.lib().orc_ref_poisson_gid(128) == 0
    assert O.lib().orc_ref_poisson_gid(2048) == 128
    assert O.lib().orc_ref_poisson_gid(2048 + 130) == 130
