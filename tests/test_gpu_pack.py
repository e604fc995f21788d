"""Packs (several networks per launch) and graph-captured chunks.

In deterministic mode a network's raster must be bit-identical alone, packed
with others (static and plastic, different sizes, seeds and delays), replayed
through captured HIP graphs, and in the CPU oracle.
"""
import numpy as np
import pytest

import oracle as O
from helpers import first_divergence, oracle_of
from naer_amd import _abi, sweep
from naer_amd.pack import Pack

pytestmark = pytest.mark.gpu

STEPS = 300


def _specs():
    specs = sweep.sweep_specs(6, n_min=300, n_max=3000, seed=1234)
    return specs


def _alone(spec, steps=STEPS, graph=0):
    net = sweep.make_network(spec, _abi.MODE_DETERMINISTIC, sweep.expected_record(spec, steps))
    net.compile()
    if graph:
        net.set_graph(graph)
    net.run(steps)
    sp = net.spikes()
    net.close()
    return sp


def test_pack_equals_alone_and_oracle():
    specs = _specs()
    alone = [_alone(s) for s in specs]
    assert sum(len(a) for a in alone) > 1000
    nets = [sweep.make_network(s, _abi.MODE_DETERMINISTIC, sweep.expected_record(s, STEPS)) for s in specs]
    p = Pack(nets)
    p.build()
    p.run(STEPS)
    packed = p.spikes()
    p.close()
    for i, s in enumerate(specs):
        assert first_divergence(packed[i], alone[i]) is None, (i, s)
    for i in (0, 1):
        o = oracle_of(sweep.make_network(specs[i], _abi.MODE_DETERMINISTIC, sweep.expected_record(specs[i], STEPS)))
        o.run(STEPS)
        assert first_divergence(alone[i], o.spikes()) is None, i


@pytest.mark.parametrize("refine_split", [0, 3])
def test_pack_placements_equal_alone(refine_split):
    """Packs place directly by default (refine_split 0: events straight into
    their tile buckets, no refine pass); refine_split 3: the two-level
    placement with three refine blocks per group bucket.  Against each
    network alone (the single-net default, two levels)."""
    specs = _specs()
    nets = [sweep.make_network(s, _abi.MODE_DETERMINISTIC, sweep.expected_record(s, STEPS)) for s in specs]
    p = Pack(nets)
    p.build()
    sc = p.schedule()
    assert sc["refine_split"] == 0  # the packs' default
    p.set_schedule(sc["task_synapses"], sc["segments_per_block"], refine_split)
    assert p.schedule()["refine_split"] == refine_split
    p.run(STEPS)
    packed = p.spikes()
    p.close()
    for i, s in enumerate(specs):
        assert first_divergence(packed[i], _alone(s)) is None, (i, s)


@pytest.mark.parametrize("plastic_packs", [1, 2])
def test_split_packs_with_graphs_equal_alone(plastic_packs):
    specs = _specs()
    res = sweep.run_packed(specs, STEPS, _abi.MODE_DETERMINISTIC, chunk=100, record=True, graphs=True,
                           plastic_packs=plastic_packs)
    for s in specs:
        assert first_divergence(res["rasters"][s["index"]], _alone(s)) is None, s


def test_graph_replay_equals_plain_launches():
    s = _specs()[3]
    plain = _alone(s, 250)
    graphed = _alone(s, 250, graph=100)  # two replays + 50 plain steps
    assert first_divergence(plain, graphed) is None
