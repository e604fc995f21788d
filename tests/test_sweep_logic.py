"""Host logic of the batched sweep (naer_amd/sweep.py): the workload's
parameters and the split of its networks into packs (no GPU)."""
from naer_amd import sweep


def test_sweep_specs_are_fixed_and_in_range():
    a = sweep.sweep_specs(64)
    b = sweep.sweep_specs(64)
    assert a == b
    assert all(1_000 <= s["n"] <= 50_000 for s in a)
    assert sum(s["plastic"] for s in a) == 32  # every even-numbered network
    assert [s["index"] for s in a] == list(range(64))


def test_pack_groups_partition_by_kind_and_balance_sizes():
    specs = sweep.sweep_specs(256)
    for ppk, pp in ((1, None), (1, 2), (2, 2), (2, 3)):
        groups = sweep.pack_groups(specs, ppk, pp)
        assert len(groups) == ppk + (pp or ppk)
        seen = sorted(s["index"] for g in groups for s in g)
        assert seen == list(range(256))  # every network in exactly one pack
        for g in groups:
            assert len({s["plastic"] for s in g}) == 1  # kinds never share a pack
        for plastic in (False, True):
            sizes = [sum(s["n"] for s in g) for g in groups if g[0]["plastic"] == plastic]
            assert max(sizes) <= 1.25 * min(sizes)  # every k-th by size: similar work


def test_pack_groups_drop_empty_packs():
    specs = [s for s in sweep.sweep_specs(8) if not s["plastic"]]
    groups = sweep.pack_groups(specs, 1, 2)
    assert len(groups) == 1 and len(groups[0]) == 4
