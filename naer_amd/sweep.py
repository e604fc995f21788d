"""Batched parameter sweep: many independent Brunel networks per launch.

The sweep workload: n_nets balanced networks with sizes log-uniform in
[n_min, n_max], in-degree 10% of N (80% from E), J_E = 0.1 mV, g uniform in
[3, 7], Poisson drive uniform in [10, 25] kHz x 0.1 mV, delay 1.5 ms, seed
base+i, every even-numbered network with pair STDP on E->E; each runs
`steps` steps of 0.1 ms.

Two ways to run it:
  packed      all static networks in one pack and the plastic networks in two
              more (a flush of pending potentiation runs over plastic packs
              only; a plastic network steps about 1.6x slower, so its work is
              split to end with the static pack's), every pack on its own
              stream, steps enqueued in graph-captured chunks (100 steps
              rounded up to whole batches: 105 at the sweep's 15-step
              batches), delivery schedules from the autotuner;
  sequential  the plain loop: one network per launch sequence, one after the
              other.
In deterministic mode every network's raster is bit-identical either way.

    python -m naer_amd.sweep --nets 256 --steps 5000 [--mode det|fast]
"""
import argparse
import json
import math
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _abi, models
from .pack import Pack

SWEEP_SEED = 0x5EED5EED


def sweep_specs(n_nets=256, n_min=1_000, n_max=50_000, seed=SWEEP_SEED):
    """Parameters of every network of the sweep (host-side draws, fixed seed)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_nets):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        specs.append(dict(index=i, n=n, indegree=max(10, int(round(0.1 * n))), g=float(rng.uniform(3.0, 7.0)),
                          drive_hz=float(rng.uniform(10_000.0, 25_000.0)), plastic=(i % 2 == 0),
                          seed=(seed * 1000003 + i) & 0xFFFFFFFFFFFFFFFF))
    return specs


def make_network(spec, mode=_abi.MODE_DETERMINISTIC, record_capacity=0):
    return models.balanced(spec["n"], spec["indegree"], 0.1, spec["g"], spec["drive_hz"], 0.1,
                           delay_ms=1.5, plastic=spec["plastic"], seed=spec["seed"], mode=mode,
                           record_capacity=record_capacity)


def expected_record(spec, steps, rate_bound_hz=500.0):
    """Record buffer large enough for a refractory-limited rate."""
    return int(spec["n"] * rate_bound_hz * steps * 1e-4) + 1024


def _progress(msg):
    print("# " + msg, file=sys.stderr, flush=True)


def pack_groups(specs, packs_per_kind=1, plastic_packs=None):
    """Networks per pack: static and plastic networks never share a pack (a
    flush of pending potentiation then runs over plastic packs only); packs of
    a kind take every k-th network by size, so they carry similar work.  Empty
    groups are dropped."""
    groups = []
    for plastic in (False, True):
        kind = sorted((s for s in specs if s["plastic"] == plastic), key=lambda s: s["n"])
        k = plastic_packs if (plastic and plastic_packs) else packs_per_kind
        groups += [kind[i::k] for i in range(k)]
    return [g for g in groups if g]


def run_packed(specs, steps, mode=_abi.MODE_DETERMINISTIC, chunk=100, record=True, graphs=True, verbose=False,
               autotune=True, packs_per_kind=1, plastic_packs=2):
    # each pack runs on its own stream
    groups = pack_groups(specs, packs_per_kind, plastic_packs)
    packs = []
    t_build = time.perf_counter()
    for g in groups:
        nets = [make_network(s, mode, expected_record(s, steps) if record else 0) for s in g]
        p = Pack(nets)
        p.build()
        if graphs:
            p.set_graph(chunk)
        packs.append((g, p))
        if verbose:
            _progress(f"packed: built {len(g)} networks ({time.perf_counter() - t_build:.1f} s)")
    for _, p in packs:
        p.sync()
    build_s = time.perf_counter() - t_build
    t0 = time.perf_counter()
    # the autotuner's trial batches are real steps of the run (results do not
    # depend on the schedule); each pack then runs the rest of its steps
    done = [0] * len(packs)
    schedules = []
    # the tuning trials take up to 2 passes x 19 schedules x 2 batches x 16 steps
    if autotune and steps >= 2000:
        # packs the tuner does not apply to (ring delivery) get all their
        # steps enqueued first; the others tune side by side, one host thread
        # each (the library's calls release the interpreter), so the device
        # stays busy with both
        tunable = [p.schedule()["bucketed"] == 1 for _, p in packs]
        for k, (_, p) in enumerate(packs):
            if not tunable[k]:
                p.run(steps)
                done[k] = steps

        def tune(p):
            p.run(min(steps // 10, 500))  # past the initial transient
            return p.autotune(2), int(p.stats().step)

        picked = [(k, p) for k, (_, p) in enumerate(packs) if tunable[k]]
        with ThreadPoolExecutor(max_workers=max(1, len(picked))) as ex:
            for (k, _), (sched, step) in zip(picked, ex.map(lambda kp: tune(kp[1]), picked)):
                schedules.append(sched)
                done[k] = step
                if verbose:
                    _progress(f"packed: schedule {sched}")
    # a captured chunk holds whole batches (naer_net_set_graph rounds 100
    # steps up to 105 at 15-step batches): stepping in that unit replays the
    # graph, and no run ends inside a batch (a plastic pack would deliver
    # such a partial batch through the ring's device atomics)
    unit = [-(-chunk // int(p.stats().batch)) * int(p.stats().batch) for _, p in packs]
    last = t0
    while min(done) < steps:
        for k, (_, p) in enumerate(packs):
            c = min(unit[k], steps - done[k])
            if c > 0:
                p.run(c)
                done[k] += c
        if verbose and time.perf_counter() - last > 10:
            for _, p in packs:
                p.sync()
            last = time.perf_counter()
            _progress(f"packed: {min(done)}/{steps} steps ({last - t0:.1f} s)")
    for _, p in packs:
        p.sync()
    run_s = time.perf_counter() - t0
    rasters = {}
    stats = []
    for g, p in packs:
        st = p.stats()
        stats.append(dict(nets=len(g), neurons=int(st.n_neurons), synapses=int(st.n_synapses),
                          spikes=int(st.spikes), events=int(st.events), batch=int(st.batch)))
        if record:
            for s, r in zip(g, p.spikes()):
                rasters[s["index"]] = r
        p.close()
    return dict(build_s=build_s, run_s=run_s, packs=stats, rasters=rasters, schedules=schedules)


def run_sequential(specs, steps, mode=_abi.MODE_DETERMINISTIC, record=True, verbose=False):
    rasters = {}
    build_s = run_s = 0.0
    events = 0
    for s in specs:
        t = time.perf_counter()
        net = make_network(s, mode, expected_record(s, steps) if record else 0)
        net.compile()
        net.finish()
        build_s += time.perf_counter() - t
        t = time.perf_counter()
        net.run(steps)
        net.finish()
        run_s += time.perf_counter() - t
        events += int(net.stats().events)
        if record:
            rasters[s["index"]] = net.spikes()
        net.close()
        if verbose and (s["index"] % 16 == 15):
            _progress(f"sequential: {s['index'] + 1}/{len(specs)} networks (run {run_s:.1f} s)")
    return dict(build_s=build_s, run_s=run_s, events=events, rasters=rasters)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nets", type=int, default=256)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--n-max", type=int, default=50_000)
    ap.add_argument("--mode", choices=["det", "fast"], default="det")
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--packs-per-kind", type=int, default=1)
    ap.add_argument("--plastic-packs", type=int, default=2)
    a = ap.parse_args()
    mode = _abi.MODE_DETERMINISTIC if a.mode == "det" else _abi.MODE_FAST
    specs = sweep_specs(a.nets, n_max=a.n_max)
    packed = run_packed(specs, a.steps, mode, graphs=not a.no_graphs, verbose=True, autotune=not a.no_autotune,
                        packs_per_kind=a.packs_per_kind, plastic_packs=a.plastic_packs)
    out = dict(nets=a.nets, steps=a.steps, mode=a.mode, packed_run_s=packed["run_s"],
               packed_build_s=packed["build_s"], packed_jobs_per_s=a.nets / packed["run_s"],
               packs=packed["packs"], schedules=packed["schedules"])
    if not a.no_sequential:
        seq = run_sequential(specs, a.steps, mode, verbose=True)
        ident = sum(int(np.array_equal(packed["rasters"][i], seq["rasters"][i])) for i in range(a.nets))
        out.update(seq_run_s=seq["run_s"], seq_build_s=seq["build_s"], seq_jobs_per_s=a.nets / seq["run_s"],
                   speedup=seq["run_s"] / packed["run_s"], rasters_identical=ident)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
