"""Serving: a request loop with continuous batching of simulation jobs.

Jobs arrive as a Poisson process.  Each job is a balanced LIF network
(N log-uniform in [2k, 200k], in-degree uniform in [500, 2000], duration
uniform in [0.2, 2] s of simulated time, 30% with pair STDP on E->E; J_E =
0.1 mV, g = 5, 20 kHz x 0.1 mV drive, delay 1.5 ms).

The server keeps two arenas (static jobs, plastic jobs) whose capacities split
the HBM budget, each on its own stream, and services each arena as soon as its
own chunk is done (naer_net_idle): it drains the recorded spikes, runs the
watchdog, retires the jobs that have reached their duration, admits waiting
jobs that fit (first come first served, later jobs may pass a head that does
not fit yet), and enqueues the next chunk of `chunk` steps (rounded to whole
batches of the arena) of every resident job, the newly admitted ones
included.  The two arenas thus step independently and their work overlaps on
the device.  A job's raster depends only
on its own description: alone or in any mix it is the same, bit for bit in
deterministic mode.

Failure isolation: a job that fails is retired or rejected on its own and
the others continue unchanged.  Every job ends with a status: "done";
"rejected" (it can never fit an arena, or its admission failed, e.g. an
allocation error; the message is kept); "cancelled" (the client cancelled
it, `cancel_at`); "timeout" (not done `deadline_s` after its arrival); or
"failed" (the watchdog found a potential that is not finite after a chunk).
(The reference's convention is to report and carry on, src/main.cpp:106-110;
a server needs it per job.)

    python -m naer_amd.serving --rates 2,5,10,20 --jobs 48
"""
import argparse
import collections
import json
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi, models

DELAY_STEPS = 15       # 1.5 ms at dt = 0.1 ms
MAX_ROW = 8192
BYTES_PER_SYNAPSE = 8  # uint32 target/delay word + fp32 weight
REFRACTORY_STEPS = int(round(models.BRUNEL_LIF["t_ref_ms"] / 0.1))  # of every job's LIF populations


@dataclass
class JobSpec:
    index: int
    arrival_s: float
    n: int
    indegree: int
    duration_s: float
    plastic: bool
    seed: int

    def steps(self, dt_ms=0.1):
        return int(round(self.duration_s * 1000.0 / dt_ms))

    def n_synapses(self):
        return self.n * self.indegree

    def network(self, mode=_abi.MODE_DETERMINISTIC, record_capacity=0):
        return models.balanced(self.n, self.indegree, 0.1, 5.0, 20_000.0, 0.1, delay_ms=1.5,
                               plastic=self.plastic, seed=self.seed, mode=mode,
                               record_capacity=record_capacity)


def served_jobs(n_jobs, rate_per_s, seed=0x5E11, n_range=(2_000, 200_000), k_range=(500, 2000),
                dur_range=(0.2, 2.0), stdp_fraction=0.3):
    """Job stream with Poisson arrivals at rate_per_s jobs per second."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        t += rng.exponential(1.0 / rate_per_s) if rate_per_s > 0 else 0.0
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        jobs.append(JobSpec(index=i, arrival_s=t, n=n, indegree=int(rng.integers(k_range[0], k_range[1] + 1)),
                            duration_s=float(rng.uniform(*dur_range)), plastic=bool(rng.random() < stdp_fraction),
                            seed=(seed * 7919 + i) & 0xFFFFFFFFFFFFFFFF))
    return jobs


def neuron_bytes(mode, plastic):
    """Device bytes per neuron of arena capacity: ring (delay+1 slots), state,
    row pointers, build scratch and two spike-list segments (B entries each)."""
    ring = (DELAY_STEPS + 1) * (8 if mode == _abi.MODE_DETERMINISTIC else 4)
    lists = 2 * DELAY_STEPS * 16  # B = 15 steps in both arenas
    state = 4 * 5 + 8 + 4 + 8 + 16  # V ref last pnext pcnt, row_ptr, deg, cursor, counters
    if plastic:
        state += 16 + 32 + 4 * 32  # STDP: pre/post records of the newest spikes, ring of 32 spike steps
    return ring + lists + state


def arena_neurons(job):
    """Neurons a job takes in its arena: its own, one more (the end of its
    last row), rounded up to the arena's 256-neuron alignment."""
    return (job.n + 1 + 255) // 256 * 256


def footprint(job, mode=_abi.MODE_DETERMINISTIC):
    """Device bytes a job occupies in its arena."""
    nb = neuron_bytes(mode, job.plastic)
    return arena_neurons(job) * nb + job.n_synapses() * BYTES_PER_SYNAPSE


@dataclass
class Capacity:
    neurons: int
    synapses: int


def split_budget(budget_bytes, mode=_abi.MODE_DETERMINISTIC, stdp_fraction=0.3, synapses_per_neuron=700):
    """Capacities of the static and the plastic arena for a total budget,
    shared in proportion to the expected footprint of the job mix."""
    per_syn = [BYTES_PER_SYNAPSE + neuron_bytes(mode, plastic) / synapses_per_neuron for plastic in (False, True)]
    w = [(1.0 - stdp_fraction) * per_syn[0], stdp_fraction * per_syn[1]]
    caps = []
    for k in (0, 1):
        syn = int(budget_bytes * w[k] / sum(w) / per_syn[k])
        caps.append(Capacity(neurons=max(256, syn // synapses_per_neuron), synapses=syn))
    return caps[0], caps[1]


@dataclass
class JobResult:
    index: int
    arrival_s: float
    admitted_s: float = -1.0
    done_s: float = -1.0
    slot: int = -1
    arena: int = -1
    spikes: int = 0
    raster: list = field(default_factory=list)
    status: str = "waiting"  # done, rejected, cancelled, timeout, failed
    error: str = ""

    @property
    def latency_s(self):
        return self.done_s - self.arrival_s


class Server:
    """Continuous-batching request loop over two arenas (static, plastic).

    arena_factory(capacity, plastic) builds an arena; the default is the device
    Arena, tests pass a host stand-in to exercise admission logic without a GPU.
    """

    def __init__(self, budget_bytes=200e9, mode=_abi.MODE_DETERMINISTIC, chunk=100, graphs=True,
                 keep_rasters=False, record_per_chunk=1 << 22, arena_factory=None, dt_ms=0.1,
                 synapses_per_neuron=700, stdp_fraction=0.3):
        self.mode = mode
        self.chunk = chunk
        self.dt_ms = dt_ms
        self.keep = keep_rasters
        # the two arenas share the budget as the expected job mix does
        static_cap, plastic_cap = split_budget(budget_bytes, mode, stdp_fraction=stdp_fraction,
                                               synapses_per_neuron=synapses_per_neuron)
        if arena_factory is None:
            from .arena import Arena

            def arena_factory(cap, plastic):
                return Arena(cap.neurons, cap.synapses, dt_ms=dt_ms, mode=mode,
                             delay_steps=(DELAY_STEPS, DELAY_STEPS), max_row=MAX_ROW, plastic=plastic,
                             stdp=models.PAIR_STDP if plastic else None,
                             refractory_min=REFRACTORY_STEPS if plastic else 0,
                             record_capacity=record_per_chunk)
        self.arenas = [arena_factory(static_cap, False), arena_factory(plastic_cap, True)]
        self.capacity = (static_cap, plastic_cap)
        # chunks of whole batches of every arena: an arena delivers a partial
        # batch through the delay ring's device atomics (it admits and retires
        # jobs between chunks, so it does not defer), several times slower
        # than its buckets
        batch = 1
        for a in self.arenas:
            if hasattr(a, "stats"):
                batch = math.lcm(batch, max(1, int(a.stats().batch)))
        self.chunk = max(batch, int(round(chunk / batch)) * batch)
        chunk = self.chunk
        if graphs:
            for a in self.arenas:
                if hasattr(a, "set_graph"):
                    a.set_graph(chunk)

    def close(self):
        for a in self.arenas:
            a.close()

    def serve(self, jobs, clock=time.perf_counter, sleep=time.sleep, cancel_at=None, deadline_s=None,
              watchdog=True):
        """Serve the job stream (arrival times relative to the start) and
        return per-job results and the summary.  cancel_at: {job index:
        time} at which the client cancels a job; deadline_s: the longest a
        job may take from its arrival; watchdog: after every chunk, retire
        the jobs whose state is no longer finite."""
        jobs = sorted(jobs, key=lambda j: j.arrival_s)
        res = {j.index: JobResult(j.index, j.arrival_s) for j in jobs}
        cancel_at = dict(cancel_at or {})
        pending = collections.deque(jobs)
        waiting = []
        active = [dict(), dict()]  # per arena: slot -> (job, steps done)
        rejected = []
        t0 = clock()

        def finish(job, status, error=""):
            r = res[job.index]
            r.status, r.error, r.done_s = status, error, clock() - t0
            if status == "rejected":
                rejected.append(job.index)

        def expired(job, now):
            if job.index in cancel_at and now >= cancel_at[job.index]:
                return "cancelled"
            if deadline_s is not None and now - job.arrival_s >= deadline_s:
                return "timeout"
            return None

        # per arena: the slots that run in its enqueued chunk (None: no chunk
        # in flight).  Each arena is serviced as soon as its own chunk is done
        # (naer_net_idle), so the two step independently on the device.
        in_flight = [None, None]

        def admit(ai, now):
            """Waiting jobs of arena ai, first come first served (a job that
            does not fit yet waits while later ones may pass it); the arena
            has no chunk in flight, so an admission does not wait for one."""
            nonlocal waiting
            still = []
            for job in waiting:
                why = expired(job, now)
                if why:
                    finish(job, why)
                    continue
                if (1 if job.plastic else 0) != ai:
                    still.append(job)
                    continue
                cap = self.capacity[ai]
                if arena_neurons(job) > cap.neurons or job.n_synapses() > cap.synapses or \
                        job.indegree * 4 > MAX_ROW:
                    finish(job, "rejected", "larger than the arena")  # can never fit
                    continue
                try:
                    slot = self.arenas[ai].admit(job.network(self.mode))
                except _abi.NaerError as e:  # never admissible, or the build failed
                    finish(job, "rejected", str(e))
                    continue
                if slot is None:
                    still.append(job)
                    continue
                active[ai][slot] = [job, 0]
                r = res[job.index]
                r.admitted_s, r.slot, r.arena, r.status = clock() - t0, slot, ai, "running"
            waiting = still

        def collect(ai):
            """Arena ai's chunk is done: its spikes, the watchdog, the jobs
            that reached their duration or were cancelled retire."""
            arena = self.arenas[ai]
            ran = in_flight[ai]
            in_flight[ai] = None
            drained = arena.drain()
            bad = set(arena.check()) if watchdog and hasattr(arena, "check") else set()
            now = clock() - t0
            done = []
            for slot in ran:
                job, steps = active[ai][slot]
                total = job.steps(self.dt_ms)
                sp = drained.get(slot)
                if sp is not None and len(sp):
                    sp = sp[sp["step"] < total]
                    res[job.index].spikes += len(sp)
                    if self.keep:
                        res[job.index].raster.append(sp)
                active[ai][slot][1] = steps + self.chunk
                if slot in bad:
                    done.append((slot, "failed", "a potential is not finite"))
                elif steps + self.chunk >= total:
                    done.append((slot, "done", ""))
                else:
                    why = expired(job, now)
                    if why:
                        done.append((slot, why, ""))
            for slot, status, error in done:
                job = active[ai].pop(slot)[0]
                arena.retire(slot)
                finish(job, status, error)

        while pending or waiting or any(active):
            now = clock() - t0
            while pending and pending[0].arrival_s <= now:
                waiting.append(pending.popleft())
            serviced = False
            for ai, arena in enumerate(self.arenas):
                if in_flight[ai] is not None:
                    if hasattr(arena, "idle") and not arena.idle():
                        continue
                    collect(ai)
                    serviced = True
                admit(ai, clock() - t0)
                if active[ai]:
                    # the next chunk of every resident job, admitted ones included
                    arena.run(self.chunk)
                    in_flight[ai] = set(active[ai])
                    serviced = True
            if serviced:
                continue
            if any(f is not None for f in in_flight):
                sleep(0.0002)  # a chunk is running: poll again shortly
            elif waiting and not pending:
                # the arenas are empty and still cannot take them
                for job in waiting:
                    finish(job, "rejected", "does not fit an empty arena")
                waiting = []
            elif pending:
                sleep(max(0.0, min(0.05, pending[0].arrival_s - (clock() - t0))))
            elif waiting:
                sleep(0.001)
        makespan = clock() - t0
        finished = [r for r in res.values() if r.status == "done"]
        lat = np.array([r.latency_s for r in finished]) if finished else np.zeros(1)
        for r in finished:
            if self.keep:
                r.raster = (_abi.sort_spikes(np.concatenate(r.raster))
                            if r.raster else np.zeros(0, dtype=_abi.SPIKE_DTYPE))
        statuses = collections.Counter(r.status for r in res.values())
        return dict(results=res, rejected=rejected, makespan_s=makespan, completed=len(finished),
                    statuses=dict(statuses),
                    jobs_per_s=len(finished) / makespan if makespan > 0 else 0.0,
                    p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                    mean_s=float(lat.mean()))


def run_alone(job, mode=_abi.MODE_DETERMINISTIC, dt_ms=0.1):
    """The job simulated as a single network (reference for mix-vs-alone)."""
    steps = job.steps(dt_ms)
    net = job.network(mode, record_capacity=int(job.n * steps * 0.05) + 4096)
    net.compile()
    net.run(steps)
    sp = net.spikes()
    net.close()
    return sp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="2,5,10,20")
    ap.add_argument("--jobs", type=int, default=48)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--mode", choices=["det", "fast"], default="det")
    ap.add_argument("--chunk", type=int, default=100)
    a = ap.parse_args()
    mode = _abi.MODE_DETERMINISTIC if a.mode == "det" else _abi.MODE_FAST
    srv = Server(a.budget_gb * 1e9, mode=mode, chunk=a.chunk)
    out = []
    for rate in [float(x) for x in a.rates.split(",")]:
        r = srv.serve(served_jobs(a.jobs, rate))
        out.append(dict(rate=rate, jobs=a.jobs, completed=r["completed"], jobs_per_s=r["jobs_per_s"],
                        p50_s=r["p50_s"], p99_s=r["p99_s"], makespan_s=r["makespan_s"],
                        rejected=len(r["rejected"])))
        print(json.dumps(out[-1]), flush=True)
    srv.close()


if __name__ == "__main__":
    main()
