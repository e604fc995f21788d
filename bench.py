"""Benchmark of the simulation step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--n N] [--mode fast|det] [--workload W]

Workload (default): the 1M-neuron balanced LIF network (800k E / 200k I, in-degree 2,000,
2e9 synapses, delays uniform over whole ms in 1-8 ms, J_E = 0.05 mV, g = 5,
Poisson drive for ~5 Hz), fp32 state and weights, no plasticity.  The
connectivity is built on the device before the timed region; inputs are
resident in HBM.

A "step" is one aligned batch of B steps of dt = 0.1 ms (B = 9 on the 1M
network: 0.9 ms of biological time): one update launch over all neurons for
the B steps and the delivery of the batch's spikes.  W batches of warm-up
(with --autotune the delivery schedule is tuned on real steps inside it) are followed by
exactly K timed batches, so the timed region starts and ends on batch
boundaries and holds K updates and K deliveries.

Prints one JSON line: synaptic events/s (value), realtime factor, the roofline
of the spike delivery (the kernels that carry a batch's spikes to their
targets) and of the whole batch, measured with HIP events on the net's stream
over whole batches, per-kernel times, a reduced batched sweep (packed against
one network per launch sequence, rasters compared), the served stream at one
arrival rate (p50/p99 latency), and the CPU oracle timed on a bounded sample
of the same network on all host cores.

    python bench.py --copy-probe   # a plain device copy of known size, the
                                   # calibration run for rocprofv3 --pmc counters
    python bench.py --workload plastic_100k   # the plasticity network: N = 1e5,
                                   # in-degree 1000 (1e8 synapses), pair STDP on E->E
    python bench.py --workload balanced_10k   # the small network: N = 1e4, in-degree 1000
    python bench.py --workload brunel_event   # the reference's event-driven step; a step
                                   # is one 1500 us window
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes: a static synaptic event = col 4 + weight 4 + accumulate 4;
# a neuron-step = V r/w 8 + refractory r/w 8 + input slot read+clear 8
BYTES_PER_EVENT = 12
BYTES_PER_NEURON_STEP = 24
PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_large_1m.md")
COPY_PROBE_BYTES = 512 << 20


def measured_traffic():
    """HBM bytes per delivered batch of the delivery kernels, from the
    calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE summary (None if absent)."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_batch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def copy_probe():
    """A plain copy of COPY_PROBE_BYTES on the device, five times: what
    FETCH_SIZE / WRITE_SIZE read for it calibrates the counters."""
    import torch
    src = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES, "launches": 5}), flush=True)


WORKLOADS = {
    "large_1m": ("balanced LIF, N=1,000,000 (80/20 E/I), in-degree 2000, 2e9 synapses, delays 1-8 ms, "
                 "J_E=0.05 mV, g=5"),
    "plastic_100k": ("balanced LIF, N=100,000, in-degree 1000 (1e8 synapses), delays 1-5 ms, pair STDP on E->E "
                     "(w0 uniform in [0, 0.3 mV])"),
    "balanced_10k": "balanced LIF, N=10,000, in-degree 1000, J_E=0.1 mV, J_I=-0.5 mV, delay 1.5 ms",
    "brunel_event": ("the reference's brunel model (src/main.cpp:32-50) at 1e8 synapses (N=44,720) in the "
                     "event-driven mode: the reference's 1500 us epochs and LIF step"),
}


WORKLOAD_N = {"large_1m": 1_000_000, "plastic_100k": 100_000, "balanced_10k": 10_000, "brunel_event": 44_720}
EVENT_WINDOW_US = 1500


def bench_event(args):
    """brunel in the event-driven mode: a step is one 1500 us window; the
    oracle's event-driven restatement on brunel at 1e6 synapses is the CPU
    baseline."""
    from naer_amd import _abi, models
    synapses = args.n or 100_000_000
    t = time.perf_counter()
    net = models.brunel(synapses, mode=_abi.MODE_EVENT)
    net.compile()
    build_s = time.perf_counter() - t
    windows = max(1, args.steps)
    net.run(max(1, args.warmup))
    net.finish()
    s0 = net.stats()
    t0 = time.perf_counter()
    net.run(windows)
    net.finish()
    wall = time.perf_counter() - t0
    s1 = net.stats()
    events = s1.events - s0.events
    bio_s = windows * EVENT_WINDOW_US * 1e-6
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        small = models.brunel(1_000_000, mode=_abi.MODE_EVENT)
        o = O.OracleNet(small.description(), O.FIXED, threads=1)
        o.run_events(20)
        e0 = o.event_count
        tc = time.perf_counter()
        o.run_events(200)
        dtc = time.perf_counter() - tc
        cpu = {"value": (o.event_count - e0) / dtc, "unit": "synaptic events/s", "cores": 1, "kind": "port",
               "sample": f"oracle event-driven restatement (single thread) on brunel at 1e6 synapses "
                         f"(N={o.N}), 200 windows after 20, {dtc:.1f} s"}
        o.close()
    out = {
        "metric": "synaptic events/s & realtime factor, brunel_event",
        "value": round(events / wall, 1), "unit": "synaptic events/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / windows, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 decay products)",
        "data": "synthetic (device-built Bernoulli connectivity and the reference's Poisson trains)",
        "config": {"workload": f"brunel_event: {WORKLOADS['brunel_event']}, {windows} windows = {bio_s:g} s",
                   "step": "one 1500 us window (the reference's Network::step)",
                   "n_neurons": net.getNeurons(), "n_synapses": int(net.getSynapses()), "mode": "event"},
        "realtime_factor": round(bio_s / wall, 3),
        "rate_hz": round((s1.spikes - s0.spikes) / net.getNeurons() / bio_s, 3),
        "events": int(events), "build_s": round(build_s, 2),
        "roofline": None,
        "roofline_note": "the event-driven step kernel is bound by its in-order per-target arithmetic; see "
                         "profiles/kernel_stats_event_brunel_1e8.md",
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def make_workload(name, n=None, indegree=None, **kw):
    from naer_amd import models
    size = {} if n is None else {"n": n}
    if name == "large_1m":
        return models.large_1m(**size, indegree=indegree or 2000, **kw)
    if name == "plastic_100k":
        return models.plastic_100k(**size, indegree=indegree or 1000, **kw)
    return models.balanced_10k(**size, **kw)


def host_cores():
    """Cores this process may run on (the box's share), and the CPU model."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_baseline(args):
    """The oracle (clean-room C restatement, OpenMP on every host core) on a
    bounded sample: the same network parameters at N = args.cpu_n, stepped in
    chunks of 1000 steps until args.cpu_seconds have passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads, model = host_cores()
    net = make_workload(args.workload, n=min(args.cpu_n, args.n or WORKLOAD_N[args.workload]),
                        indegree=args.indegree if args.workload == "large_1m" else None)
    o = O.OracleNet(net.description(), O.FIXED, threads=threads)
    o.run(200)  # leave the initial transient
    e0 = o.event_count
    t = time.perf_counter()
    steps = 0
    while steps < 100_000 and (time.perf_counter() - t < args.cpu_seconds or steps == 0):
        o.run(1000)
        steps += 1000
    dt = time.perf_counter() - t
    ev = o.event_count - e0
    o.close()
    return {"value": ev / dt, "unit": "synaptic events/s", "cores": threads, "kind": "port",
            "cpu": model,
            "sample": f"oracle (fp32 + fixed-point ring, OpenMP, {threads} threads) on the same network at "
                      f"N={net.getNeurons()}, {steps} steps after 200 warm-up steps, {dt:.1f} s",
            "rtf": steps * 1e-4 / dt}


def sweep_sample(args):
    """The batched sweep, args.sweep_nets networks x args.sweep_steps steps
    (default: the whole sweep, 256 networks x 0.5 s) in deterministic mode:
    packed (graph-captured chunks, one stream per pack, schedules from the
    autotuner on the run's own steps) against the plain
    one-network-per-launch-sequence loop, jobs/s of each (build excluded) and
    the count of bit-identical rasters."""
    from naer_amd import sweep
    specs = sweep.sweep_specs(args.sweep_nets)
    packed = sweep.run_packed(specs, args.sweep_steps)
    seq = sweep.run_sequential(specs, args.sweep_steps)
    ident = sum(int(np.array_equal(packed["rasters"][s["index"]], seq["rasters"][s["index"]])) for s in specs)
    return {"nets": len(specs), "steps": args.sweep_steps, "mode": "det",
            "packed_jobs_per_s": round(len(specs) / packed["run_s"], 2),
            "sequential_jobs_per_s": round(len(specs) / seq["run_s"], 2),
            "speedup": round(seq["run_s"] / packed["run_s"], 3), "rasters_identical": ident,
            "schedules": packed["schedules"],
            "sample": f"the sweep's networks (N log-uniform 1k-50k, K = 0.1 N, g in [3, 7], drive 10-25 kHz, "
                      f"half with STDP), {len(specs)} networks x {args.sweep_steps} steps"}


def serving_sample(args):
    """The served stream at one arrival rate: args.serve_jobs jobs at
    args.serve_rate jobs/s through the continuous-batching server (200 GB
    budget, deterministic mode)."""
    from naer_amd import serving
    srv = serving.Server(200e9)
    try:
        r = srv.serve(serving.served_jobs(args.serve_jobs, args.serve_rate))
    finally:
        srv.close()
    return {"rate_jobs_per_s": args.serve_rate, "jobs": args.serve_jobs, "completed": r["completed"],
            "jobs_per_s": round(r["jobs_per_s"], 3), "p50_s": round(r["p50_s"], 3), "p99_s": round(r["p99_s"], 3),
            "makespan_s": round(r["makespan_s"], 2), "rejected": len(r["rejected"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1111, help="timed batches (1111 x 0.9 ms = 1 s on the 1M network)")
    ap.add_argument("--warmup", type=int, default=120, help="warm-up batches (with --autotune, the autotuner runs inside)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="large_1m")
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="biological time run before the warm-up to leave the initial transient")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--indegree", type=int, default=None)
    ap.add_argument("--mode", choices=["fast", "det"], default="fast")
    ap.add_argument("--cpu-n", type=int, default=50_000)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=200, help="batches of the profiled stretch (>= 20)")
    ap.add_argument("--reserve", type=int, choices=[0, 1], default=None,
                    help="delivery placement: 1 reserved (no counting passes), 0 counted two-level (default: the "
                         "library's for the network)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner pick the delivery schedule during the warm-up (default: the "
                         "library's schedule for the network)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--sweep-nets", type=int, default=256)
    ap.add_argument("--sweep-steps", type=int, default=5000)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=10.0)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--graph-steps", type=int, default=0,
                    help="batches per HIP-graph-captured chunk of the timed run (0: plain launches)")
    args = ap.parse_args()
    t_start = time.perf_counter()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the simulation step of one network is not "
              "split across GPUs in this build (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.workload == "brunel_event":
        bench_event(args)
        return

    from naer_amd import _abi

    mode = _abi.MODE_FAST if args.mode == "fast" else _abi.MODE_DETERMINISTIC
    t = time.perf_counter()
    net = make_workload(args.workload, n=args.n, indegree=args.indegree, mode=mode)
    net.compile()
    build_s = time.perf_counter() - t
    B = int(net.stats().batch)
    print(f"# built N={net.getNeurons()} S={net.getSynapses()} in {build_s:.1f} s, batches of {B} steps",
          file=sys.stderr, flush=True)

    # the network first leaves its initial transient (V0 uniform below
    # threshold: the rate starts near 2 Hz and settles at ~5 Hz within
    # ~50 ms), so that any --warmup/--steps measure the steady state; then
    # the warm-up: half of it, (--autotune) the autotuner picking the
    # delivery schedule on real (aligned) batches, the rest; it ends on a
    # batch boundary
    settle = -(-int(round(args.settle_ms / net.dt_ms)) // B) * B
    net.run(settle)
    net.run(B * (args.warmup // 2))
    if args.reserve is not None:
        s = net.schedule()
        parts = min(4, s["parts_max"]) if args.reserve else s["refine_split"] or 1
        net.set_schedule(s["task_synapses"], s["segments_per_block"], parts, reserve=args.reserve)
    schedule = net.autotune(2) if args.warmup > 0 and args.autotune else net.schedule()
    done = int(net.stats().step)
    target = max(B * args.warmup, -(-done // B) * B)
    if target > done:
        net.run(target - done)
    net.finish()
    if args.graph_steps:
        # the timed batches replay captured chunks; the capture happens on the
        # first chunk, so run one before timing
        net.set_graph(B * args.graph_steps)
        net.run(B * args.graph_steps)
        net.finish()
    s0 = net.stats()
    warmup_run = int(s0.step)
    assert warmup_run % B == 0
    t0 = time.perf_counter()
    net.run(B * args.steps)
    net.finish()
    wall = time.perf_counter() - t0
    s1 = net.stats()

    steps = args.steps
    events = s1.events - s0.events
    spikes = s1.spikes - s0.spikes
    dt_ms = net.dt_ms
    bio_s = steps * B * dt_ms * 1e-3
    value = events / wall
    rtf = bio_s / wall
    N = net.getNeurons()

    roofline = None
    roofline_batch = None
    kernels = {}
    if not args.no_profile:
        # kernel times: a further profiled stretch of whole batches (HIP
        # events around every launch on the net's stream), kept out of the
        # timed region above
        prof = max(20, min(steps, args.profile_steps))
        net.set_graph(0)  # per-launch HIP events need plain launches
        net.set_profiling(True)
        p0 = net.stats()
        net.run(B * prof)
        net.finish()
        p1 = net.stats()
        net.set_profiling(False)
        prof_events = p1.events - p0.events

        def part(k):
            ms = p1.kernel_ms[k] - p0.kernel_ms[k]
            n = p1.kernel_calls[k] - p0.kernel_calls[k]
            return ms, n

        up_ms, up_n = part(_abi.K_UPDATE)
        dl_ms, dl_n = part(_abi.K_SCATTER)  # delivered batches, all delivery kernels
        kernels["k_update"] = {"avg_us": round(up_ms * 1e3 / max(up_n, 1), 2), "launches": int(up_n),
                               "alg_bytes_per_launch": N * B * BYTES_PER_NEURON_STEP}
        names = {_abi.K_BUCKET_HIST: "k_bucket_hist", _abi.K_BUCKET_SCAN: "k_bucket_colscan+k_bucket_starts",
                 _abi.K_BUCKET_PLACE: "k_bucket_place",
                 _abi.K_BUCKET_REFINE: "k_bucket_refine_res" if schedule.get("reserve") else "k_bucket_refine"}
        bucketed = False
        for k, name in names.items():
            ms, n = part(k)
            if n:
                bucketed = True
                kernels[name] = {"avg_us": round(ms * 1e3 / n, 2), "launches": int(n)}
        if not bucketed:
            kernels["k_scatter"] = {"avg_us": round(dl_ms * 1e3 / max(dl_n, 1), 2), "launches": int(dl_n)}
        dl_bytes = prof_events * BYTES_PER_EVENT / max(dl_n, 1)  # per delivered batch
        dl_us = dl_ms * 1e3 / max(dl_n, 1)
        ach = dl_bytes / (dl_us * 1e-6) / 1e9 if dl_us > 0 else 0.0
        traffic = measured_traffic() if args.workload == "large_1m" and args.n is None else None
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": traffic,
                    "kernel": ("spike delivery: " + " + ".join(k for k in kernels if k != "k_update")
                               .replace("+k_bucket_starts", " + k_bucket_starts")) if bucketed else "k_scatter",
                    "avg_us": round(dl_us, 2), "bytes_per_launch": round(dl_bytes),
                    "events_per_launch": round(prof_events / max(dl_n, 1)),
                    "profiled_batches": int(dl_n),
                    "unit_of_work": f"{BYTES_PER_EVENT} B per synaptic event, one launch = one aligned batch of "
                                    f"{B} steps"}
        # the whole batch: update + delivery, per-step algorithmic bytes of
        # events and neuron-steps over the summed kernel time
        b_bytes = dl_bytes + N * B * BYTES_PER_NEURON_STEP
        b_us = dl_us + up_ms * 1e3 / max(up_n, 1)
        roofline_batch = {"achieved": round(b_bytes / (b_us * 1e-6) / 1e9, 1), "unit": "GB/s",
                          "frac": round(b_bytes / (b_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                          "avg_us": round(b_us, 2), "bytes_per_batch": round(b_bytes),
                          "unit_of_work": f"{BYTES_PER_EVENT} B per event + {BYTES_PER_NEURON_STEP} B per "
                                          f"neuron-step, summed kernel time of one batch"}

    schedule_out = dict(schedule)
    net.close()  # the samples below need the device memory

    sweep = serving = None
    if args.workload == "large_1m" and not args.no_sweep:
        print(f"# sweep sample ({time.perf_counter() - t_start:.0f} s)", file=sys.stderr, flush=True)
        try:
            sweep = sweep_sample(args)
        except Exception as e:  # a sample, reported beside the headline
            sweep = {"error": str(e)}
    if args.workload == "large_1m" and not args.no_serving:
        print(f"# serving sample ({time.perf_counter() - t_start:.0f} s)", file=sys.stderr, flush=True)
        try:
            serving = serving_sample(args)
        except Exception as e:
            serving = {"error": str(e)}

    cpu = None
    if not args.no_cpu:
        print(f"# cpu baseline ({time.perf_counter() - t_start:.0f} s)", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline(args)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": str(e)}

    out = {
        "metric": "synaptic events/s & realtime factor, 1M LIF net; sweep jobs/s + p99 latency"
                  if args.workload == "large_1m" else f"synaptic events/s & realtime factor, {args.workload}",
        "value": round(value, 1),
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-built connectivity and Poisson drive from fixed seeds)",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}, {bio_s:g} s at dt={dt_ms:g} ms",
                   "step": f"one aligned batch of {B} steps of {dt_ms:g} ms (update + delivery)",
                   "state": f"steady state: timed after {settle * dt_ms:g} ms of settling and {args.warmup} "
                            f"warm-up batches",
                   "batch_steps": B, "n_neurons": N, "n_synapses": int(s1.n_synapses), "mode": args.mode,
                   "dt_ms": dt_ms},
        "realtime_factor": round(rtf, 3),
        "rate_hz": round(spikes / N / bio_s, 3),
        "events": int(events),
        "build_s": round(build_s, 2),
        "schedule": schedule_out,
        "warmup_steps_run": warmup_run,
        "settle_steps": settle,
        "graph_steps": args.graph_steps,
        "roofline": roofline,
        "roofline_batch": roofline_batch,
        "kernels": kernels,
        "sweep": sweep,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
