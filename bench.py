"""Benchmark of the simulation step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--n N] [--mode fast|det]

Workload: the 1M-neuron balanced LIF network (800k E / 200k I, in-degree 2,000,
2e9 synapses, delays uniform over whole ms in 1-8 ms, J_E = 0.05 mV, g = 5,
Poisson drive for ~5 Hz), fp32 state and weights, no plasticity.  The
connectivity is built on the device before the timed region; inputs are
resident in HBM.  A "step" is one dt = 0.1 ms update of all neurons plus the
scatter of their spikes; K steps are timed (default 10,000 = 1 s biological).

Prints one JSON line: synaptic events/s (value), realtime factor, the roofline
of the scatter kernel measured with HIP events on the net's stream, and the
CPU oracle timed on a bounded sample of the same network on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SCATTER_BYTES_PER_EVENT = 12  # col 4 + weight 4 + ring accumulate 4 (fast mode)


def cpu_baseline(args):
    """The oracle (clean-room C restatement, OpenMP) on a bounded sample: the
    same network parameters at N = args.cpu_n, args.cpu_steps steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from naer_amd import models

    threads = min(16, os.cpu_count() or 1)
    net = models.large_1m(n=args.cpu_n)
    o = O.OracleNet(net.description(), O.FIXED, threads=threads)
    o.run(200)  # leave the initial transient
    e0 = o.event_count
    t = time.perf_counter()
    o.run(args.cpu_steps)
    dt = time.perf_counter() - t
    ev = o.event_count - e0
    o.close()
    return {"value": ev / dt, "unit": "synaptic events/s", "cores": threads, "kind": "port",
            "sample": f"oracle (fp32 + fixed-point ring, OpenMP) on the same network at "
                      f"N={args.cpu_n} (in-degree 2000), {args.cpu_steps} steps after 200 "
                      f"warm-up steps, {dt:.1f} s",
            "rtf": args.cpu_steps * 1e-4 / dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10_000)
    ap.add_argument("--warmup", type=int, default=1_000)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--indegree", type=int, default=2000)
    ap.add_argument("--mode", choices=["fast", "det"], default="fast")
    ap.add_argument("--cpu-n", type=int, default=50_000)
    ap.add_argument("--cpu-steps", type=int, default=2_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the simulation step of one network is not "
              "split across GPUs in this build (--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    from naer_amd import _abi, models

    mode = _abi.MODE_FAST if args.mode == "fast" else _abi.MODE_DETERMINISTIC
    t = time.perf_counter()
    net = models.large_1m(n=args.n, indegree=args.indegree, mode=mode)
    net.compile()
    build_s = time.perf_counter() - t
    print(f"# built N={net.getNeurons()} S={net.getSynapses()} in {build_s:.1f} s", file=sys.stderr,
          flush=True)

    net.run(args.warmup)
    net.finish()
    s0 = net.stats()
    if not args.no_profile:
        net.set_profiling(True)
    t0 = time.perf_counter()
    net.run(args.steps)
    net.finish()
    wall = time.perf_counter() - t0
    s1 = net.stats()

    steps = args.steps
    events = s1.events - s0.events
    spikes = s1.spikes - s0.spikes
    dt_ms = net.dt_ms
    bio_s = steps * dt_ms * 1e-3
    value = events / wall
    rtf = bio_s / wall
    N = net.getNeurons()

    roofline = None
    kernels = {}
    if not args.no_profile:
        sc_ms = s1.kernel_ms[_abi.K_SCATTER] - s0.kernel_ms[_abi.K_SCATTER]
        sc_n = s1.kernel_calls[_abi.K_SCATTER] - s0.kernel_calls[_abi.K_SCATTER]
        up_ms = s1.kernel_ms[_abi.K_UPDATE] - s0.kernel_ms[_abi.K_UPDATE]
        up_n = s1.kernel_calls[_abi.K_UPDATE] - s0.kernel_calls[_abi.K_UPDATE]
        ring_b = 8 if mode == _abi.MODE_FAST else 16
        B = s1.batch
        upd_bytes = N * steps * ring_b + N * (steps / B) * 24  # ring r+w per step, V/ref/last r+w per launch
        sc_bytes = events * (SCATTER_BYTES_PER_EVENT + (4 if mode == _abi.MODE_DETERMINISTIC else 0))
        ach = sc_bytes / (sc_ms * 1e-3) / 1e9 if sc_ms > 0 else 0.0
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_scatter", "avg_us": round(sc_ms * 1e3 / max(sc_n, 1), 2),
                    "bytes_per_launch": round(sc_bytes / max(sc_n, 1))}
        kernels = {
            "update": {"avg_us": round(up_ms * 1e3 / max(up_n, 1), 2), "launches": int(up_n),
                       "ms": round(up_ms, 2),
                       "gbs": round(upd_bytes / (up_ms * 1e-3) / 1e9, 1) if up_ms > 0 else None},
            "scatter": {"avg_us": round(sc_ms * 1e3 / max(sc_n, 1), 2), "launches": int(sc_n),
                        "ms": round(sc_ms, 2)},
        }

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": str(e)}

    out = {
        "metric": "synaptic events/s & realtime factor, 1M LIF net",
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
        "config": {"workload": "large_1m: balanced LIF, N=1,000,000 (80/20 E/I), in-degree 2000, "
                               "2e9 synapses, delays 1-8 ms, J_E=0.05 mV, g=5, 1 s at dt=0.1 ms",
                   "n_neurons": N, "n_synapses": int(net.getSynapses()), "mode": args.mode,
                   "dt_ms": dt_ms},
        "realtime_factor": round(rtf, 3),
        "rate_hz": round(spikes / N / bio_s, 3),
        "events": int(events),
        "build_s": round(build_s, 2),
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
