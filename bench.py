"""Throughput of the slime-mold step on one MI355X.

Default workload (the plain-FP32-at-scale case): 4096x4096 f32 trail, 2^24
agents initialised from seed 1, the presets Default -> Sponge -> Firecracker
Trees run back to back (the K timed steps split in three), steps enqueued as
captured hipGraphs of 100 steps.  A step is one agent pass plus one fused
deposit + decay + diffuse pass over the whole grid.

Prints one JSON line: value = agent-updates/s over the timed region, plus the
roofline of the dominant kernel (per-kernel device time from HIP events on the
library's stream, eager launches right after the timed region), the diffuse
kernel's achieved bandwidth, and the CPU oracle timed on a bounded sample of
the same state on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "slime-mold_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "agent-updates/s + diffuse GB/s (% of HBM peak), plain FP32 vs batched BF16"
HBM_PEAK_GBPS = 8000.0
PRESETS = ["Default", "Sponge", "Firecracker Trees"]


def agent_bytes(storage_bytes):
    # 16 B state read + 16 B state write + 12 sensor taps + deposit read and write
    return 32 + 14 * storage_bytes


def trail_bytes(storage_bytes):
    # one read and one write of every cell
    return 2 * storage_bytes


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=4096)
    ap.add_argument("--agents-log2", type=int, default=24)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--storage", default="f32", choices=["f32", "bf16", "f16"])
    ap.add_argument("--resort", type=int, default=20, help="Morton re-sort interval (0: off)")
    ap.add_argument("--graph-chunk", type=int, default=100)
    ap.add_argument("--sort-cell-log2", type=int, default=4, help="re-sort bin edge 2^k cells")
    ap.add_argument("--timing-steps", type=int, default=30,
                    help="eager steps per preset for the per-kernel event timing")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(orc, agents, trail, settings, seconds, threads):
    """Oracle (C, OpenMP) on the same state: one step to size the sample, then
    as many whole steps as fit in `seconds`."""
    a, t = agents.copy(), trail.copy()
    t0 = time.perf_counter()
    orc.step(a, t, settings, 1, nthreads=threads)
    one = time.perf_counter() - t0
    n = max(1, min(200, int(seconds / max(one, 1e-6))))
    t0 = time.perf_counter()
    orc.step(a, t, settings, n, nthreads=threads)
    dt = time.perf_counter() - t0
    return n, dt


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        print("bench.py: only --gpus 1 is supported (single-GPU build)", file=sys.stderr)
        return 2
    import slimemold

    if slimemold.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    pm = slimemold.init_preset_manager()
    presets = [pm.get_preset(n).settings for n in PRESETS]
    W, H, N = args.width, args.height, 1 << args.agents_log2
    sb = 4 if args.storage == "f32" else 2
    K = args.steps
    seg = [K // 3, K // 3, K - 2 * (K // 3)]

    sim = slimemold.Simulation(W, H, N, seed=args.seed, params=presets[0], storage=args.storage,
                               resort_interval=args.resort, graph_chunk=args.graph_chunk,
                               sort_cell_log2=args.sort_cell_log2)
    print(f"[bench] {W}x{H} agents={N} storage={args.storage} resort={args.resort} "
          f"graph_chunk={args.graph_chunk} warmup={args.warmup} steps={K}", file=sys.stderr,
          flush=True)
    # snapshot of the initial state for the CPU sample
    agents0 = sim.read_agents() if args.cpu_seconds > 0 else None
    trail0 = sim.read_trail() if args.cpu_seconds > 0 else None

    sim.step(args.warmup)
    sim.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    for p, n in zip(presets, seg):
        sim.set_params(p)
        sim.step(n)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"[bench] timed {K} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    # per-kernel device time, eager launches with events on the library's stream
    tim = {"agents_ms": 0.0, "trail_ms": 0.0, "sort_ms": 0.0, "agents": 0, "trail": 0, "sort": 0}
    for p in presets:
        sim.set_params(p)
        t = sim.step_timed(args.timing_steps)
        tim["agents_ms"] += t.agents_ms
        tim["trail_ms"] += t.trail_ms
        tim["sort_ms"] += t.sort_ms
        tim["agents"] += t.agents_launches
        tim["trail"] += t.trail_launches
        tim["sort"] += t.sort_launches
    agents_avg_s = tim["agents_ms"] / max(1, tim["agents"]) / 1e3
    trail_avg_s = tim["trail_ms"] / max(1, tim["trail"]) / 1e3
    sort_avg_s = tim["sort_ms"] / max(1, tim["sort"]) / 1e3 if tim["sort"] else 0.0
    mass = sim.total_mass()
    sim.close()

    a_bytes = agent_bytes(sb) * N
    t_bytes = trail_bytes(sb) * W * H
    a_gbps = a_bytes / agents_avg_s / 1e9
    t_gbps = t_bytes / trail_avg_s / 1e9
    dominant = "agents" if agents_avg_s >= trail_avg_s else "trail"
    roof = {
        "kernel": "agents_kernel" if dominant == "agents" else "trail_strip_kernel",
        "bound": "hbm",
        "achieved": round(a_gbps if dominant == "agents" else t_gbps, 1),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round((a_gbps if dominant == "agents" else t_gbps) / HBM_PEAK_GBPS, 4),
        "traffic": None,
        "bytes_per_launch": a_bytes if dominant == "agents" else t_bytes,
        "avg_launch_us": round((agents_avg_s if dominant == "agents" else trail_avg_s) * 1e6, 2),
    }
    out = {
        "metric": METRIC,
        "value": N * K / elapsed,
        "unit": "agent-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic (seeded SplitMix64 agent init, zero trail)",
        "config": {
            "workload": (f"plain {args.storage} trail {W}x{H}, 2^{args.agents_log2} agents, presets "
                         f"{' -> '.join(PRESETS)} back to back ({'/'.join(map(str, seg))} steps), "
                         f"seed {args.seed}, hipGraph {args.graph_chunk}-step chunks"),
            "width": W, "height": H, "agents": N, "storage": args.storage, "seed": args.seed,
            "presets": PRESETS, "graph_chunk": args.graph_chunk, "resort_interval": args.resort,
            "sort_cell_log2": args.sort_cell_log2,
        },
        "roofline": roof,
        "diffuse": {
            "kernel": "trail_strip_kernel (deposit + decay + 3x3 diffuse)",
            "achieved_gbps": round(t_gbps, 1),
            "pct_hbm_peak": round(100 * t_gbps / HBM_PEAK_GBPS, 2),
            "cells_per_s": W * H / trail_avg_s,
            "avg_launch_us": round(trail_avg_s * 1e6, 2),
            "bytes_per_launch": t_bytes,
            "note": "trail of this size fits the 256 MB Infinity Cache" if t_bytes <= 256e6 else "",
        },
        "agents_kernel": {
            "avg_launch_us": round(agents_avg_s * 1e6, 2),
            "achieved_gbps": round(a_gbps, 1),
            "agent_updates_per_s": N / agents_avg_s,
            "bytes_per_launch": a_bytes,
        },
        "sort_avg_launch_us": round(sort_avg_s * 1e6, 2),
        "final_total_mass": mass,
        "cpu_baseline": None,
    }
    if args.cpu_seconds > 0:
        import oracle

        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        n, dt = cpu_baseline(oracle, agents0, trail0, presets[0], args.cpu_seconds, threads)
        out["cpu_baseline"] = {
            "value": N * n / dt,
            "unit": "agent-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": (f"C oracle (OpenMP, {threads} threads) on the same {W}x{H}, {N}-agent "
                       f"initial state, Default preset, {n} whole steps in {dt:.1f} s"),
        }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
