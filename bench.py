"""Throughput of the slime-mold step on one MI355X.

Plain leg (`value`): 4096x4096 f32 trail, 2^24 agents initialised from seed 1,
the presets Default -> Sponge -> Firecracker Trees run back to back (the K timed
steps split in three), steps enqueued as captured hipGraphs of 100 steps.  A
step is one agent pass plus one fused deposit + decay + diffuse pass over the
whole grid.

Batched leg: the 256-instance mix (grids 256^2..2048^2, 2^14..2^22 agents,
parameters swept over the preset range) with a bf16 trail, all instances
advanced together (autotuned into instance groups on four streams).  Beside
it the same batch with an f32 trail ("batched_f32", the mix as configured:
the plain FP32 path on the batched workload, in one pool;
"batched_vs_batched_f32").  Separate legs: the same instances as 256 separate
f32 simulations with the batch's tuning, stepped one simulation at a time
("batched_vs_separate_serial") and all enqueued at once
("batched_vs_separate_concurrent").

The plain leg runs a fixed launch tuning (flags; by default 16-row trail
tiles, a re-sort every 15 steps into 8x8-cell bins, 1024 agents per agent-pass
workgroup and non-temporal trail-pass loads and stores; the library's
autotuner found the bins, and sweeps over 20, 300 and 2000 steps chose the
interval, the workgroup size and the streaming mode).
Before its warmup the batched (or combined) leg lets the library's autotuner
choose its tuning (re-sort interval and bin size, agents per workgroup, trail
tile rows, tiled sensing, instance groups, streaming) on its own
state, untimed; each leg reports its tuning under "tuning".

Prints one JSON line: value = plain agent-updates/s over the timed region; the
roofline of the plain leg's dominant kernel, whose average launch time comes
from HIP events around every kernel over the same steps replayed in a second
simulation from the same seed, tuning and warmup (the step is deterministic,
so each launch meets the state it met in the timed region); the
diffuse kernel's achieved bandwidth; the batched leg; and the CPU oracle timed
on a bounded sample of the plain leg's state on the host cores.  Beside them,
the library's Gaussian-diffuse extension at the presets' blur radius 2 (the
plain workload as "plain_gaussian", the pass alone as "gaussian_hbm"); the
headline stays the reference's 3x3 box.
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
STORAGE_BYTES = {"f32": 4, "bf16": 2, "f16": 2}

# Counter calibration (tests/counter_calibration.py under rocprofv3 on MI355X:
# a device-to-device copy of known size): FETCH_SIZE reads half the bytes a
# plain copy moves, WRITE_SIZE all of them.
FETCH_SIZE_SCALE = 2.0
WRITE_SIZE_SCALE = 1.0

# HBM bytes per launch of the plain leg's dominant kernel, measured with
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch, averaged over
# the dispatches of the timed steps, scaled as above) for exactly the run a
# line describes: keyed by the workload, the launch tuning (re-sort interval,
# bin log2, trail rows, agents per group, streaming, re-sort key) and the timed and warmup
# step counts.  A line whose run has no entry reports traffic null.  Sources:
# profiles/agents_kernel_plain_traffic.md (20 timed steps after 5 warmup, and
# 2000 after 100).
MEASURED_TRAFFIC = {
    # (kernel, W, H, N, storage, tuning, steps, warmup): (FETCH KiB, WRITE KiB)
    ("agents_kernel", 4096, 4096, 1 << 24, "f32", (15, 3, 16, 1024, 3, 0), 20, 5):
        (181524, 278540),
    ("agents_kernel", 4096, 4096, 1 << 24, "f32", (15, 3, 16, 1024, 3, 0), 2000, 100):
        (169235, 219605),
}


def measured_traffic(kernel, args):
    key = (kernel, args.width, args.height, 1 << args.agents_log2, args.storage,
           (args.resort, args.sort_cell_log2, args.trail_rows, args.agents_per_group,
            args.streaming, args.sort_key), args.steps, args.warmup)
    kib = MEASURED_TRAFFIC.get(key)
    if kib is None or args.tiled or args.window or args.autotune_plain or args.f64_positions:
        return None
    return int((kib[0] * FETCH_SIZE_SCALE + kib[1] * WRITE_SIZE_SCALE) * 1024)


def agent_bytes(sb, pos64=False):
    # 16 B state read + 16 B state write + 12 sensor taps + deposit read and write
    return 32 + 14 * sb + (16 if pos64 else 0)


def trail_bytes(sb):
    # one read and one write of every cell
    return 2 * sb


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="plain", choices=["plain", "combined"],
                    help="plain: plain + batched legs; combined: 1024-instance mix + "
                         "16384^2 / 2^28-agent instance, bf16")
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=4096)
    ap.add_argument("--agents-log2", type=int, default=24)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--storage", default="f32", choices=["f32", "bf16", "f16"])
    ap.add_argument("--f64-positions", action="store_true",
                    help="plain leg: f64 agent positions (SLM_FLAG_F64_POSITIONS)")
    ap.add_argument("--resort", type=int, default=15,
                    help="plain leg's Morton re-sort interval (0: off)")
    ap.add_argument("--batch-resort", type=int, default=20,
                    help="batched/combined legs' re-sort interval before autotuning")
    ap.add_argument("--graph-chunk", type=int, default=100)
    ap.add_argument("--sort-cell-log2", type=int, default=3, help="re-sort bin edge 2^k cells")
    ap.add_argument("--sort-key", type=lambda v: int(v, 0), default=0,
                    help="plain leg's re-sort key refinement (slm_tuning.sort_key: heading "
                         "classes log2 in bits 0-2, 0x10 = bin at the centre sensor)")
    ap.add_argument("--tiled", action="store_true",
                    help="plain leg: LDS-staged sensing from per-tile work lists")
    ap.add_argument("--window", action="store_true",
                    help="plain leg: LDS-staged sensing through per-workgroup windows")
    ap.add_argument("--batch-tuning", default="",
                    help="batched / combined legs: a fixed tuning instead of the autotuner, as "
                         "comma-separated field=value pairs of slm_tuning (e.g. "
                         "instance_groups=128,resort_interval=20,sort_cell_log2=3,trail_rows=8)")
    ap.add_argument("--persistent", action="store_true",
                    help="batched / combined legs: persistent agent and trail passes (one "
                         "launch each per step over every instance; no instance groups)")
    ap.add_argument("--trail-rows", type=int, default=16,
                    help="plain leg: rows per trail-pass tile (0: chosen from the grid)")
    ap.add_argument("--agents-per-group", type=int, default=1024,
                    help="plain leg: agents per workgroup of the agent pass (256..4096)")
    ap.add_argument("--streaming", type=int, default=3, choices=[0, 1, 2, 3],
                    help="plain leg: non-temporal trail-pass stores (1), loads (2) or both (3)")
    ap.add_argument("--autotune", type=int, default=80,
                    help="autotune trial steps before the batched/combined leg's warmup (0: off)")
    ap.add_argument("--autotune-plain", action="store_true",
                    help="also autotune the plain leg (otherwise it runs the flags' tuning)")
    ap.add_argument("--batch-steps", type=int, default=1000, help="batched leg steps (0: skip)")
    ap.add_argument("--diffuse-steps", type=int, default=500,
                    help="steps of each diffuse-only leg (8192^2 f32, 16384^2 bf16; 0: skip)")
    ap.add_argument("--gaussian-steps", type=int, default=300,
                    help="steps of each Gaussian-diffuse leg, same maps (0: skip)")
    ap.add_argument("--plain-gaussian-steps", type=int, default=2000,
                    help="steps of the plain workload with the Gaussian diffuse (0: skip)")
    ap.add_argument("--batch-f32-steps", type=int, default=-1,
                    help="steps of the batched workload with an f32 trail, autotuned the same "
                         "way (-1: as many as the batched leg, 0: skip)")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--separate-steps", type=int, default=300,
                    help="steps of the batched leg's instances run as separate f32 "
                         "simulations, one at a time and all at once (-1: as many as the "
                         "batched leg, 0: skip)")
    ap.add_argument("--batch-storage", default="bf16", choices=["f32", "bf16", "f16"])
    ap.add_argument("--rounding", default="stochastic", choices=["nearest", "stochastic"],
                    help="bf16 rounding of the batched, combined and bf16 diffuse legs, and "
                         "of the plain legs when --storage bf16")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the oracle baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-single-thread", type=int, default=1,
                    help="also time one whole step of the CPU baseline on one thread (0: skip)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name and the CPUs this process may run on."""
    name = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name}, {len(os.sched_getaffinity(0))} CPUs available"


def cpu_baseline(orc, agents, trail, presets, seconds, threads):
    """Oracle (C, OpenMP) on the same state and the same preset sequence as
    the GPU leg: one step to size the sample, then as many whole steps as fit
    in `seconds`, split in thirds over the presets in order."""
    a, t = agents.copy(), trail.copy()
    t0 = time.perf_counter()
    orc.step(a, t, presets[0], 1, nthreads=threads)
    one = time.perf_counter() - t0
    n = max(3, min(300, int(seconds / max(one, 1e-6))))
    seg = [n // 3, n // 3, n - 2 * (n // 3)]
    a, t = agents.copy(), trail.copy()
    t0 = time.perf_counter()
    for p, k in zip(presets, seg):
        orc.step(a, t, p, k, nthreads=threads)
    dt = time.perf_counter() - t0
    return n, seg, dt


def kernel_times(handle, presets, segments, set_params):
    """Average device time per launch of each kernel, from HIP events recorded
    inside the same graph chunks slm_step replays (slm_step_timed), over the
    given presets and step counts."""
    tot = {"agents_ms": 0.0, "trail_ms": 0.0, "sort_ms": 0.0, "agents": 0, "trail": 0, "sort": 0}
    for p, steps in zip(presets, segments):
        if set_params is not None:
            set_params(p)
        t = handle.step_timed(steps)
        tot["agents_ms"] += t.agents_ms
        tot["trail_ms"] += t.trail_ms
        tot["sort_ms"] += t.sort_ms
        tot["agents"] += t.agents_launches
        tot["trail"] += t.trail_launches
        tot["sort"] += t.sort_launches
    return (tot["agents_ms"] / max(1, tot["agents"]) / 1e3,
            tot["trail_ms"] / max(1, tot["trail"]) / 1e3,
            tot["sort_ms"] / tot["sort"] / 1e3 if tot["sort"] else 0.0)


def roofline(name, bytes_per_launch, seconds):
    """Algorithmic bytes per launch over the average launch time, against the
    HBM peak (nothing in the step is a dense contraction, so HBM is the
    roofline every kernel here is priced against)."""
    gbps = bytes_per_launch / seconds / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4), "traffic": None,
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": round(seconds * 1e6, 2)}


# What the counters show limits each kernel (rocprofv3 TA_BUSY, SQ_INSTS_VALU,
# SQ_BUSY_CYCLES, FETCH_SIZE / WRITE_SIZE; profiles/agents_trail_c1_counters.md).
LIMITER = {
    "agents_kernel": ("vector-L1 misses in flight against L2 latency on the divergent sensor "
                      "gathers (the L1 waits on L2 data 62% of the time, the texture addresser "
                      "waits on the L1 59%; profiles/agents_kernel_l1_counters.md), then VALU; "
                      "DRAM traffic is well below the algorithmic bytes because the taps hit L2, "
                      "so the kernel is not HBM-bound"),
    "trail_kernel": "HBM bandwidth",
}


def graph_split(segments, chunk):
    """(graph steps, eager steps) of slm_step over these segments: whole
    chunks replay captured graphs, the rest of a segment launches eagerly."""
    if chunk <= 0:
        return 0, sum(segments)
    g = sum(n // chunk * chunk for n in segments)
    return g, sum(segments) - g


def plain_sr(args):
    """Stochastic rounding of the plain leg's stores: bf16 storage with --rounding stochastic."""
    return args.storage == "bf16" and args.rounding == "stochastic"


def plain_leg(slm, args):
    presets = [slm.workloads.preset(n) for n in slm.workloads.PLAIN_PRESETS]
    W, H, N = args.width, args.height, 1 << args.agents_log2
    sb = STORAGE_BYTES[args.storage]
    K = args.steps
    seg = [K // 3, K // 3, K - 2 * (K // 3)]

    def make():
        sim = slm.Simulation(W, H, N, seed=args.seed, params=presets[0], storage=args.storage,
                             resort_interval=args.resort, graph_chunk=args.graph_chunk,
                             sort_cell_log2=args.sort_cell_log2, tiled_sensing=args.tiled,
                             window_sensing=args.window, f64_positions=args.f64_positions,
                             stochastic_rounding=plain_sr(args))
        sim.set_tuning(trail_rows=args.trail_rows, agents_per_group=args.agents_per_group,
                       streaming=args.streaming, sort_key=args.sort_key)
        return sim

    sim = make()
    log(f"plain {W}x{H} agents={N} storage={args.storage} resort={args.resort} "
        f"graph_chunk={args.graph_chunk} warmup={args.warmup} steps={K}")
    state0 = ((sim.read_agents(), sim.read_trail())
              if args.cpu_seconds > 0 and args.storage == "f32" and not args.f64_positions
              else None)
    tuning = autotune(sim, args.autotune if args.autotune_plain else 0, "plain")
    sim.step(args.warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    for p, n in zip(presets, seg):
        sim.set_params(p)
        sim.step(n)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"plain: timed {K} steps in {elapsed:.3f} s")
    mass = sim.total_mass()
    sim.close()
    # The kernels' own times: the same K steps again, from the same seed,
    # tuning and warmup (the step is deterministic, so every launch sees the
    # state it saw in the timed region), with HIP events around each kernel
    # on the library's stream.
    rep = make()
    if tuning.get("autotuned"):
        rep.set_tuning(**{f: tuning[f] for f in TUNING_FIELDS})
    rep.step(args.warmup)
    rep.synchronize()
    t0 = time.perf_counter()
    a_s, t_s, s_s = kernel_times(rep, presets, seg, rep.set_params)
    m_elapsed = time.perf_counter() - t0
    rep.close()
    log(f"plain: the same {K} steps again with per-kernel events in {m_elapsed:.3f} s")
    a_bytes, t_bytes = agent_bytes(sb, args.f64_positions) * N, trail_bytes(sb) * W * H
    return {
        "value": N * K / elapsed, "elapsed": elapsed, "steps": K, "segments": seg,
        "measure_value": N * K / m_elapsed,
        "agents_s": a_s, "trail_s": t_s, "sort_s": s_s, "mass": mass,
        "agents_roof": roofline("agents_kernel", a_bytes, a_s),
        "trail_roof": roofline("trail_kernel", t_bytes, t_s),
        "cells": W * H, "agents": N, "state0": state0, "presets": presets, "tuning": tuning,
    }


TUNING_FIELDS = ("resort_interval", "sort_cell_log2", "agents_per_group", "trail_rows",
                 "tiled_sensing", "instance_groups", "streaming", "sort_key")


def plain_gaussian_leg(slm, args, steps):
    """The plain leg's workload and tuning with the Gaussian diffuse in place of
    the 3x3 box, at the presets' blur radius 2 and sigma 1: wall time of
    `steps` steps (presets in thirds, as the plain leg)."""
    presets = [slm.workloads.preset(n) for n in slm.workloads.PLAIN_PRESETS]
    W, H, N = args.width, args.height, 1 << args.agents_log2
    seg = [steps // 3, steps // 3, steps - 2 * (steps // 3)]
    sim = slm.Simulation(W, H, N, seed=args.seed, params=presets[0], storage=args.storage,
                         resort_interval=args.resort, graph_chunk=args.graph_chunk,
                         sort_cell_log2=args.sort_cell_log2, gaussian=True,
                         f64_positions=args.f64_positions, stochastic_rounding=plain_sr(args))
    sim.set_tuning(trail_rows=args.trail_rows, agents_per_group=args.agents_per_group,
                   streaming=args.streaming, sort_key=args.sort_key)
    sim.step(args.warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    for p, n in zip(presets, seg):
        sim.set_params(p)
        sim.step(n)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    sim.close()
    log(f"plain gaussian: timed {steps} steps in {elapsed:.3f} s")
    return {"agent_updates_per_s": N * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
            "steps": steps, "radius": 2, "sigma": 1.0, "storage": args.storage}


def autotune(handle, trial_steps, label):
    """Let the library pick the launch tuning (untimed; the state is kept)."""
    if trial_steps <= 0:
        return dict(handle.get_tuning().as_dict(), autotuned=False)
    t0 = time.perf_counter()
    t = handle.autotune(trial_steps)
    d = t.as_dict()
    d.update(autotuned=True, seconds=round(time.perf_counter() - t0, 2))
    log(f"{label}: autotune {d}")
    return d


def diffuse_leg(slm, side, storage, steps, graph_chunk, tune_steps, rounding="nearest",
                gaussian=False):
    """The trail pass alone on a map larger than the 256 MB Infinity Cache:
    `side`^2 cells and no agents, so every step is one trail-kernel launch;
    graph chunks, wall time of `steps` steps between synchronisations (launch
    gaps included, so a lower bound on the kernel's own bandwidth), after the
    autotuner has picked the tile rows (the 3x3 box pass; the Gaussian pass,
    at the Default preset's blur radius 2 and sigma 1, has no tuning)."""
    sb = STORAGE_BYTES[storage]
    sr = storage == "bf16" and rounding == "stochastic"
    sim = slm.Simulation(side, side, 0, seed=3, params=slm.workloads.preset("Default"),
                         storage=storage, graph_chunk=graph_chunk, stochastic_rounding=sr,
                         gaussian=gaussian)
    kind = "gaussian" if gaussian else "diffuse"
    tuning = autotune(sim, 0 if gaussian else tune_steps, f"{kind} {side}^2 {storage}")
    sim.step(2 * graph_chunk)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    dt = time.perf_counter() - t0
    sim.close()
    cells = side * side
    gbps = trail_bytes(sb) * cells * steps / dt / 1e9
    log(f"{kind} {side}^2 {storage}: {dt / steps * 1e6:.1f} us/step, {gbps:.0f} GB/s")
    out = {"side": side, "storage": storage, "map_bytes": cells * sb, "steps": steps,
           "us_per_step": round(dt / steps * 1e6, 2), "achieved_gbps": round(gbps, 1),
           "pct_hbm_peak": round(100 * gbps / HBM_PEAK_GBPS, 2),
           "cells_per_s": cells * steps / dt, "rounding": "stochastic" if sr else "nearest"}
    if gaussian:
        out.update(kernel="gauss_kernel (deposit + decay + separable Gaussian)", radius=2,
                   sigma=1.0)
    else:
        out["trail_rows"] = tuning["trail_rows"]
    return out


def batch_leg(slm, specs, storage, steps, warmup, resort, graph_chunk, cell_log2, label,
              tiled=False, tune_steps=0, rounding="nearest", persistent=False):
    sb = STORAGE_BYTES[storage]
    log(f"{label}: {len(specs)} instances storage={storage} persistent={persistent} creating")
    sr = storage == "bf16" and rounding == "stochastic"
    b = slm.Batch(specs, storage=storage, resort_interval=resort, graph_chunk=graph_chunk,
                  sort_cell_log2=cell_log2, tiled_sensing=tiled, stochastic_rounding=sr,
                  persistent=persistent)
    info = b.info()
    log(f"{label}: {info.total_agents} agents, {info.total_cells} cells, "
        f"{info.device_bytes / 2**30:.1f} GiB; warmup {warmup}")
    if FIXED_TUNING:
        b.set_tuning(**FIXED_TUNING)
        tuning = dict(b.get_tuning().as_dict(), autotuned=False)
        log(f"{label}: fixed tuning {tuning}")
    else:
        tuning = autotune(b, tune_steps, label)
    b.step(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"{label}: timed {steps} steps in {elapsed:.3f} s")
    M = max(1, steps // 5)
    a_s, t_s, s_s = kernel_times(b, [None], [M], None)
    masses = b.total_mass()
    b.close()
    n_agents, n_cells = info.total_agents, info.total_cells
    return {
        "instances": len(specs), "storage": storage, "steps": steps, "warmup": warmup,
        "total_agents": n_agents, "total_cells": n_cells,
        "agent_updates_per_s": n_agents * steps / elapsed,
        "cell_updates_per_s": n_cells * steps / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "agents_kernel_us": round(a_s * 1e6, 2),
        "trail_kernel_us": round(t_s * 1e6, 2),
        "sort_us": round(s_s * 1e6, 2),
        "diffuse_gbps": round(trail_bytes(sb) * n_cells / t_s / 1e9, 1),
        "diffuse_pct_hbm_peak": round(100 * trail_bytes(sb) * n_cells / t_s / 1e9 / HBM_PEAK_GBPS, 2),
        "agents_gbps": round(agent_bytes(sb) * n_agents / a_s / 1e9, 1),
        "graph_chunk": graph_chunk, "tuning": tuning, "measure_steps": M,
        "rounding": "stochastic" if sr else "nearest",
        "persistent": persistent,
        "total_mass_sum": float(masses.sum()),
    }


def separate_leg(slm, specs, storage, steps, warmup, graph_chunk, tuning, serial=False):
    """The batched leg's instances as separate single simulations (one handle,
    stream and graph per instance), with the batch's tuning: the plain path on
    the same workload.  Each round enqueues one graph chunk per simulation;
    the streams share the device's hardware queues.  serial: each simulation's
    chunk finishes before the next one's is enqueued (one simulation at a
    time, as a caller stepping one handle after another)."""
    fields = ("resort_interval", "sort_cell_log2", "agents_per_group", "trail_rows",
              "tiled_sensing", "sort_key")
    log(f"separate: {len(specs)} simulations storage={storage} creating")
    sims = []
    try:
        for sp in specs:
            sim = slm.Simulation(sp.width, sp.height, sp.agent_count, seed=sp.seed,
                                 params=sp.params, storage=storage, graph_chunk=graph_chunk)
            sim.set_tuning(**{f: tuning[f] for f in fields})
            sims.append(sim)
        chunk = max(1, graph_chunk)

        def run(n):
            done = 0
            while done < n:
                k = min(chunk, n - done)
                for sim in sims:
                    sim.step(k)
                    if serial:
                        sim.synchronize()
                done += k
            for sim in sims:
                sim.synchronize()

        run(warmup)
        t0 = time.perf_counter()
        run(steps)
        elapsed = time.perf_counter() - t0
        log(f"separate: timed {steps} steps in {elapsed:.3f} s")
        n_agents = sum(sp.agent_count for sp in specs)
        n_cells = sum(sp.width * sp.height for sp in specs)
    finally:
        for sim in sims:
            sim.close()
    return {"instances": len(specs), "storage": storage, "steps": steps, "warmup": warmup,
            "serial": serial,
            "agent_updates_per_s": n_agents * steps / elapsed,
            "cell_updates_per_s": n_cells * steps / elapsed,
            "ms_per_step": elapsed / steps * 1e3,
            "tuning": {f: tuning[f] for f in fields}}


FIXED_TUNING = {}


def main():
    args = parse()
    for pair in filter(None, args.batch_tuning.split(",")):
        k, v = pair.split("=")
        FIXED_TUNING[k.strip()] = int(v)
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs one simulation per GPU on one GPU; "
                                   "--gpus must be 1",
                          "n_gpus": args.gpus}))
        print("bench.py: only --gpus 1 is supported (the library runs on one GPU; instances "
              "of a batch would shard across GPUs by cost, which is not built)", file=sys.stderr)
        return 2
    import slimemold as slm
    import slimemold.workloads  # noqa: F401

    if slm.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        return 1

    if args.workload == "combined":
        specs = slm.workloads.combined()
        res = batch_leg(slm, specs, "bf16", args.steps, args.warmup, args.batch_resort,
                        args.graph_chunk, args.sort_cell_log2, "combined", args.tiled,
                        args.autotune, args.rounding, args.persistent)
        out = {
            "metric": METRIC, "value": res["agent_updates_per_s"], "unit": "agent-updates/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16", "data": "synthetic (seeded)",
            "config": {"workload": "combined: 1024 mixed instances + 16384x16384 / 2^28 agents, "
                                   "bf16 trail, one agent + one trail launch per step"},
            "combined": res, "cpu_baseline": None, "roofline": None,
        }
        print(json.dumps(out))
        return 0

    plain = plain_leg(slm, args)
    W, H, N = args.width, args.height, 1 << args.agents_log2
    dominant = plain["agents_roof"] if plain["agents_s"] >= plain["trail_s"] else plain["trail_roof"]
    dominant = dict(dominant)
    dominant["limiter"] = LIMITER[dominant["kernel"]]
    traffic = measured_traffic(dominant["kernel"], args)
    dominant["traffic"] = traffic
    if traffic is not None:
        dram = traffic / (dominant["avg_launch_us"] * 1e-6) / 1e9
        dominant["dram_gbps"] = round(dram, 1)
        dominant["dram_frac"] = round(dram / HBM_PEAK_GBPS, 4)
        dominant["traffic_source"] = "profiles/agents_kernel_plain_traffic.md (same command)"
    graph_steps, eager_steps = graph_split(plain["segments"], args.graph_chunk)
    out = {
        "metric": METRIC,
        "value": plain["value"],
        "unit": "agent-updates/s",
        "n_gpus": 1,
        "steps": plain["steps"],
        "warmup": args.warmup,
        "ms_per_step": plain["elapsed"] / plain["steps"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic (seeded SplitMix64 agent init, zero trail)",
        "config": {
            "workload": (f"plain {args.storage} trail {W}x{H}, 2^{args.agents_log2} agents, "
                         f"presets {' -> '.join(slm.workloads.PLAIN_PRESETS)} back to back "
                         f"({'/'.join(map(str, plain['segments']))} steps), seed {args.seed}, "
                         f"{graph_steps} of the {plain['steps']} timed steps replayed from "
                         f"captured {args.graph_chunk}-step hipGraphs, {eager_steps} launched "
                         f"eagerly (a segment's steps beyond its last whole chunk)"),
            "width": W, "height": H, "agents": N, "storage": args.storage, "seed": args.seed,
            "f64_positions": args.f64_positions,
            "rounding": "stochastic" if plain_sr(args) else "nearest",
            "presets": slm.workloads.PLAIN_PRESETS, "graph_chunk": args.graph_chunk,
            "graph_steps": graph_steps, "eager_steps": eager_steps,
            "tuning": plain["tuning"],
        },
        "roofline": dominant,
        "diffuse": {
            "kernel": "trail_kernel (deposit + decay + 3x3 diffuse)",
            "achieved_gbps": plain["trail_roof"]["achieved"],
            "pct_hbm_peak": round(100 * plain["trail_roof"]["frac"], 2),
            "cells_per_s": W * H / plain["trail_s"],
            "avg_launch_us": plain["trail_roof"]["avg_launch_us"],
            "bytes_per_launch": plain["trail_roof"]["bytes_per_launch"],
            "note": ("trail of this size fits the 256 MB Infinity Cache; event records "
                     "around each kernel add a few microseconds to short kernels; see "
                     "diffuse_hbm for maps larger than the cache"
                     if plain["trail_roof"]["bytes_per_launch"] <= 256e6 else ""),
        },
        "agents_kernel": {
            "avg_launch_us": plain["agents_roof"]["avg_launch_us"],
            "achieved_gbps": plain["agents_roof"]["achieved"],
            "agent_updates_per_s": N / plain["agents_s"],
            "bytes_per_launch": plain["agents_roof"]["bytes_per_launch"],
        },
        "sort_avg_launch_us": round(plain["sort_s"] * 1e6, 2),
        "kernel_timing": (f"HIP events around every kernel on the library's stream, over the "
                          f"timed region's {plain['steps']} steps replayed from the same seed, "
                          f"tuning and warmup in a second simulation (the same launches on the "
                          f"same states; event records inside the graph chunks, eager steps "
                          f"each with their own); the replay ran at "
                          f"{plain['measure_value']:.4g} agent-updates/s"),
        "cell_updates_per_s": plain["cells"] * plain["steps"] / plain["elapsed"],
        "final_total_mass": plain["mass"],
        "cpu_baseline": None,
    }
    if args.diffuse_steps > 0:
        out["diffuse_hbm"] = [diffuse_leg(slm, 8192, "f32", args.diffuse_steps, args.graph_chunk,
                                          args.autotune),
                              diffuse_leg(slm, 16384, "bf16", args.diffuse_steps,
                                          args.graph_chunk, args.autotune, args.rounding)]
    if args.plain_gaussian_steps > 0:
        out["plain_gaussian"] = plain_gaussian_leg(slm, args, args.plain_gaussian_steps)
    if args.gaussian_steps > 0:
        out["gaussian_hbm"] = [
            diffuse_leg(slm, 8192, "f32", args.gaussian_steps, args.graph_chunk, 0,
                        gaussian=True),
            diffuse_leg(slm, 16384, "bf16", args.gaussian_steps, args.graph_chunk, 0,
                        args.rounding, gaussian=True)]
    if args.batch_steps > 0:
        specs = slm.workloads.batch_mix(args.batch_count)
        bres = batch_leg(slm, specs, args.batch_storage, args.batch_steps, 20, args.batch_resort,
                         args.graph_chunk, args.sort_cell_log2, "batched", args.tiled,
                         args.autotune, args.rounding, args.persistent)
        out["batched"] = bres
        out["batched_vs_plain"] = round(bres["agent_updates_per_s"] / plain["value"], 3)
        f32_steps = args.batch_steps if args.batch_f32_steps < 0 else args.batch_f32_steps
        if f32_steps > 0:
            fres = batch_leg(slm, specs, "f32", f32_steps, 20, args.batch_resort,
                             args.graph_chunk, args.sort_cell_log2, "batched_f32", args.tiled,
                             args.autotune, persistent=args.persistent)
            out["batched_f32"] = fres
            out["batched_vs_batched_f32"] = round(
                bres["agent_updates_per_s"] / fres["agent_updates_per_s"], 3)
        sep_steps = args.batch_steps if args.separate_steps < 0 else args.separate_steps
        if sep_steps > 0:
            out["separate"] = {}
            for mode in ("serial", "concurrent"):
                sres = separate_leg(slm, specs, "f32", sep_steps, 20, args.graph_chunk,
                                    bres["tuning"], serial=mode == "serial")
                out["separate"][mode] = sres
                out[f"batched_vs_separate_{mode}"] = round(
                    bres["agent_updates_per_s"] / sres["agent_updates_per_s"], 3)
    if args.cpu_seconds > 0 and (args.storage != "f32" or args.f64_positions):
        out["cpu_baseline_note"] = ("the CPU sample runs the f32 plain path only; this line's "
                                    "storage or positions differ, so none is reported")
    elif args.cpu_seconds > 0:
        import oracle

        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        a0, t0 = plain["state0"]
        n, cseg, dt = cpu_baseline(oracle, a0, t0, plain["presets"], args.cpu_seconds, threads)
        out["cpu_baseline"] = {
            "value": N * n / dt,
            "unit": "agent-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": (f"C oracle (OpenMP, {threads} threads) on the plain leg's {W}x{H}, "
                       f"{N}-agent initial state, presets "
                       f"{' -> '.join(slm.workloads.PLAIN_PRESETS)} for "
                       f"{'/'.join(map(str, cseg))} whole steps ({n} in {dt:.1f} s)"),
            "host": cpu_model(),
        }
        if args.cpu_single_thread:
            a1, t1 = a0.copy(), t0.copy()
            t_start = time.perf_counter()
            oracle.step(a1, t1, plain["presets"][0], 1, nthreads=1)
            dt1 = time.perf_counter() - t_start
            out["cpu_baseline_1thread"] = {
                "value": N / dt1, "unit": "agent-updates/s", "cores": 1, "kind": "port",
                "sample": (f"C oracle on 1 thread, the same state and preset, 1 whole step "
                           f"in {dt1:.1f} s"),
            }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
