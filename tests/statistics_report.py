"""Ensemble statistics of the small headless configuration (1024x1024 f32
trail, 2^18 agents, 1000 steps) per preset: the HIP path (its shared sin/cos)
against the oracle switched to the C library's sinf/cosf - a stand-in for a
reference whose transcendental bits differ, as lavapipe's would.  Single runs
of the two diverge chaotically and differ by several percent in total mass
between seeds, so the comparison is of ensemble means: for each preset, the
mean total trail mass and the mean power of each octave of the radially
averaged spectrum over `--seeds` seeds on each side, their relative
difference, and its standard error.  The bar compared with is 1% in
expectation.

  python tests/statistics_report.py --side gpu  --out gpu.json     (MI355X)
  python tests/statistics_report.py --side libm --out libm.json    (CPU oracle)
  python tests/statistics_report.py --compare gpu.json libm.json   (markdown;
        each side may list several files, comma-separated, whose runs are pooled)

Measurement infrastructure, not a test (pytest does not collect this file).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT, os.path.join(ROOT, "slime-mold_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from test_gpu_statistics import octave_power, radial_spectrum  # noqa: E402

W = H = 1024
N = 1 << 18
STEPS = 1000
OCTAVES = [(1, 4), (4, 16), (16, 64), (64, 256), (256, 512)]


def stats_of(trail):
    prof = radial_spectrum(trail)
    return {"mass": float(trail.astype(np.float64).sum()),
            "octaves": [octave_power(prof, lo, hi) for lo, hi in OCTAVES]}


def run_gpu(presets, seeds):
    import slimemold as slm
    out = {}
    for name in presets:
        s = slm.init_preset_manager().get_preset(name).settings
        rows = []
        for seed in seeds:
            with slm.Simulation(W, H, N, seed=seed, params=s, graph_chunk=100) as sim:
                sim.step(STEPS)
                rows.append(stats_of(sim.read_trail()))
            print(f"[gpu] {name} seed {seed}: mass {rows[-1]['mass']:.6e}", flush=True)
        out[name] = rows
    return out


def run_libm(presets, seeds, threads):
    import oracle as orc
    import slimemold as slm
    out = {}
    orc.set_libm_trig(True)
    try:
        for name in presets:
            s = slm.init_preset_manager().get_preset(name).settings
            rows = []
            for seed in seeds:
                t0 = time.perf_counter()
                a = orc.init_agents(seed, W, H, N, s.agent_speed_min, s.agent_speed_max)
                t = np.zeros((H, W), np.float32)
                orc.step(a, t, s, STEPS, nthreads=threads)
                rows.append(stats_of(t))
                print(f"[libm] {name} seed {seed}: mass {rows[-1]['mass']:.6e} "
                      f"({time.perf_counter() - t0:.0f} s)", flush=True)
            out[name] = rows
    finally:
        orc.set_libm_trig(False)
    return out


def mean_se(v):
    v = np.asarray(v, np.float64)
    return float(v.mean()), float(v.std(ddof=1) / math.sqrt(len(v))) if len(v) > 1 else 0.0


def rel_diff(a, b):
    """(mean_a - mean_b) / mean_b and its standard error (first order)."""
    ma, sa = mean_se(a)
    mb, sb = mean_se(b)
    r = ma / mb - 1.0
    se = (ma / mb) * math.sqrt((sa / ma) ** 2 + (sb / mb) ** 2) if ma and mb else float("nan")
    return r, se


def compare(gpu, libm):
    lines = ["| preset | seeds | mass: GPU mean (rel. spread) | mass: rel. diff ± s.e. | "
             + " | ".join(f"k∈[{lo},{hi}) rel. diff ± s.e." for lo, hi in OCTAVES)
             + " | within 1% in expectation |",
             "|---|---|---|---|" + "---|" * len(OCTAVES) + "---|"]
    for name in gpu:
        if name not in libm:
            continue
        g, l = gpu[name], libm[name]
        gm = [r["mass"] for r in g]
        lm = [r["mass"] for r in l]
        m, se = rel_diff(gm, lm)
        spread = float(np.std(gm, ddof=1) / np.mean(gm)) if len(gm) > 1 else 0.0
        cells = []
        ok = abs(m) <= 0.01 or abs(m) <= 2 * se  # 1%, or not distinguishable from 0
        for i, (lo, hi) in enumerate(OCTAVES):
            r, s = rel_diff([x["octaves"][i] for x in g], [x["octaves"][i] for x in l])
            cells.append(f"{100 * r:+.2f}% ± {100 * s:.2f}%")
            if lo >= 16 and hi <= 256:
                ok = ok and (abs(r) <= 0.01 or abs(r) <= 2 * s)
        lines.append(f"| {name} | {len(g)}/{len(l)} | {np.mean(gm):.4e} ({100 * spread:.1f}%) | "
                     f"{100 * m:+.2f}% ± {100 * se:.2f}% | " + " | ".join(cells)
                     + f" | {'yes' if ok else 'no'} |")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", choices=["gpu", "libm"])
    ap.add_argument("--out")
    ap.add_argument("--compare", nargs=2, metavar=("GPU_JSONS", "LIBM_JSONS"))
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--first-seed", type=int, default=100)
    ap.add_argument("--presets", default="Default,Sponge,Firecracker Trees,Threads,Curls,Waves,"
                                          "Snake,Mesh")
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.compare:
        def merged(paths):
            runs = {}
            for path in paths.split(","):
                for name, rows in json.load(open(path))["runs"].items():
                    runs.setdefault(name, []).extend(rows)
            return runs
        print(compare(merged(args.compare[0]), merged(args.compare[1])))
        return 0
    presets = [p for p in args.presets.split(",") if p]
    seeds = list(range(args.first_seed, args.first_seed + args.seeds))
    t0 = time.perf_counter()
    runs = run_gpu(presets, seeds) if args.side == "gpu" else run_libm(presets, seeds,
                                                                         args.threads)
    doc = {"side": args.side, "grid": [W, H], "agents": N, "steps": STEPS, "seeds": seeds,
           "octaves": OCTAVES, "seconds": time.perf_counter() - t0, "runs": runs}
    with open(args.out, "w") as f:
        json.dump(doc, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
