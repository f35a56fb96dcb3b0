"""Trail precision of the 16-bit storage modes against the FP64 oracle.

For each storage format the simulation runs from the same seeded state; every
`every` steps it measures one step's error: the GPU steps once from state S,
and the FP64 oracle applies the trail pass (deposit, decay, 3x3 diffuse) to the
same S with the deposit cells the GPU's agents chose, so the comparison
isolates storage rounding and f32 arithmetic (agent trajectories diverge
chaotically between precisions after a few steps, which is why whole runs are
compared only through their mass).  Reported per checkpoint: max and RMS cell
error on a window, relative mass error of the whole grid, and the run's total
mass (f64 on the device).

Usage (writes JSON; the default is the 8192^2 / 2^26-agent configuration;
--combined reports the mixed-instance bf16 workload per grid-size class):
    python tests/precision_report.py --out precision.json
    python tests/precision_report.py --combined --out precision_combined.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "slime-mold_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)


def one_step_error(sim, orc, settings, window, threads):
    """GPU step from the current state vs the FP64 trail pass of that state."""
    t0 = sim.read_trail().astype(np.float64)
    sim.step(1)
    t1 = sim.read_trail().astype(np.float64)
    agents = sim.read_agents()
    h, w = t0.shape
    marks = orc.deposit_marks(agents, w, h)
    ref = orc.trail_pass_f64(t0, settings, marks, nthreads=threads)
    y0, x0 = (h - window) // 2, (w - window) // 2
    d = t1[y0:y0 + window, x0:x0 + window] - ref[y0:y0 + window, x0:x0 + window]
    m_gpu, m_ref = float(t1.sum()), float(ref.sum())
    return {
        "window_max_abs": float(np.abs(d).max()),
        "window_rms": float(np.sqrt(np.mean(d * d))),
        "mass_rel_err": abs(m_gpu - m_ref) / m_ref if m_ref else 0.0,
        "mass_gpu": m_gpu,
        "mass_fp64": m_ref,
    }


def run(slm, orc, side=8192, agents_log2=26, steps=5000, every=500, window=512, seed=7,
        storages=("bf16", "bf16sr", "f16", "f32"), preset="Default", f64_positions=False,
        threads=0,
        resort=20, log=print):
    settings = slm.init_preset_manager().get_preset(preset).settings
    out = {"side": side, "agents": 1 << agents_log2, "steps": steps, "every": every,
           "window": window, "seed": seed, "preset": preset, "f64_positions": f64_positions,
           "storages": {}}
    for storage in storages:
        t_start = time.perf_counter()
        rows = []
        with slm.Simulation(side, side, 1 << agents_log2, seed=seed, params=settings,
                            storage="bf16" if storage == "bf16sr" else storage,
                            stochastic_rounding=storage == "bf16sr", resort_interval=resort,
                            graph_chunk=100, f64_positions=f64_positions) as sim:
            done = 0
            for c in range(0, steps + 1, every):
                if c > done:
                    sim.step(c - done)
                    done = c
                mass = sim.total_mass()
                err = one_step_error(sim, orc, settings, window, threads)
                done += 1
                err.update({"step": c, "mass": mass})
                rows.append(err)
                log(f"[precision] {storage} step {c}: mass {mass:.6e} one-step mass err "
                    f"{err['mass_rel_err']:.3e} window max {err['window_max_abs']:.3e} "
                    f"rms {err['window_rms']:.3e}")
        out["storages"][storage] = {
            "checkpoints": rows,
            "max_mass_rel_err": max(r["mass_rel_err"] for r in rows),
            "max_window_abs": max(r["window_max_abs"] for r in rows),
            "max_window_rms": max(r["window_rms"] for r in rows),
            "seconds": time.perf_counter() - t_start,
        }
    if "f32" in out["storages"]:
        ref = [r["mass"] for r in out["storages"]["f32"]["checkpoints"]]
        for storage, res in out["storages"].items():
            res["run_mass_rel_to_f32"] = [abs(r["mass"] - m) / m if m else 0.0
                                          for r, m in zip(res["checkpoints"], ref)]
    return out


def run_combined(slm, orc, steps=200, per_class=3, threads=0, log=print, count=1024,
                 stochastic_rounding=False):
    """The combined workload (mixed instances + one 16384^2 / 2^28-agent
    instance, bf16): after `steps` steps, one step's mass error against the
    FP64 trail pass for up to `per_class` instances of every grid size."""
    specs = slm.workloads.combined(count=count)
    classes = {}
    for k, sp in enumerate(specs):
        if sp.agent_count > 0:
            classes.setdefault(sp.width, []).append(k)
    picked = {side: ks[:per_class] for side, ks in sorted(classes.items())}
    out = {"steps": steps, "instances": len(specs), "classes": {}}
    t_start = time.perf_counter()
    out["stochastic_rounding"] = stochastic_rounding
    with slm.Batch(specs, storage="bf16", resort_interval=20, graph_chunk=100,
                   stochastic_rounding=stochastic_rounding) as b:
        b.step(steps)
        before = {k: b.read_trail(k).astype(np.float64) for ks in picked.values() for k in ks}
        b.step(1)
        for side, ks in picked.items():
            rows = []
            for k in ks:
                sp = specs[k]
                t1 = b.read_trail(k).astype(np.float64)
                marks = orc.deposit_marks(b.read_agents(k), sp.width, sp.height)
                ref = orc.trail_pass_f64(before.pop(k), sp.params, marks, nthreads=threads)
                m_gpu, m_ref = float(t1.sum()), float(ref.sum())
                rows.append({"instance": k, "agents": sp.agent_count,
                             "mass_rel_err": abs(m_gpu - m_ref) / m_ref if m_ref else 0.0,
                             "max_abs": float(np.abs(t1 - ref).max())})
            out["classes"][f"{side}x{side}"] = {
                "instances_in_class": len(classes[side]), "sampled": rows,
                "max_mass_rel_err": max(r["mass_rel_err"] for r in rows),
                "max_abs": max(r["max_abs"] for r in rows)}
            log(f"[precision] combined {side}^2: max one-step mass err "
                f"{out['classes'][f'{side}x{side}']['max_mass_rel_err']:.3e} over {len(rows)}")
    out["seconds"] = time.perf_counter() - t_start
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=8192)
    ap.add_argument("--agents-log2", type=int, default=26)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--every", type=int, default=500)
    ap.add_argument("--window", type=int, default=512)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--f64-positions", action="store_true")
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--out", default="")
    ap.add_argument("--combined", action="store_true",
                    help="per-size-class report on the combined bf16 workload instead")
    ap.add_argument("--combined-steps", type=int, default=200)
    ap.add_argument("--stochastic-rounding", action="store_true",
                    help="combined report with stochastic bf16 rounding")
    args = ap.parse_args()
    import oracle
    import slimemold
    import slimemold.workloads  # noqa: F401

    if args.combined:
        res = run_combined(slimemold, oracle, args.combined_steps, threads=args.threads,
                           log=lambda m: print(m, flush=True),
                           stochastic_rounding=args.stochastic_rounding)
        print(json.dumps({c: {k: v for k, v in r.items() if k != "sampled"}
                          for c, r in res["classes"].items()}))
        if args.out:
            with open(args.out, "w") as f:
                json.dump(res, f, indent=1)
        return

    res = run(slimemold, oracle, args.side, args.agents_log2, args.steps, args.every,
              args.window, args.seed, f64_positions=args.f64_positions, threads=args.threads,
              log=lambda m: print(m, flush=True))
    summary = {s: {k: v for k, v in r.items() if k != "checkpoints"}
               for s, r in res["storages"].items()}
    print(json.dumps(summary))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
