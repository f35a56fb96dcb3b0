"""Shared helpers for the GPU parity tests: run the HIP path and the oracle from
the same initial state and compare bit for bit."""
import numpy as np

PRESET_NAMES = ["Default", "Sponge", "Firecracker Trees", "Threads", "Curls", "Waves", "Snake",
                "Mesh"]


def preset(slm, name):
    return slm.init_preset_manager().get_preset(name).settings


def describe_mismatch(got, want, what):
    bad = got != want
    if not np.any(bad):
        return ""
    idx = np.argwhere(bad)
    first = tuple(int(v) for v in idx[0])
    diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
    return (f"{what}: {int(bad.sum())} of {bad.size} differ, max |diff| {diff.max():.3e}, "
            f"first at {first}: got {got[first]!r} want {want[first]!r}")


def assert_bitwise(got, want, what):
    assert got.shape == want.shape, (got.shape, want.shape)
    msg = describe_mismatch(got, want, what)
    assert not msg, msg


def run_parity(slm, orc, w, h, n, seed, settings, steps, storage="f32", trail0=None,
               agents0=None, **simkw):
    """Step the GPU and the oracle `steps` times from the same state.
    Returns (gpu_agents, gpu_trail, oracle_agents, oracle_trail)."""
    print(f"[parity] {w}x{h} n={n} seed={seed} steps={steps} storage={storage} {simkw}", flush=True)
    if storage == "bf16sr":  # bf16 storage, stochastic rounding in the trail pass
        simkw = dict(simkw, stochastic_rounding=True)
    with slm.Simulation(w, h, n, seed=seed, params=settings,
                        storage="bf16" if storage == "bf16sr" else storage, **simkw) as sim:
        if agents0 is not None:
            sim.write_agents(agents0)
        if trail0 is not None:
            sim.write_trail(trail0)
        a = sim.read_agents()
        t = sim.read_trail()
        sim.step(steps)
        ga, gt = sim.read_agents(), sim.read_trail()
    orc.step(a, t, settings, steps, storage=storage)
    return ga, gt, a, t


def display_point(orc, agents, trail, settings, storage="f32", gaussian=False):
    """One oracle step from (agents, trail), in place, returning the trail as
    the reference's frame displays it: after the agents' deposits and the
    decay, before the diffuse (src/main.rs:1183-1217).  That is the trail pass
    with diffusion rate 0 (mix(T2, avg, 0) = T2) over the cells the moved
    agents deposit into, kept in f32 (the display samples the decayed values
    before any storage rounding)."""
    before = trail.copy()
    orc.step(agents, trail, settings, 1, storage=storage, gaussian=gaussian)
    h, w = trail.shape
    op = orc.params(settings)
    d = {n: getattr(op, n) for n, _ in op._fields_}
    d["diffusion_rate"] = 0.0
    return orc.trail_pass(before, d, mark=orc.deposit_marks(agents, w, h), storage="f32")
