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
