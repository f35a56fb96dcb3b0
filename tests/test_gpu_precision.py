"""16-bit trail storage against the FP64 oracle (reduced size; the full
8192^2 / 2^26-agent report is tests/precision_report.py).

Bars: one step's whole-grid mass error vs FP64 <= 1e-3 for bf16 (the target)
and <= 1e-4 for f16; cell errors within a few storage ulps of values <= 1
(bf16 ulp 2^-8 at 1, f16 2^-11); f32 storage within f32 rounding.
"""
import pytest

import precision_report

pytestmark = pytest.mark.gpu


def test_storage_precision_vs_fp64(slm, orc):
    res = precision_report.run(slm, orc, side=1024, agents_log2=18, steps=300, every=100,
                               window=256, seed=7, log=lambda m: print(m, flush=True))
    bf, sr, fh, f32 = (res["storages"][s] for s in ("bf16", "bf16sr", "f16", "f32"))
    assert bf["max_mass_rel_err"] <= 1e-3
    assert fh["max_mass_rel_err"] <= 1e-4
    assert f32["max_mass_rel_err"] <= 1e-6
    assert bf["max_window_abs"] <= 2.0 ** -7
    # stochastic rounding: each cell within one bf16 ulp (2^-7 at values < 1),
    # the mass unbiased
    assert sr["max_window_abs"] <= 2.0 ** -7
    assert sr["max_mass_rel_err"] <= 1e-4
    assert fh["max_window_abs"] <= 2.0 ** -10
    assert f32["max_window_abs"] <= 1e-6
    # whole runs diverge chaotically, but the macroscopic mass stays close
    for s in ("bf16", "bf16sr", "f16"):
        assert max(res["storages"][s]["run_mass_rel_to_f32"]) < 0.05
