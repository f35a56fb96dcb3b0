"""CPU-side checks of the boundary: the library loads and exports every entry
point include/slimemold.h declares, the header is valid C, argument errors are
reported without a device, and the host mirror's Settings/presets equal the
reference's (tests/golden/presets.json, extracted from src/settings.rs and
src/presets.rs)."""
import ctypes
import json
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "slimemold.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(slm_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(slm):
    L = slm.lib()
    names = declared_functions()
    assert len(names) >= 21
    for n in names:
        assert hasattr(L, n), n


def test_header_is_plain_c_and_links():
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no gcc")
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        with open(c, "w") as f:
            f.write('#include "slimemold.h"\n#include <stdio.h>\n'
                    "int main(void){slm_config c; slm_params p; slm_config_default(&c);"
                    "slm_params_default(&p); printf(\"%u %s\\n\", c.width, slm_last_error());"
                    "return 0;}\n")
        libdir = os.path.join(ROOT, "slime-mold_amd", "slimemold")
        subprocess.check_call([cc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                               "-I", os.path.join(ROOT, "include"), c, "-o",
                               os.path.join(d, "t"), "-L", libdir, "-lslimemold_hip",
                               f"-Wl,-rpath,{libdir}"])
        out = subprocess.check_output([os.path.join(d, "t")]).decode()
        assert out.startswith("1600 ")


def test_defaults_match_reference(slm):
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "presets.json")))["defaults"]
    p = slm.default_params()
    assert p.decay_factor == ref["pheromone_decay_factor"]
    assert p.agent_jitter == ref["agent_jitter"]
    assert p.agent_speed_min == ref["agent_speed_min"]
    assert p.agent_speed_max == ref["agent_speed_max"]
    assert p.agent_turn_speed == pytest.approx(ref["agent_turn_speed"], rel=1e-7)
    assert p.agent_sensor_angle == pytest.approx(ref["agent_sensor_angle"], rel=1e-7)
    assert p.agent_sensor_distance == ref["agent_sensor_distance"]
    assert p.diffusion_rate == ref["pheromone_diffusion_rate"]
    assert p.pheromone_deposition_amount == ref["pheromone_deposition_amount"]
    assert p.blur_radius == ref["blur_radius"] and p.blur_sigma == ref["blur_sigma"]
    c = slm.default_config()
    assert (c.width, c.height) == (ref["window_width"], ref["window_height"])


def test_host_presets_match_reference(slm):
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "presets.json")))["presets"]
    mine = slm.init_preset_manager()
    assert mine.get_preset_names() == [p["name"] for p in ref]
    for r in ref:
        s = mine.get_preset(r["name"]).settings
        for field, val in r["settings"].items():
            assert getattr(s, field) == val, (r["name"], field)
        # the parameter block carries them into the step as f32
        p = s.to_params()
        assert p.agent_turn_speed == pytest.approx(s.agent_turn_speed, rel=1e-7)
        assert p.decay_factor == pytest.approx(s.pheromone_decay_factor, rel=1e-7)


def _create(slm, **over):
    cfg = slm.default_config()
    cfg.width, cfg.height, cfg.agent_count = 64, 64, 100
    for k, v in over.items():
        setattr(cfg, k, v)
    p = slm.default_params()
    h = ctypes.c_void_p()
    rc = slm.lib().slm_create(ctypes.byref(cfg), ctypes.byref(p), ctypes.byref(h))
    return rc, slm.lib().slm_last_error().decode(), h


@pytest.mark.parametrize("over,needle", [
    (dict(width=0), "width and height"),
    (dict(height=1 << 25), "2^24"),
    (dict(width=1 << 16, height=1 << 16), "2^31"),
    (dict(storage=7), "storage"),
    (dict(agent_count=1 << 40), "agent_count"),
    (dict(sort_cell_log2=20), "sort_cell_log2"),
    (dict(flags=0x80), "flags"),
])
def test_create_rejects_bad_config_before_touching_a_device(slm, over, needle):
    rc, msg, h = _create(slm, **over)
    assert rc == -1 and needle in msg and not h.value


def test_create_rejects_bad_params(slm):
    cfg = slm.default_config()
    cfg.agent_count = 10
    p = slm.default_params()
    p.agent_turn_speed = float("nan")
    h = ctypes.c_void_p()
    assert slm.lib().slm_create(ctypes.byref(cfg), ctypes.byref(p), ctypes.byref(h)) == -1
    p = slm.default_params()
    p.agent_sensor_distance = 1e9
    assert slm.lib().slm_create(ctypes.byref(cfg), ctypes.byref(p), ctypes.byref(h)) == -1
    assert slm.lib().slm_create(None, ctypes.byref(p), ctypes.byref(h)) == -1
    assert slm.lib().slm_create(ctypes.byref(cfg), ctypes.byref(p), None) == -1


def test_null_handle_calls_fail_cleanly(slm):
    L = slm.lib()
    assert L.slm_step(None, 1) == -1
    assert L.slm_clear(None) == -1
    assert L.slm_destroy(None) == -1
    assert L.slm_resize(None, 4, 4) == -1
    assert L.slm_total_mass(None, None) == -1
    assert "NULL" in L.slm_last_error().decode()


def test_no_device_is_reported_not_crashed(slm):
    if slm.device_count() > 0:
        pytest.skip("a GPU is visible")
    rc, msg, h = _create(slm)
    assert rc == -4 and not h.value
