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


@pytest.mark.parametrize("counts", [
    [(1 << 64) - 1, 1],          # the sum wraps to 0 in 64 bits
    [(1 << 63), (1 << 63)],      # so does this one
    [1 << 40],                   # one instance beyond the slot range
    [0xffffff00, 1],             # the total just beyond it
])
def test_batch_create_rejects_agent_totals_before_any_launch(slm, counts):
    descs = (slm._native.InstanceDesc * len(counts))()
    for d, n in zip(descs, counts):
        d.width, d.height, d.agent_count, d.seed = 16, 16, n, 1
        d.params = slm.default_params()
    cfg = slm.default_batch_config()
    h = ctypes.c_void_p()
    rc = slm.lib().slm_batch_create(ctypes.byref(cfg), descs, len(counts), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert "total agents" in slm.lib().slm_last_error().decode()


def test_null_handle_calls_fail_cleanly(slm):
    L = slm.lib()
    assert L.slm_step(None, 1) == -1
    assert L.slm_clear(None) == -1
    assert L.slm_destroy(None) == -1
    assert L.slm_resize(None, 4, 4) == -1
    assert L.slm_total_mass(None, None) == -1
    t = slm.Tuning()
    assert L.slm_get_tuning(None, ctypes.byref(t)) == -1
    assert L.slm_set_tuning(None, ctypes.byref(t)) == -1
    assert L.slm_autotune(None, 5, ctypes.byref(t)) == -1
    assert L.slm_batch_autotune(None, 5, None) == -1
    assert "NULL" in L.slm_last_error().decode()


def test_no_device_is_reported_not_crashed(slm):
    if slm.device_count() > 0:
        pytest.skip("a GPU is visible")
    rc, msg, h = _create(slm)
    assert rc == -4 and not h.value


CPP_PROBE = r"""
#include <cstdio>
#include "slimemold.hpp"
int main() {
  auto m = slime_mold::init_preset_manager();
  std::printf("[");
  bool first = true;
  for (const auto& n : m.get_preset_names()) {
    const auto& s = m.get_preset(n)->settings;
    std::printf("%s{\"name\": \"%s\", \"settings\": {\"agent_count\": %zu, "
                "\"agent_jitter\": %.9g, \"agent_sensor_angle\": %.9g, "
                "\"agent_sensor_distance\": %.9g, \"agent_speed_max\": %.9g, "
                "\"agent_speed_min\": %.9g, \"agent_turn_speed\": %.9g, \"blur_radius\": %.9g, "
                "\"blur_sigma\": %.9g, \"pheromone_decay_factor\": %.9g, "
                "\"pheromone_deposition_amount\": %.9g, \"pheromone_diffusion_rate\": %.9g, "
                "\"window_height\": %u, \"window_width\": %u}}",
                first ? "" : ", ", n.c_str(), s.agent_count, s.agent_jitter,
                s.agent_sensor_angle, s.agent_sensor_distance, s.agent_speed_max,
                s.agent_speed_min, s.agent_turn_speed, s.blur_radius, s.blur_sigma,
                s.pheromone_decay_factor, s.pheromone_deposition_amount,
                s.pheromone_diffusion_rate, s.window_height, s.window_width);
    first = false;
  }
  std::printf("]\n");
  if (m.get_preset("No such preset") != nullptr) return 3;
  std::vector<uint8_t> raw(768);
  for (int i = 0; i < 768; i++) raw[i] = static_cast<uint8_t>(i % 256);
  auto lut = slime_mold::LutData::from_bytes("ramp", raw);
  lut.reverse();
  if (lut.bytes()[0] != 255 || lut.bytes()[256] != 255 || lut.bytes()[767] != 0) return 4;
  int code = 0;
  try {  // with no device this must throw, never crash
    slime_mold::Simulation sim(64, 64, 16, 1, slime_mold::Settings{});
    sim.step(1);
    sim.synchronize();
  } catch (const slime_mold::Error& e) {
    code = e.code;
  }
  std::fprintf(stderr, "create: %d\n", code);
  return 0;
}
"""


def test_cpp_mirror_compiles_and_matches_reference_presets(slm):
    """slime-mold_amd/cpp/slimemold.hpp: C++17 header over the C ABI; its
    preset table must equal the reference's (src/presets.rs:45-155)."""
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    import tempfile
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "presets.json")))["presets"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        with open(src, "w") as f:
            f.write(CPP_PROBE)
        libdir = os.path.join(ROOT, "slime-mold_amd", "slimemold")
        subprocess.check_call([cxx, "-std=c++17", "-Wall", "-Wextra", "-Werror",
                               "-I", os.path.join(ROOT, "slime-mold_amd", "cpp"), src, "-o",
                               os.path.join(d, "t"), "-L", libdir, "-lslimemold_hip",
                               f"-Wl,-rpath,{libdir}"])
        r = subprocess.run([os.path.join(d, "t")], capture_output=True, text=True, check=True)
    got = json.loads(r.stdout)
    assert [p["name"] for p in got] == [p["name"] for p in ref]
    f32 = lambda v: float(ctypes.c_float(v).value)  # noqa: E731
    for g, e in zip(got, ref):
        for k, v in e["settings"].items():
            if isinstance(v, float):  # %.9g round-trips a float32 exactly
                assert f32(g["settings"][k]) == f32(v), (g["name"], k)
            else:
                assert g["settings"][k] == v, (g["name"], k)
    code = int(r.stderr.split("create:")[1])
    if slm.device_count() == 0:
        assert code == -4
    else:
        assert code == 0


def test_snapshot_format_is_checked_before_any_device_use(slm):
    """A snapshot built by an independent encoder of the documented layout
    passes the library's checks (no device here: -4 only after them); every
    corruption is rejected as an invalid argument first."""
    import snapshot_util as su
    import numpy as np

    L = slm.lib()
    w, h, n = 8, 5, 3
    agents = np.array([[1.0, 2.0, 0.5, 30.0], [7.5, 4.9, 3.0, 40.0], [0.0, 0.0, 6.0, 50.0]],
                      np.float32)
    trail = np.linspace(0, 1, w * h, dtype=np.float32)
    good = su.build(w, h, agents, trail, steps=12, reassigns=1, seed=9)

    def create(data):
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        out = ctypes.c_void_p()
        rc = L.slm_create_from_snapshot(buf, len(data), 0, ctypes.byref(out))
        if rc == 0:
            L.slm_destroy(out)
        return rc, L.slm_last_error().decode()

    rc, msg = create(good)
    assert rc in (0, -4), msg
    flipped = bytearray(good)
    flipped[200] ^= 1
    assert create(bytes(flipped)) == (-1, "snapshot checksum mismatch")
    assert create(b"XLMSNAP1" + good[8:])[0] == -1
    assert create(good[:-8])[0] == -1
    assert create(su.build(w, h, agents, trail, version=2))[0] == -5
    bad_trail = su.build(w, h, agents, trail[:-1])  # fewer cells than width*height
    assert create(bad_trail)[0] == -1
    assert create(su.build(w, h, agents, trail, flags=4))[0] == -1
    assert create(su.build(w, h, agents, np.zeros(w * h, np.uint16), storage=3))[0] == -1


def test_ctypes_mirror_matches_the_c_layout(slm):
    """Every public struct's size and field offsets, as the C compiler lays them
    out from include/slimemold.h, equal the Python binding's ctypes mirror."""
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no gcc")
    N = slm._native
    structs = {"slm_params": N.Params, "slm_config": N.Config, "slm_info": N.Info,
               "slm_timing": N.Timing, "slm_instance_desc": N.InstanceDesc,
               "slm_batch_config": N.BatchConfig, "slm_batch_info": N.BatchInfo,
               "slm_tuning": N.Tuning}
    lines = ['#include "slimemold.h"', "#include <stddef.h>", "#include <stdio.h>",
             "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        with open(c, "w") as f:
            f.write("\n".join(lines) + "\n")
        subprocess.check_call([cc, "-std=c99", "-Wall", "-Werror", "-I",
                               os.path.join(ROOT, "include"), c, "-o", os.path.join(d, "layout")])
        out = subprocess.check_output([os.path.join(d, "layout")]).decode().split("\n")
    got = {}
    for line in filter(None, out):
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_python_flag_constants_match_the_header(slm):
    import slimemold._native as N
    text = open(HEADER).read()
    defs = dict(re.findall(r"#define SLM_(FLAG_[A-Z0-9_]+) (0x[0-9a-f]+)u", text))
    assert len(defs) >= 5
    for name, value in defs.items():
        assert getattr(N, name) == int(value, 16), name
