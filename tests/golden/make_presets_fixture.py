"""Extract the reference's default settings and presets as data.

Reads src/settings.rs (constants, lines 3-27) and src/presets.rs (presets,
lines 45-155) of the reference checkout as text and writes presets.json:
{"defaults": {field: value}, "presets": [{"name": ..., "settings": {field: value}}]}
with every field resolved (`..Settings::default()` filled from the defaults).
Run where the reference is mounted:
    python tests/golden/make_presets_fixture.py /root/reference
"""
import json
import os
import re
import sys

CONST_TO_FIELD = {
    "AGENT_COUNT": "agent_count", "AGENT_SPEED_MIN": "agent_speed_min",
    "AGENT_SPEED_MAX": "agent_speed_max", "AGENT_TURN_SPEED": "agent_turn_speed",
    "DEPOSITION_AMOUNT": "pheromone_deposition_amount", "AGENT_JITTER": "agent_jitter",
    "AGENT_SENSOR_ANGLE": "agent_sensor_angle", "AGENT_SENSOR_DISTANCE": "agent_sensor_distance",
    "DECAY_FACTOR": "pheromone_decay_factor", "DIFFUSION_RATE": "pheromone_diffusion_rate",
    "BLUR_RADIUS": "blur_radius", "BLUR_SIGMA": "blur_sigma", "DEFAULT_WIDTH": "window_width",
    "DEFAULT_HEIGHT": "window_height",
}


def num(tok):
    tok = tok.replace("_", "")
    return float(tok) if ("." in tok or "e" in tok) else int(tok)


def main(ref):
    settings = open(os.path.join(ref, "src", "settings.rs")).read()
    defaults = {}
    for name, val in re.findall(r"pub const (\w+): \w+ = ([0-9_.]+);", settings):
        if name in CONST_TO_FIELD:
            defaults[CONST_TO_FIELD[name]] = num(val)
    presets_src = open(os.path.join(ref, "src", "presets.rs")).read()
    presets = []
    for m in re.finditer(r'Preset::new\(\s*"([^"]+)"\.to_string\(\),\s*(Settings::default\(\)|Settings \{(.*?)\})', presets_src, re.S):
        s = dict(defaults)
        if m.group(3):
            for field, val in re.findall(r"(\w+):\s*([0-9_.]+),", m.group(3)):
                s[field] = num(val)
        presets.append({"name": m.group(1), "settings": s})
    out = {"source": "src/settings.rs:3-27, src/presets.rs:45-155", "defaults": defaults,
           "presets": presets}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "presets.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {path}: {len(presets)} presets")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
