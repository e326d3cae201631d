"""The light and metallic-roughness pins catch the mistakes they are meant to
catch.  CPU only.

Each case plants one plausible mistake in a copy of oracle/oracle.c (a text
replacement that must match exactly), builds it with the Makefile's flags
into a temporary library and runs the named checks of
tests/test_oracle_light_pins.py against it: every one of them must fail.
The unmodified oracle passes them (test_oracle_light_pins.py)."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import test_oracle_light_pins as pins

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
CFLAGS = ["-std=c11", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-D_POSIX_C_SOURCE=200809L", "-shared"]

F_D = "spec f_d = smul(smulf(ssub(S1(1.0f), F), 1.0f - b->metallic), smulf(b->color, orc_inv_pi));"

# name -> (checks that must fail, [(old text, new text, occurrences)])
PLANTS = {
    "spot cone angles swapped": (
        [pins.check_spot_light],
        [("smooth_step(x->cos_max, x->cos_min, cos_t)", "smooth_step(x->cos_min, x->cos_max, cos_t)", 1)]),
    "spot falloff linear": (
        [pins.check_spot_light],
        [("smooth_step(x->cos_max, x->cos_min, cos_t)",
          "clampf_((cos_t - x->cos_max) / (x->cos_min - x->cos_max), 0.0f, 1.0f)", 1)]),
    "spot frame not negated": (
        [pins.check_spot_light],
        [("float cos_t = -vdot(wi, x->z);", "float cos_t = vdot(wi, x->z);", 1)]),
    "spot intensity over d, not d^2": (
        [pins.check_spot_light],
        [("ls->Li = sdivf(l, distance * distance);", "ls->Li = sdivf(l, distance);", 1)]),
    "area sample pdf without the square": (
        [pins.check_area_light_sampling, pins.check_area_light_hit_pdf, pins.check_area_light_direct],
        [("pdf *= dist2 / cosine;", "pdf *= sqrtf(dist2) / cosine;", 1)]),
    "area sample pdf without the cosine": (
        [pins.check_area_light_sampling, pins.check_area_light_direct],
        [("pdf *= dist2 / cosine;", "pdf *= dist2;", 1)]),
    "area sample pdf without the area": (
        [pins.check_area_light_sampling, pins.check_area_light_direct],
        [("float pdf = 1.0f / area;", "float pdf = 1.0f;", 1)]),
    "triangle sample not folded into the triangle": (
        [pins.check_area_light_sampling],
        [("if (uu + vv > 1.0f) {", "if (0) {", 1)]),
    "emitter-hit pdf without the square": (
        [pins.check_area_light_hit_pdf, pins.check_area_light_direct],
        [("float distance_squared = is->t * is->t * vlen2(d);", "float distance_squared = is->t * vlen2(d);", 1)]),
    "emitter-hit pdf without the area": (
        [pins.check_area_light_hit_pdf, pins.check_area_light_direct],
        [("return distance_squared / (cosine * triangle_area(s, tri));", "return distance_squared / cosine;", 1)]),
    "metallic-roughness f0 lerp reversed": (
        [pins.check_metallic_roughness_f],
        [("return sadd(smulf(S1(0.04f), 1.0f - b->metallic), smulf(b->color, b->metallic));",
          "return sadd(smulf(S1(0.04f), b->metallic), smulf(b->color, 1.0f - b->metallic));", 1)]),
    "metallic-roughness diffuse without (1 - F)": (
        [pins.check_metallic_roughness_f],
        [(F_D, "spec f_d = smulf(smulf(b->color, orc_inv_pi), 1.0f - b->metallic);", 2)]),
    "metallic-roughness diffuse without (1 - metallic)": (
        [pins.check_metallic_roughness_f],
        [(F_D, "spec f_d = smul(ssub(S1(1.0f), F), smulf(b->color, orc_inv_pi));", 2)]),
    "metallic-roughness Schlick exponent 4": (
        [pins.check_metallic_roughness_f],
        [("float p = o_powf(1.0f - c, 5.0f);", "float p = o_powf(1.0f - c, 4.0f);", 1)]),
}


def _build(name, edits, out_dir):
    with open(os.path.join(ORACLE, "oracle.c")) as f:
        src = f.read()
    for old, new, count in edits:
        assert src.count(old) == count, f"{name}: '{old}' occurs {src.count(old)} times, not {count}"
        src = src.replace(old, new)
    tag = "".join(c if c.isalnum() else "_" for c in name)
    c_path = os.path.join(out_dir, f"planted_{tag}.c")
    so_path = os.path.join(out_dir, f"liboracle_{tag}.so")
    with open(c_path, "w") as f:
        f.write(src)
    subprocess.run(["gcc", *CFLAGS, "-I", ORACLE, "-o", so_path, c_path, "-lm", "-lpthread"], check=True,
                   capture_output=True)
    return so_path


@pytest.fixture(scope="module")
def planted_libs(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("planted"))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = {name: ex.submit(_build, name, edits, out) for name, (_, edits) in PLANTS.items()}
        return {name: f.result() for name, f in futs.items()}


@pytest.mark.parametrize("name", list(PLANTS))
def test_planted_mistake_fails_its_pins(planted_libs, name):
    from oracle import load_oracle_from

    lib = load_oracle_from(planted_libs[name])
    for check in PLANTS[name][0]:
        with pytest.raises(AssertionError):
            check(lib)
