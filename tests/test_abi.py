"""The C-ABI boundary on the CPU: the library loads, exports every entry point
include/gargantua.h declares, and the Python mirror's record layouts match
the header as a C compiler lays them out. No compute calls (no GPU here)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from gargantua_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gargantua.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|uint32_t|uint64_t|const char\*)\s+(garg_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = api.load()
    names = declared_symbols()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(api.SIGNATURES), set(names) ^ set(api.SIGNATURES)


def test_pure_entry_points_without_a_device():
    lib = api.load()
    assert lib.garg_version().startswith(b"gargantua")
    assert lib.garg_request_rays(1920, 1080) == 240 * 135 * 64
    assert lib.garg_request_rays(9, 9) == 4 * 64
    assert lib.garg_bake_bytes(960, 540, 96) >= 960 * 540 * (96 * 8 + 16)


def test_null_arguments_are_refused_with_a_sentence():
    lib = api.load()
    assert lib.garg_ctx_create(0, 0, None) == 1
    assert b"out is required" in lib.garg_last_error()
    assert lib.garg_march_batch(None, None, None, 0, None, None) == 1
    assert lib.garg_last_error()


LAYOUT_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "gargantua.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
    S(garg_particle) S(garg_geodesic_settings) S(garg_march_uniforms) S(garg_ray_probe)
    S(garg_ray_probe_result) S(garg_march_counters) S(garg_device_info) S(garg_march_request) S(garg_batch_config)
    S(garg_ray_probe_fp64_result) S(garg_render_camera_fp64)
    F(garg_render_camera_fp64, forward) F(garg_render_camera_fp64, tan_half_fov)
    F(garg_march_uniforms, tile_origin) F(garg_march_uniforms, star_seed) F(garg_march_uniforms, sphere_radius)
    F(garg_march_uniforms, spin) F(garg_march_uniforms, volume_half_extent_z) F(garg_march_uniforms, geodesic)
    F(garg_march_counters, steps) F(garg_march_request, targets) F(garg_march_request, ray_begin)
    F(garg_device_info, total_bytes) F(garg_ray_probe_fp64_result, steps) F(garg_ray_probe_fp64_result, outcome)
    return 0;
}
"""


def test_record_layouts_match_the_header():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(src, "w").write(LAYOUT_PROBE)
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    c = {k: int(v) for k, v in (line.split() for line in out.splitlines())}
    py = {"garg_particle": api.Particle, "garg_geodesic_settings": api.GeodesicSettings,
          "garg_march_uniforms": api.MarchUniforms, "garg_ray_probe": api.RayProbe,
          "garg_ray_probe_result": api.RayProbeResult, "garg_march_counters": api.MarchCounters,
          "garg_device_info": api.DeviceInfo, "garg_march_request": api.MarchRequest,
          "garg_batch_config": api.BatchConfig, "garg_ray_probe_fp64_result": api.RayProbeFp64Result,
          "garg_render_camera_fp64": api.RenderCameraFp64}
    for name, cls in py.items():
        assert C.sizeof(cls) == c[name], name
    for key, value in c.items():
        if "." in key:
            struct, field = key.split(".")
            assert getattr(py[struct], field).offset == value, key
    # The reference's records keep their sizes (ShaderTypes.h:14-20, 79-92).
    assert c["garg_particle"] == 32 and c["garg_ray_probe"] == 24 and c["garg_ray_probe_result"] == 44


def test_oracle_is_not_linked_by_the_product():
    out = subprocess.run(["ldd", api.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    for root, _, files in os.walk(os.path.join(ROOT, "gargantua_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(root, f)).read()
                assert "import oracle" not in text and "oracle.h" not in text, f
