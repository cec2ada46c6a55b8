"""The gargantua command line: option grammar and exit codes on the CPU
(Options.swift:64-147, Exit.swift:5-17); still, validate, bench and serve on
the GPU."""
import json
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gargantua_amd", "bin", "gargantua")


def run(*args, timeout=600):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout, cwd=ROOT)


@pytest.mark.parametrize("args,code,sentence", [
    (["version"], 0, ""),
    (["--volume", "100"], 64, "volume must be 96, 128, 192 or 256."),
    (["--spin", "0.999"], 64, "spin must be between 0 and 0.998."),
    (["--preset", "huge"], 64, "unknown preset huge"),
    (["--strategy", "walk"], 64, "unknown strategy walk"),
    (["--fps-cap", "45"], 64, "fps cap must be 30 or 60."),
    (["--camera", "1,2,70"], 64, "camera needs azimuth,elevation,distance"),
    (["--seed", "-3"], 64, "seed must be an unsigned 64 bit integer"),
    (["--particles"], 64, "--particles needs a value."),
    (["--bogus"], 64, "unknown argument --bogus."),
    (["serve", "--policy", "lifo"], 64, "unknown policy lifo; choose srpt or fifo."),
    (["serve", "--bake-share", "1.5"], 64, "bake-share needs a number between 0 and 1."),
    (["serve", "--max-wait", "0"], 64, "max-wait needs a positive number."),
    (["--kerr-schild", "--spin", "1.5"], 64, "spin must be between 0 and 0.998."),
    ([], 2, "macOS only"),
])
def test_option_grammar_and_exit_codes(args, code, sentence):
    r = run(*args)
    assert r.returncode == code, r.stderr
    assert sentence in (r.stderr if code else r.stdout)


def test_constants_command_lists_the_table():
    r = run("constants")
    assert r.returncode == 0
    assert "#define B_CRITICAL" in r.stdout and "5.196152422706632" in r.stdout


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, width = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(kind + body)
        if kind == b"IHDR":
            width, height = struct.unpack(">II", body[:8])
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(height, width * 4 + 1)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(height, width, 4)


@pytest.mark.gpu
def test_still_writes_a_png(tmp_path):
    out = tmp_path / "still.png"
    r = run("--preset", "small", "--still", str(out), "--width", "320", "--height", "180", "--sequence", "2")
    assert r.returncode == 0, r.stderr
    img = read_png(out)
    assert img.shape == (180, 320, 4) and img[..., :3].max() > 0 and np.all(img[..., 3] == 255)
    for f in range(2):
        assert read_png(tmp_path / f"seq-{f:03d}.png").shape == (540, 960, 4)


@pytest.mark.gpu
def test_still_in_kerr_schild_form(tmp_path):
    out = tmp_path / "still.png"
    r = run("--preset", "small", "--spin", "0.9", "--kerr-schild", "--still", str(out), "--width", "320",
            "--height", "180")
    assert r.returncode == 0, r.stderr
    img = read_png(out)
    assert img.shape == (180, 320, 4) and img[..., :3].max() > 0


@pytest.mark.gpu
def test_validate_passes_every_judged_golden():
    r = run("--preset", "small", "validate")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" pass") >= 8


@pytest.mark.gpu
@pytest.mark.parametrize("tick_rays,policy,kerr_schild", [(4 << 20, "srpt", False), (50_000, "srpt", False),
                                                     (50_000, "fifo", False), (50_000, "srpt", True)])
def test_serve_results_equal_alone_and_in_the_mix(tmp_path, tick_rays, policy, kerr_schild):
    """Every request re-rendered alone equals its render in the mix; a small
    per-tick ray budget splits frames over many ticks and leaves requests
    waiting for budget, which each policy hands out in its own order; the
    spinning requests also in the Kerr-Schild form."""
    report = tmp_path / "serve.json"
    r = run("serve", "--preset", "small", "--rate", "6", "--seconds", "3", "--frames", "3", "--verify", "1",
            "--tick-rays", str(tick_rays), "--policy", policy, "--report", str(report),
            *(["--kerr-schild"] if kerr_schild else []))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = json.loads(report.read_text())
    assert rep["completed"] > 0 and rep["verified"] > 0 and rep["mismatched"] == 0
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0
    assert rep["policy"] == policy and rep["kerr_schild"] == kerr_schild


@pytest.mark.gpu
def test_serve_bakes_static_stills_and_verifies_them_alone(tmp_path):
    """With --strategy bake every still of this stream is a static
    Schwarzschild still (--bake-share 1): it bakes once with the still policy
    and walks; re-baked and walked alone it gives the same bits."""
    report = tmp_path / "serve.json"
    r = run("serve", "--preset", "small", "--strategy", "bake", "--bake-share", "1", "--rate", "6", "--seconds", "3",
            "--frames", "3", "--verify", "1", "--report", str(report))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = json.loads(report.read_text())
    assert rep["baked"] > 0 and rep["verified_baked"] > 0 and rep["mismatched"] == 0
    assert rep["completed"] == rep["requests"] - rep["refused"] - rep["shed"]


@pytest.mark.gpu
def test_serve_max_preset_at_twenty_requests_per_second(tmp_path):
    """The serving configuration on its own disk state, shortened: the max
    preset (4M particles, 256^3) advancing every tick, Poisson arrivals at 20
    req/s for 10 s, stills and 8-frame orbits of every resolution, spin and
    budget, a quarter of the stills baked (up to 96 samples per ray, so some
    rays truncate at 256^3's finer spacing), every fifth request re-rendered
    alone and compared bit for bit."""
    report = tmp_path / "serve.json"
    r = run("serve", "--preset", "max", "--strategy", "bake", "--rate", "20", "--seconds", "10", "--frames", "8",
            "--verify", "5", "--report", str(report))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = json.loads(report.read_text())
    assert rep["requests"] > 100 and rep["completed"] == rep["requests"] and rep["unfinished"] == 0
    assert rep["verified"] >= rep["requests"] // 5 and rep["mismatched"] == 0
    assert rep["baked"] > 0 and rep["truncated_rays"] > 0
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0 and rep["steps_per_s"] > 1e10


@pytest.mark.gpu
def test_serve_sheds_what_it_cannot_finish_in_time(tmp_path):
    """240-frame orbits at 20 req/s offer more than the GPU marches; with
    --max-wait the loop sheds at admission what would wait longer, and
    finishes everything it admits."""
    report = tmp_path / "serve.json"
    r = run("serve", "--preset", "small", "--rate", "20", "--seconds", "4", "--frames", "240", "--max-wait", "2",
            "--report", str(report))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = json.loads(report.read_text())
    assert rep["shed"] > 0 and rep["unfinished"] == 0
    assert rep["completed"] + rep["shed"] + rep["refused"] == rep["requests"]
