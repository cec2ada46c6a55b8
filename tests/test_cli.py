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


def run(*args, timeout=600, cwd=ROOT, env=None):
    full = dict(os.environ, **(env or {}))
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd, env=full)


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
    (["--form", "boyer-lindquist"], 64, "form must be fast, kerr-schild or polar."),
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
    rows = {line.split()[0]: line.split()[-1] for line in r.stdout.splitlines()[1:] if line.strip()}
    # All eleven judged in the product: parity and sky against the library's
    # FP64 GPU integrator, beaming against its FP64 GPU render.
    assert len(rows) == 11, rows
    assert set(rows.values()) == {"pass"}, rows


@pytest.mark.gpu
@pytest.mark.parametrize("tick_rays,policy,form", [(4 << 20, "srpt", "polar"), (50_000, "srpt", "polar"),
                                              (50_000, "fifo", "polar"), (50_000, "srpt", "kerr-schild"),
                                              (50_000, "srpt", "fast")])
def test_serve_results_equal_alone_and_in_the_mix(tmp_path, tick_rays, policy, form):
    """Every request re-rendered alone equals its render in the mix; a small
    per-tick ray budget splits frames over many ticks and leaves requests
    waiting for budget, which each policy hands out in its own order; the
    spinning requests in each of the three forms (polar is the default)."""
    report = tmp_path / "serve.json"
    r = run("serve", "--preset", "small", "--rate", "6", "--seconds", "3", "--frames", "3", "--verify", "1",
            "--tick-rays", str(tick_rays), "--policy", policy, "--report", str(report),
            *(["--form", form] if form != "polar" else []))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = json.loads(report.read_text())
    assert rep["completed"] > 0 and rep["verified"] > 0 and rep["mismatched"] == 0
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0
    assert rep["policy"] == policy and rep["form"] == form


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
    --max-wait the loop sheds at admission what would wait longer (and
    what the HBM budget or the batch's slots hold back longer than that),
    and finishes everything it admits."""
    report = tmp_path / "serve.json"
    r = run("serve", "--preset", "small", "--rate", "20", "--seconds", "4", "--frames", "240", "--max-wait", "2",
            "--report", str(report))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = json.loads(report.read_text())
    assert rep["shed"] > 0 and rep["unfinished"] == 0
    assert rep["completed"] + rep["shed"] + rep["refused"] == rep["requests"]


# The report of Bench.swift:236-264 (bench/results/2026-08-23-204508.json):
# its keys, the three statistics of every pass, and the two frame figures.
REPORT_KEYS = {"bake", "date", "device", "frames", "iterations", "march", "os", "output", "particles", "passes",
               "preset", "verdict", "version", "volume"}


@pytest.mark.gpu
def test_bench_writes_the_reference_report(tmp_path):
    r = run("bench", "--preset", "small", cwd=str(tmp_path))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    reports = list((tmp_path / "bench" / "results").glob("*.json"))
    assert len(reports) == 1
    rep = json.loads(reports[0].read_text())
    assert REPORT_KEYS <= set(rep), REPORT_KEYS - set(rep)
    for name in ("sim", "splat", "march", "bake", "walk", "present"):
        assert set(rep["passes"][name]) == {"mean_ms", "p50_ms", "max_ms"}
        assert 0 < rep["passes"][name]["p50_ms"] <= rep["passes"][name]["max_ms"]
    assert set(rep["frames"]) == {"march_ms", "bake_ms"} and set(rep["bake"]) == {"bytes", "overflow_rays"}
    assert rep["iterations"] == 30 and rep["preset"] == "small" and rep["march"] == [640, 360]
    assert rep["verdict"].startswith("march frame") and "bake frame" in rep["verdict"]


@pytest.mark.gpu
def test_bench_sweep_is_the_reference_sweep():
    """Bench.swift:49-124: fifteen step policies (march time, probe drift,
    exhausted probes, image difference from the spec policy), the splat's
    phases, then the launch shapes."""
    r = run("bench", "--preset", "small", "--sweep")
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [l for l in r.stdout.splitlines() if l.startswith(("fixed", "prop"))]
    assert len(rows) == 15
    spec = [l for l in rows if "(spec)" in l][0].split()
    assert float(spec[-1]) == 0.0  # the spec policy against itself
    assert "bin " in r.stdout and "resolve " in r.stdout and r.stdout.count("waves/SIMD  march") == 9


@pytest.mark.gpu
def test_bench_sweep_batch_writes_the_table_the_autotuner_reads(tmp_path, ctx):
    from gargantua_amd import autotune
    r = run("bench", "--preset", "small", "--sweep-batch", env={"GARG_AUTOTUNE_DIR": str(tmp_path)})
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    arch = ctx.info().arch.decode()
    entry = json.loads((tmp_path / arch / "headline.json").read_text())
    assert entry["ms"] > 0 and entry["config"]["kerr_schild"] == 2
    old = os.environ.get("GARG_AUTOTUNE_DIR")
    os.environ["GARG_AUTOTUNE_DIR"] = str(tmp_path)
    try:
        cfg = autotune.best(ctx, "headline")
    finally:
        if old is None:
            del os.environ["GARG_AUTOTUNE_DIR"]
        else:
            os.environ["GARG_AUTOTUNE_DIR"] = old
    for k, v in entry["config"].items():
        assert getattr(cfg, k) == v, k
    # Threshold 0 (whole tiles) and five waves per SIMD are in the grid.
    assert " 0  " in r.stdout and any(l.split()[:2] == ["256", "5"] for l in r.stdout.splitlines() if l[:3].strip().isdigit())
