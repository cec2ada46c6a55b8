"""bench.py's command-line contract on the CPU: one GPU only, so more than
one is refused with a sentence before anything touches a device."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and r.stdout == ""


def test_autotune_table_round_trip(tmp_path, monkeypatch):
    """The autotuner's table: one file per (arch, workload) under
    GARG_AUTOTUNE_DIR, the same schema `gargantua bench --sweep-batch` writes."""
    from gargantua_amd import autotune
    monkeypatch.setenv("GARG_AUTOTUNE_DIR", str(tmp_path))
    cfg = {"block": 128, "waves_per_simd": 5, "compaction_period": 16, "refill_threshold": 0, "kerr_schild": 2}
    autotune.save("gfx950:test", "headline", cfg, 3.5)
    assert autotune.load_entry("gfx950:test", "headline")["config"] == cfg
    assert autotune.load_entry("gfx950:test", "mixed") is None


def test_autotune_brick_table_round_trip(tmp_path, monkeypatch):
    """The exact splat's LDS brick shape in the same table: the saved shape
    is read back, and a missing entry or a shape the library does not build
    falls back to the default brick."""
    from gargantua_amd import api, autotune
    monkeypatch.setenv("GARG_AUTOTUNE_DIR", str(tmp_path))
    assert autotune.brick_for("gfx950:test", "splat_1M_128") == api.Volume.BRICKS[0]
    autotune.save("gfx950:test", "splat_1M_128", {"brick": [32, 8, 4]}, 0.21)
    assert autotune.brick_for("gfx950:test", "splat_1M_128") == (32, 8, 4)
    autotune.save("gfx950:test", "splat_4M_256", {"brick": [4, 4, 4]}, 0.5)
    assert autotune.brick_for("gfx950:test", "splat_4M_256") == api.Volume.BRICKS[0]


@pytest.mark.gpu
def test_measured_ceilings_are_physical(ctx):
    """The ceilings bench.py states beside its roofline: non-packed FP32
    fused multiply-adds in each operand form cannot exceed MI355X's FP32
    vector peak (157.3 TFLOP/s) and reach at least a fifth of the 78.6
    TFLOP/s of a wave64 FMA every four cycles per SIMD; no-return atomics run
    at a positive rate. Bad arguments are refused, not run."""
    from gargantua_amd import api
    for form in (api.VALU_FMA_LITERAL, api.VALU_FMA_VGPR, api.VALU_FMA_SGPR):
        tflops = ctx.valu_rate(form, 256, 1 << 18) / 1e12
        assert 15.0 < tflops <= 157.3, (form, tflops)
    assert ctx.atomic_rate(1 << 20, 4, 1 << 16) > 0.0
    with pytest.raises(api.GargError):
        ctx.valu_rate(api.VALU_FMA_LITERAL, 0, 1 << 18)
    with pytest.raises(api.GargError):
        ctx.valu_rate(3, 16, 1 << 18)
