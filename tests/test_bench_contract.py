"""bench.py's command-line contract on the CPU: one GPU only, so more than
one is refused with a sentence before anything touches a device."""
import os
import subprocess
import sys

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
