#!/usr/bin/env python3
"""The full golden suite on this GPU, judged against the FP64 CPU oracle.

`gargantua validate` judges all eleven goldens in the product (parity and
sky against the library's FP64 GPU integrator, beaming against its FP64 GPU
render). This runner judges the same eleven (Validation/Validate.swift:17)
against the CPU oracle instead, by running the golden tests of tests/ on
the device and printing one row per golden. Exit 0 when every golden
passes, 1 otherwise. Test infrastructure: it imports the oracle.
"""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Golden -> the test that holds the GPU to it (fast integrator, compensated
# where the still policy applies).
TESTS = {
    "isco": "tests/test_gpu_sim_splat.py::test_isco_golden",
    "determinism": "tests/test_gpu_sim_splat.py::test_determinism_golden",
    "shadow": "tests/test_gpu_probes.py::test_shadow_golden[fast-compensated]",
    "deflection": "tests/test_gpu_probes.py::test_deflection_golden[fast-compensated]",
    "conservation_interactive": "tests/test_gpu_probes.py::test_conservation_goldens[fast-compensated-interactive]",
    "conservation_still": "tests/test_gpu_probes.py::test_conservation_goldens[fast-compensated-still]",
    "parity": "tests/test_gpu_probes.py::test_parity_golden[fast-compensated]",
    "beaming": "tests/test_gpu_march.py::test_beaming_golden",
    "sky": "tests/test_gpu_probes.py::test_sky_golden[fast-compensated]",
    "jump": "tests/test_gpu_march.py::test_jump_golden",
    "shadow_kerr": "tests/test_gpu_probes.py::test_shadow_kerr_golden[fast-compensated]",
}


def main() -> int:
    cmd = [sys.executable, "-m", "pytest", "-q", "-s", "-p", "no:cacheprovider", "-rA", *TESTS.values()]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    passed = set(re.findall(r"^PASSED (\S+)", r.stdout, re.M))
    failed = set(re.findall(r"^FAILED (\S+)", r.stdout, re.M))
    ok = True
    print(f"{'golden':26s} status")
    for name, test in TESTS.items():
        status = "pass" if test in passed else ("FAIL" if any(f.startswith(test) for f in failed) else "error")
        ok &= status == "pass"
        print(f"{name:26s} {status}")
    if not ok:
        print(r.stdout[-4000:])
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
