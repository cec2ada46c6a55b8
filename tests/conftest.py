"""Shared fixtures. Tests marked `gpu` need a HIP device and the built
library; everything else runs on the CPU against the FP64 oracle."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libgargantua_hip.so")


def load_golden(name: str) -> dict:
    with open(os.path.join(GOLDEN_DIR, f"{name}.json")) as f:
        return json.load(f)


def golden_passes(g: dict, measured: float) -> bool:
    """Golden.passes (Oracle/Golden.swift:45-51)."""
    if g["kind"] == "relative":
        return abs(measured - g["expected"]) / abs(g["expected"]) <= g["tolerance"]
    if g["kind"] == "absolute":
        return abs(measured - g["expected"]) <= g["tolerance"]
    return measured == g["expected"]


@pytest.fixture(scope="session")
def ctx():
    from gargantua_amd import api
    c = api.Context()
    print(f"device {c.info().name.decode()} {c.info().arch.decode()}", flush=True)
    yield c
    c.close()
