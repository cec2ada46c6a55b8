"""gargantua for MI355X: the per-frame compute passes of the gargantua black
hole renderer as hand-written gfx950 HIP kernels behind a C ABI
(include/gargantua.h, libgargantua_hip.so), with this Python mirror of the
host interface in gargantua_amd.api."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8) -> str:
    """Compile libgargantua_hip.so for gfx950 (hipcc, in-tree)."""
    subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}"], check=True)
    return os.path.join(HERE, "libgargantua_hip.so")
