"""Builds the native libraries in-tree (gfx950 only)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8) -> None:
    """The product (library, C++ facade, BTEcmd), then the profiling probe (staticbtesolver_amd/profile.sh)."""
    subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}", "all", "probe"], check=True)
