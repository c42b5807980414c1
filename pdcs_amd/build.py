"""Build libpdcs_hip.so in-tree with hipcc for gfx950 (no CMake, no JIT cache)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> str:
    """Compile pdcs_amd/csrc into pdcs_amd/libpdcs_hip.so; returns its path."""
    csrc = os.path.join(_HERE, "csrc")
    out = os.path.join(_HERE, "libpdcs_hip.so")
    # always rebuild: remove the target so make cannot skip
    if os.path.exists(out):
        os.remove(out)
    r = subprocess.run(["make", "-C", csrc], capture_output=not verbose, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc build failed:\n{r.stdout}\n{r.stderr}")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
