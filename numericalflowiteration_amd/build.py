"""Build step of the package: compiles libnufi_hip.so for gfx950 with hipcc.

Runs the Makefile next to the HIP sources (numericalflowiteration_amd/csrc);
every HIP translation unit is compiled with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")


def build(jobs: int = 4, force: bool = False) -> str:
    env = dict(os.environ)
    env.setdefault("HIPCC", "/opt/rocm/bin/hipcc")
    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True, env=env)
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", CSRC], check=True, env=env)
    return os.path.join(PKG_DIR, "libnufi_hip.so")


if __name__ == "__main__":
    print(build())
