"""Build libqcu.so (hipcc, gfx950) and the CPU oracle."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)


def build_library(force: bool = True, jobs: int = 8) -> str:
    cmd = ["make", f"-j{jobs}", "-C", os.path.join(_HERE, "csrc")]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return os.path.join(_HERE, "libqcu.so")


def build_oracle() -> None:
    subprocess.run(["make", "-B", "-s", "-C", os.path.join(_ROOT, "oracle")], check=True)
