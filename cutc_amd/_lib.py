"""Locating and building the package's native libraries (in-tree, cutc_amd/lib/)."""
from __future__ import annotations

import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")


def lib_path(name: str) -> str:
    """Absolute path of a built library; raises if it has not been built."""
    p = os.path.join(LIB_DIR, name)
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build it with `make -C {PKG_DIR}` (or __graft_entry__.build())")
    return p


def build(jobs: int = 8, verbose: bool = False) -> None:
    """Compile libcutcint.so (hipcc, gfx950) and libcutcsynth.so (gcc)."""
    cmd = ["make", "-C", PKG_DIR, f"-j{jobs}"]
    r = subprocess.run(cmd, capture_output=not verbose, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{r.stdout}\n{r.stderr}")
