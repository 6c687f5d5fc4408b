"""Recipe for oracle/_ref (reference sources compiled directly), when buildable.

The reference's Feynman–Kac path (Solvers/FeynmanKacSolver.cpp) includes
<eigen3/Eigen/Core> and <boost/random.hpp>.  Neither library is installed in
this image and there is no network, so the path cannot be compiled from its own
sources without writing stand-ins for those headers — which this project does
not do.  The recipe therefore only probes and reports; the CPU restatement in
oracle/fk_oracle.c is pinned by the checks listed in tests/test_oracle.py.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

REF = Path("/root/reference")


def probe() -> str:
    """Return why the reference path is not buildable here (empty if it is)."""
    if not REF.exists():
        return "reference tree absent"
    gxx = shutil.which("g++")
    if gxx is None:
        return "g++ absent"
    src = REF / "Solvers" / "FeynmanKacSolver.cpp"
    r = subprocess.run([gxx, "-std=c++14", "-fsyntax-only", str(src)], capture_output=True, text=True)
    if r.returncode != 0:
        first = (r.stderr.strip().splitlines() or ["unknown error"])[0]
        return f"FeynmanKacSolver.cpp does not compile here: {first}"
    return ""


def build_if_present() -> None:
    reason = probe()
    if reason:
        print(f"oracle/_ref not built: {reason}", flush=True)
