"""Build recipe for the native parts of spac_amd.

Outputs (in-tree, git-ignored, shipped with the source tree):
  spac_amd/lib/libspac_fk.so   HIP kernels (gfx950) + C ABI + C++ plugin surface
  oracle/_build/libfk_oracle.so  CPU restatement used only by tests and bench

Every floating-point translation unit is compiled with -ffp-contract=off so that
the CPU restatement and the device code perform the same IEEE operations, and
device code with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "spac_amd" / "csrc"
CPP = ROOT / "spac_amd" / "cpp"
LIBDIR = ROOT / "spac_amd" / "lib"
ORACLE = ROOT / "oracle"
ORACLE_BUILD = ORACLE / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPAC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    f"-I{CPP}",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def hip_sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def cpp_sources() -> list[Path]:
    return sorted(CPP.glob("*.cpp")) if CPP.exists() else []


def build_library() -> Path:
    """Compile every .hip/.cpp source into spac_amd/lib/libspac_fk.so."""
    LIBDIR.mkdir(parents=True, exist_ok=True)
    objdir = ROOT / "_build" / "obj"
    objdir.mkdir(parents=True, exist_ok=True)
    objs = []
    for src in hip_sources():
        obj = objdir / (src.stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", src, "-o", obj])
        objs.append(obj)
    for src in cpp_sources():
        obj = objdir / (src.stem + ".cpp.o")
        _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
        objs.append(obj)
    out = LIBDIR / "libspac_fk.so"
    tmp = LIBDIR / "libspac_fk.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    """Compile the CPU restatement (C, OpenMP) into oracle/_build/libfk_oracle.so."""
    ORACLE_BUILD.mkdir(parents=True, exist_ok=True)
    out = ORACLE_BUILD / "libfk_oracle.so"
    tmp = ORACLE_BUILD / "libfk_oracle.so.tmp"
    srcs = sorted(ORACLE.glob("*.c"))
    _run(["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-Wall", "-Wno-unused-function", *srcs, "-o", tmp, "-lm"])
    os.replace(tmp, out)
    return out


def build_surface_driver() -> Path:
    """The C++ driver of tests/cpp (links the library like a reference caller would)."""
    out_dir = ROOT / "tests" / "cpp" / "_bin"
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / "surface_driver"
    _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{CPP}", f"-I{CSRC}", f"-I{ROOT / 'include'}",
          ROOT / "tests" / "cpp" / "surface_driver.cpp", f"-L{LIBDIR}", "-lspac_fk", f"-Wl,-rpath,{LIBDIR}",
          "-Wl,-rpath,$ORIGIN/../../../spac_amd/lib", "-o", out])
    return out


def build_all() -> None:
    build_oracle()
    build_library()
    build_surface_driver()
    from oracle import build_ref  # noqa: F401  (oracle/_ref: reference sources, when present)
    build_ref.build_if_present()


if __name__ == "__main__":
    sys.path.insert(0, str(ROOT))
    what = sys.argv[1:] or ["all"]
    if "oracle" in what:
        build_oracle()
    if "lib" in what:
        build_library()
    if "driver" in what:
        build_surface_driver()
    if "all" in what:
        build_all()
