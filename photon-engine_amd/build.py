"""Build script for libphoton_hip.so (gfx950) and, beside it, the oracle libraries used by the tests.

Products:
  photon-engine_amd/libphoton_hip.so    HIP kernels + C-ABI (include/photon_hip.h), built with hipcc for gfx950
Test infrastructure (never linked by the product):
  oracle/liboracle.so                   C++ restatement of the reference hot path (g++)
  oracle/_ref/libphoton_ref*.so         the reference's own sources, compiled in place (oracle/Makefile.ref),
                                        only where /root/reference exists
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libphoton_hip.so"
ORACLE = ROOT / "oracle"
REFERENCE = Path(os.environ.get("PHOTON_REFERENCE", "/root/reference"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
# Files whose arithmetic must match the reference float for float (no a*b+c contraction; for display.hip the
# device code too).
STRICT_HOST = {"scene_upload.cpp", "tables.cpp", "scene_prep.cpp", "light_bvh.cpp", "display.hip"}

HIP_SOURCES = ["capi.hip", "wavefront.hip", "integrator.hip", "serve.hip", "guide.hip", "display.hip"]
HOST_SOURCES = ["bvh_build.cpp", "scene_upload.cpp", "tables.cpp", "scene_prep.cpp", "light_bvh.cpp"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    flags = list(HIP_FLAGS)
    if src.name in STRICT_HOST:
        flags.append("-ffp-contract=off")
    if src.suffix == ".cpp":
        flags = ["-x", "hip"] + flags
    _run([HIPCC, *flags, "-c", str(src), "-o", str(out)])
    return out


def build_hip(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES + HOST_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lpthread",
          "-Wl,-soname,libphoton_hip.so"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    out = ORACLE / "liboracle.so"
    srcs = sorted(str(p) for p in ORACLE.glob("oracle_*.cpp"))
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
          f"-I{ROOT / 'include'}", f"-I{ORACLE}", *srcs, "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_reference(jobs: int = 8) -> bool:
    """Compile the reference's own sources into oracle/_ref (only where the reference tree exists)."""
    if not (REFERENCE / "src").is_dir():
        return False
    _run(["make", "-f", "oracle/Makefile.ref", f"-j{jobs}", f"REF={REFERENCE}"], cwd=ROOT)
    return True


def build_all(verbose: bool = True) -> None:
    jobs = min(8, os.cpu_count() or 4)
    lib = build_hip(jobs)
    if verbose:
        print(f"built {lib}", flush=True)
    orc = build_oracle()
    if verbose:
        print(f"built {orc}", flush=True)
    if build_reference(jobs) and verbose:
        print("built oracle/_ref (reference sources)", flush=True)


if __name__ == "__main__":
    build_all()
