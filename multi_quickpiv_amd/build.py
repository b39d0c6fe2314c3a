"""Build the in-tree native libraries.

    python -m multi_quickpiv_amd.build          # libmqpiv.so (HIP, gfx950)
    python -m multi_quickpiv_amd.build --oracle # also oracle/liborc.so (CPU checker)

libmqpiv.so is compiled with hipcc for gfx950 only.  Packed-FP32 instructions
are disabled for all device code (-packed-fp32-ops).
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
BUILD = PKG / "_build"
LIB = PKG / "libmqpiv.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]

SOURCES = [
    # longest transforms first so the parallel build finishes sooner
    "mqpiv_k2d_wave.hip",
    "mqpiv_k2d_f64_128.hip",
    "mqpiv_k2d_f32_128.hip",
    "mqpiv_k2d_f32_96.hip",
    "mqpiv_k2d_f64_96.hip",
    "mqpiv_k2d_f32_64.hip",
    "mqpiv_k2d_f64_64.hip",
    "mqpiv_k2d_f32_48.hip",
    "mqpiv_k2d_f64_48.hip",
    "mqpiv_k2d_f32_32.hip",
    "mqpiv_k2d_f64_32.hip",
    "mqpiv_k2d_f32_16.hip",
    "mqpiv_k2d_f64_16.hip",
    "mqpiv_k3d_f32_64.hip",
    "mqpiv_k3d_f64_64.hip",
    "mqpiv_k3d_f32_48.hip",
    "mqpiv_k3d_f64_48.hip",
    "mqpiv_k3d_f32_32.hip",
    "mqpiv_k3d_f64_32.hip",
    "mqpiv_k3d_f32_16.hip",
    "mqpiv_k3d_f64_16.hip",
    "mqpiv_kernels2d.hip",
    "mqpiv_kernels3d.hip",
    "mqpiv_sat.hip",
    "mqpiv_multipass.hip",
    "mqpiv_skip.hip",
    "mqpiv_means3d.hip",
    "mqpiv_api.cpp",
    "mqpiv_server.cpp",
    "mqpiv_post.hip",
    "mqpiv_post_api.cpp",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def _compile(src: str, incremental: bool = False, build: Path = BUILD, extra=()) -> Path:
    obj = build / (Path(src).stem + ".o")
    if incremental and obj.exists():
        deps = [CSRC / src, *CSRC.glob("*.hpp"), *CSRC.glob("*.h"), ROOT / "include" / "mqpiv.h"]
        if obj.stat().st_mtime > max(d.stat().st_mtime for d in deps):
            return obj
    cmd = [HIPCC, *DEVICE_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *DEVICE_FLAGS, *extra, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_lib(jobs: int = 8, incremental: bool = False, out: Path = LIB, extra=()) -> Path:
    """Compiles every source into ``out``.  ``extra`` adds compiler flags (a
    tuning build, e.g. ``-DMQPIV_OCC_CAP=3``); such builds use their own
    object directory next to ``out``."""
    build = BUILD if out == LIB and not extra else out.parent / (out.stem + "_build")
    build.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, incremental, build, tuple(extra)), SOURCES))
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    odir = ROOT / "oracle"
    _run(["make", "-s", "-C", str(odir)])
    return odir / "liborc.so"


def main(argv: list[str]) -> int:
    build_lib(incremental="--incremental" in argv)
    print(f"built {LIB}")
    if "--oracle" in argv:
        print(f"built {build_oracle()}")
    if "--clean" in argv:
        shutil.rmtree(BUILD, ignore_errors=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
