"""Builds libqarp_hip.so in-tree with hipcc for gfx950.

    python -m openqarp_amd.build            # compile every source, link
    python -m openqarp_amd.build --jobs 8

Objects go to openqarp_amd/_build/, the library to openqarp_amd/_lib/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB_DIR = PKG / "_lib"
LIB = LIB_DIR / "libqarp_hip.so"
ARCH = os.environ.get("QHIP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    # No packed-FP32 instructions anywhere in device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Matrix-core accumulators in ordinary VGPRs: the results feed VALU adds
    # and LDS stores directly instead of through v_accvgpr_read copies.
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def sources() -> list[Path]:
    return sorted(list(SRC.glob("*.hip")) + list(SRC.glob("*.cpp")))


def compile_one(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    cmd = [hipcc(), *COMMON, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".hip":
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def stale(src: Path) -> bool:
    """An object is out of date when its source or any header is newer."""
    obj = OBJ / (src.name + ".o")
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    headers = list(SRC.glob("*.h")) + list((PKG.parent / "include").glob("*.h"))
    return any(p.stat().st_mtime > t for p in [src, *headers])


def build(jobs: int = 0, verbose: bool = False, changed_only: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    all_srcs = sources()
    srcs = [s for s in all_srcs if stale(s)] if changed_only else all_srcs
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(compile_one, srcs))
    objs = [OBJ / (src.name + ".o") for src in all_srcs]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", "-pthread", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--changed", action="store_true", help="recompile only sources newer than their objects")
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True, changed_only=a.changed)
    return 0


if __name__ == "__main__":
    sys.exit(main())
