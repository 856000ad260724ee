"""Builds fastload_amd/libfastload.so in-tree (csrc/: HIP sources with hipcc for gfx950, host-only
sources with g++, linked by hipcc) and the measurement-only comparator library
fastload_amd/libfastload_compare.so (csrc_compare/, linked with rocSPARSE). Run `python fastload_amd/build.py` (or __graft_entry__.build());
the script does not import the package, so a stale library never blocks a rebuild.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libfastload.so"
CSRC_COMPARE = PKG / "csrc_compare"
LIB_COMPARE = PKG / "libfastload_compare.so"
ARCH = os.environ.get("FASTLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

# -packed-fp32-ops: no v_pk_*_f32 generation (those forms are unreliable on these boxes).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-march=x86-64-v3"]
# per-file additions: the host SpMV must round like the reference's host build (no FMA contraction)
FILE_FLAGS = {"fl_cpu.cpp": ["-ffp-contract=off", "-pthread"]}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(map(str, cmd))}\n{out}")
    return out


def _compile(src: Path) -> Path:
    obj = OBJ / (src.parent.name + "_" + src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = [CXX, *CXX_FLAGS, *FILE_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def _link(out: Path, objs, libs):
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), *libs])
    os.replace(tmp, out)


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    cmp_srcs = sorted(CSRC_COMPARE.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs) + len(cmp_srcs))) as ex:
        objs = list(ex.map(_compile, srcs + cmp_srcs))
    _link(LIB, objs[:len(srcs)], ["-lpthread"])
    # the comparators (measurement only) live in their own library, which libfastload.so never loads
    _link(LIB_COMPARE, objs[len(srcs):], ["-L/opt/rocm/lib", "-lrocsparse"])
    if verbose:
        print(f"built {LIB} and {LIB_COMPARE}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
