"""Build the HIP library ``libqsketch_amd.so`` (gfx950) in-tree.

Usage: ``python -m qsketch_amd.build`` or ``qsketch_amd.build.build()``.
hipcc cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqsketch_amd.so")
SOURCES = ["qs_sketch.hip", "qs_insert.hip", "qs_sort.hip", "qs_gen.hip", "qs_learned.hip", "qs_single.hip",
           "qs_calib.hip", "qs_hot.hip", "qs_sf.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QS_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: never let the compiler form packed FP32 instructions.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# Per-source additions.  The learned forward reads its MFMA accumulators with
# VALU instructions (ReLU and BF16 packing), which take no AGPR
# operands: accumulators in VGPRs save one v_accvgpr_read per value.
SOURCE_FLAGS = {
    "qs_learned.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build(verbose: bool = False) -> str:
    """Compile every source to an object, then link the shared library."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *FLAGS, *SOURCE_FLAGS.get(src, []), "-c", path, "-o", obj])
        objs.append(obj)
    for cmd in cmds if verbose else []:
        print(" ".join(cmd), flush=True)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:  # one hipcc process per source
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_native_tests(verbose: bool = False) -> None:
    """C++ drivers used by the GPU tests: the include/qsketch.hpp interface
    against the oracle, the exhaustive fastmod check and the item packing check."""
    root = os.path.dirname(HERE)
    nat = os.path.join(root, "tests", "native")
    cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           os.path.join(nat, "cpp_api_test.cpp"), "-L" + HERE, "-lqsketch_amd",
           "-L" + os.path.join(root, "oracle"), "-loracle",
           "-Wl,-rpath,$ORIGIN/../../qsketch_amd", "-Wl,-rpath,$ORIGIN/../../oracle",
           "-o", os.path.join(nat, "cpp_api_test")]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    cmd = [HIPCC, "-O2", "-std=c++17", os.path.join(nat, "fastmod_check.cpp"),
           "-o", os.path.join(nat, "fastmod_check")]
    _run(cmd)
    cmd = [HIPCC, "-O2", "-std=c++17", os.path.join(nat, "pack_check.cpp"),
           "-o", os.path.join(nat, "pack_check")]
    _run(cmd)


if __name__ == "__main__":
    print(build(verbose=True))
