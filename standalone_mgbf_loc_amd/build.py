"""Builds libmgbf.so (HIP kernels + C ABI) for gfx950, in-tree.

    python -m standalone_mgbf_loc_amd.build

The library lands in standalone_mgbf_loc_amd/lib/libmgbf.so; nothing is
installed outside the repository.
"""
from __future__ import annotations

import json
import os
import re
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libmgbf.so")
INCLUDE = os.path.join(ROOT, "include")

SOURCES = ["mgbf_coef.cpp", "mgbf_capi.cpp", "mgbf_kernels.hip", "mgbf_noalloc.hip"]
RADIAL_TU = "mgbf_radial_inst.hip"
RADIAL_H = range(1, 11)
BALL_TU = "mgbf_ball3.hip"
BALL_H = range(1, 5)
ARCH = os.environ.get("MGBF_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list:
    return [
        "-O3", "-std=c++17", "-fPIC",
        f"--offload-arch={ARCH}",
        # packed FP32 (v_pk_*_f32) code generation is disabled on purpose
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # matrix-core results stay in VGPRs: the ring adds then read them
        # directly instead of copying out of AGPRs first
        "-mllvm", "-amdgpu-mfma-vgpr-form",
        "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
        f"-I{INCLUDE}", f"-I{CSRC}",
    ]


# the compiler's per-function resource report (-Rpass-analysis=kernel-resource-usage)
_REMARK = re.compile(r"remark:\s+(Function Name|ScratchSize \[bytes/lane\]|VGPRs|AGPRs|VGPRs Spill|SGPRs Spill|"
                     r"Occupancy \[waves/SIMD\]|LDS Size \[bytes/block\]):\s+(\S+)")
RESOURCES = os.path.join(LIBDIR, "kernel_resources.json")


def _resources(text: str) -> dict:
    out, cur = {}, None
    for m in _REMARK.finditer(text):
        key, val = m.group(1), m.group(2)
        if key == "Function Name":
            cur = val
            out[cur] = {}
        elif cur is not None:
            out[cur][key] = int(val) if val.isdigit() else val
    return out


def _compile(job):
    """(error text or None, {function: resources})."""
    src, obj, extra, verbose = job
    cmd = [hipcc()] + flags() + extra + ["-Rpass-analysis=kernel-resource-usage", "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        return " ".join(cmd) + "\n" + res.stdout + res.stderr, {}
    return None, _resources(res.stderr)


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles every translation unit (in parallel) and links libmgbf.so."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    work = []
    for s in SOURCES:
        work.append((os.path.join(CSRC, s), os.path.join(objdir, s + ".o"), [], verbose))
    for h in RADIAL_H:
        work.append((os.path.join(CSRC, RADIAL_TU), os.path.join(objdir, f"radial_h{h}.o"),
                     [f"-DMGBF_RADIAL_H={h}"], verbose))
    for h in BALL_H:
        work.append((os.path.join(CSRC, BALL_TU), os.path.join(objdir, f"ball3_h{h}.o"),
                     [f"-DMGBF_BALL_H={h}"], verbose))
    jobs = jobs or min(len(work), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, work))
    errs = [e for e, _ in results if e]
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("hipcc failed building libmgbf.so")
    # every kernel keeps its state in registers and LDS: a kernel with
    # scratch memory (an array the compiler could not keep in registers,
    # such as a ring indexed in a loop left rolled) runs several times
    # slower, so it fails the build
    res = {}
    for _, r in results:
        res.update(r)
    with open(RESOURCES, "w") as f:
        json.dump(res, f, indent=0, sort_keys=True)
    scratch = sorted(k for k, v in res.items() if v.get("ScratchSize [bytes/lane]", 0) > 0)
    if scratch:
        raise RuntimeError("kernels with scratch memory: " + ", ".join(scratch[:10]))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}"] + [w[1] for w in work] + ["-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError("linking libmgbf.so failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
