"""Build libngbem_amd.so (HIP, gfx950) in-tree: ngbem_amd/_lib/libngbem_amd.so.

Host geometry/tree code (bem_host.cpp) is compiled by g++ with
-ffp-contract=off so the cluster tree is reproducible; the HIP sources are
compiled by hipcc for gfx950 only.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT, "libngbem_amd.so")
ARCH = os.environ.get("NGBEM_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-variable"]


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found")
    return h


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    bdir = os.path.join(OUT, "obj")
    os.makedirs(bdir, exist_ok=True)
    hipcc = _hipcc()
    jobs = [
        ["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-c",
         os.path.join(CSRC, "bem_host.cpp"), "-o", os.path.join(bdir, "bem_host.o")],
        [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "bem_device.hip"), "-o", os.path.join(bdir, "bem_device.o")],
        [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "bem_capi.hip"), "-o", os.path.join(bdir, "bem_capi.o")],
        [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "bem_krylov.hip"), "-o", os.path.join(bdir, "bem_krylov.o")],
    ]
    with cf.ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp,
          os.path.join(bdir, "bem_host.o"), os.path.join(bdir, "bem_device.o"), os.path.join(bdir, "bem_capi.o"),
          os.path.join(bdir, "bem_krylov.o")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
