"""Build libtbec.so (hipcc, gfx950) in-tree.

The stage kernel's instantiations are split over translation units, one per
(precision, tile shape) pair (csrc/stage_inst.hip compiled with -DINST_PREC/-DINST_TILE),
and compiled in parallel, then linked with the host-side sources.

    python -m trapbec_amd.build [-v]
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtbec.so")
SOURCES = ["tbec.hip", "tbec_step.hip", "tbec_batch.hip", "tbec_vortex.hip", "tbec_mlp.hip",
           "tbec_ic.hip", "tbec_diag.hip", "tbec_phys.hip", "tbec_refic.hip",
           "tbec_noalloc.hip", "tbec_stream.hip", "tbec_prolong.hip"]
INSTANCES = [(p, t) for p in (32, 64) for t in range(5)]

# Packed FP32 instructions are disabled for every HIP file: hipcc would otherwise emit
# them from ordinary float code.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]


def hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise FileNotFoundError("hipcc not found")


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)}\nfailed:\n{res.stderr[-6000:]}")
    return cmd


def build_lib(verbose: bool = False, jobs: int = 0, extra=(), out: str = None) -> str:
    """Compile libtbec for gfx950 (to `out`, default trapbec_amd/libtbec.so); `extra` are
    additional compiler flags."""
    cc = hipcc()
    extra = list(extra)
    lib = out or LIB
    jobs = jobs or min(8, os.cpu_count() or 1)
    tmpd = tempfile.mkdtemp(prefix="tbec_build_")
    try:
        cmds = []
        for s in SOURCES:
            src = os.path.join(CSRC, s)
            if not os.path.exists(src):
                raise FileNotFoundError(f"libtbec source {src} is missing")
            cmds.append([cc, *HIPFLAGS, *extra, "-c", src, "-o", os.path.join(tmpd, s + ".o")])
        inst = os.path.join(CSRC, "stage_inst.hip")
        for p, t in INSTANCES:
            cmds.append([cc, *HIPFLAGS, *extra, f"-DINST_PREC={p}", f"-DINST_TILE={t}", "-c", inst,
                         "-o", os.path.join(tmpd, f"stage_p{p}_t{t}.o")])
        with ThreadPoolExecutor(jobs) as ex:
            for cmd in ex.map(_run, cmds):
                if verbose:
                    print(" ".join(cmd), flush=True)
        objs = [c[-1] for c in cmds]
        tmp = lib + ".tmp"
        _run([cc, "--offload-arch=gfx950", "-shared", *objs, "-o", tmp])
        os.replace(tmp, lib)
    finally:
        shutil.rmtree(tmpd, ignore_errors=True)
    return lib


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv))
