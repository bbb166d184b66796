"""Build ``libtqsv.so`` (all HIP kernels + the C ABI) for gfx950 with hipcc.

Usage: ``python -m torchquantum_amd.csrc.build`` or ``build()`` from Python.
The library is written in-tree next to this package (``torchquantum_amd/libtqsv.so``)
so that it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
LIB = PKG / "libtqsv.so"

SOURCES = [
    "tqsv_gates.hip",
    "tqsv_stage.hip",
    "tqsv_expval.hip",
    "tqsv_adjoint.hip",
    "tqsv_devgates.hip",
    "tqsv_mlp.hip",
    "tqsv_sample.hip",
    "tqsv_jit.hip",
]

ARCH = os.environ.get("TQSV_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-munsafe-fp-atomics",
    # no packed-FP32 (v_pk_*_f32) instructions anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build libtqsv.so)")


def build(verbose: bool = False, jobs: int = 6) -> Path:
    """Compile every source to an object, then link the shared library."""
    cc = hipcc()
    srcs = [HERE / s for s in SOURCES]
    missing = [str(p) for p in srcs if not p.exists()]
    if missing:
        raise RuntimeError(f"libtqsv.so sources missing: {missing}")
    tmp = Path(tempfile.mkdtemp(prefix="tqsv_build_"))
    try:
        procs = []
        objs = []
        for src in srcs:
            obj = tmp / (src.stem + ".o")
            cmd = [cc, *FLAGS, "-I", str(HERE), "-c", str(src), "-o", str(obj)]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
            objs.append(obj)
            while sum(p.poll() is None for _, p in procs) >= jobs:
                procs[[p.poll() is None for _, p in procs].index(True)][1].wait()
        failed = []
        for src, p in procs:
            out = p.communicate()[0].decode(errors="replace")
            out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
            if p.returncode != 0:
                failed.append(f"--- {src.name}\n{out}")
            elif verbose and out.strip():
                print(out)
        if failed:
            raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
        tmp_lib = tmp / "libtqsv.so"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp_lib),
               "-lhiprtc"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
        shutil.move(str(tmp_lib), str(LIB))
        return LIB
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
