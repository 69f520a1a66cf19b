"""Build recipe for libferric_hip.so (gfx950) and the test-only MFMA probe.

Every HIP source under ferric_amd/csrc is compiled by hipcc for gfx950 and
linked into ferric_amd/lib/libferric_hip.so (in-tree, so the built library
travels with the source tree). Packed-FP32 VALU instructions are disabled for
all device code.

    python -m ferric_amd.build            # build (always recompiles)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libferric_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]

PROBE_SRC = os.path.join(ROOT, "tests", "hip", "mfma_probe.hip")
PROBE_LIB = os.path.join(ROOT, "tests", "hip", "build", "libmfma_probe.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd[-3:])}")
    return r


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJDIR, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = jobs or min(8, len(srcs))

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *CFLAGS, "-c", src, "-o", obj])
        if verbose:
            print("compiled", os.path.basename(src), flush=True)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    # rocBLAS carries the plain dense GEMMs of the dressing stage and the PNO
    # transforms; rocSOLVER the PNO eigensolves and the domain fits' Cholesky.
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    if os.path.exists(PROBE_SRC):
        os.makedirs(os.path.dirname(PROBE_LIB), exist_ok=True)
        _run([HIPCC, *CFLAGS, "-shared", "-o", PROBE_LIB, PROBE_SRC])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
