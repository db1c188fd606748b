"""Build recipe for the in-tree HIP library ``jaxtrace_amd/_jt.so`` (gfx950).

Every ``csrc/*.hip`` and ``csrc/*.cpp`` file is compiled by ``hipcc`` to an
object under ``jaxtrace_amd/_build/`` (in parallel), then linked into one
shared library that ``jaxtrace_amd._lib`` loads with ctypes.

Run as ``python -m jaxtrace_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_jt.so")
BUILD_DIR = os.path.join(HERE, "_build")

ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17",
    f"--offload-arch={ARCH}",
    # one IEEE operation at a time, in source order: parity with the
    # reference's float32 op sequence
    "-ffp-contract=off",
    # packed-FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]

# per-file extra flags: hardware float atomics for the adjoint's scatter
EXTRA = {"adjoint.hip": ["-munsafe-fp-atomics"]}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found; the jaxtrace_amd HIP library cannot be built")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), "-c", src, "-o", obj] + COMMON + EXTRA.get(os.path.basename(src), [])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
