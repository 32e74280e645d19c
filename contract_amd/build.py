"""Build libcontract.so (the drop-in C ABI + extension API) for gfx950.

    python -m contract_amd.build          # or __graft_entry__.build()

Every ``csrc/*.hip`` file is compiled with hipcc in parallel and linked into
``contract_amd/lib/libcontract.so``.  Packed FP32 instructions are disabled
(``-packed-fp32-ops``): the kernels' FP32 path uses scalar FMA only.
"""

from __future__ import annotations

import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcontract.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CONTRACT_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", *HIP_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(jobs, 1)) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    build_probe()
    if verbose:
        print(f"built {LIB} from {len(sources)} sources, and {PROBE_LIB}")
    return LIB


PROBE_SRC = os.path.join(HERE, "probe", "hbm_probe.hip")
PROBE_LIB = os.path.join(LIBDIR, "libcontract_probe.so")


def build_probe() -> str:
    """The benchmark's HBM ceiling probes (read-only stream, device copy), a
    library of their own: nothing in libcontract.so depends on them."""
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC, "-x", "hip", *HIP_FLAGS, "-shared", PROBE_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {PROBE_SRC}:\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, PROBE_LIB)
    return PROBE_LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
