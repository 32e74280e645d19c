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
    prefix = install()
    build_c_check(prefix)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources, {PROBE_LIB}, installed to {prefix}, built {C_CHECK}")
    return LIB


DEFAULT_PREFIX = os.path.join(ROOT, "build", "install")


def install(prefix: str = DEFAULT_PREFIX) -> str:
    """Install layout of the reference's build (CMakeLists.txt:44-51):
    PREFIX/lib/libcontract.so and PREFIX/include/{contract.h,
    contract_enum.h, contract_ext.h}, so a binding generator or a C build
    points -I/-L at one prefix."""
    os.makedirs(os.path.join(prefix, "lib"), exist_ok=True)
    os.makedirs(os.path.join(prefix, "include"), exist_ok=True)
    shutil.copy2(LIB, os.path.join(prefix, "lib", "libcontract.so"))
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        shutil.copy2(h, os.path.join(prefix, "include", os.path.basename(h)))
    return prefix


C_CHECK_SRC = os.path.join(ROOT, "tests", "c", "abi_check.c")
C_CHECK = os.path.join(ROOT, "tests", "c", "_build", "abi_check")


def build_c_check(prefix: str = DEFAULT_PREFIX) -> str:
    """The C99 consumer of the installed library (tests/c/abi_check.c),
    compiled with gcc against PREFIX/include and linked with -lcontract from
    PREFIX/lib (run path relative to the binary, so the tree can move)."""
    os.makedirs(os.path.dirname(C_CHECK), exist_ok=True)
    rel = os.path.relpath(os.path.join(prefix, "lib"), os.path.dirname(C_CHECK))
    cmd = [os.environ.get("CC", "gcc"), "-std=c99", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(prefix, "include"), "-I", "/opt/rocm/include", C_CHECK_SRC, "-o", C_CHECK + ".tmp",
           "-L", os.path.join(prefix, "lib"), "-lcontract", "-L", "/opt/rocm/lib", "-lamdhip64", "-lm",
           f"-Wl,-rpath,$ORIGIN/{rel}:/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed for {C_CHECK_SRC}:\n{r.stdout}\n{r.stderr}")
    shutil.move(C_CHECK + ".tmp", C_CHECK)
    return C_CHECK


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
