"""Build libcgocn.so (HIP, gfx950) in-tree with hipcc.

Each translation unit is compiled on its own (``hipcc -c``, parallel) and
linked into ``cgenie_amd/libcgocn.so``:

* k_tstep_f64.hip, k_bgc.hip, k_mom.hip, k_sfx.hip  -ffp-contract=off: operations round
  as the reference writes them (the FP64 path is bit-comparable with the oracle)
* k_tstep_f32.hip              -ffp-contract=fast: the FP32 form
* cgocn_api.hip                host side of the C ABI

All HIP code is built with packed FP32 arithmetic disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcgocn.so")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("CGO_OFFLOAD_ARCH", "gfx950")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cgocn.h")


def abi_stamp() -> str:
    """Hash of include/cgocn.h.  The library is built with it (cgo_abi_info)
    and the Python binding compares it with the header it was written
    against, so a library built from another version of the header is refused
    when it is loaded instead of being called with mismatched structs."""
    with open(HEADER, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
UNITS = {
    "k_tstep_f64.hip": ["-ffp-contract=off"],
    "k_bgc.hip": ["-ffp-contract=off"],
    "k_mom.hip": ["-ffp-contract=off"],
    "k_sfx.hip": ["-ffp-contract=off"],
    "k_tstep_f32.hip": ["-ffp-contract=fast"],
    "cgocn_api.hip": ["-ffp-contract=off"],
}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def _compile(src: str, flags: list[str]) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *COMMON, *flags, f'-DCGO_ABI_STAMP="{abi_stamp()}"', "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
