"""Build the native parts of fir_amd in-tree.

* ``fir_amd/libfir_hip.so`` -- gfx950 kernels + the C-ABI of
  include/fir_hip.h (hipcc).
* ``fir_amd/bin/FIR_ForwardProj``, ``FIR_BackProj``, ``FIR_OSEM`` and
  ``fir_amd/libfir_host.so`` -- the C++ host layer (key files, Interfile
  I/O, VolData/ProjData/ScannerData/LORCache, operators) over the C-ABI.

Usage: ``python -m fir_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ARCH = os.environ.get("FIR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the walk and the FP32 accumulations must round every
# multiply and add separately, as the reference's x86-64 build does.
# -munsafe-fp-atomics: FP32 atomicAdd becomes global_atomic_add_f32.
# packed FP32 instructions are disabled for all HIP code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_hip(verbose=True):
    srcs = [os.path.join(HERE, "csrc", f)
            for f in ("fir_hip.hip", "path_cache.hip", "listmode.hip")]
    out = os.path.join(HERE, "libfir_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_host():
    hdir = os.path.join(HERE, "host")
    if not os.path.isdir(hdir) or not os.path.exists(
            os.path.join(hdir, "Makefile")):
        return None
    jobs = os.environ.get("MAX_JOBS", "8")
    _run(["make", "-C", hdir, f"-j{min(int(jobs), 16)}"])
    return os.path.join(HERE, "bin")


def build_all():
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
