"""Compile the gfx950 HIP library ``dpt/_lib/libdpt_amd.so`` in-tree.

Run as ``python -m dpt.build_native``.  hipcc cross-compiles without a
GPU.  Flags:

* ``-ffp-contract=off``: every fused multiply-add in the kernels is explicit,
  so the sampling arithmetic is a fixed function of the inputs;
* ``-fno-gpu-flush-denormals-to-zero``: FP32 subnormals are kept (the
  reference flushes only its scatter increments, which the kernels do
  explicitly);
* ``-fhip-fp32-correctly-rounded-divide-sqrt``: IEEE division for the clip
  interval and step, as in the reference;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ALU
  instructions.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
SOURCES = [
    PACKAGE / "csrc" / "projector.hip",
    PACKAGE / "csrc" / "gather.hip",
    PACKAGE / "csrc" / "fit.hip",
    PACKAGE / "csrc" / "transport.hip",
    PACKAGE / "csrc" / "probe.hip",
    PACKAGE / "csrc" / "pose.hip",
]
OUTPUT = PACKAGE / "_lib" / "libdpt_amd.so"
ARCH = os.environ.get("DPT_AMD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(found).exists():
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build dpt")
    return found


def build(verbose: bool = False) -> Path:
    """Compile every HIP source into one shared library (always recompiles)."""
    OUTPUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, "-I", str(PACKAGE.parent / "include"), *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
