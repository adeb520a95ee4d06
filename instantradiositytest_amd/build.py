"""Build the gfx950 library (libirt.so) and the synthetic-scene generator.

Plain hipcc / g++ command lines, no build system: the HIP translation unit is
compiled for gfx950 only, with packed-FP32 instruction selection disabled
(`-Xclang -target-feature -Xclang -packed-fp32-ops`).  Outputs land in
``instantradiositytest_amd/_lib`` (git-ignored, shipped with the tree).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "_lib"
ROOT = PKG.parent
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # One IEEE rounding per source operation, as in the FP32 restatement:
    # no multiply-add contraction.  Hot loops write fmaf() explicitly where
    # they deviate on purpose.
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# The mixed profile's march (irt_march_contracted.hip): multiply-add
# contraction inside each expression and the hardware's 1-ulp division and
# square root instead of the exact contract above.
CONTRACTED_FLAGS = [f for f in HIP_FLAGS
                    if f not in ("-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt")]
CONTRACTED_FLAGS += ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]


def build_irt(verbose_asm: bool = False) -> Path:
    """Compile libirt.so (HIP kernels + C ABI + host precompute)."""
    LIB_DIR.mkdir(exist_ok=True)
    out = LIB_DIR / "libirt.so"
    tmp = LIB_DIR / "libirt.so.tmp"
    extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_asm else []
    objs = []
    for src, flags in (("irt_api.hip", HIP_FLAGS), ("irt_host.cpp", HIP_FLAGS),
                       ("irt_march_contracted.hip", CONTRACTED_FLAGS)):
        obj = LIB_DIR / (Path(src).stem + ".o")
        _run([HIPCC, *flags, *extra, "-c", f"-I{INCLUDE}", f"-I{CSRC}", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_scenegen() -> Path:
    """Compile libirt_scenegen.so (CPU synthetic scene / G-buffer generator)."""
    LIB_DIR.mkdir(exist_ok=True)
    out = LIB_DIR / "libirt_scenegen.so"
    tmp = LIB_DIR / "libirt_scenegen.so.tmp"
    gxx = shutil.which("g++") or "g++"
    _run([gxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", f"-I{INCLUDE}",
          str(CSRC / "scenegen.cpp"), str(CSRC / "irt_host.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_scenegen()
    build_irt()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
