"""Build recipe for the engine library (proteindf_amd/_lib/libpdfxc.so).

hipcc compiles the host engine and the gfx950 kernels into one shared library
with a plain C ABI (include/pdfxc.h).  Packed FP32 instructions are disabled
(-packed-fp32-ops); FP64 atomics use the hardware instruction
(-munsafe-fp-atomics); MFMA accumulators stay in VGPRs
(-amdgpu-mfma-vgpr-form) so the accumulate loops do not copy them to AGPRs.

    python -m proteindf_amd.build
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
OUT = PKG / "_lib" / "libpdfxc.so"
SOURCES = [SRC / "pdfxc_host.cpp", SRC / "pdfxc_io.cpp", SRC / "pdfxc_kernels.hip", SRC / "pdfxc_split.hip", SRC / "pdfxc_grad.hip"]
HEADERS = [SRC / "pdfxc_internal.hpp", SRC / "pdfxc_device.hpp", SRC / "xc_functional.hpp", ROOT / "include" / "pdfxc.h"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-mllvm", "-amdgpu-mfma-vgpr-form=1",
           "-Wall", "-Wno-unused-result", f"-I{ROOT / 'include'}", "-o", str(tmp)]
    cmd += [str(s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    print(OUT)
    sys.exit(0)
