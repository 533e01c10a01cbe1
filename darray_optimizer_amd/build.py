"""Build libdarray_hip.so (gfx950) in-tree with hipcc, and the CPU oracle.

    python -m darray_optimizer_amd.build            # library + oracle
    python -m darray_optimizer_amd.build --lib-only

Objects go to darray_optimizer_amd/_build/; the shared library to
darray_optimizer_amd/libdarray_hip.so, where the Python bindings load it.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libdarray_hip.so"
ORACLE = ROOT / "oracle"

ARCH = os.environ.get("DARRAY_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed-FP32 VALU forms are switched off for all device code; hipcc's host
# pass warns that it does not know the feature, which is harmless.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-parameter", "-Wno-unused-function"]

# Per-file device flags.  The FP64 / FP32 MFMA kernels keep their accumulators
# in VGPRs: left to the heuristics, the fused FP64 pass allocated them in AGPRs
# and copied about 290 registers between the register files every 16-point
# chunk (a 510-register wave), and the GEMM tiles copied theirs too (the
# 128 x 128 FP64 tile took 356 registers, 238 in VGPR form).
_VGPR_MFMA = ["-mllvm", "--amdgpu-mfma-vgpr-form=1"]
FILE_FLAGS = {f: _VGPR_MFMA for f in ("rbf_f64.hip", "dense.hip", "small_spd.hip", "batch.hip")}

SOURCES = ["handle.hip", "dense.hip", "small_spd.hip", "solve.hip", "rbf_f64.hip", "rbf_bf16.hip",
           "rbf16_dk1.hip", "rbf16_dk2.hip", "rbf16_dk4.hip", "rbf16_dk8.hip",
           "rbf16_fp8.hip", "rbf16_persistent.hip", "lra.hip", "batch.hip", "noalloc.hip", "nystrom.hip", "svm.hip", "gen.cpp", "libsvm_io.cpp", "shim.cpp"]


def _compile(src: str) -> Path:
    s = CSRC / src
    o = OBJ / (s.stem + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, *FILE_FLAGS.get(src, []), "-c", str(s), "-o", str(o)]
    if s.suffix == ".hip":
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return o


def build_lib(jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


CPP_TESTS = ROOT / "tests" / "cpp"
CPP_BIN = CPP_TESTS / "bin"


def build_cpp_drivers() -> list[Path]:
    """C++ programs written against the DArray header mirror (include/darray_hip/),
    linked to libdarray_hip.so with an $ORIGIN-relative rpath."""
    CPP_BIN.mkdir(exist_ok=True)
    cxx = shutil.which("g++") or "c++"
    rpath = "-Wl,-rpath,$ORIGIN/../../../darray_optimizer_amd"
    out = []
    for src in sorted(CPP_TESTS.glob("*.cpp")):
        exe = CPP_BIN / src.stem
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
               f"-L{PKG}", "-ldarray_hip", rpath]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        out.append(exe)
    return out


def build_oracle() -> Path:
    r = subprocess.run(["make", "-C", str(ORACLE), "-B", "liboracle.so"], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    return ORACLE / "liboracle.so"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    print("built", build_lib(a.jobs), flush=True)
    if not a.lib_only:
        print("built", *build_cpp_drivers(), flush=True)
        print("built", build_oracle(), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
