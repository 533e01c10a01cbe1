"""The DArray C++ header mirror (include/darray_hip/) and its C++ drivers.

CPU: every driver in tests/cpp compiles against the headers (syntax only, no
library needed).  GPU: the drivers built by the package build run their
known-answer checks (Cholesky of min(i+1,j+1) and of m m^T, views through
matrix_multiply / triangular_solve, LRA and Knorm against host sums, the
range finder on a low-rank kernel) and print PASS.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CPP = ROOT / "tests" / "cpp"
BIN = CPP / "bin"
DRIVERS = sorted(p.stem for p in CPP.glob("*.cpp"))


@pytest.mark.parametrize("name", DRIVERS)
def test_driver_compiles_against_headers(name):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{ROOT / 'include'}", str(CPP / f"{name}.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_headers_declare_reference_entry_points():
    text = "".join((ROOT / "include" / "darray_hip" / f).read_text()
                   for f in ("darray.h", "dkernel.h", "blas_like.h", "lapack_like.h"))
    text += "".join((ROOT / "include" / "darray_hip" / f).read_text()
                    for f in ("dbarrier.h", "read_input.h"))
    for sym in ("class Grid", "class LMatrix", "class DMatrix", "struct DMatrixDescriptor",
                "void LRA(", "double Knorm(", "void matrix_multiply(", "void triangular_solve(",
                "void cholesky_factorize(", "void cholesky_factorize_recursive(",
                "void multiply_matrix(", "void symmetric_update(", "void rbf_kernel(",
                "void tc_gemm(", "void tc_sgemm(", "LMatrix<T> svd(", "void read_input_file(",
                "void writemodel(", "int pre_conjgrad("):
        assert sym in text, sym


def _run(args, timeout=120):
    exe = BIN / args[0]
    if not exe.exists():
        pytest.fail(f"{exe} is not built: run __graft_entry__.build()")
    print("running", " ".join(args), flush=True)
    r = subprocess.run([str(exe), *args[1:]], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "PASS"


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["chol_ones", "1024", "256", "blocked", "double"],
    ["chol_ones", "1000", "128", "blocked", "float"],
    ["chol_ones", "777", "32", "recursive", "double"],
    ["chol_ones", "512", "32", "recursive", "float"],
    ["chol_ones", "4096", "256", "blocked", "double"],
])
def test_cholesky_known_answer(args):
    _run(args)


@pytest.mark.gpu
def test_cholesky_mmT_known_answer():
    _run(["chol_mmt"])


@pytest.mark.gpu
def test_views_multiply_and_solve():
    _run(["blas_views"])


@pytest.mark.gpu
def test_lra_knorm_and_range_finder():
    _run(["lra_knorm"], timeout=300)


@pytest.mark.gpu
def test_reference_seam_rbf_kernel_tc_gemm_tc_sgemm():
    """rbf_kernel / tc_gemm / tc_sgemm through their C++ linkage with host
    layouts, padded leading dimensions, beta, inner dimension > 8192 and all
    four transposes, against host sums in double."""
    _run(["seam"], timeout=300)


@pytest.mark.gpu
def test_svd_and_nystrom_factor_with_knorm():
    """DArray::svd known answers and the range finder -> Nystrom Z -> Knorm flow
    of testsvm.cpp:86-177 through the header mirror."""
    _run(["nystrom_knorm"], timeout=300)


@pytest.mark.gpu
def test_svm_pipeline_libsvm_to_model(tmp_path):
    """testsvm.cpp's flow: LIBSVM file -> reader -> labelled range finder ->
    Nystrom Z + Knorm -> barrier SVM -> LIBSVM model, with the training
    accuracy of the written model's decision values."""
    _run(["svm_pipeline", str(tmp_path)], timeout=300)
