"""The C-ABI library loads and exports every entry point include/*.h declares,
plus the C++-linkage drop-ins of dkernel.o.  No GPU calls."""
import re
import subprocess
from pathlib import Path

import darray_optimizer_amd as dh
from darray_optimizer_amd import _capi

ROOT = Path(__file__).resolve().parent.parent


def declared_c_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        if 'extern "C"' not in text:
            continue
        names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(dh_[a-z0-9_]+)\s*\(", text,
                                re.M))
    return names


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(_capi.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_loads():
    assert dh.lib() is not None


def test_every_declared_entry_point_is_exported():
    decl = declared_c_functions()
    assert len(decl) >= 20
    exp = exported_symbols()
    missing = sorted(decl - exp)
    assert not missing, missing


def test_bindings_cover_the_header():
    assert declared_c_functions() == set(_capi.SIGNATURES)


def test_dkernel_drop_in_symbols():
    # Itanium-mangled names of the C++ entry points of dkernel.h:6-8 and
    # dkernel.cu:28,38, as the reference's drivers link them.
    exp = exported_symbols()
    for sym in ("_Z10rbf_kerneliiiiPfiS_iS_S_S_if", "_Z7tc_gemmiiiiPfiS_iS_iff",
                "_Z8tc_sgemmicciiiPfiS_iS_iff", "_Z9cuda_initv", "_Z13cuda_finalizev"):
        assert sym in exp, sym


def test_bad_arguments_are_reported_without_a_gpu():
    # Argument checks run before any device work.
    import ctypes as C
    st = dh.lib().dh_gemm_f64(None, b"N", b"N", 1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1)
    assert st == -1 and "NULL handle" in _capi.last_error()
    st = dh.lib().dh_fill_uniform(0, 0.0, 1.0, 3, 2, None, 3)
    assert st == -6
