"""The C-ABI library loads, exports every symbol include/asgard_amd.h declares,
and its host-side index construction matches the oracle bit for bit.
No GPU work is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import asgard_amd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "asgard_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(akm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = asgard_amd.load_library()
    names = declared_symbols()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(asgard_amd.EXPORTS)


def test_version_and_error_channel():
    lib = asgard_amd.load_library()
    assert b"gfx950" in lib.akm_version()
    # an invalid argument is reported through the status code and message
    with pytest.raises(ValueError):
        asgard_amd.connect_1d(0)
    assert "max_level" in asgard_amd.last_error()


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("full", [True, False])
def test_connect_1d_bit_exact(L, full):
    p, i, dg = asgard_amd.connect_1d(L, full)
    po, io, do = oracle.connect_1d(L, full)
    assert np.array_equal(p, po)
    assert np.array_equal(i, io)
    assert np.array_equal(dg, do)


def test_null_handle_is_rejected():
    lib = asgard_amd.load_library()
    rc = lib.akm_apply(None, 1.0, None, 0.0, None, 1)
    assert rc == asgard_amd.AKM_ERR_INVALID


def test_create_rejects_bad_indices_without_gpu_work():
    # validation happens on the host before any device allocation
    pntr = np.array([0, 1], np.int64)
    indx = np.array([3], np.int32)  # column 3 of a 1-column operator
    iA = np.zeros(2, np.int32)
    vA = np.ones(4)
    with pytest.raises(ValueError):
        asgard_amd.KronmultMatrix.sparse(2, 2, 1, 1, 1, pntr, indx, iA, vA)


# ---- prototypes: header parameter lists against the ctypes declarations ----

_SCALARS = {"int": ctypes.c_int, "int64_t": ctypes.c_int64, "uint64_t": ctypes.c_uint64,
            "double": ctypes.c_double, "size_t": ctypes.c_size_t, "akm_prec": ctypes.c_int,
            "akm_status": ctypes.c_int}
_HANDLES = {"akm_matrix", "akm_precond", "akm_batch"}


def header_prototypes():
    """{name: [parameter type strings]} of every akm_ function in the header."""
    text = open(os.path.join(ROOT, "include", "asgard_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(?:akm_status|int|const char \*)\s*(akm_[a-z0-9_]+)\s*\(([^)]*)\)\s*;",
                         text):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        types = []
        for p in params:
            p = re.sub(r"\s+", " ", p)
            # drop the parameter name (the last identifier)
            t = re.sub(r"\s*\b[a-zA-Z_][a-zA-Z0-9_]*$", "", p).strip()
            types.append(t)
        protos[m.group(1)] = types
    return protos


def _compatible(ctype_decl, ctypes_type):
    t = ctype_decl.replace("const ", "").strip()
    pointer = t.endswith("*")
    base = t.rstrip("* ").strip()
    if not pointer:
        if base in _HANDLES:
            return ctypes_type is ctypes.c_void_p
        if base == "akm_imex_hook":
            return ctypes_type is asgard_amd.IMEX_HOOK
        return ctypes_type is _SCALARS[base]
    # pointers: an untyped pointer, a char buffer, or POINTER(matching base)
    if ctypes_type in (ctypes.c_void_p, ctypes.c_char_p):
        return True
    target = getattr(ctypes_type, "_type_", None)
    if base in _SCALARS:
        return target is _SCALARS[base]
    if base in _HANDLES:
        return target is ctypes.c_void_p
    if base == "akm_dg_partial":
        return target is asgard_amd.DgPartial
    if base == "akm_solve_stats":
        return target is asgard_amd.SolveStats
    return False


def test_ctypes_prototypes_match_header():
    lib = asgard_amd.load_library()
    protos = header_prototypes()
    assert set(protos) == set(asgard_amd.EXPORTS)
    for name, types in protos.items():
        argtypes = getattr(lib, name).argtypes
        if argtypes is None:
            assert types == [], name
            continue
        assert len(argtypes) == len(types), f"{name}: {len(argtypes)} ctypes vs {len(types)}"
        for k, (decl, ct) in enumerate(zip(types, argtypes)):
            assert _compatible(decl, ct), f"{name} parameter {k}: header '{decl}' vs {ct}"


def test_cpp_surface_compiles(tmp_path):
    """include/asgard_amd.hpp and the consumer program compile against the
    header (syntax and types, no link)."""
    import shutil
    import subprocess

    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    run = subprocess.run(
        [cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
         "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
         os.path.join(ROOT, "tests", "cpp", "consumer.cpp")], capture_output=True, text=True)
    assert run.returncode == 0, run.stderr
