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
