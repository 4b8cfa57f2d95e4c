"""The C ABI library loads and exports every entry point include/bcm_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np

import oracle
from bcmtools_amd import _lib as L

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "bcm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void\*|size_t|double)\s+\*?\s*(bcm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    assert len(names) >= 40
    for must in ("bcm_mesh_create", "bcm_ghost_fill", "bcm_rbsor_sweep", "bcm_jacobi7", "bcm_bicgstab",
                 "bcm_solve_mixed", "bcm_tree_build"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(L.SIGNATURES) == set(declared_functions())


def test_struct_layouts():
    assert L.BLOCK_DTYPE.itemsize == 256
    assert ctypes.sizeof(L.TreeParams) == 4 * 3 + 4 * 3 + 8 * 3 + 8 + 4 * 4 + 8 * 5 * 8 + 4 * 3 + 4 + 8 + 4 * 6


def test_oracle_loads():
    lib = oracle.load()
    assert hasattr(lib, "oracle_ghost_fill_d")


def test_field_api_rejects_bad_arguments_without_gpu():
    lib = L.load()
    assert lib.bcm_mesh_create(None, 0, 16, 1, ctypes.byref(ctypes.c_void_p())) == L.ERR_ARG
    blocks = np.zeros(1, dtype=L.BLOCK_DTYPE)
    # odd block size, vc > n/2 (include/BlockBase.h:75-78, include/DataClass.h:53-56)
    assert lib.bcm_mesh_create(blocks.ctypes.data, 1, 15, 1, ctypes.byref(ctypes.c_void_p())) == L.ERR_ARG
    assert lib.bcm_mesh_create(blocks.ctypes.data, 1, 4, 3, ctypes.byref(ctypes.c_void_p())) == L.ERR_ARG
    blocks["nb_id"][0][0][0] = 7  # neighbour id out of range
    assert lib.bcm_mesh_create(blocks.ctypes.data, 1, 4, 1, ctypes.byref(ctypes.c_void_p())) == L.ERR_ARG
    assert lib.bcm_ghost_fill(None, 7) == L.ERR_ARG
    assert lib.bcm_dot(None, None, None) == L.ERR_ARG
