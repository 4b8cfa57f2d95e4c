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
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void\*|size_t|double)\s+\*?\s*(bcm_\w+)\s*\(", text,
                                 flags=re.M)))


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


# Entry points that report a count return -BCM_ERR_ARG for a NULL handle;
# the destroy calls accept NULL like free(); every other int entry point
# returns BCM_ERR_ARG, except the process-wide setting that takes no
# handle (0 is its default).
_SETTINGS = {"bcm_tree_set_threads"}
_COUNTS = {"bcm_tree_num_blocks", "bcm_sor_get_variant", "bcm_server_poll", "bcm_comm_face_size"}
_DESTROYS = {"bcm_tree_destroy", "bcm_mesh_destroy", "bcm_field_destroy", "bcm_server_destroy", "bcm_vfield_destroy",
             "bcm_host_free", "bcm_comm_list_destroy"}


def test_every_entry_point_rejects_null_handles_without_gpu():
    lib = L.load()
    ints = (ctypes.c_int, ctypes.c_long, ctypes.c_longlong, ctypes.c_size_t, ctypes.c_uint, ctypes.c_int64,
            ctypes.c_uint64)
    checked = 0
    for name, (res, args) in sorted(L.SIGNATURES.items()):
        if res not in (ctypes.c_int, ctypes.c_int64):
            continue
        vals = [0 if a in ints else 0.0 if a in (ctypes.c_double, ctypes.c_float) else
                ctypes.cast(None, a) if a is L.DIVIDER_FN else None for a in args]
        got = getattr(lib, name)(*vals)
        if name in _DESTROYS or name in _SETTINGS:
            assert got == L.OK, name
        elif name in _COUNTS:
            assert got == -L.ERR_ARG, name
        else:
            assert got == L.ERR_ARG, (name, got)
        checked += 1
    assert checked >= 70


def fortran_declarations():
    """name -> argument count of every declaration in include/bcm_fortran.h."""
    text = open(os.path.join(ROOT, "include", "bcm_fortran.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*void\s+(\w+_)\s*\(([^;]*)\)\s*;", text, flags=re.M):
        out[m.group(1)] = len([a for a in m.group(2).split(",") if a.strip()])
    return out


def test_fortran_header_bindings_and_exports_agree():
    decl = fortran_declarations()
    assert len(decl) == 71
    assert decl == L.FORTRAN_ARGC
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in decl if not hasattr(lib, n)]
    assert not missing, missing


def test_fortran_entry_points_exit_on_bad_sizes_without_gpu():
    """Sizes are checked before anything touches the device; a bad one
    prints the entry point's name and exits 1, as the reference's Exit()."""
    import subprocess
    import sys
    code = ("import ctypes; from bcmtools_amd import _lib as L; "
            "sz=(ctypes.c_int*3)(4,0,4); g=ctypes.c_int(1); a=ctypes.c_double(1.0); "
            "L.fortran('fill_')(None, ctypes.byref(a), sz, ctypes.byref(g))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert out.returncode == 1
    assert "bcm: fill_: block size must be positive" in out.stderr
    code = ("import ctypes; from bcmtools_amd import _lib as L; "
            "sz=(ctypes.c_int*3)(4,4,4); g=ctypes.c_int(1); a=ctypes.c_double(1.0); "
            "L.fortran('bc_x1_d_')(None, ctypes.byref(a), sz, ctypes.byref(g))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert out.returncode == 1
    assert "bcm: bc_x1_d_: needs at least 2 ghost layers (g = 1)" in out.stderr


def test_server_checkpoint_and_restart_checks_without_gpu(tmp_path):
    """bcm_server_submit_restart reads the restart header at submission
    (missing file: BCM_ERR_IO; other edge, ghost width or block count:
    BCM_ERR_ARG), and a checkpoint names a pending request; none of these
    checks touches the device."""
    from oracle import restart as R
    lib = L.load()
    p = L.ServerParams(1e9, 1, 16, 100, 2, 0.8, 4, 0, 0, 0.0)
    h = ctypes.c_void_p()
    assert lib.bcm_server_create(ctypes.byref(p), ctypes.byref(h)) == L.OK
    blocks = np.zeros(3, dtype=L.BLOCK_DTYPE)
    for b in range(3):
        blocks["nb_id"][b] = -1
    d = str(tmp_path)

    def submit(label):
        return lib.bcm_server_submit_restart(h, blocks.ctypes.data, 3, None, 1, ctypes.c_double(1e-8), 0, d.encode(),
                                             label.encode(), 0, 7, None, 1)

    assert submit("missing") == L.ERR_IO
    R.dump2(np.zeros((2, 18, 18, 18)), 16, 1, R.name2(d, "two", 0, 7))  # 2 blocks, not 3
    assert submit("two") == L.ERR_ARG
    R.dump2(np.zeros((3, 8, 8, 8)), 6, 1, R.name2(d, "other", 0, 7))  # edge 6: neither 16 nor 16/2
    assert submit("other") == L.ERR_ARG
    with open(R.name2(d, "short", 0, 7), "wb") as fp:
        fp.write(np.array([16, 16], dtype=np.int32).tobytes())
    assert submit("short") == L.ERR_IO
    it = ctypes.c_int(-1)
    assert lib.bcm_server_checkpoint(h, 1, d.encode(), b"x", 0, 0, ctypes.byref(it)) == L.ERR_ARG
    assert lib.bcm_server_checkpoint(h, 1, None, b"x", 0, 0, ctypes.byref(it)) == L.ERR_ARG
    n_active, n_queued = ctypes.c_int(-1), ctypes.c_int(-1)
    assert lib.bcm_server_stats(h, None, ctypes.byref(n_active), ctypes.byref(n_queued)) == L.OK
    assert (n_active.value, n_queued.value) == (0, 0)
    assert lib.bcm_server_destroy(h) == L.OK
