"""Vector fields (three interleaved components per cell) and the per-face
exchange, through the C ABI, against the oracle.

A vector field's ghost fill equals the oracle's scalar ghost fill of each
component (the reference's Vector3DUpdater is a stub; the component-wise
scalar rules are the definition), bit for bit, FP64 and FP32, one and two
ghost layers, across level changes.  bcm_ghost_fill_face over every face
(and subface) of every block equals the whole-mesh fill.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from bcmtools_amd import Mesh
from bcmtools_amd import _lib as L

from conftest import load_golden

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,dtype", [("ghost_sphere_atonce_d", np.float64), ("ghost_sphere_vc2_d", np.float64),
                                        ("ghost_sphere_atonce_f", np.float32)])
def test_vector_ghost_fill_is_the_scalar_fill_of_each_component(name, dtype):
    g = load_golden(name)
    blocks, n, vc = g["blocks"], g["n"], g["vc"]
    case(f"vector ghost fill {name} nb={len(blocks)}")
    rng = np.random.default_rng(7)
    w = n + 2 * vc
    comps = [rng.standard_normal((len(blocks), w, w, w)).astype(dtype) for _ in range(3)]
    vec = np.stack(comps, axis=-1).copy()  # [block][k][j][i][component]
    mesh = Mesh(blocks, n, vc)
    lib = mesh.lib
    h = ctypes.c_void_p()
    L.check(lib.bcm_vfield_create(mesh.handle, L.F64 if dtype == np.float64 else L.F32, ctypes.byref(h)),
            "bcm_vfield_create")
    L.check(lib.bcm_vfield_upload(h, vec.ctypes.data, vec.nbytes), "bcm_vfield_upload")
    L.check(lib.bcm_vfield_ghost_fill(h, L.GF_ALL), "bcm_vfield_ghost_fill")
    out = np.empty_like(vec)
    L.check(lib.bcm_vfield_download(h, out.ctypes.data, out.nbytes), "bcm_vfield_download")
    for c in range(3):
        ref = comps[c].copy()
        oracle.ghost_fill(blocks, n, vc, ref, L.GF_ALL)
        np.testing.assert_array_equal(out[..., c], ref, err_msg=f"component {c}")
    # Block-range copies round trip one block.
    one = np.empty_like(vec[3:4])
    L.check(lib.bcm_vfield_download_blocks(h, 3, 1, one.ctypes.data, one.nbytes), "bcm_vfield_download_blocks")
    np.testing.assert_array_equal(one[0], out[3])
    L.check(lib.bcm_vfield_destroy(h), "bcm_vfield_destroy")


@pytest.mark.parametrize("name", ["ghost_sphere_atonce_d", "ghost_sphere_vc2_d", "ghost_periodic_x_d"])
def test_face_by_face_exchange_equals_the_whole_fill(name):
    g = load_golden(name)
    blocks, n, vc = g["blocks"], g["n"], g["vc"]
    case(f"per-face exchange {name} nb={len(blocks)}")
    mesh = Mesh(blocks, n, vc)
    f = mesh.field(g["fin"].dtype, g["fin"])
    calls = 0
    for b in range(len(blocks)):
        for face in range(6):
            d = int(blocks["nb_level_diff"][b][face])
            subs = range(4) if d == 1 else [0]
            for s in subs:
                if blocks["nb_id"][b][face][s] < 0:
                    continue
                L.check(mesh.lib.bcm_ghost_fill_face(f.handle, b, face, s), "bcm_ghost_fill_face")
                calls += 1
    assert calls > 0
    np.testing.assert_array_equal(f.download(), g["fout"])
    # A face without a neighbour is refused.
    outer = [(b, face) for b in range(len(blocks)) for face in range(6) if blocks["nb_id"][b][face][0] < 0
             and blocks["nb_level_diff"][b][face] != 1]
    if outer:
        b, face = outer[0]
        assert mesh.lib.bcm_ghost_fill_face(f.handle, b, face, 0) == L.ERR_ARG
