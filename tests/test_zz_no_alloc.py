"""bcm_sor_residual_noalloc (partial sums in the library's fixed device
array, no hipMalloc) against bcm_sor_residual through the C ABI: the same
result bits on two calls in a row, the second on a smaller mesh, and the
oracle's relative residual within 1e-12."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from bcmtools_amd import Mesh
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def residual_pair(mesh, f, s):
    a, b = ctypes.c_double(), ctypes.c_double()
    L.check(mesh.lib.bcm_sor_residual_noalloc(f.handle, s.handle, ctypes.byref(a)), "bcm_sor_residual_noalloc")
    L.check(mesh.lib.bcm_sor_residual(f.handle, s.handle, ctypes.byref(b)), "bcm_sor_residual")
    return a.value, b.value


def field_pair(blocks, n, seed):
    rng = np.random.default_rng(seed)
    s = -W.rhs_field(blocks, n, 1, seed=seed)
    f = rng.standard_normal(s.shape)
    return f, s


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_residual_equals_allocating_entry_point(order):
    cases = [(W.shell_tree(4).blocks, 32, 1), (W.shell_tree(3).blocks, 16, 2)]
    if order == "small_then_large":
        cases = cases[::-1]
    for blocks, n, seed in cases:
        case(f"noalloc residual nb={len(blocks)} n={n} seed={seed}")
        f, s = field_pair(blocks, n, seed)
        mesh = Mesh(blocks, n, 1)
        gf, gs = mesh.field(np.float64, f), mesh.field(np.float64, s)
        got, ref = residual_pair(mesh, gf, gs)
        assert got == ref, (got, ref)
        want = oracle.sor_residual(blocks, n, 1, f, s)
        assert abs(got - want) <= 1e-12 * want, (got, want)
        # A second call on the same mesh after the other entry point ran.
        got2, _ = residual_pair(mesh, gf, gs)
        assert got2 == ref
        mesh.close()


def test_noalloc_residual_rejects_other_dtypes():
    blocks = W.shell_tree(3).blocks
    mesh = Mesh(blocks, 8, 1)
    f32 = mesh.field(np.float32)
    s32 = mesh.field(np.float32)
    r = ctypes.c_double()
    assert mesh.lib.bcm_sor_residual_noalloc(f32.handle, s32.handle, ctypes.byref(r)) == L.ERR_ARG
    assert mesh.lib.bcm_sor_residual_noalloc(None, s32.handle, ctypes.byref(r)) == L.ERR_ARG
