"""Masked SolverTest kernels and solvers on the GPU (blas.f90:753-1113,
ILS3D.h:846-1097), against the oracle's masked forms: bit for bit for the
kernels and the Jacobi sweeps, to tolerance for the sums and BiCGSTAB_Mask."""
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


def problem(dtype, seed):
    t = W.shell_tree(3)
    blocks, n, vc = t.blocks, 8, 1
    f = W.rhs_field(blocks, n, vc, seed=seed)
    A = [np.zeros_like(f) for _ in range(7)]
    b = np.zeros_like(f)
    oracle.poisson_assemble(blocks, n, vc, A, b, -f, [0.0])
    rng = np.random.default_rng(seed)
    mask = np.zeros(f.shape)
    mask[:, vc:vc + n, vc:vc + n, vc:vc + n] = rng.random((len(blocks), n, n, n)) < 0.7
    cast = lambda a: a.astype(dtype)
    return blocks, n, vc, [cast(a) for a in A], cast(b), cast(mask), rng


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_masked_kernels_bit_exact(dtype):
    blocks, n, vc, A, b, mask, rng = problem(dtype, 3)
    nb = len(blocks)
    case(f"masked kernels {np.dtype(dtype).name} nb={nb}")
    x = rng.standard_normal(b.shape).astype(dtype)
    y = rng.standard_normal(b.shape).astype(dtype)
    out0 = rng.standard_normal(b.shape).astype(dtype)
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(dtype, a) for a in A]
    gb, gx, gy, gm = (mesh.field(dtype, a) for a in (b, x, y, mask))

    def fresh():
        return mesh.field(dtype, out0)

    o = fresh(); mesh.jacobi7(o, gx, gA, gb, 0.9, mask=gm)
    r = out0.copy(); oracle.jacobi7(nb, n, vc, r, x, A, b, 0.9, mask=mask)
    np.testing.assert_array_equal(o.download(), r)
    o = fresh(); mesh.calc_ax(o, gA, gx, mask=gm)
    r = out0.copy(); oracle.calc_ax(nb, n, vc, r, A, x, mask=mask)
    np.testing.assert_array_equal(o.download(), r)
    o = fresh(); mesh.calc_r(o, gA, gx, gb, mask=gm)
    r = out0.copy(); oracle.calc_r(nb, n, vc, r, A, x, b, mask=mask)
    np.testing.assert_array_equal(o.download(), r)
    for op in ("axpy", "xpay", "axpyz", "axpbypz", "copy"):
        o = fresh()
        if op == "axpy":
            mesh.axpy(o, gx, 0.7, mask=gm)
        elif op == "xpay":
            mesh.xpay(o, gx, 0.7, mask=gm)
        elif op == "axpyz":
            mesh.axpyz(o, gx, gy, 0.7, mask=gm)
        elif op == "axpbypz":
            mesh.axpbypz(o, gx, gy, 0.7, -1.3, mask=gm)
        else:
            mesh.copy_mask(o, gx, gm)
        r = out0.copy(); oracle.blas(nb, n, vc, op, r, x, y, 0.7, -1.3, mask=mask)
        np.testing.assert_array_equal(o.download(), r, err_msg=op)
    # Sums: the summation orders differ (per-thread then a fixed tree here,
    # per block sequentially in the oracle), so compare to a tolerance on
    # the sum of |terms|: 1e-12 in FP64, 1e-6 in FP32 (about 10 ulp).
    tol = 1e-12 if dtype == np.float64 else 1e-6
    on = mask == 1
    xy = x.astype(np.float64) * y.astype(np.float64)
    assert abs(mesh.dot(gx, gy, mask=gm) - oracle.dot(nb, n, vc, x, y, mask=mask)) <= tol * np.abs(xy[on]).sum()
    r = np.zeros_like(x)
    oracle.calc_r(nb, n, vc, r, A, x, b)
    r2 = r.astype(np.float64) ** 2
    assert abs(mesh.calc_r2(gA, gx, gb, mask=gm) - oracle.calc_r2(nb, n, vc, A, x, b, mask=mask)) <= \
        tol * r2[on].sum()


def test_bicgstab_mask_vs_oracle():
    blocks, n, vc, A, b, mask, _ = problem(np.float64, 4)
    case(f"bicgstab_mask nb={len(blocks)}")
    x = np.zeros_like(b)
    work = [np.zeros_like(b) for _ in range(9)]
    st, cnt, res = oracle.bicgstab(blocks, n, vc, x, A, b, work, 0.8, 2, 2000, 1e-13, mask=mask)
    assert st == 0
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(np.float64, a) for a in A]
    gx = mesh.field(np.float64)
    gw = [mesh.field(np.float64) for _ in range(9)]
    rc, gcnt, gres = mesh.bicgstab(gx, gA, mesh.field(np.float64, b), gw, omega=0.8, count_pre=2, count_max=2000,
                                   epsilon=1e-13, mask=mesh.field(np.float64, mask))
    assert rc == L.OK
    assert abs(gcnt - cnt) <= max(3, cnt // 10)
    xi, ri = W.interior(gx.download(), vc), W.interior(x, vc)
    assert np.abs(xi - ri).max() <= 1e-10 * np.abs(ri).max()
    assert np.all(W.interior(gx.download(), vc)[W.interior(mask, vc) == 0.0] == 0.0)


def test_jacobi_solve_mask_vs_oracle():
    blocks, n, vc, A, b, mask, _ = problem(np.float64, 5)
    case(f"jacobi_solve_mask nb={len(blocks)}")
    x, x0 = np.zeros_like(b), np.zeros_like(b)
    st, cnt, res, out = oracle.jacobi_solve(blocks, n, vc, x, A, b, x0, 1.0, 60, 1e-14, mask=mask)
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(np.float64, a) for a in A]
    gx, g0 = mesh.field(np.float64), mesh.field(np.float64)
    rc, gcnt, gres = mesh.jacobi_solve(gx, gA, mesh.field(np.float64, b), g0, omega=1.0, count_max=60,
                                       epsilon=1e-14, mask=mesh.field(np.float64, mask))
    assert gcnt == cnt == 60
    np.testing.assert_array_equal(W.interior(gx.download(), vc), W.interior(out, vc))


@pytest.mark.parametrize("dtype,vc", [(np.float64, 1), (np.float32, 2)])
def test_setup_mask_on_device(dtype, vc):
    """bcm_setup_mask writes the oracle's setup_mask exactly, and the masked
    kernels with it equal the plain kernels bit for bit."""
    blocks, n, _, A, b, _, rng = problem(dtype, 6)
    case(f"setup_mask {np.dtype(dtype).name} vc={vc} nb={len(blocks)}")
    nx0 = n + 2 * vc
    shape = (len(blocks), nx0, nx0, nx0)
    mesh = Mesh(blocks, n, vc)
    gm = mesh.field(dtype, rng.standard_normal(shape).astype(dtype))
    mesh.setup_mask(gm)
    np.testing.assert_array_equal(gm.download(), oracle.setup_mask(shape, vc, dtype))
    x = rng.standard_normal(shape).astype(dtype)
    out0 = rng.standard_normal(shape).astype(dtype)
    gx = mesh.field(dtype, x)
    o1, o2 = mesh.field(dtype, out0), mesh.field(dtype, out0)
    mesh.axpy(o1, gx, 0.7)
    mesh.axpy(o2, gx, 0.7, mask=gm)
    np.testing.assert_array_equal(o1.download(), o2.download())
    assert mesh.dot(gx, gx) == mesh.dot(gx, gx, mask=gm)
