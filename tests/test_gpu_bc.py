"""Outer-face Dirichlet and Neumann conditions of SolverTest on the GPU.

bcm_impose_bc (BlockScalar3D::ImposeBoundaryCondition) and the Neumann and
Dirichlet folds of bcm_poisson_assemble are bit-exact against the oracle;
the matrix-free kernels equal the stored-coefficient kernels bit for bit
on folded Neumann faces; the fused BiCGSTAB engine and the host-driven
ILS3D BiCGSTAB solve a mixed Dirichlet/Neumann problem, checked by the
oracle's residual and against the oracle's BiCGSTAB.
"""
import sys

import numpy as np
import pytest

import oracle
from bcmtools_amd import Mesh
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W

pytestmark = pytest.mark.gpu

D, N = L.BC_DIRICHLET, L.BC_NEUMANN
TYPES = [N, D, N, D, D, N]
VALS = [0.7, -1.3, 2.1, 0.4, -0.9, 1.6]


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def bc_tree(types, level, adaptive):
    if adaptive:
        return W.Tree(root_n=(1, 1, 1), divider=L.DIV_SHELL, min_level=1, max_level=level,
                      shells=[(0.5, 0.5, 0.5, 0.3, 0.3)], balance=1, bc_setter=L.BCSET_PER_FACE,
                      face_bc_type=types)
    return W.Tree(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=level, bc_setter=L.BCSET_PER_FACE,
                  face_bc_type=types)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,vc,adaptive", [(8, 1, False), (8, 2, True), (16, 1, True)])
def test_impose_bc_bit_exact(dtype, n, vc, adaptive):
    blocks = bc_tree(TYPES, 2, adaptive).blocks
    case(f"impose_bc n={n} vc={vc} adaptive={adaptive} {np.dtype(dtype).name} nb={len(blocks)}")
    rng = np.random.default_rng(n + vc)
    nx0 = n + 2 * vc
    data = rng.standard_normal((len(blocks), nx0, nx0, nx0)).astype(dtype)
    mesh = Mesh(blocks, n, vc)
    f = mesh.field(dtype, data)
    mesh.impose_bc(f, VALS)
    ref = data.copy()
    oracle.impose_bc(blocks, n, vc, ref, VALS)
    np.testing.assert_array_equal(f.download(), ref)


@pytest.mark.parametrize("adaptive", [False, True])
def test_neumann_assembly_and_matrix_free_kernels(adaptive):
    n, vc = 8, 1
    blocks = bc_tree(TYPES, 2, adaptive).blocks
    case(f"neumann assembly adaptive={adaptive} nb={len(blocks)}")
    rhs = -W.rhs_field(blocks, n, vc, seed=8)
    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, vc, A, b, rhs, VALS)
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(np.float64) for _ in range(7)]
    gb = mesh.field(np.float64)
    mesh.poisson_assemble(gA, gb, mesh.field(np.float64, rhs), VALS)
    inner = np.pad(np.ones((n,) * 3, bool), vc)
    for i in range(7):
        np.testing.assert_array_equal(gA[i].download()[:, inner], A[i][:, inner])
    np.testing.assert_array_equal(gb.download()[:, inner], b[:, inner])
    rng = np.random.default_rng(3)
    x = rng.standard_normal(rhs.shape)
    gx = mesh.field(np.float64, x)
    o1, o2 = mesh.field(np.float64), mesh.field(np.float64)
    mesh.jacobi7(o1, gx, gA, gb, 0.9)
    mesh.jacobi7_poisson(o2, gx, gb, 0.9)
    np.testing.assert_array_equal(o1.download(), o2.download())
    ref = np.zeros_like(x)
    oracle.jacobi7(len(blocks), n, vc, ref, x, A, b, 0.9)
    np.testing.assert_array_equal(W.interior(o2.download(), vc), W.interior(ref, vc))
    mesh.calc_ax(o1, gA, gx)
    mesh.calc_ax_poisson(o2, gx)
    np.testing.assert_array_equal(o1.download(), o2.download())
    mesh.calc_r(o1, gA, gx, gb)
    mesh.calc_r_poisson(o2, gx, gb)
    np.testing.assert_array_equal(o1.download(), o2.download())


def oracle_relres(blocks, n, x, rhs, vals):
    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, 1, A, b, rhs, list(vals))
    r = np.zeros_like(rhs)
    oracle.calc_r(len(blocks), n, 1, r, A, x, b)
    ri, bi = W.interior(r, 1), W.interior(b, 1)
    return float(np.sqrt(np.sum(ri * ri)) / np.sqrt(np.sum(bi * bi)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_engine_solves_mixed_dirichlet_neumann(precision):
    n = 16
    blocks = bc_tree(TYPES, 3, True).blocks
    case(f"engine neumann {precision} nb={len(blocks)}")
    rhs = -W.rhs_field(blocks, n, 1, seed=9)
    mesh = Mesh(blocks, n, 1)
    gx = mesh.field(np.float64)
    st = mesh.solve(gx, mesh.field(np.float64, rhs), VALS, precision=precision, tol=1e-10, max_iter=5000)[0]
    assert st["status"] == "converged", st
    rr = oracle_relres(blocks, n, gx.download(), rhs, VALS)
    assert rr <= 1e-10 * (1 + 1e-6), (rr, st)


def test_host_bicgstab_neumann_vs_oracle():
    n, vc = 8, 1
    blocks = bc_tree(TYPES, 2, True).blocks
    case(f"bicgstab neumann nb={len(blocks)}")
    rhs = -W.rhs_field(blocks, n, vc, seed=10)
    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, vc, A, b, rhs, VALS)
    x = np.zeros_like(rhs)
    work = [np.zeros_like(rhs) for _ in range(9)]
    st, cnt, res = oracle.bicgstab(blocks, n, vc, x, A, b, work, 0.8, 2, 2000, 1e-13)
    assert st == 0
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(np.float64, a) for a in A]
    gb = mesh.field(np.float64, b)
    gx = mesh.field(np.float64)
    gw = [mesh.field(np.float64) for _ in range(9)]
    rc, gcnt, gres = mesh.bicgstab(gx, gA, gb, gw, omega=0.8, count_pre=2, count_max=2000, epsilon=1e-13)
    assert rc == L.OK
    xg = gx.download()
    xi, ri = W.interior(xg, vc), W.interior(x, vc)
    assert np.abs(xi - ri).max() <= 1e-10 * np.abs(ri).max()
    r = np.zeros_like(x)
    oracle.calc_r(len(blocks), n, vc, r, A, xg, b)
    assert np.linalg.norm(W.interior(r, vc)) <= 1e-10 * np.linalg.norm(W.interior(b, vc))


@pytest.mark.parametrize("method", [L.VC_AT_ONCE, L.VC_SEPARATE_XYZ, L.VC_SEPARATE_XYZ_SEPARATE_LEVEL_DIFF])
@pytest.mark.parametrize("dtype,n,vc", [(np.float64, 8, 1), (np.float32, 8, 2), (np.float64, 16, 1)])
def test_update_virtual_cells_and_impose_boundary_condition(method, dtype, n, vc):
    """LocalScalar3D::UpdateVirtualCells / ImposeBoundaryCondition with each
    VCUpdateMethod: the exchange equals the oracle's ghost fill, and with the
    outer-face conditions after it the oracle's ghost fill + impose_bc, bit
    for bit."""
    blocks = bc_tree(TYPES, 3, True).blocks
    case(f"update_virtual_cells method={method} n={n} vc={vc} {np.dtype(dtype).name} nb={len(blocks)}")
    rng = np.random.default_rng(7 + method)
    nx0 = n + 2 * vc
    data = rng.standard_normal((len(blocks), nx0, nx0, nx0)).astype(dtype)
    mesh = Mesh(blocks, n, vc)
    f = mesh.field(dtype, data)
    mesh.update_virtual_cells(f, method)
    ref = data.copy()
    oracle.ghost_fill(blocks, n, vc, ref)
    np.testing.assert_array_equal(f.download(), ref)
    g = mesh.field(dtype, data)
    mesh.impose_boundary_condition(g, VALS, method)
    oracle.impose_bc(blocks, n, vc, ref, VALS)
    np.testing.assert_array_equal(g.download(), ref)


def test_update_virtual_cells_rejects_an_unknown_method():
    blocks = bc_tree(TYPES, 2, False).blocks
    case(f"update_virtual_cells bad method nb={len(blocks)}")
    mesh = Mesh(blocks, 8, 1)
    f = mesh.field(np.float64)
    with pytest.raises(L.BCMError):
        mesh.update_virtual_cells(f, 3)
    with pytest.raises(L.BCMError):
        mesh.impose_boundary_condition(f, VALS, -1)
