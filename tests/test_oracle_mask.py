"""Pin the oracle's masked SolverTest kernels and solvers (blas.f90:700-1113,
ILS3D.h:846-1097).

The Fortran cannot be compiled here.  The masked forms are pinned by what
they must equal: with setup_mask's mask (1 inside, 0 in the ghosts) every
masked kernel and the masked solvers give the unmasked results bit for bit;
with any 0/1 mask a masked kernel writes the unmasked value where the mask
is 1 and leaves every other cell alone (sums skip those cells); and
BiCGSTAB_Mask with a solid region masked out solves the Poisson system
restricted to the open cells, with the solid cells held at 0, which a
direct sparse solve of that restricted system confirms.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W

from test_oracle import assemble, uniform_poisson_matrix


def system(seed=1, n=6, vc=1):
    t = W.Tree(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=1)
    blocks = t.blocks
    rng = np.random.default_rng(seed)
    f = W.rhs_field(blocks, n, vc, seed=seed)
    A, b = assemble(blocks, n, vc, -f, [0.0])
    x = rng.standard_normal(f.shape)
    y = rng.standard_normal(f.shape)
    return blocks, n, vc, A, b, x, y


def kernels(blocks, n, vc, A, b, x, y, out0, mask):
    nb = len(blocks)
    res = {}
    o = out0.copy(); oracle.jacobi7(nb, n, vc, o, x, A, b, 0.9, mask=mask); res["jacobi"] = o
    o = out0.copy(); oracle.calc_ax(nb, n, vc, o, A, x, mask=mask); res["ax"] = o
    o = out0.copy(); oracle.calc_r(nb, n, vc, o, A, x, b, mask=mask); res["r"] = o
    for op in ("axpy", "xpay", "axpyz", "axpbypz", "copy"):
        o = out0.copy(); oracle.blas(nb, n, vc, op, o, x, y, 0.7, -1.3, mask=mask); res[op] = o
    res["r2"] = oracle.calc_r2(nb, n, vc, A, x, b, mask=mask)
    res["dot"] = oracle.dot(nb, n, vc, x, y, mask=mask)
    return res


def test_setup_mask_equals_unmasked():
    blocks, n, vc, A, b, x, y = system()
    out0 = np.random.default_rng(5).standard_normal(x.shape)
    plain = kernels(blocks, n, vc, A, b, x, y, out0, None)
    masked = kernels(blocks, n, vc, A, b, x, y, out0, oracle.setup_mask(x.shape, vc))
    for k in plain:
        np.testing.assert_array_equal(masked[k], plain[k], err_msg=k)


def test_random_mask_writes_only_open_cells():
    blocks, n, vc, A, b, x, y = system(seed=2)
    rng = np.random.default_rng(3)
    mask = oracle.setup_mask(x.shape, vc) * (rng.random(x.shape) < 0.6)
    out0 = rng.standard_normal(x.shape)
    plain = kernels(blocks, n, vc, A, b, x, y, out0, None)
    masked = kernels(blocks, n, vc, A, b, x, y, out0, mask)
    on = mask == 1.0
    for k in ("jacobi", "ax", "r", "axpy", "xpay", "axpyz", "axpbypz", "copy"):
        np.testing.assert_array_equal(masked[k][on], plain[k][on], err_msg=k)
        np.testing.assert_array_equal(masked[k][~on], out0[~on], err_msg=k)
    r = np.zeros_like(x)
    oracle.calc_r(len(blocks), n, vc, r, A, x, b)
    assert masked["r2"] == pytest.approx(float(np.sum(r[on] ** 2)), rel=1e-12)
    assert masked["dot"] == pytest.approx(float(np.sum(x[on] * y[on])), rel=1e-12)


def test_bicgstab_mask_with_setup_mask_equals_bicgstab():
    blocks, n, vc, A, b, _, _ = system(seed=4, n=8)
    runs = []
    for mask in (None, oracle.setup_mask(b.shape, vc)):
        x = np.zeros_like(b)
        work = [np.zeros_like(b) for _ in range(9)]
        st = oracle.bicgstab(blocks, n, vc, x, A, b, work, 0.8, 2, 200, 1e-10, mask=mask)
        runs.append((st, x))
    assert runs[0][0] == runs[1][0]
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def test_bicgstab_mask_solves_the_open_cells():
    """A sphere of solid cells (mask 0) inside a 16^3 grid: x stays 0 there and
    the open cells solve the Poisson system with the solid cells held at 0."""
    t = W.Tree(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=1)
    blocks, n, vc = t.blocks, 8, 1
    f = W.rhs_field(blocks, n, vc, seed=6)
    A, b = assemble(blocks, n, vc, -f, [0.0])
    m = 16
    h = 1.0 / m
    mask = oracle.setup_mask(b.shape, vc)
    gidx = np.full(b.shape, -1, dtype=np.int64)
    for bk, bi in enumerate(blocks):
        o = np.rint(bi["origin"] / (n * h)).astype(int)
        k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        gk, gj, gi = k + o[2] * n, j + o[1] * n, i + o[0] * n
        c = np.stack([(gi + 0.5) * h, (gj + 0.5) * h, (gk + 0.5) * h])
        solid = np.sum((c - 0.5) ** 2, axis=0) < 0.2 ** 2
        mask[bk, vc:vc + n, vc:vc + n, vc:vc + n][solid] = 0.0
        gidx[bk, vc:vc + n, vc:vc + n, vc:vc + n] = (gk * m + gj) * m + gi
    x = np.zeros_like(b)
    work = [np.zeros_like(b) for _ in range(9)]
    st, cnt, res = oracle.bicgstab(blocks, n, vc, x, A, b, work, 0.8, 2, 2000, 1e-12, mask=mask)
    assert st == 0, (st, cnt, res)
    open_ = mask == 1.0
    inner = gidx >= 0
    assert np.all(x[inner & ~open_] == 0.0)
    M = uniform_poisson_matrix(m, h)
    keep = np.sort(gidx[open_])
    rhs_g = np.zeros(m ** 3)
    rhs_g[gidx[inner]] = b[inner]
    u = spla.spsolve(M[keep][:, keep].tocsc(), rhs_g[keep])
    xg = np.zeros(m ** 3)
    xg[gidx[open_]] = x[open_]
    np.testing.assert_allclose(xg[keep], u, rtol=0, atol=1e-9 * np.abs(u).max())
