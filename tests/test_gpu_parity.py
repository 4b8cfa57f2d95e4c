"""HIP path vs the oracle and the reference's own fixtures, through the C ABI.

Bit-exact: ghost fill (copy/F2C/C2F, FP64 and FP32), Dirichlet mirror,
red-black SOR (sor1 and sor2 forms, fused and generic kernels), Jacobi7,
calc_ax, calc_r, BLAS-1, Poisson assembly, SOR to tolerance.  Tolerance
(reduction order differs from the reference's OpenMP order): dot, calc_r2,
SOR residual (rel 1e-12); BiCGSTAB and the mixed-precision solve (solved
past 1e-10, then solution difference and the oracle's relative residual
<= 1e-10).
"""
import sys

import numpy as np
import pytest

import oracle
from bcmtools_amd import Mesh
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W

from conftest import load_golden

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def stored_relres(n, vc, A, x, b):
    """||b - A x|| / ||b|| over the interior, by the oracle's calc_r."""
    r = np.zeros_like(x)
    oracle.calc_r(len(x), n, vc, r, A, x, b)
    ri, bi = W.interior(r, vc), W.interior(b, vc)
    return float(np.sqrt(np.sum(ri * ri)) / np.sqrt(np.sum(bi * bi)))


def poisson_relres(blocks, n, x, rhs, bvalues=(0.0,)):
    """The same for the folded Poisson system of rhs (vc = 1)."""
    A = [np.zeros_like(rhs) for _ in range(7)]
    b = np.zeros_like(rhs)
    oracle.poisson_assemble(blocks, n, 1, A, b, rhs, list(bvalues))
    return stored_relres(n, 1, A, x, b)


def gpu_update(mesh, fld, method):
    if method == 0:
        mesh.update_vc(fld)
    elif method == 1:
        for ax in range(3):
            mesh.update_vc_xyz(fld, ax)
    else:
        for kind in (L.GF_F2C, L.GF_FLAT, L.GF_C2F):
            for ax in range(3):
                mesh.update_vc_xyz(fld, ax, kind)


@pytest.mark.parametrize("name", ["ghost_sphere_atonce_d", "ghost_sphere_xyz_d", "ghost_sphere_leveldiff_d",
                                  "ghost_sphere_atonce_f", "ghost_simple_vc2_d", "ghost_periodic_x_d",
                                  "ghost_sphere_vc2_d", "ghost_sphere_vc2_leveldiff_f"])
def test_ghost_fill_vs_reference_fixture(name):
    g = load_golden(name)
    case(f"ghost fill {name} nb={len(g['blocks'])} n={g['n']} vc={g['vc']}")
    mesh = Mesh(g["blocks"], g["n"], g["vc"])
    fld = mesh.field(g["fin"].dtype, g["fin"])
    gpu_update(mesh, fld, g["method"])
    np.testing.assert_array_equal(fld.download(), g["fout"])


@pytest.mark.parametrize("name", ["sor1_log4", "sor1_adaptive_y", "sor1_inner3_z"])
def test_sor1_driver_vs_reference_fixture(name):
    g = load_golden(name)
    blocks, n, vc = g["blocks"], g["n"], g["vc"]
    case(f"sor1 {name} nb={len(blocks)} n={n} iters={g['iters']} inner={g['inner']}")
    mesh = Mesh(blocks, n, vc)
    f = mesh.field(np.float64)
    vals = [g["b0"], g["b1"]]
    mesh.update_vc(f)
    mesh.dirichlet_mirror(f, vals)
    mesh.sor_iterate(f, None, g["omega"], vals, g["iters"], g["inner"])
    out = f.download()
    mask = W.face_ghost_mask(n, vc)
    np.testing.assert_array_equal(out[:, mask], g["fout"][:, mask])


@pytest.mark.parametrize("n,level,form", [(16, 3, 2), (32, 2, 2), (16, 3, 1), (32, 2, 1), (8, 3, 2), (4, 3, 1)])
def test_rbsor_vs_oracle(n, level, form):
    t = W.shell_tree(level)
    blocks = t.blocks
    case(f"rbsor n={n} level={level} form=sor{form} nb={len(blocks)}")
    rng = np.random.default_rng(n + level)
    nx0 = n + 2
    f0 = rng.standard_normal((len(blocks), nx0, nx0, nx0))
    s = W.rhs_field(blocks, n, 1, seed=3) if form == 2 else None
    mesh = Mesh(blocks, n, 1)
    f = mesh.field(np.float64, f0)
    sf = mesh.field(np.float64, s) if s is not None else None
    vals = [0.0]
    mesh.sor_iterate(f, sf, 1.3, vals, 3)
    ref = f0.copy()
    oracle.dirichlet(blocks, n, 1, ref, vals)
    oracle.sor_iterate(blocks, n, 1, ref, s, 1.3, vals, 3)
    np.testing.assert_array_equal(f.download(), ref)


def random_A(shape, rng, dtype):
    A = [rng.uniform(0.1, 1.0, shape).astype(dtype) for _ in range(7)]
    A[0] = (-rng.uniform(6.0, 8.0, shape)).astype(dtype)
    return A


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_solvertest_kernels_vs_oracle(dtype):
    t = W.shell_tree(3)
    blocks, n, vc = t.blocks, 16, 1
    nb = len(blocks)
    case(f"jacobi7/calc_ax/calc_r/blas dtype={np.dtype(dtype).name} nb={nb}")
    rng = np.random.default_rng(11)
    shape = (nb, n + 2, n + 2, n + 2)
    A = random_A(shape, rng, dtype)
    x = rng.standard_normal(shape).astype(dtype)
    b = rng.standard_normal(shape).astype(dtype)
    y = rng.standard_normal(shape).astype(dtype)
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(dtype, a) for a in A]
    gx, gb, gy = mesh.field(dtype, x), mesh.field(dtype, b), mesh.field(dtype, y)
    out = mesh.field(dtype, np.zeros(shape, dtype))
    # jacobi_smoother
    mesh.jacobi7(out, gx, gA, gb, 0.8)
    ref = np.zeros(shape, dtype)
    oracle.jacobi7(nb, n, vc, ref, x, A, b, 0.8)
    np.testing.assert_array_equal(out.download(), ref)
    # calc_ax
    mesh.calc_ax(out, gA, gx)
    oracle.calc_ax(nb, n, vc, ref, A, x)
    np.testing.assert_array_equal(out.download(), ref)
    # calc_r
    mesh.calc_r(out, gA, gx, gb)
    oracle.calc_r(nb, n, vc, ref, A, x, b)
    np.testing.assert_array_equal(out.download(), ref)
    # calc_r2 and dot: reduction order differs -> tolerance
    rtol = 1e-12 if dtype == np.float64 else 2e-5
    assert mesh.calc_r2(gA, gx, gb) == pytest.approx(oracle.calc_r2(nb, n, vc, A, x, b), rel=rtol)
    assert mesh.dot(gx, gy) == pytest.approx(oracle.dot(nb, n, vc, x, y), rel=rtol * 10)
    # BLAS-1
    for op, call in (("axpy", lambda z: mesh.axpy(z, gx, 0.3)), ("xpay", lambda z: mesh.xpay(z, gx, 0.3)),
                     ("axpyz", lambda z: mesh.axpyz(z, gx, gy, 0.3)),
                     ("axpbypz", lambda z: mesh.axpbypz(z, gx, gy, 0.3, -1.7))):
        z0 = rng.standard_normal(shape).astype(dtype)
        gz = mesh.field(dtype, z0)
        call(gz)
        zr = z0.copy()
        oracle.blas(nb, n, vc, op, zr, x, y, 0.3, -1.7)
        np.testing.assert_array_equal(gz.download(), zr, err_msg=op)
    mesh.copy(out, gx)
    np.testing.assert_array_equal(out.download(), x)
    mesh.fill(out, 2.5)
    assert np.all(out.download() == dtype(2.5))


def test_poisson_assembly_and_matrix_free_kernels_vs_oracle():
    t = W.shell_tree(3)
    blocks, n, vc = t.blocks, 16, 1
    nb = len(blocks)
    case(f"poisson assemble + matrix-free nb={nb}")
    f = W.rhs_field(blocks, n, vc, seed=9)
    vals = [0.25]
    mesh = Mesh(blocks, n, vc)
    grhs = mesh.field(np.float64, -f)
    gA = [mesh.field(np.float64) for _ in range(7)]
    gb = mesh.field(np.float64)
    mesh.poisson_assemble(gA, gb, grhs, vals)
    A = [np.zeros_like(f) for _ in range(7)]
    b = np.zeros_like(f)
    oracle.poisson_assemble(blocks, n, vc, A, b, -f, vals)
    interior = np.pad(np.ones((n,) * 3, bool), vc)
    for i in range(7):
        np.testing.assert_array_equal(gA[i].download()[:, interior], A[i][:, interior])
    np.testing.assert_array_equal(gb.download()[:, interior], b[:, interior])
    # Matrix-free kernels equal the stored-coefficient kernels bit for bit.
    rng = np.random.default_rng(2)
    x = rng.standard_normal(f.shape)
    gx = mesh.field(np.float64, x)
    o1, o2 = mesh.field(np.float64), mesh.field(np.float64)
    mesh.jacobi7(o1, gx, gA, gb, 0.9)
    mesh.jacobi7_poisson(o2, gx, gb, 0.9)
    np.testing.assert_array_equal(o1.download(), o2.download())
    mesh.calc_ax(o1, gA, gx)
    mesh.calc_ax_poisson(o2, gx)
    np.testing.assert_array_equal(o1.download(), o2.download())
    mesh.calc_r(o1, gA, gx, gb)
    mesh.calc_r_poisson(o2, gx, gb)
    np.testing.assert_array_equal(o1.download(), o2.download())


@pytest.mark.parametrize("matrix_free", [False, True])
def test_config1_100_jacobi_sweeps_bit_exact(matrix_free):
    """Config 1: 64 x 16^3, Dirichlet u = 0, 100 x (Jacobi sweep + ghost fill)."""
    t = W.config1_tree()
    blocks, n, vc = t.blocks, 16, 1
    nb = len(blocks)
    case(f"config1 jacobi x100 matrix_free={matrix_free}")
    f = W.rhs_field(blocks, n, vc, seed=1)
    A = [np.zeros_like(f) for _ in range(7)]
    b = np.zeros_like(f)
    oracle.poisson_assemble(blocks, n, vc, A, b, -f, [0.0])
    x, x0 = np.zeros_like(f), np.zeros_like(f)
    for _ in range(100):
        oracle.jacobi7(nb, n, vc, x0, x, A, b, 1.0)
        x, x0 = x0, x
        oracle.ghost_fill(blocks, n, vc, x, L.GF_ALL)
    mesh = Mesh(blocks, n, vc)
    gx, g0 = mesh.field(np.float64), mesh.field(np.float64)
    gb = mesh.field(np.float64, b)
    gA = [mesh.field(np.float64, a) for a in A]
    for _ in range(100):
        if matrix_free:
            mesh.jacobi7_poisson(g0, gx, gb, 1.0)
        else:
            mesh.jacobi7(g0, gx, gA, gb, 1.0)
        gx, g0 = g0, gx
        mesh.update_vc(gx)
    np.testing.assert_array_equal(gx.download(), x)


def test_bicgstab_vs_oracle():
    t = W.shell_tree(3)
    blocks, n, vc = t.blocks, 8, 1
    case(f"bicgstab fp64 nb={len(blocks)}")
    f = W.rhs_field(blocks, n, vc, seed=4)
    A = [np.zeros_like(f) for _ in range(7)]
    b = np.zeros_like(f)
    oracle.poisson_assemble(blocks, n, vc, A, b, -f, [0.0])
    x = np.zeros_like(f)
    work = [np.zeros_like(f) for _ in range(9)]
    st, cnt, res = oracle.bicgstab(blocks, n, vc, x, A, b, work, 0.8, 2, 1000, 1e-13)
    assert st == 0
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(np.float64, a) for a in A]
    gb = mesh.field(np.float64, b)
    gx = mesh.field(np.float64)
    gw = [mesh.field(np.float64) for _ in range(9)]
    rc, gcnt, gres = mesh.bicgstab(gx, gA, gb, gw, omega=0.8, count_pre=2, count_max=1000, epsilon=1e-13)
    assert rc == L.OK
    assert abs(gcnt - cnt) <= max(3, cnt // 10)
    # Both solved to 1e-13: solutions within 1e-10, the oracle's residual of
    # the GPU solution within 1e-10.
    xg = gx.download()
    xi, ri = W.interior(xg, vc), W.interior(x, vc)
    assert np.abs(xi - ri).max() <= 1e-10 * np.abs(ri).max()
    assert stored_relres(n, vc, A, xg, b) <= 1e-10


def test_jacobi_solve_vs_oracle():
    t = W.config1_tree()
    blocks, n, vc = t.blocks, 8, 1
    case("jacobi solve fp64")
    f = W.rhs_field(blocks, n, vc, seed=6)
    A = [np.zeros_like(f) for _ in range(7)]
    b = np.zeros_like(f)
    oracle.poisson_assemble(blocks, n, vc, A, b, -f, [0.0])
    x, x0 = np.zeros_like(f), np.zeros_like(f)
    st, cnt, res, out = oracle.jacobi_solve(blocks, n, vc, x, A, b, x0, 1.0, 300, 1e-14)
    mesh = Mesh(blocks, n, vc)
    gA = [mesh.field(np.float64, a) for a in A]
    gb = mesh.field(np.float64, b)
    gx, g0 = mesh.field(np.float64), mesh.field(np.float64)
    rc, gcnt, gres = mesh.jacobi_solve(gx, gA, gb, g0, omega=1.0, count_max=300, epsilon=1e-14)
    assert gcnt == cnt == 300
    # Sweeps are bit-exact; only the residual reduction order differs.
    np.testing.assert_array_equal(gx.download(), out)
    assert gres == pytest.approx(res, rel=1e-10)


def test_mixed_precision_reaches_1e10_and_matches_fp64():
    """FP32 inner BiCGSTAB in FP64 iterative refinement on the level-4 shell
    (456 x 16^3): both it and the FP64 path solved to 1e-11; solutions
    within 1e-10 of each other, the oracle's residual of each <= 1e-10."""
    t = W.shell_tree(4)
    blocks, n, vc = t.blocks, 16, 1
    case(f"mixed precision nb={len(blocks)}")
    f = W.rhs_field(blocks, n, vc, seed=8)
    mesh = Mesh(blocks, n, vc)
    rhs = mesh.field(np.float64, -f)
    x1 = mesh.field(np.float64)
    rc, outer, inner, rr = mesh.solve_mixed(x1, rhs, [0.0], tol=1e-11, inner_tol=1e-4, fp32_strict=True)
    assert rc == L.OK and rr <= 1e-11 and outer >= 2
    x2 = mesh.field(np.float64)
    rc2, it2, rr2 = mesh.solve_fp64(x2, rhs, [0.0], tol=1e-11)
    assert rc2 == L.OK and rr2 <= 1e-11
    a, b = W.interior(x1.download(), vc), W.interior(x2.download(), vc)
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    for xx in (x1, x2):
        assert poisson_relres(blocks, n, xx.download(), -f) <= 1e-10


@pytest.mark.parametrize("seed,inner_max", [(9, 500), (3, 1000)])
def test_mixed_precision_level6_recovers_from_fp32_failure(seed, inner_max):
    """On the level-6 mid-surface mesh (6,728 x 16^3) an FP32 Jacobi-BiCGSTAB
    correction can fail to lower the true residual (its recurrence drifts
    from it while the residual first rises).  The refinement discards such
    a correction instead of failing with NaN: strict mode ends with OK or
    BCM_ERR_NOCONV, a finite x and a residual no larger than at the start;
    otherwise the later corrections are solved in FP64 and the solve reaches
    1e-10 by the oracle's residual, within 1e-10 of the FP64 path's
    solution."""
    blocks, n = W.config3_tree(6).blocks, 16
    case(f"mixed precision level 6 nb={len(blocks)} seed={seed} inner_max={inner_max}")
    f = W.rhs_field_fast(blocks, n, 1, seed=seed)
    mesh = Mesh(blocks, n, 1)
    rhs = mesh.field(np.float64, -f)
    xs = mesh.field(np.float64)
    rc, outer, inner, rr = mesh.solve_mixed(xs, rhs, [0.0], tol=1e-10, inner_max=inner_max, fp32_strict=True)
    case(f"strict: rc={rc} outer={outer} inner={inner} relres={rr:.3e}")
    assert rc in (L.OK, L.ERR_NOCONV) and np.isfinite(rr) and rr <= 1.0
    assert rc != L.OK or rr <= 1e-10
    assert np.all(np.isfinite(xs.download()))
    x1 = mesh.field(np.float64)
    rc, outer, inner, rr = mesh.solve_mixed(x1, rhs, [0.0], tol=1e-10, inner_max=inner_max)
    case(f"with FP64 corrections: rc={rc} outer={outer} inner={inner} relres={rr:.3e}")
    assert rc == L.OK and rr <= 1e-10, (rc, outer, inner, rr)
    x2 = mesh.field(np.float64)
    rc2, it2, rr2 = mesh.solve_fp64(x2, rhs, [0.0], tol=1e-10)
    assert rc2 == L.OK
    a, b = W.interior(x1.download(), 1), W.interior(x2.download(), 1)
    assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    assert poisson_relres(blocks, n, x1.download(), -f) <= 1e-10 * (1 + 1e-6)


def test_full_size_config2_ghost_fill_vs_oracle():
    """Config 2 at full size (3,816 x 32^3): GPU ghost fill equals the oracle
    bit for bit; a second fill changes nothing (idempotence)."""
    t = W.config2_tree()
    blocks, n, vc = t.blocks, 32, 1
    case(f"full-size config2 ghost fill nb={len(blocks)}")
    f = W.rhs_field(blocks, n, vc, seed=12)
    mesh = Mesh(blocks, n, vc)
    gf = mesh.field(np.float64, f)
    mesh.update_vc(gf)
    out = gf.download()
    oracle.ghost_fill(blocks, n, vc, f, L.GF_ALL)
    assert np.array_equal(out, f)
    mesh.update_vc(gf)
    assert np.array_equal(gf.download(), out)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_large_mesh_of_small_blocks_ghost_fill_vs_oracle(dtype):
    """A complete fill of a mesh of 1,024 or more blocks of 16^3 runs block-
    wise (one work-group gathers a block's six faces); it equals the oracle
    bit for bit, every cell, and the face-task fill of the same field."""
    t = W.shell_tree(5)
    blocks, n, vc = t.blocks, 16, 1
    assert len(blocks) >= 1024
    case(f"large small-block ghost fill {np.dtype(dtype).name} nb={len(blocks)}")
    f = W.rhs_field(blocks, n, vc, seed=13).astype(dtype)
    mesh = Mesh(blocks, n, vc)
    gf = mesh.field(dtype, f)
    mesh.update_vc(gf)
    out = gf.download()
    gt = mesh.field(dtype, f)
    for ax in (L.GF_X, L.GF_Y, L.GF_Z):
        mesh.update_vc(gt, L.GF_ALL | ax)  # per axis: always the face tasks
    np.testing.assert_array_equal(out, gt.download())
    oracle.ghost_fill(blocks, n, vc, f, L.GF_ALL)
    np.testing.assert_array_equal(out, f)


@pytest.mark.parametrize("n,maker,form", [
    (16, lambda: W.shell_tree(3), 2),
    (32, lambda: W.shell_tree(2), 1),
    (16, lambda: W.Tree(root_n=(1, 1, 1), periodic=(0, 1, 1), divider=L.DIV_SIMPLE, min_level=1, max_level=3,
                        bc_setter=L.BCSET_SOR1_X), 1),
    (16, lambda: W.random_tree(5), 2),
])
def test_halo_gathering_sor_iterations_vs_oracle(n, maker, form):
    """Iterations after the first run the halo-gathering sweep (ghosts built
    in the kernel from the previous iterate); 6 iterations must equal the
    reference-order oracle bit for bit, every cell."""
    t = maker()
    blocks = t.blocks
    vals = [0.25, 1.5] if blocks["bc_id"].max() >= 1 else [0.25]
    case(f"halo sor n={n} form=sor{form} nb={len(blocks)}")
    rng = np.random.default_rng(n + len(blocks))
    f0 = rng.standard_normal((len(blocks), n + 2, n + 2, n + 2))
    s = W.rhs_field(blocks, n, 1, seed=4) if form == 2 else None
    mesh = Mesh(blocks, n, 1)
    f = mesh.field(np.float64, f0)
    sf = mesh.field(np.float64, s) if s is not None else None
    mesh.sor_iterate(f, sf, 1.25, vals, 6)
    ref = f0.copy()
    oracle.dirichlet(blocks, n, 1, ref, vals)
    oracle.sor_iterate(blocks, n, 1, ref, s, 1.25, vals, 6)
    np.testing.assert_array_equal(f.download(), ref)
    # A second call continues from the stored state.
    mesh.sor_iterate(f, sf, 1.25, vals, 3)
    oracle.sor_iterate(blocks, n, 1, ref, s, 1.25, vals, 3)
    np.testing.assert_array_equal(f.download(), ref)


def test_sor_residual_vs_oracle():
    """bcm_sor_residual after a few outer iterations against the oracle's
    ||s - lap_h f|| / ||s|| (FP64 reductions in another order: rel 1e-12)."""
    blocks, n = W.shell_tree(3).blocks, 16
    case(f"sor residual nb={len(blocks)}")
    s = -W.rhs_field(blocks, n, 1, seed=11)
    mesh = Mesh(blocks, n, 1)
    f = mesh.field(np.float64)
    sf = mesh.field(np.float64, s)
    mesh.dirichlet_mirror(f, [0.0])
    for it in (0, 3, 40):
        if it:
            mesh.sor_iterate(f, sf, 1.4, [0.0], it)
        got = mesh.sor_residual(f, sf)
        want = oracle.sor_residual(blocks, n, 1, f.download(), s)
        assert abs(got - want) <= 1e-12 * want, (it, got, want)


def test_sor_to_tolerance_vs_oracle():
    """The sor2 outer iteration run to relres 1e-8 at omega 1.4 (omega 1.6
    diverges with ghosts frozen between exchanges) on the level-3 shell mesh
    with 8^3 blocks: the GPU iterate after the same number of sweeps equals
    the oracle's bit for bit, and the oracle's residual of it is <= 1e-8."""
    blocks, n, om = W.shell_tree(3).blocks, 8, 1.4
    case(f"sor to tolerance nb={len(blocks)} n={n}")
    s = -W.rhs_field(blocks, n, 1, seed=12)
    mesh = Mesh(blocks, n, 1)
    f = mesh.field(np.float64)
    sf = mesh.field(np.float64, s)
    mesh.dirichlet_mirror(f, [0.0])
    sweeps, chunk = 0, 250
    while True:
        mesh.sor_iterate(f, sf, om, [0.0], chunk)
        sweeps += chunk
        rr = mesh.sor_residual(f, sf)
        if rr <= 1e-8 or sweeps >= 20000:
            break
    assert rr <= 1e-8, (sweeps, rr)
    ref = np.zeros_like(s)
    oracle.dirichlet(blocks, n, 1, ref, [0.0])
    oracle.sor_iterate(blocks, n, 1, ref, s, om, [0.0], sweeps)
    np.testing.assert_array_equal(f.download(), ref)
    assert oracle.sor_residual(blocks, n, 1, ref, s) <= 1e-8


def test_full_size_config2_sor_vs_oracle():
    """Config 2 at full size: 4 outer iterations (1 classic + 3 halo-gathering) bit-exact vs the oracle."""
    t = W.config2_tree()
    blocks, n = t.blocks, 32
    case(f"full-size config2 sor nb={len(blocks)}")
    s = -W.rhs_field(blocks, n, 1, seed=0)
    mesh = Mesh(blocks, n, 1)
    f = mesh.field(np.float64)
    sf = mesh.field(np.float64, s)
    mesh.dirichlet_mirror(f, [0.0])
    mesh.sor_iterate(f, sf, 1.6, [0.0], 4)
    out = f.download()
    ref = np.zeros_like(s)
    oracle.dirichlet(blocks, n, 1, ref, [0.0])
    oracle.sor_iterate(blocks, n, 1, ref, s, 1.6, [0.0], 4)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("n,maker", [(16, lambda: W.shell_tree(3)), (32, lambda: W.shell_tree(2)),
                                     (8, lambda: W.random_tree(6, max_level=3))])
def test_every_sor_schedule_matches_the_oracle_and_autotune(n, maker):
    """All outer-iteration schedules (fused + task fill, fused + block-wise
    gather, halo-gathering, two colour passes) give the reference-order
    result bit for bit; the autotuner leaves f untouched, picks a supported
    schedule, and the mesh then iterates with it."""
    blocks = maker().blocks
    vals = [0.5]
    case(f"sor schedules n={n} nb={len(blocks)}")
    rng = np.random.default_rng(3 * n + len(blocks))
    f0 = rng.standard_normal((len(blocks), n + 2, n + 2, n + 2))
    s = W.rhs_field(blocks, n, 1, seed=8)
    ref = f0.copy()
    oracle.dirichlet(blocks, n, 1, ref, vals)
    oracle.sor_iterate(blocks, n, 1, ref, s, 1.3, vals, 5)
    mesh = Mesh(blocks, n, 1)
    sf = mesh.field(np.float64, s)
    ran = []
    for v in range(1, L.SOR_NUM_VARIANTS):
        try:
            mesh.set_sor_variant(v)
        except L.BCMError:
            continue  # not supported on this mesh
        f = mesh.field(np.float64, f0)
        mesh.sor_iterate(f, sf, 1.3, vals, 5)
        np.testing.assert_array_equal(f.download(), ref, err_msg=f"variant {v}")
        ran.append(v)
    assert L.SOR_FUSED_TASKS in ran and L.SOR_COLOUR_TASKS in ran
    f = mesh.field(np.float64, f0)
    best, times = mesh.sor_autotune(f, sf, 1.3, vals, iters=3)
    assert best in ran and set(times) == set(ran)
    np.testing.assert_array_equal(f.download(), f0)
    assert mesh.sor_variant() == best
    mesh.sor_iterate(f, sf, 1.3, vals, 5)
    np.testing.assert_array_equal(f.download(), ref)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("iters", [1, 2, 3])
@pytest.mark.parametrize("n,level", [(16, 3), (32, 2)])
def test_halo_schedule_equals_classic_schedule(dtype, iters, n, level):
    """The halo-gathering schedule (compact iterates; its last sweep writes
    the field) gives the classic schedule's bits for 1, 2 and 3 iterations,
    FP64 and FP32, sor1 and sor2 forms, blocks of 16^3 and 32^3 (every
    instance of the first, middle and last sweep kernels)."""
    blocks = W.shell_tree(level).blocks
    vals = [0.25]
    case(f"halo vs classic {np.dtype(dtype).name} n={n} iters={iters} nb={len(blocks)}")
    rng = np.random.default_rng(iters)
    f0 = rng.standard_normal((len(blocks), n + 2, n + 2, n + 2)).astype(dtype)
    s = W.rhs_field(blocks, n, 1, seed=5).astype(dtype)
    for src in (None, s):
        out = []
        for variant in (L.SOR_HALO, L.SOR_FUSED_TASKS):
            mesh = Mesh(blocks, n, 1)
            mesh.set_sor_variant(variant)
            f = mesh.field(dtype, f0)
            sf = mesh.field(dtype, src) if src is not None else None
            mesh.sor_iterate(f, sf, 1.25, vals, iters)
            out.append(f.download())
        np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("variant", [L.SOR_AUTO, L.SOR_HALO, L.SOR_FUSED_TASKS, L.SOR_FUSED_GATHER,
                                     L.SOR_COLOUR_TASKS])
@pytest.mark.parametrize("inner", [2, 3])
def test_sor2_inner_loop_vs_oracle(variant, inner):
    """nLoopInner > 1 (examples/sor2/Solver.cpp:106-109): `inner` sweeps +
    in-block BC per block between exchanges, sor2 form with a source, on
    every schedule (the halo one hands over to the fused sweeps)."""
    n, level = 16, 3
    blocks = W.shell_tree(level).blocks
    case(f"sor2 inner={inner} variant={variant} n={n} nb={len(blocks)}")
    rng = np.random.default_rng(inner)
    f0 = rng.standard_normal((len(blocks), n + 2, n + 2, n + 2))
    s = W.rhs_field(blocks, n, 1, seed=9)
    vals = [0.75]
    mesh = Mesh(blocks, n, 1)
    mesh.set_sor_variant(variant)
    f = mesh.field(np.float64, f0)
    mesh.sor_iterate(f, mesh.field(np.float64, s), 1.4, vals, 4, inner)
    ref = f0.copy()
    oracle.dirichlet(blocks, n, 1, ref, vals)
    oracle.sor_iterate(blocks, n, 1, ref, s, 1.4, vals, 4, inner)
    np.testing.assert_array_equal(f.download(), ref)


def test_sor_iterate_inner_rejects_a_loop_count_below_one():
    blocks = W.shell_tree(2).blocks
    mesh = Mesh(blocks, 16, 1)
    f = mesh.field(np.float64)
    v = np.array([0.0])
    assert mesh.lib.bcm_sor_iterate_inner(f.handle, None, 1.2, L.dptr(v), 1, 1, 0) == L.ERR_ARG


@pytest.mark.parametrize("variant,inner", [(L.SOR_HALO, 1), (L.SOR_FUSED_TASKS, 1), (L.SOR_HALO, 2)])
def test_sor_iterate_on_a_batch_equals_each_mesh_alone(variant, inner):
    """Several independent meshes in one device mesh (segments): one
    sor_iterate over the batch gives every segment the bits of that mesh
    iterated alone, with the halo schedule and the classic one, and with
    two inner sweeps per exchange."""
    n = 16
    tables = [W.random_tree(s, max_level=lv).blocks for s, lv in ((31, 3), (32, 2), (33, 3))]
    case(f"sor batch variant={variant} inner={inner} segments={[len(t) for t in tables]}")
    rng = np.random.default_rng(variant)
    f0 = [rng.standard_normal((len(t), n + 2, n + 2, n + 2)) for t in tables]
    s0 = [W.rhs_field(t, n, 1, seed=40 + i) for i, t in enumerate(tables)]
    mesh = Mesh.batch(tables, n)
    mesh.set_sor_variant(variant)
    f = mesh.field(np.float64, np.concatenate(f0))
    sf = mesh.field(np.float64, np.concatenate(s0))
    mesh.sor_iterate(f, sf, 1.3, [0.5], 5, inner)
    out = f.download()
    for i, t in enumerate(tables):
        first, cnt = mesh.segment(i)
        alone = Mesh(t, n, 1)
        alone.set_sor_variant(variant)
        g = alone.field(np.float64, f0[i])
        alone.sor_iterate(g, alone.field(np.float64, s0[i]), 1.3, [0.5], 5, inner)
        np.testing.assert_array_equal(out[first:first + cnt], g.download())
        ref = f0[i].copy()
        oracle.dirichlet(t, n, 1, ref, [0.5])
        oracle.sor_iterate(t, n, 1, ref, s0[i], 1.3, [0.5], 5, inner)
        np.testing.assert_array_equal(g.download(), ref)
