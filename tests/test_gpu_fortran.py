"""SolverTest's per-block Fortran kernels through the library's
Fortran-convention entry points (include/bcm_fortran.h), on device arrays,
against the numpy restatement in oracle/solvertest.py.

Element-wise and stencil results are bit for bit (the same statements in
the same order, no fused multiply-add); sums (dot, calc_r2, dotx2,
dot_mask) to 1e-13 relative, since the device sums by a tree.  Block sizes
are not cubic, one and two ghost layers; sup.f90's boxes are driven as
examples/SolverTest/Scalar3DUpdater1.cpp lays them and must reproduce the
compiled reference library's ghost fill (tests/golden fixtures).  Calls
from several host threads at once (an OpenMP block loop) give the results
of the calls made one by one.  A bad array argument is caught on the host
before any kernel runs.
"""
import ctypes
import subprocess
import sys
import threading

import numpy as np
import pytest

from bcmtools_amd import _lib as L
from oracle import solvertest as S

from conftest import ROOT, load_golden
from hipmem import DeviceArray

pytestmark = pytest.mark.gpu

SZ = (6, 5, 4)
RTOL_SUM = 1e-13


def case(msg):
    print(f"[gpu case] {msg}", flush=True)


def F(name):
    return L.fortran(name)


def ivec(v):
    return (ctypes.c_int * 3)(*v)


def iref(v):
    return ctypes.byref(ctypes.c_int(v))


def dref(v):
    return ctypes.byref(ctypes.c_double(v))


def dev(a):
    return DeviceArray(a)


def rand_coef(rng, sz, g):
    A = [rng.uniform(0.1, 1.0, S.shape(sz, g)) for _ in range(7)]
    A[0] = -(2.0 + rng.uniform(0.5, 1.0, S.shape(sz, g)))
    return A


@pytest.mark.parametrize("g", [1, 2])
def test_blas_elementwise_bit_exact(g):
    case(f"fortran blas element-wise sz={SZ} g={g}")
    rng = np.random.default_rng(40 + g)
    x, y, z = (rng.standard_normal(S.shape(SZ, g)) for _ in range(3))
    sz, gg = ivec(SZ), iref(g)
    ops = [
        ("axpy_", lambda Z, X, Y: S.axpy(Z, X, 0.3, SZ, g), lambda Z, X, Y: F("axpy_")(Z, X, dref(0.3), sz, gg)),
        ("xpay_", lambda Z, X, Y: S.xpay(Z, X, -1.7, SZ, g), lambda Z, X, Y: F("xpay_")(Z, X, dref(-1.7), sz, gg)),
        ("axpyz_", lambda Z, X, Y: S.axpyz(Z, X, Y, 2.5, SZ, g),
         lambda Z, X, Y: F("axpyz_")(Z, X, Y, dref(2.5), sz, gg)),
        ("axpbypz_", lambda Z, X, Y: S.axpbypz(Z, X, Y, 0.7, -0.2, SZ, g),
         lambda Z, X, Y: F("axpbypz_")(Z, X, Y, dref(0.7), dref(-0.2), sz, gg)),
        ("add_", lambda Z, X, Y: S.add(Z, X, Y, SZ, g), lambda Z, X, Y: F("add_")(Z, X, Y, sz, gg)),
        ("triad_", lambda Z, X, Y: S.triad(Z, X, Y, 1.3, SZ, g),
         lambda Z, X, Y: F("triad_")(Z, X, Y, dref(1.3), sz, gg)),
        ("scal_", lambda Z, X, Y: S.scal(Z, 0.9, SZ, g), lambda Z, X, Y: F("scal_")(Z, dref(0.9), sz, gg)),
    ]
    for name, ref, call in ops:
        want = z.copy()
        ref(want, x, y)
        Z, X, Y = dev(z), dev(x), dev(y)
        call(Z.ptr, X.ptr, Y.ptr)
        np.testing.assert_array_equal(Z.get(), want, err_msg=name)
        np.testing.assert_array_equal(X.get(), x, err_msg=name + " changed an input")


def test_fill_copy_and_vector_fill():
    case("fortran fill/copy")
    rng = np.random.default_rng(3)
    g = 2
    sz, gg = ivec(SZ), iref(g)
    x = rng.standard_normal(S.shape(SZ, g))
    X = dev(x)
    F("fill_")(X.ptr, dref(-3.25), sz, gg)
    assert np.all(X.get() == -3.25)
    Y = dev(np.zeros_like(x))
    X.put(x)
    F("copy_")(Y.ptr, X.ptr, sz, gg)
    np.testing.assert_array_equal(Y.get(), x)
    xi = rng.integers(-9, 9, S.shape(SZ, g)).astype(np.int32)
    XI, YI = dev(xi), dev(np.zeros_like(xi))
    F("copy_integer_")(YI.ptr, XI.ptr, sz, gg)
    np.testing.assert_array_equal(YI.get(), xi)
    v = dev(np.zeros((3,) + S.shape(SZ, g)))
    F("fill_vf3d_")(v.ptr, dref(7.5), sz, gg, iref(3))
    assert np.all(v.get() == 7.5)


@pytest.mark.parametrize("g", [1, 2])
def test_stencils_bit_exact_and_sums(g):
    case(f"fortran stencils sz={SZ} g={g}")
    rng = np.random.default_rng(50 + g)
    A = rand_coef(rng, SZ, g)
    x0, b, y = (rng.standard_normal(S.shape(SZ, g)) for _ in range(3))
    sz, gg = ivec(SZ), iref(g)
    DA = [dev(a) for a in A]
    AP = [d.ptr for d in DA]
    X0, B, Y = dev(x0), dev(b), dev(y)
    out0 = rng.standard_normal(S.shape(SZ, g))  # ghost cells of outputs must stay as they were

    want = out0.copy()
    S.jacobi_smoother(want, x0, A, b, 0.8, SZ, g)
    X1 = dev(out0)
    F("jacobi_smoother_")(X1.ptr, X0.ptr, *AP, B.ptr, dref(0.8), sz, gg)
    np.testing.assert_array_equal(X1.get(), want)

    P = dev(S.pack7(A))
    X1.put(out0)
    F("jacobi_smoother2_")(X1.ptr, X0.ptr, P.ptr, B.ptr, dref(0.8), sz, gg)
    np.testing.assert_array_equal(X1.get(), want)

    want = out0.copy()
    S.calc_ax(want, A, x0, SZ, g)
    AX = dev(out0)
    F("calc_ax_")(AX.ptr, *AP, X0.ptr, sz, gg)
    np.testing.assert_array_equal(AX.get(), want)

    want = out0.copy()
    S.calc_r(want, A, x0, b, SZ, g)
    R = dev(out0)
    F("calc_r_")(R.ptr, *AP, X0.ptr, B.ptr, sz, gg)
    np.testing.assert_array_equal(R.get(), want)

    rr = ctypes.c_double(-1.0)
    F("calc_r2_")(ctypes.byref(rr), *AP, X0.ptr, B.ptr, sz, gg)
    assert rr.value == pytest.approx(S.calc_r2(A, x0, b, SZ, g), rel=RTOL_SUM)

    xy, xz = ctypes.c_double(), ctypes.c_double()
    F("dot_")(ctypes.byref(xy), Y.ptr, X0.ptr, sz, gg)
    assert xy.value == pytest.approx(S.dot(x0, y, SZ, g), rel=RTOL_SUM)
    F("dot_all_")(ctypes.byref(xz), Y.ptr, X0.ptr, sz, gg)
    assert xz.value == xy.value
    F("dotx2_")(ctypes.byref(xy), ctypes.byref(xz), X0.ptr, Y.ptr, B.ptr, sz, gg)
    assert xy.value == pytest.approx(S.dot(x0, y, SZ, g), rel=RTOL_SUM)
    assert xz.value == pytest.approx(S.dot(x0, b, SZ, g), rel=RTOL_SUM)


def test_masked_forms_over_the_flat_range():
    case("fortran masked forms")
    rng = np.random.default_rng(60)
    g = 1
    A = rand_coef(rng, SZ, g)
    x0, b = (rng.standard_normal(S.shape(SZ, g)) for _ in range(2))
    sz, gg = ivec(SZ), iref(g)
    M, MK = dev(np.full(S.shape(SZ, g), 9.0)), dev(np.full(S.shape(SZ, g), 9, dtype=np.int32))
    F("setup_mask_")(M.ptr, MK.ptr, sz, gg)
    m, mk = S.setup_mask(SZ, g)
    np.testing.assert_array_equal(M.get(), m)
    np.testing.assert_array_equal(MK.get(), mk)
    rand_mask = (rng.uniform(size=S.shape(SZ, g)) < 0.6).astype(np.int32)
    DA = [dev(a) for a in A]
    AP = [d.ptr for d in DA]
    X0, B = dev(x0), dev(b)
    for label, mask in (("setup", mk), ("random", rand_mask)):
        MASK = dev(mask)
        out0 = rng.standard_normal(S.shape(SZ, g))
        want = out0.copy()
        S.jacobi_smoother_mask(want, x0, A, b, mask, 1.1, SZ, g)
        X1 = dev(out0)
        F("jacobi_smoother_mask_")(X1.ptr, X0.ptr, *AP, B.ptr, MASK.ptr, dref(1.1), sz, gg)
        np.testing.assert_array_equal(X1.get(), want, err_msg=label)
        want = out0.copy()
        S.calc_ax_mask(want, A, x0, mask, SZ, g)
        AX = dev(out0)
        F("calc_ax_mask_")(AX.ptr, *AP, X0.ptr, MASK.ptr, sz, gg)
        np.testing.assert_array_equal(AX.get(), want, err_msg=label)
        want = out0.copy()
        S.copy_mask(want, x0, mask, SZ, g)
        Y = dev(out0)
        F("copy_mask_")(Y.ptr, X0.ptr, MASK.ptr, sz, gg)
        np.testing.assert_array_equal(Y.get(), want, err_msg=label)
        xy = ctypes.c_double()
        F("dot_mask_")(ctypes.byref(xy), B.ptr, X0.ptr, MASK.ptr, sz, gg)
        assert xy.value == pytest.approx(S.dot_mask(x0, b, mask, SZ, g), rel=RTOL_SUM), label


@pytest.mark.parametrize("name", ["ghost_sphere_atonce_d", "ghost_sphere_vc2_d", "ghost_periodic_x_d"])
def test_sup_boxes_reproduce_the_reference_ghost_fill(name):
    gd = load_golden(name)
    blocks, n, vc = gd["blocks"], gd["n"], gd["vc"]
    case(f"fortran sup boxes {name} nb={len(blocks)}")
    fin = gd["fin"]
    devs = [dev(fin[b]) for b in range(len(blocks))]
    sz, gg = ivec((n, n, n)), iref(vc)
    # Through a buffer for the first block of each kind, directly otherwise.
    via_buffer = set()
    want = fin.copy()
    for b in range(len(blocks)):
        for face in range(6):
            d = int(blocks["nb_level_diff"][b][face])
            ids = blocks["nb_id"][b][face]
            subs = [(s, int(ids[s])) for s in range(4)] if d == 1 else \
                [(int(blocks["nb_subface"][b][face]) if d == -1 else 0, int(ids[0]))]
            for sub, nid in subs:
                if nid < 0:
                    continue
                i1d, i1s, szc, op = S.updater1_boxes(face, d, sub, n, vc)
                S.region_store(want[b], i1d, S.region_values(fin[nid], i1s, szc, (n, n, n), vc, op), vc)
                suffix = {"copy": "", "c2f": "c2f_", "f2c": "f2c_"}[op]
                if op not in via_buffer:
                    via_buffer.add(op)
                    buf = dev(np.zeros(int(np.prod(szc))))
                    F("sup_copy_to_buffer_" + suffix)(buf.ptr, devs[nid].ptr, ivec(i1s), ivec(szc), sz, gg)
                    F("sup_copy_from_buffer_")(devs[b].ptr, ivec(i1d), buf.ptr, ivec(szc), sz, gg)
                else:
                    F("sup_copy_from_neighbor_" + suffix)(devs[b].ptr, ivec(i1d), devs[nid].ptr, ivec(i1s),
                                                          ivec(szc), sz, gg)
    got = np.stack([d.get() for d in devs])
    np.testing.assert_array_equal(got, want)
    # Same-level and C2F ghosts are the reference library's; F2C to rounding.
    np.testing.assert_allclose(got, gd["fout"], rtol=4e-16, atol=4e-16)


@pytest.mark.parametrize("face", range(6))
def test_face_conditions_bit_exact(face):
    case(f"fortran bcx/bca face {face}")
    rng = np.random.default_rng(70 + face)
    g = 2
    sz, gg = ivec(SZ), iref(g)
    names = {0: "3", 1: "1", 2: "4", 3: "2", 4: "6", 5: "5"}
    letters = "wesnbt"
    x = rng.standard_normal(S.shape(SZ, g))
    for kind in ("d", "n", "d_f"):
        want = x.copy()
        S.bcx(want, face, kind, 0.61, SZ, g)
        X = dev(x)
        F(f"bc_x{names[face]}_{kind}_")(X.ptr, dref(0.61), sz, gg)
        np.testing.assert_array_equal(X.get(), want, err_msg=kind)
    for kind in ("d", "n"):
        A = rand_coef(rng, SZ, g)
        b = rng.standard_normal(S.shape(SZ, g))
        Ap, An, bw = A[0].copy(), A[face + 1].copy(), b.copy()
        S.bca(Ap, An, bw, face, kind, -0.45, SZ, g)
        DAp, DAn, DB = dev(A[0]), dev(A[face + 1]), dev(b)
        F(f"bc_a{letters[face]}_{kind}_")(DAp.ptr, DAn.ptr, DB.ptr, dref(-0.45), sz, gg)
        np.testing.assert_array_equal(DAp.get(), Ap, err_msg=kind)
        np.testing.assert_array_equal(DAn.get(), An, err_msg=kind)
        np.testing.assert_array_equal(DB.get(), bw, err_msg=kind)
    if face % 2:
        axis = face // 2
        want = x.copy()
        S.bcx_periodic(want, axis, SZ, g)
        X = dev(x)
        F(f"bc_x{names[face]}_p_")(X.ptr, dref(0.0), sz, gg)
        np.testing.assert_array_equal(X.get(), want)


def test_poiseuille_inflow():
    case("fortran poiseuille")
    rng = np.random.default_rng(80)
    sz_t, g = (4, 8, 8), 2
    sz, gg = ivec(sz_t), iref(g)
    center, org, bsize, csize = (0.0, 0.5, 0.5), (0.0, 0.0, 0.0), (0.5, 1.0, 1.0), (0.125, 0.125, 0.125)
    cv = lambda v: (ctypes.c_double * 3)(*v)
    x = rng.standard_normal(S.shape(sz_t, g))
    for name, ref in (("bc_x3_poiseuille_u_", S.bc_x3_poiseuille_u), ("bc_x3_poiseuille_p_", S.bc_x3_poiseuille_p)):
        want = x.copy()
        ref(want, 1.3, sz_t, g, center, org, csize)
        X = dev(x)
        F(name)(X.ptr, dref(1.3), sz, gg, cv(center), cv(org), cv(bsize), cv(csize))
        np.testing.assert_array_equal(X.get(), want, err_msg=name)
    for name, ref in (("bc_aw_poiseuille_u_", S.bc_aw_poiseuille_u), ("bc_aw_poiseuille_p_", S.bc_aw_poiseuille_p)):
        Ap, Aw, b = (rng.standard_normal(S.shape(sz_t, g)) for _ in range(3))
        wAp, wAw, wb = Ap.copy(), Aw.copy(), b.copy()
        ref(wAp, wAw, wb, 1.3, sz_t, g, center, org, csize)
        DAp, DAw, DB = dev(Ap), dev(Aw), dev(b)
        F(name)(DAp.ptr, DAw.ptr, DB.ptr, dref(1.3), sz, gg, cv(center), cv(org), cv(bsize), cv(csize))
        np.testing.assert_array_equal(DAp.get(), wAp, err_msg=name)
        np.testing.assert_array_equal(DAw.get(), wAw, err_msg=name)
        np.testing.assert_array_equal(DB.get(), wb, err_msg=name)
    tot = ctypes.c_double()
    for name in ("allreduce_", "allreduce_max_", "allreduce_min_"):
        F(name)(ctypes.byref(tot), dref(4.5))
        assert tot.value == 4.5


def test_concurrent_host_threads_match_serial_calls():
    """Eight host threads, each running Jacobi sweeps and dots on its own
    block, as an OpenMP block loop calls the kernels."""
    case("fortran concurrent host threads")
    rng = np.random.default_rng(90)
    g, nb, sweeps = 1, 8, 20
    sz_t = (16, 16, 16)
    sz, gg = ivec(sz_t), iref(g)
    A = rand_coef(rng, sz_t, g)
    DA = [dev(a) for a in A]
    AP = [d.ptr for d in DA]
    blocks = [(rng.standard_normal(S.shape(sz_t, g)), rng.standard_normal(S.shape(sz_t, g))) for _ in range(nb)]

    def run(x, b, out):
        X0, X1, B = dev(x), dev(x), dev(b)
        dots = []
        for _ in range(sweeps):
            F("jacobi_smoother_")(X1.ptr, X0.ptr, *AP, B.ptr, dref(0.8), sz, gg)
            X0, X1 = X1, X0
            v = ctypes.c_double()
            F("dot_")(ctypes.byref(v), X0.ptr, X0.ptr, sz, gg)
            dots.append(v.value)
        out.append((X0.get(), dots))

    serial = []
    for x, b in blocks:
        run(x, b, serial)
    results = [[] for _ in range(nb)]
    threads = [threading.Thread(target=run, args=(blocks[i][0], blocks[i][1], results[i])) for i in range(nb)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
        assert not t.is_alive()
    for i in range(nb):
        np.testing.assert_array_equal(results[i][0][0], serial[i][0])
        assert results[i][0][1] == serial[i][1]


def test_bad_arrays_are_caught_before_any_kernel():
    """An array too small for the stated sizes, and a host array, end the
    process with the entry point's message; nothing is launched."""
    case("fortran bad arrays (host-side checks)")
    code = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, 'tests')
from hipmem import DeviceArray
from bcmtools_amd import _lib as L
small = DeviceArray(np.zeros(10))
sz = (ctypes.c_int * 3)(4, 4, 4); g = ctypes.c_int(1)
which = sys.argv[1]
if which == 'small':
    L.fortran('fill_')(small.ptr, ctypes.byref(ctypes.c_double(1.0)), sz, ctypes.byref(g))
else:
    host = np.zeros(6 ** 3)
    L.fortran('fill_')(ctypes.c_void_p(host.ctypes.data), ctypes.byref(ctypes.c_double(1.0)), sz, ctypes.byref(g))
print('returned')
"""
    for which, msg in (("small", "run past the end of its allocation"),
                       ("host", "is not device memory of the current device")):
        out = subprocess.run([sys.executable, "-c", code, which], cwd=ROOT, capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 1, (which, out.stdout, out.stderr)
        assert "bcm: fill_: argument x" in out.stderr and msg in out.stderr, out.stderr
        assert "returned" not in out.stdout
