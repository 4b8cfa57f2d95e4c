"""ctypes binding of the CPU oracle (oracle/bcm_oracle.c -> liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path
(bcmtools_amd, libbcm_hip.so) never imports it.

Fields are numpy arrays of shape (nblocks, n+2vc, n+2vc, n+2vc) in the
Index3DS order of include/Index3D.h:35-59 (x fastest, i.e. C order k, j, i).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "oracle_ghost_fill_d": [_P, _I, _I, _I, _P, _I],
        "oracle_ghost_fill_f": [_P, _I, _I, _I, _P, _I],
        "oracle_dirichlet_d": [_P, _I, _I, _I, _P, _P],
        "oracle_dirichlet_f": [_P, _I, _I, _I, _P, _P],
        "oracle_impose_bc_d": [_P, _I, _I, _I, _P, _P],
        "oracle_impose_bc_f": [_P, _I, _I, _I, _P, _P],
        "oracle_sor_sweep_d": [_P, _I, _I, _I, _P, _P, _D],
        "oracle_sor_iterate_d": [_P, _I, _I, _I, _P, _P, _D, _P, _I],
        "oracle_sor_iterate_inner_d": [_P, _I, _I, _I, _P, _P, _D, _P, _I, _I],
        "oracle_jacobi7_d": [_I, _I, _I, _P, _P, _P, _P, _D],
        "oracle_jacobi7_f": [_I, _I, _I, _P, _P, _P, _P, _D],
        "oracle_calc_ax_d": [_I, _I, _I, _P, _P, _P],
        "oracle_calc_ax_f": [_I, _I, _I, _P, _P, _P],
        "oracle_calc_r_d": [_I, _I, _I, _P, _P, _P, _P],
        "oracle_calc_r_f": [_I, _I, _I, _P, _P, _P, _P],
        "oracle_blas_d": [_I, _I, _I, _I, _P, _P, _P, _D, _D],
        "oracle_blas_f": [_I, _I, _I, _I, _P, _P, _P, _D, _D],
        "oracle_poisson_assemble_d": [_P, _I, _I, _I, _P, _P, _P, _P],
        "oracle_bicgstab_d": [_P, _I, _I, _I, _P, _P, _P, _P, _D, _I, _I, _D, _P, _P],
        "oracle_jacobi_solve_d": [_P, _I, _I, _I, _P, _P, _P, _P, _D, _I, _D, _P, _P],
        "oracle_bicgstab_mask_d": [_P, _I, _I, _I, _P, _P, _P, _P, _D, _I, _I, _D, _P, _P, _P],
        "oracle_jacobi_solve_mask_d": [_P, _I, _I, _I, _P, _P, _P, _P, _D, _I, _D, _P, _P, _P],
    }
    for suf in ("d", "f"):
        sig["oracle_jacobi7_mask_" + suf] = sig["oracle_jacobi7_" + suf] + [_P]
        sig["oracle_calc_ax_mask_" + suf] = sig["oracle_calc_ax_" + suf] + [_P]
        sig["oracle_calc_r_mask_" + suf] = sig["oracle_calc_r_" + suf] + [_P]
        sig["oracle_blas_mask_" + suf] = sig["oracle_blas_" + suf] + [_P]
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = None
    for name in ("oracle_calc_r2_d", "oracle_calc_r2_f", "oracle_dot_d", "oracle_dot_f"):
        for nm, extra in ((name, []), (name[:-2] + "_mask" + name[-2:], [_P])):
            fn = getattr(lib, nm)
            fn.argtypes = ([_I, _I, _I, _P, _P, _P] if "r2" in nm else [_I, _I, _I, _P, _P]) + extra
            fn.restype = _D
    lib.oracle_bicgstab_mask_d.restype = _I
    lib.oracle_jacobi_solve_mask_d.restype = _I
    lib.oracle_bicgstab_d.restype = _I
    lib.oracle_jacobi_solve_d.restype = _I
    _lib = lib
    return lib


def _suf(a):
    return "d" if a.dtype == np.float64 else "f"


def _ptr(a):
    return a.ctypes.data if a is not None else None


def _arr7(A):
    arr = (ctypes.c_void_p * 7)()
    for i, a in enumerate(A):
        arr[i] = a.ctypes.data
    return arr


def ghost_fill(blocks, n, vc, data, mask=7):
    """In-place ghost fill of `data` (float64 or float32)."""
    lib = load()
    getattr(lib, "oracle_ghost_fill_" + _suf(data))(blocks.ctypes.data, len(blocks), n, vc, data.ctypes.data, mask)


def impose_bc(blocks, n, vc, data, values):
    lib = load()
    v = np.ascontiguousarray(values, dtype=np.float64)
    getattr(lib, "oracle_impose_bc_" + _suf(data))(blocks.ctypes.data, len(blocks), n, vc, data.ctypes.data,
                                                   v.ctypes.data)


def dirichlet(blocks, n, vc, data, values):
    lib = load()
    v = np.ascontiguousarray(values, dtype=np.float64)
    getattr(lib, "oracle_dirichlet_" + _suf(data))(blocks.ctypes.data, len(blocks), n, vc, data.ctypes.data,
                                                   v.ctypes.data)


def sor_sweep(blocks, n, vc, f, s, omega):
    load().oracle_sor_sweep_d(blocks.ctypes.data, len(blocks), n, vc, f.ctypes.data, _ptr(s), float(omega))


def sor_iterate(blocks, n, vc, f, s, omega, values, iters, inner=1):
    """`iters` outer iterations of examples/sor{1,2}/Solver.cpp's run() with
    nLoopInner = `inner` sweeps per block between exchanges."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    load().oracle_sor_iterate_inner_d(blocks.ctypes.data, len(blocks), n, vc, f.ctypes.data, _ptr(s), float(omega),
                                      v.ctypes.data, int(iters), int(inner))


def sor_residual(blocks, n, vc, f, s):
    """Relative residual of the sor2 system lap_h(f) = s over the interior
    cells: ||s - lap_h f||_2 / ||s||_2, lap_h f = cx((fW + fE) - 2 fc) +
    cy((fS + fN) - 2 fc) + cz((fB + fT) - 2 fc) with c = 1/h^2 per axis of
    each block (the operator whose sweep examples/sor2/Solver.cpp:136-160
    relaxes; the reference itself reports no residual).  Plain numpy."""
    nb = len(blocks)
    v = vc
    c = f[:, v:v + n, v:v + n, v:v + n]
    h = blocks["block_size"] / n
    lam = 1.0 / (h * h)
    cx, cy, cz = (lam[:, a][:, None, None, None] for a in range(3))
    lap = ((cx * ((f[:, v:v + n, v:v + n, v - 1:v + n - 1] + f[:, v:v + n, v:v + n, v + 1:v + n + 1]) - 2.0 * c)
            + cy * ((f[:, v:v + n, v - 1:v + n - 1, v:v + n] + f[:, v:v + n, v + 1:v + n + 1, v:v + n]) - 2.0 * c))
           + cz * ((f[:, v - 1:v + n - 1, v:v + n, v:v + n] + f[:, v + 1:v + n + 1, v:v + n, v:v + n]) - 2.0 * c))
    si = s[:, v:v + n, v:v + n, v:v + n]
    r = si - lap
    ss = float(np.sum(si * si))
    rr = float(np.sum(r * r))
    assert nb == f.shape[0]
    return float(np.sqrt(rr / ss)) if ss > 0.0 else float(np.sqrt(rr))


def setup_mask(shape, vc, dtype=np.float64):
    """setup_mask (examples/SolverTest/blas.f90:700-751): every cell 0, then
    the interior cells 1.  `shape` is (nblocks, n+2vc, n+2vc, n+2vc)."""
    m = np.zeros(shape, dtype=dtype)
    n = shape[-1] - 2 * vc
    m[:, vc:vc + n, vc:vc + n, vc:vc + n] = 1
    return m


def _m(name, arr, mask):
    """Symbol of the plain or the masked (mask != None) variant."""
    return getattr(load(), name + ("_mask_" if mask is not None else "_") + _suf(arr))


def _mask_args(mask):
    return (mask.ctypes.data,) if mask is not None else ()


def jacobi7(nb, n, vc, x1, x0, A, b, omega, mask=None):
    _m("oracle_jacobi7", x1, mask)(nb, n, vc, x1.ctypes.data, x0.ctypes.data, _arr7(A), b.ctypes.data, float(omega),
                                   *_mask_args(mask))


def calc_ax(nb, n, vc, ax, A, x, mask=None):
    _m("oracle_calc_ax", ax, mask)(nb, n, vc, ax.ctypes.data, _arr7(A), x.ctypes.data, *_mask_args(mask))


def calc_r(nb, n, vc, r, A, x, b, mask=None):
    _m("oracle_calc_r", r, mask)(nb, n, vc, r.ctypes.data, _arr7(A), x.ctypes.data, b.ctypes.data,
                                 *_mask_args(mask))


def calc_r2(nb, n, vc, A, x, b, mask=None):
    return _m("oracle_calc_r2", x, mask)(nb, n, vc, _arr7(A), x.ctypes.data, b.ctypes.data, *_mask_args(mask))


def dot(nb, n, vc, x, y, mask=None):
    return _m("oracle_dot", x, mask)(nb, n, vc, x.ctypes.data, y.ctypes.data, *_mask_args(mask))


BLAS_OPS = {"axpy": 0, "xpay": 1, "axpyz": 2, "axpbypz": 3, "copy": 4}


def blas(nb, n, vc, op, z, x, y=None, a=0.0, b=0.0, mask=None):
    _m("oracle_blas", z, mask)(nb, n, vc, BLAS_OPS[op], z.ctypes.data, x.ctypes.data, _ptr(y), float(a), float(b),
                               *_mask_args(mask))


def poisson_assemble(blocks, n, vc, A, b, rhs, values):
    v = np.ascontiguousarray(values, dtype=np.float64)
    load().oracle_poisson_assemble_d(blocks.ctypes.data, len(blocks), n, vc, _arr7(A) if A is not None else None,
                                     b.ctypes.data, rhs.ctypes.data, v.ctypes.data)


def bicgstab(blocks, n, vc, x, A, b, work, omega, count_pre, count_max, eps, mask=None):
    w = (ctypes.c_void_p * 9)()
    for i, a in enumerate(work):
        w[i] = a.ctypes.data
    cnt = ctypes.c_int()
    res = ctypes.c_double()
    fn = load().oracle_bicgstab_mask_d if mask is not None else load().oracle_bicgstab_d
    st = fn(blocks.ctypes.data, len(blocks), n, vc, x.ctypes.data, _arr7(A), b.ctypes.data, w, float(omega),
            int(count_pre), int(count_max), float(eps), ctypes.byref(cnt), ctypes.byref(res), *_mask_args(mask))
    return st, cnt.value, res.value


def jacobi_solve(blocks, n, vc, x, A, b, x0, omega, count_max, eps, mask=None):
    """Returns (status, count, residual, result array) -- the result may live in x or x0 after swaps."""
    px = ctypes.c_void_p(x.ctypes.data)
    p0 = ctypes.c_void_p(x0.ctypes.data)
    cnt = ctypes.c_int()
    res = ctypes.c_double()
    fn = load().oracle_jacobi_solve_mask_d if mask is not None else load().oracle_jacobi_solve_d
    st = fn(blocks.ctypes.data, len(blocks), n, vc, ctypes.byref(px), _arr7(A), b.ctypes.data, ctypes.byref(p0),
            float(omega), int(count_max), float(eps), ctypes.byref(cnt), ctypes.byref(res), *_mask_args(mask))
    out = x if px.value == x.ctypes.data else x0
    return st, cnt.value, res.value, out
