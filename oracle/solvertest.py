"""numpy restatement of examples/SolverTest's per-block Fortran kernels
(blas.f90, sup.f90, bca.f90, bcx.f90, bcax.f90) with real = double.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
library's Fortran-convention entry points; the product path never imports
it.

An array is one block's values, a numpy array of shape
(sz(3)+2g, sz(2)+2g, sz(1)+2g) in C order (x fastest), the memory of
blas.f90's x(1-g:sz(1)+g, 1-g:sz(2)+g, 1-g:sz(3)+g).  Fortran index
(i, j, k) is element [k-1+g, j-1+g, i-1+g].  Every statement keeps the
Fortran expression's operation order, so element-wise results are the
IEEE results of the reference's statements; sums (dot, calc_r2) are
numpy sums, equal to the reference's sequential sums only to rounding.
Each function cites the Fortran it restates.
"""
import numpy as np


def shape(sz, g):
    return (sz[2] + 2 * g, sz[1] + 2 * g, sz[0] + 2 * g)


def interior(sz, g):
    return (slice(g, g + sz[2]), slice(g, g + sz[1]), slice(g, g + sz[0]))


def _shift(sz, g, di, dj, dk):
    return (slice(g + dk, g + dk + sz[2]), slice(g + dj, g + dj + sz[1]), slice(g + di, g + di + sz[0]))


# ---------------------------------------------------------------- blas.f90
def axpy(y, x, a, sz, g):  # blas.f90:242-274
    I = interior(sz, g)
    y[I] = y[I] + a * x[I]


def xpay(y, x, a, sz, g):  # blas.f90:276-308
    I = interior(sz, g)
    y[I] = x[I] + a * y[I]


def axpyz(z, x, y, a, sz, g):  # blas.f90:310-343
    I = interior(sz, g)
    z[I] = a * x[I] + y[I]


def axpbypz(z, x, y, a, b, sz, g):  # blas.f90:345-377
    I = interior(sz, g)
    z[I] = (a * x[I] + b * y[I]) + z[I]


def add(C, A, B, sz, g):  # blas.f90:142-173
    I = interior(sz, g)
    C[I] = A[I] + B[I]


def triad(C, A, B, d, sz, g):  # blas.f90:175-207
    I = interior(sz, g)
    C[I] = A[I] + d * B[I]


def scal(y, a, sz, g):  # blas.f90:209-240
    I = interior(sz, g)
    y[I] = a * y[I]


def dot(x, y, sz, g):  # blas.f90:379-412
    I = interior(sz, g)
    return float(np.sum(x[I] * y[I]))


def jacobi_smoother(x1, x0, A, b, omega, sz, g):
    """blas.f90:452-502; A = (Ap, Aw, Ae, As, An, Ab, At)."""
    Ap, Aw, Ae, As, An, Ab, At = A
    I = interior(sz, g)
    S = lambda di, dj, dk: x0[_shift(sz, g, di, dj, dk)]
    r = b[I] - Aw[I] * S(-1, 0, 0)
    r = r - Ae[I] * S(1, 0, 0)
    r = r - As[I] * S(0, -1, 0)
    r = r - An[I] * S(0, 1, 0)
    r = r - Ab[I] * S(0, 0, -1)
    r = r - At[I] * S(0, 0, 1)
    xt = r / Ap[I]
    x1[I] = x0[I] - omega * (x0[I] - xt)


def calc_ax(ax, A, x, sz, g):  # blas.f90:556-600
    Ap, Aw, Ae, As, An, Ab, At = A
    I = interior(sz, g)
    S = lambda di, dj, dk: x[_shift(sz, g, di, dj, dk)]
    v = Aw[I] * S(-1, 0, 0)
    v = v + Ae[I] * S(1, 0, 0)
    v = v + As[I] * S(0, -1, 0)
    v = v + An[I] * S(0, 1, 0)
    v = v + Ab[I] * S(0, 0, -1)
    v = v + At[I] * S(0, 0, 1)
    ax[I] = v + Ap[I] * x[I]


def _residual(A, x, b, sz, g):  # blas.f90:632-639
    Ap, Aw, Ae, As, An, Ab, At = A
    I = interior(sz, g)
    S = lambda di, dj, dk: x[_shift(sz, g, di, dj, dk)]
    r = b[I] - Aw[I] * S(-1, 0, 0)
    r = r - Ae[I] * S(1, 0, 0)
    r = r - As[I] * S(0, -1, 0)
    r = r - An[I] * S(0, 1, 0)
    r = r - Ab[I] * S(0, 0, -1)
    r = r - At[I] * S(0, 0, 1)
    return r - Ap[I] * x[I]


def calc_r(r, A, x, b, sz, g):  # blas.f90:602-648
    r[interior(sz, g)] = _residual(A, x, b, sz, g)


def calc_r2(A, x, b, sz, g):  # blas.f90:650-698
    r = _residual(A, x, b, sz, g)
    return float(np.sum(r * r))


def pack7(A):
    """The interleaved A(0:6, i, j, k) of jacobi_smoother2 (blas.f90:504-554)."""
    return np.stack(A, axis=-1).copy()


def setup_mask(sz, g):  # blas.f90:700-751
    m = np.zeros(shape(sz, g))
    mask = np.zeros(shape(sz, g), dtype=np.int32)
    m[interior(sz, g)] = 1.0
    mask[interior(sz, g)] = 1
    return m, mask


def _flat_range(sz, g):
    """Flat offsets of blas.f90's masked loops: ijk0 .. ijkx (1-based) -> 0-based."""
    cx, cy, cz = sz[0] + 2 * g, sz[1] + 2 * g, sz[2] + 2 * g
    ijk0 = cx * cy * g + cx * g + g + 1
    ijkx = cx * cy * cz - ijk0 + 1
    return np.arange(ijk0 - 1, ijkx), cx, cx * cy


def copy_mask(y, x, mask, sz, g):  # blas.f90:969-992
    o, _, _ = _flat_range(sz, g)
    o = o[mask.reshape(-1)[o] == 1]
    y.reshape(-1)[o] = x.reshape(-1)[o]


def dot_mask(x, y, mask, sz, g):  # blas.f90:858-883
    o, _, _ = _flat_range(sz, g)
    o = o[mask.reshape(-1)[o] == 1]
    return float(np.sum(x.reshape(-1)[o] * y.reshape(-1)[o]))


def calc_ax_mask(ax, A, x, mask, sz, g):  # blas.f90:994-1030
    o, sy, sz_ = _flat_range(sz, g)
    o = o[mask.reshape(-1)[o] == 1]
    F = [a.reshape(-1) for a in A]
    xf = x.reshape(-1)
    v = F[1][o] * xf[o - 1]
    v = v + F[2][o] * xf[o + 1]
    v = v + F[3][o] * xf[o - sy]
    v = v + F[4][o] * xf[o + sy]
    v = v + F[5][o] * xf[o - sz_]
    v = v + F[6][o] * xf[o + sz_]
    ax.reshape(-1)[o] = v + F[0][o] * xf[o]


def jacobi_smoother_mask(x1, x0, A, b, mask, omega, sz, g):  # blas.f90:915-967
    o, sy, sz_ = _flat_range(sz, g)
    o = o[mask.reshape(-1)[o] == 1]
    F = [a.reshape(-1) for a in A]
    xf = x0.reshape(-1)
    r = b.reshape(-1)[o] - F[1][o] * xf[o - 1]
    r = r - F[2][o] * xf[o + 1]
    r = r - F[3][o] * xf[o - sy]
    r = r - F[4][o] * xf[o + sy]
    r = r - F[5][o] * xf[o - sz_]
    r = r - F[6][o] * xf[o + sz_]
    xt = r / F[0][o]
    x1.reshape(-1)[o] = xf[o] - omega * (xf[o] - xt)


# ----------------------------------------------------------------- sup.f90
def intp_index(i, n):
    """sup_get_intp_index (sup.f90:18-33), vectorised over fine positions i."""
    i = np.asarray(i)
    ic = i // 2
    r = 0.25 + 0.5 * (i % 2)
    ic = np.where(i < 2, 1, np.where(i > 2 * n - 1, n - 1, ic))
    r = np.where(i < 2, -0.25, np.where(i > 2 * n - 1, 1.25, r))
    return ic, r.astype(np.float64)


def _box(i1, szc):
    k, j, i = np.meshgrid(np.arange(szc[2]), np.arange(szc[1]), np.arange(szc[0]), indexing="ij")
    return i + i1[0], j + i1[1], k + i1[2]


def _at(a, g, i, j, k):
    return a[k - 1 + g, j - 1 + g, i - 1 + g]


def region_values(src, i1_src, szc, sz, g, op):
    """The box values of sup_copy_{from_neighbor,to_buffer}[_c2f|_f2c]
    (sup.f90:35-361), shaped (szc(3), szc(2), szc(1))."""
    i0, j0, k0 = _box(i1_src, szc)
    if op == "copy":
        return _at(src, g, i0, j0, k0).copy()
    if op == "c2f":
        i1, r = intp_index(i0, sz[0])
        j1, s = intp_index(j0, sz[1])
        k1, t = intp_index(k0, sz[2])
        A = lambda di, dj, dk: _at(src, g, i1 + di, j1 + dj, k1 + dk)
        return (1.0 - t) * ((1.0 - s) * ((1.0 - r) * A(0, 0, 0) + r * A(1, 0, 0)) +
                            s * ((1.0 - r) * A(0, 1, 0) + r * A(1, 1, 0))) + \
            t * ((1.0 - s) * ((1.0 - r) * A(0, 0, 1) + r * A(1, 0, 1)) +
                 s * ((1.0 - r) * A(0, 1, 1) + r * A(1, 1, 1)))
    i1, j1, k1 = 2 * i0 - 1, 2 * j0 - 1, 2 * k0 - 1
    A = lambda di, dj, dk: _at(src, g, i1 + di, j1 + dj, k1 + dk)
    s = A(0, 0, 0) + A(1, 0, 0)
    for d in ((0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0), (1, 1, 1)):
        s = s + A(*d)
    return 0.125 * s


def region_store(dst, i1_dst, vals, g):
    """sup_copy_from_neighbor / sup_copy_from_buffer's store."""
    szc = vals.shape[::-1]
    i, j, k = _box(i1_dst, szc)
    dst[k - 1 + g, j - 1 + g, i - 1 + g] = vals


# ------------------------------------------------------------------ bcx.f90
# Faces: 0 X_M (x3), 1 X_P (x1), 2 Y_M (x4), 3 Y_P (x2), 4 Z_M (x6), 5 Z_P (x5).
def _face_index(face, sz, g, layer):
    """numpy index of Fortran layer `layer` along the face's axis (1 = the
    boundary cell of a minus face, sz of a plus face, counted outward),
    interior range on the tangential axes."""
    axis, plus = face // 2, face % 2
    n = sz[axis]
    f = (n + layer if plus else 1 - layer)  # layer 0 = boundary cell
    idx = [slice(g, g + sz[0]), slice(g, g + sz[1]), slice(g, g + sz[2])]
    idx[axis] = f - 1 + g
    return (idx[2], idx[1], idx[0])


def bcx(x, face, kind, xc, sz, g, layers=2):
    """bc_x?_d / _n / _d_f (bcx.f90:18-559).  layers=1 writes the first
    ghost layer only (a field with one ghost layer)."""
    plus = face % 2
    c = x[_face_index(face, sz, g, 0)].copy()
    g1 = _face_index(face, sz, g, 1)
    g2 = _face_index(face, sz, g, 2) if layers >= 2 else None
    if kind == "d":
        x[g1] = 2.0 * xc - c
        if g2 is not None:
            x[g2] = xc
    elif kind == "n":
        if plus:
            x[g1] = c + xc
            if g2 is not None:
                x[g2] = c + xc * 2.0
        else:
            x[g1] = c - xc
            if g2 is not None:
                x[g2] = c - xc * 2.0
    elif kind == "d_f":
        if plus:
            x[_face_index(face, sz, g, 0)] = xc
        x[g1] = xc
        if g2 is not None:
            x[g2] = xc
    else:
        raise ValueError(kind)


def bcx_periodic(x, axis, sz, g, layers=2):
    """bc_x1_p / bc_x2_p / bc_x5_p (bcx.f90:561-652): the statements in
    order, over the full tangential range (layers=1: the first ghost layer
    on each side only)."""
    n = sz[axis]
    ax = 2 - axis  # numpy axis of Fortran axis `axis`

    def sl(fidx):
        idx = [slice(None)] * 3
        idx[ax] = fidx - 1 + g
        return tuple(idx)

    x[sl(n + 1)] = x[sl(1)]
    if layers >= 2:
        x[sl(n + 2)] = x[sl(2)]
    x[sl(0)] = x[sl(n)]
    if layers >= 2:
        x[sl(-1)] = x[sl(n - 1)]


# LocalScalar3D's conditions over a whole mesh: BCM_BC_* codes.
BC_INNER, BC_DIRICHLET, BC_NEUMANN, BC_PERIODIC = 0, 1, 2, 3


def mesh_face_bc(blocks, n, vc, data, types, values):
    """BlockScalar3D::ImposeBoundaryCondition() on every block
    (BlockScalar3D.h:199-206): for each face without a neighbour, in X_M..Z_P
    order, Dirichlet / Neumann ghosts (a Neumann value times the block's
    cellSize.x, BlockScalar3D.h:98-100) or, on a plus face, the periodic
    self copy; data is (nblocks, n+2vc, n+2vc, n+2vc)."""
    sz = (n, n, n)
    layers = min(vc, 2)
    for b in range(len(blocks)):
        hx = blocks["block_size"][b][0] / n
        for face in range(6):
            if blocks["nb_id"][b][face][0] >= 0:
                continue
            t, v = types[face], values[face]
            if t == BC_DIRICHLET:
                bcx(data[b], face, "d", v, sz, vc, layers)
            elif t == BC_NEUMANN:
                bcx(data[b], face, "n", v * hx, sz, vc, layers)
            elif t == BC_PERIODIC and face % 2:
                bcx_periodic(data[b], face // 2, sz, vc, layers)


def mesh_fold_face_bc(blocks, n, vc, A, b, types, values):
    """BlockScalar3D::ImposeBoundaryCondition(Ap..b) on every block
    (BlockScalar3D.h:207-214): bca.f90's folds on each face without a
    neighbour, in X_M..Z_P order.  A is the list (Ap, Aw, Ae, As, An, Ab, At)
    of mesh arrays."""
    sz = (n, n, n)
    for k in range(len(blocks)):
        hx = blocks["block_size"][k][0] / n
        for face in range(6):
            if blocks["nb_id"][k][face][0] >= 0:
                continue
            t, v = types[face], values[face]
            if t == BC_DIRICHLET:
                bca(A[0][k], A[1 + face][k], b[k], face, "d", v, sz, vc)
            elif t == BC_NEUMANN:
                bca(A[0][k], A[1 + face][k], b[k], face, "n", v * hx, sz, vc)


# ------------------------------------------------------------------ bca.f90
def bca(Ap, Anb, b, face, kind, xc, sz, g):
    """bc_A?_d / bc_A?_n (bca.f90:12-394)."""
    c = _face_index(face, sz, g, 0)
    if kind == "d":
        Ap[c] = Ap[c] - Anb[c]
        b[c] = b[c] - Anb[c] * xc * 2.0
    else:
        Ap[c] = Ap[c] + Anb[c]
        b[c] = b[c] - Anb[c] * xc if face % 2 else b[c] + Anb[c] * xc
    Anb[c] = 0.0


# ----------------------------------------------------------------- bcax.f90
def _pois_r(sz, center, org, csize):
    """bc_get_u / bc_get_gp's distance (bcax.f90:18-84) for the X_M
    boundary cells, the reference's rx = x0(1) - x0(1) included; shaped
    (sz(3), sz(2))."""
    dx = csize[0]
    k, j = np.meshgrid(np.arange(1, sz[2] + 1), np.arange(1, sz[1] + 1), indexing="ij")
    x0y = org[1] + (j.astype(np.float64) - 0.5) * dx
    x0z = org[2] + (k.astype(np.float64) - 0.5) * dx
    rx = 0.0
    ry = x0y - center[1]
    rz = x0z - center[2]
    r2 = (rx * rx + ry * ry) + rz * rz
    return np.sqrt(r2)


def pois_u(xc, sz, center, org, csize):
    r = _pois_r(sz, center, org, csize)
    u = xc * 2.0 * (1.0 - r * r / (0.5 * 0.5))
    return np.where(r > 0.5, 0.0, u)


def pois_gp(xc, sz, center, org, csize):
    r = _pois_r(sz, center, org, csize)
    gp = np.full(r.shape, xc * csize[0])
    return np.where(r > 0.5, 0.0, gp)


def bc_x3_poiseuille_u(x, xc, sz, g, center, org, csize):  # bcax.f90:86-132
    u = pois_u(xc, sz, center, org, csize)
    c = x[_face_index(0, sz, g, 0)].copy()
    x[_face_index(0, sz, g, 1)] = 2.0 * u - c
    x[_face_index(0, sz, g, 2)] = u


def bc_aw_poiseuille_u(Ap, Aw, b, xc, sz, g, center, org, csize):  # bcax.f90:134-182
    u = pois_u(xc, sz, center, org, csize)
    c = _face_index(0, sz, g, 0)
    Ap[c] = Ap[c] - Aw[c]
    b[c] = b[c] - Aw[c] * u * 2.0
    Aw[c] = 0.0


def bc_x3_poiseuille_p(x, xc, sz, g, center, org, csize):  # bcax.f90:184-230
    gp = pois_gp(xc, sz, center, org, csize)
    c = x[_face_index(0, sz, g, 0)].copy()
    x[_face_index(0, sz, g, 1)] = c - gp
    x[_face_index(0, sz, g, 2)] = c - gp * 2.0


def bc_aw_poiseuille_p(Ap, Aw, b, xc, sz, g, center, org, csize):  # bcax.f90:232-280
    gp = pois_gp(xc, sz, center, org, csize)
    c = _face_index(0, sz, g, 0)
    Ap[c] = Ap[c] + Aw[c]
    b[c] = b[c] + Aw[c] * gp
    Aw[c] = 0.0


# ------------------------------------------- Scalar3DUpdater1's face boxes
def updater1_boxes(face, level_diff, subface, n, vc):
    """(i1_dst, i1_src, sz_c, op) that examples/SolverTest/Scalar3DUpdater1.cpp
    passes to sup.f90 for one face of a cubic block of n cells: same level
    (:40-116), coarser neighbour C2F (:262-364), finer neighbour F2C
    (:446-537).  Subface origins follow include/VCUpdater.h:251-273
    (origin0 = subface & 1, origin1 = subface >> 1); on a Y face the first
    origin runs along z and the second along x."""
    o0, o1 = subface & 1, subface >> 1
    axis, plus = face // 2, face % 2
    if level_diff == 0:
        i1d, i1s, szc = [1, 1, 1], [1, 1, 1], [n, n, n]
        i1d[axis] = 1 + n if plus else 1 - vc
        i1s[axis] = 1 if plus else 1 + n - vc
        szc[axis] = vc
        return i1d, i1s, szc, "copy"
    if level_diff == -1:
        i1d, i1s, szc = [1, 1, 1], [1, 1, 1], [n, n, n]
        i1d[axis] = 1 + n if plus else 1 - vc
        i1s[axis] = 1 if plus else 1 + 2 * n - vc
        szc[axis] = vc
        t = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[axis]  # (axis of origin0, axis of origin1)
        i1s[t[0]] = 1 + n * o0
        i1s[t[1]] = 1 + n * o1
        return i1d, i1s, szc, "c2f"
    i1d, i1s, szc = [1, 1, 1], [1, 1, 1], [n // 2, n // 2, n // 2]
    i1d[axis] = 1 + n if plus else 1 - vc
    i1s[axis] = 1 if plus else 1 + n // 2 - vc
    szc[axis] = vc
    t = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[axis]
    i1d[t[0]] = 1 + (n // 2) * o0
    i1d[t[1]] = 1 + (n // 2) * o1
    return i1d, i1s, szc, "f2c"
