"""Host-side handles over the C ABI: Tree, Mesh, Field and the kernels.

Names mirror the reference API they stand in for:
  Tree                 BCMOctree + BlockFactory (Utils/src/BCMOctree.cpp)
  Mesh                 BlockManager after registerBlock/endRegisterBlock and
                       prepareForVCUpdate (src/BlockManager.cpp:28-114)
  Field                Scalar3D<T> on every block, contiguous storage
                       (include/BlockManager.h:253-283)
  Mesh.update_vc*      BlockManager::updateVC* (include/BlockManager.h:311-440)
"""
import ctypes

import numpy as np

from . import _lib as L


class Tree:
    """Leaf-block table of an octree mesh (host only)."""

    def __init__(self, root_n=(1, 1, 1), periodic=(0, 0, 0), root_origin=(0.0, 0.0, 0.0), root_length=1.0,
                 divider=L.DIV_FLAT, min_level=0, max_level=0, shells=(), ordering=L.ORDER_Z, balance=1,
                 bc_setter=L.BCSET_ALL_DIRICHLET, shell_mode=0, shell_margin=0.0, face_bc_type=None, divide=None):
        """`divide` (with divider=DIV_USER): a callable taking a node's
        pedigree word and returning NODE_BRANCH, NODE_LEAF_ACTIVE or
        NODE_LEAF_INACTIVE, as Divider::operator() (Utils/include/Divider.h)."""
        lib = L.load()
        p = L.TreeParams()
        for a in range(3):
            p.root_n[a] = int(root_n[a])
            p.periodic[a] = int(periodic[a])
            p.root_origin[a] = float(root_origin[a])
        p.root_length = float(root_length)
        p.divider = int(divider)
        p.min_level = int(min_level)
        p.max_level = int(max_level)
        shells = list(shells)
        if len(shells) > L.MAX_SHELLS:
            raise ValueError("too many shells")
        p.n_shells = len(shells)
        for s, sh in enumerate(shells):
            for c in range(5):
                p.shells[s][c] = float(sh[c])
        p.ordering = int(ordering)
        p.balance = int(balance)
        p.bc_setter = int(bc_setter)
        p.shell_mode = int(shell_mode)
        p.shell_margin = float(shell_margin)
        if face_bc_type is not None:
            if bc_setter != L.BCSET_PER_FACE or len(face_bc_type) != 6:
                raise ValueError("face_bc_type needs bc_setter=BCSET_PER_FACE and six entries")
            for f in range(6):
                p.face_bc_type[f] = int(face_bc_type[f])
        self.params = p
        h = ctypes.c_void_p()
        if divide is not None:
            fn = L.DIVIDER_FN(lambda user, ped: int(divide(int(ped))))
            L.check(lib.bcm_tree_build_user(ctypes.byref(p), fn, None, ctypes.byref(h)), "bcm_tree_build_user")
        else:
            L.check(lib.bcm_tree_build(ctypes.byref(p), ctypes.byref(h)), "bcm_tree_build")
        try:
            nb = lib.bcm_tree_num_blocks(h)
            if nb < 0:
                raise L.BCMError(-nb, "bcm_tree_num_blocks")
            blocks = np.zeros(nb, dtype=L.BLOCK_DTYPE)
            L.check(lib.bcm_tree_get_blocks(h, blocks.ctypes.data, nb), "bcm_tree_get_blocks")
        finally:
            lib.bcm_tree_destroy(h)
        self.blocks = blocks

    def __len__(self):
        return len(self.blocks)


class Field:
    """A ghosted scalar field on every block of a mesh, resident in HBM."""

    def __init__(self, mesh, dtype=np.float64):
        self.mesh = mesh
        self.dtype = np.dtype(dtype)
        code = L.F64 if self.dtype == np.float64 else L.F32
        if self.dtype not in (np.float64, np.float32):
            raise ValueError("dtype must be float64 or float32")
        h = ctypes.c_void_p()
        L.check(mesh.lib.bcm_field_create(mesh.handle, code, ctypes.byref(h)), "bcm_field_create")
        self.handle = h

    @property
    def shape(self):
        m = self.mesh
        return (m.nblocks, m.nx0, m.nx0, m.nx0)

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=self.dtype).reshape(self.shape)
        L.check(self.mesh.lib.bcm_field_upload(self.handle, a.ctypes.data, a.nbytes), "bcm_field_upload")

    def download(self):
        a = np.empty(self.shape, dtype=self.dtype)
        L.check(self.mesh.lib.bcm_field_download(self.handle, a.ctypes.data, a.nbytes), "bcm_field_download")
        return a

    def upload_blocks(self, first, host):
        """Upload blocks first .. first+len(host)-1 (e.g. one segment of a batch)."""
        m = self.mesh
        a = np.ascontiguousarray(host, dtype=self.dtype).reshape((-1, m.nx0, m.nx0, m.nx0))
        L.check(m.lib.bcm_field_upload_blocks(self.handle, int(first), len(a), a.ctypes.data, a.nbytes),
                "bcm_field_upload_blocks")

    def download_blocks(self, first, count):
        m = self.mesh
        a = np.empty((int(count), m.nx0, m.nx0, m.nx0), dtype=self.dtype)
        L.check(m.lib.bcm_field_download_blocks(self.handle, int(first), int(count), a.ctypes.data, a.nbytes),
                "bcm_field_download_blocks")
        return a

    # -- restart files (examples/SolverTest/LocalScalar3D.cpp:84-357) ------
    def dump2(self, step, label, rank=0, directory="./BIN"):
        """LocalScalar3D::Dump2: one file per rank."""
        L.check(self.mesh.lib.bcm_field_dump2(self.handle, directory.encode(), label.encode(), int(rank), int(step)),
                "bcm_field_dump2")

    def load2(self, step, label, rank=0, directory="./BIN"):
        """LocalScalar3D::Load2 (same block edge, or half of it)."""
        L.check(self.mesh.lib.bcm_field_load2(self.handle, directory.encode(), label.encode(), int(rank), int(step)),
                "bcm_field_load2")

    def dump(self, step, label, rank=0, directory="./BIN"):
        """LocalScalar3D::Dump: one file per block."""
        L.check(self.mesh.lib.bcm_field_dump(self.handle, directory.encode(), label.encode(), int(rank), int(step)),
                "bcm_field_dump")

    def load(self, step, label, rank=0, directory="./BIN"):
        """LocalScalar3D::Load: one file per block."""
        L.check(self.mesh.lib.bcm_field_load(self.handle, directory.encode(), label.encode(), int(rank), int(step)),
                "bcm_field_load")

    def write_vtk(self, step, name, difflevel, root_origin=(0.0, 0.0, 0.0), root_length=1.0, rank=0, start_id=0,
                  directory="./VTK"):
        """VtkWriter::writeScalar: one .vti per block and a .vthb index."""
        o = np.ascontiguousarray(root_origin, dtype=np.float64)
        L.check(self.mesh.lib.bcm_field_write_vtk(self.handle, directory.encode(), name.encode(), int(step),
                                                  int(rank), int(start_id), int(difflevel), L.dptr(o),
                                                  float(root_length)), "bcm_field_write_vtk")

    def device_ptr(self):
        return self.mesh.lib.bcm_field_device_ptr(self.handle)

    def free(self):
        if self.handle:
            self.mesh.lib.bcm_field_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _handles(fields):
    arr = (ctypes.c_void_p * len(fields))()
    for i, f in enumerate(fields):
        arr[i] = f.handle.value if isinstance(f.handle, ctypes.c_void_p) else f.handle
    return arr


class Mesh:
    """Device mesh: block table, face-task lists and the mesh's HIP stream."""

    def __init__(self, blocks, n, vc=1, segments=None):
        """`segments`: block counts of independent meshes concatenated in
        `blocks` (each with neighbour ids local to its own table)."""
        self.lib = L.load()
        blocks = np.ascontiguousarray(blocks, dtype=L.BLOCK_DTYPE)
        self.blocks = blocks
        self.n = int(n)
        self.vc = int(vc)
        self.nblocks = len(blocks)
        self.nx0 = self.n + 2 * self.vc
        h = ctypes.c_void_p()
        if segments is None:
            L.check(self.lib.bcm_mesh_create(blocks.ctypes.data, self.nblocks, self.n, self.vc, ctypes.byref(h)),
                    "bcm_mesh_create")
        else:
            cnt = np.ascontiguousarray(segments, dtype=np.int32)
            L.check(self.lib.bcm_mesh_create_batch(blocks.ctypes.data, cnt.ctypes.data_as(L._IP), len(cnt), self.n,
                                                   self.vc, ctypes.byref(h)), "bcm_mesh_create_batch")
        self.handle = h
        self.nseg = self.lib.bcm_mesh_num_segments(h)

    @classmethod
    def batch(cls, block_tables, n, vc=1):
        """One device mesh holding several independent meshes (segments)."""
        tables = [np.ascontiguousarray(t, dtype=L.BLOCK_DTYPE) for t in block_tables]
        return cls(np.concatenate(tables), n, vc, segments=[len(t) for t in tables])

    def set_progress(self, seconds):
        """Progress lines on stderr from long solves (0 = off)."""
        L.check(self.lib.bcm_mesh_set_progress(self.handle, float(seconds)), "bcm_mesh_set_progress")

    def segment(self, s):
        """(first block, block count) of segment s."""
        a, b = ctypes.c_int(), ctypes.c_int()
        L.check(self.lib.bcm_mesh_segment(self.handle, int(s), ctypes.byref(a), ctypes.byref(b)), "bcm_mesh_segment")
        return a.value, b.value

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if self.handle:
            self.lib.bcm_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, dtype=np.float64, host=None):
        f = Field(self, dtype)
        if host is not None:
            f.upload(host)
        return f

    def sync(self):
        L.check(self.lib.bcm_mesh_sync(self.handle), "bcm_mesh_sync")

    def task_counts(self):
        c = (ctypes.c_int * 4)()
        L.check(self.lib.bcm_mesh_task_counts(self.handle, c), "bcm_mesh_task_counts")
        return {"flat": c[0], "f2c": c[1], "c2f": c[2], "boundary": c[3]}

    # -- timing -----------------------------------------------------------
    def set_timing(self, on=True):
        L.check(self.lib.bcm_mesh_set_timing(self.handle, int(bool(on))), "bcm_mesh_set_timing")

    def reset_timing(self):
        L.check(self.lib.bcm_mesh_reset_timing(self.handle), "bcm_mesh_reset_timing")

    def kernel_time(self, kernel):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        L.check(self.lib.bcm_mesh_kernel_time(self.handle, kernel, ctypes.byref(ms), ctypes.byref(n)),
                "bcm_mesh_kernel_time")
        return ms.value, n.value

    # -- ghost fill (BlockManager::updateVC*) ------------------------------
    def update_vc(self, f, mask=L.GF_ALL):
        L.check(self.lib.bcm_ghost_fill(f.handle, int(mask)), "bcm_ghost_fill")

    def update_vc_xyz(self, f, axis, kinds=L.GF_ALL):
        """updateVC_{X,Y,Z}[_{Flat,F2C,C2F}] (include/BlockManager.h:327-343)."""
        self.update_vc(f, kinds | (L.GF_X << axis))

    def update_virtual_cells(self, f, method=L.VC_AT_ONCE):
        """LocalScalar3D::UpdateVirtualCells (examples/SolverTest/LocalScalar3D.h:198-223)
        with the field's VCUpdateMethod."""
        L.check(self.lib.bcm_update_virtual_cells(f.handle, int(method)), "bcm_update_virtual_cells")

    def impose_boundary_condition(self, f, values, method=L.VC_AT_ONCE):
        """LocalScalar3D::ImposeBoundaryCondition (LocalScalar3D.h:225-257): the ghost
        exchange, then the outer-face conditions."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        L.check(self.lib.bcm_impose_boundary_condition(f.handle, int(method), L.dptr(v), len(v)),
                "bcm_impose_boundary_condition")

    def dirichlet_mirror(self, f, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        L.check(self.lib.bcm_dirichlet_mirror(f.handle, L.dptr(v), len(v)), "bcm_dirichlet_mirror")

    def impose_bc(self, f, values):
        """BlockScalar3D::ImposeBoundaryCondition (examples/SolverTest/BlockScalar3D.h:199-206):
        Dirichlet and Neumann ghosts on every outer face."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        L.check(self.lib.bcm_impose_bc(f.handle, L.dptr(v), len(v)), "bcm_impose_bc")

    def impose_face_bc(self, f, types, values):
        """LocalScalar3D's per-field conditions (BlockScalar3D.h:199-206): one
        BCM_BC_* type and value per domain face, on every block face without
        a neighbour."""
        t = (ctypes.c_int * 6)(*[int(v) for v in types])
        v = (ctypes.c_double * 6)(*[float(x) for x in values])
        L.check(self.lib.bcm_impose_face_bc(f.handle, t, v), "bcm_impose_face_bc")

    def fold_face_bc(self, A, b, types, values):
        """The same conditions folded into A[7] and b (BlockScalar3D.h:207-214,
        bca.f90)."""
        t = (ctypes.c_int * 6)(*[int(v) for v in types])
        v = (ctypes.c_double * 6)(*[float(x) for x in values])
        L.check(self.lib.bcm_fold_face_bc(_handles(A), b.handle, t, v), "bcm_fold_face_bc")

    # -- red-black SOR ------------------------------------------------------
    def rbsor_sweep(self, f, source=None, omega=1.0):
        L.check(self.lib.bcm_rbsor_sweep(f.handle, source.handle if source is not None else None, float(omega)),
                "bcm_rbsor_sweep")

    def sor_iterate(self, f, source, omega, bvalues, iters, inner=1):
        """`iters` outer iterations of the sor1/sor2 Solver::run, each with
        `inner` (nLoopInner) sweeps per block between ghost exchanges."""
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        src = source.handle if source is not None else None
        if inner == 1:
            L.check(self.lib.bcm_sor_iterate(f.handle, src, float(omega), L.dptr(v), len(v), int(iters)),
                    "bcm_sor_iterate")
        else:
            L.check(self.lib.bcm_sor_iterate_inner(f.handle, src, float(omega), L.dptr(v), len(v), int(iters),
                                                   int(inner)), "bcm_sor_iterate_inner")

    def set_sor_variant(self, variant):
        L.check(self.lib.bcm_sor_set_variant(self.handle, int(variant)), "bcm_sor_set_variant")

    def sor_variant(self):
        return self.lib.bcm_sor_get_variant(self.handle)

    def sor_autotune(self, f, source, omega, bvalues, iters=3):
        """Time every supported outer-iteration schedule on a scratch copy of f,
        keep the fastest; returns (best, {variant: ms per iteration})."""
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        best = ctypes.c_int()
        ms = (ctypes.c_double * L.SOR_NUM_VARIANTS)()
        L.check(self.lib.bcm_sor_autotune(f.handle, None if source is None else source.handle, float(omega),
                                          L.dptr(v), len(v), int(iters), ctypes.byref(best), ms), "bcm_sor_autotune")
        return best.value, {i: ms[i] for i in range(1, L.SOR_NUM_VARIANTS) if ms[i] >= 0}

    def sor_residual(self, f, source):
        r = ctypes.c_double()
        L.check(self.lib.bcm_sor_residual(f.handle, source.handle, ctypes.byref(r)), "bcm_sor_residual")
        return r.value

    # -- SolverTest kernels ---------------------------------------------------
    # mask: a field of the operands' type, 0/1; only cells where it is 1 are
    # visited (the *_mask forms of blas.f90:753-1113).
    def jacobi7(self, x1, x0, A, b, omega, mask=None):
        if mask is not None:
            L.check(self.lib.bcm_jacobi7_mask(x1.handle, x0.handle, _handles(A), b.handle, mask.handle, float(omega)),
                    "bcm_jacobi7_mask")
            return
        L.check(self.lib.bcm_jacobi7(x1.handle, x0.handle, _handles(A), b.handle, float(omega)), "bcm_jacobi7")

    def jacobi7_poisson(self, x1, x0, b, omega):
        L.check(self.lib.bcm_jacobi7_poisson(x1.handle, x0.handle, b.handle, float(omega)), "bcm_jacobi7_poisson")

    def jacobi_poisson_iterate(self, x, b, omega, nsweeps):
        """nsweeps x (jacobi7_poisson, then the ghost exchange) in one call."""
        L.check(self.lib.bcm_jacobi_poisson_iterate(x.handle, b.handle, float(omega), int(nsweeps)),
                "bcm_jacobi_poisson_iterate")

    def calc_ax_poisson(self, ax, x):
        L.check(self.lib.bcm_calc_ax_poisson(ax.handle, x.handle), "bcm_calc_ax_poisson")

    def calc_r_poisson(self, r, x, b):
        L.check(self.lib.bcm_calc_r_poisson(r.handle, x.handle, b.handle), "bcm_calc_r_poisson")

    def calc_ax(self, ax, A, x, mask=None):
        if mask is not None:
            L.check(self.lib.bcm_calc_ax_mask(ax.handle, _handles(A), x.handle, mask.handle), "bcm_calc_ax_mask")
            return
        L.check(self.lib.bcm_calc_ax(ax.handle, _handles(A), x.handle), "bcm_calc_ax")

    def calc_r(self, r, A, x, b, mask=None):
        if mask is not None:
            L.check(self.lib.bcm_calc_r_mask(r.handle, _handles(A), x.handle, b.handle, mask.handle),
                    "bcm_calc_r_mask")
            return
        L.check(self.lib.bcm_calc_r(r.handle, _handles(A), x.handle, b.handle), "bcm_calc_r")

    def calc_r2(self, A, x, b, mask=None):
        v = ctypes.c_double()
        if mask is not None:
            L.check(self.lib.bcm_calc_r2_mask(_handles(A), x.handle, b.handle, mask.handle, ctypes.byref(v)),
                    "bcm_calc_r2_mask")
            return v.value
        L.check(self.lib.bcm_calc_r2(_handles(A), x.handle, b.handle, ctypes.byref(v)), "bcm_calc_r2")
        return v.value

    def dot(self, x, y, mask=None):
        v = ctypes.c_double()
        if mask is not None:
            L.check(self.lib.bcm_dot_mask(x.handle, y.handle, mask.handle, ctypes.byref(v)), "bcm_dot_mask")
            return v.value
        L.check(self.lib.bcm_dot(x.handle, y.handle, ctypes.byref(v)), "bcm_dot")
        return v.value

    def axpy(self, y, x, a, mask=None):
        if mask is not None:
            L.check(self.lib.bcm_axpy_mask(y.handle, x.handle, float(a), mask.handle), "bcm_axpy_mask")
            return
        L.check(self.lib.bcm_axpy(y.handle, x.handle, float(a)), "bcm_axpy")

    def xpay(self, y, x, a, mask=None):
        if mask is not None:
            L.check(self.lib.bcm_xpay_mask(y.handle, x.handle, float(a), mask.handle), "bcm_xpay_mask")
            return
        L.check(self.lib.bcm_xpay(y.handle, x.handle, float(a)), "bcm_xpay")

    def axpyz(self, z, x, y, a, mask=None):
        if mask is not None:
            L.check(self.lib.bcm_axpyz_mask(z.handle, x.handle, y.handle, float(a), mask.handle), "bcm_axpyz_mask")
            return
        L.check(self.lib.bcm_axpyz(z.handle, x.handle, y.handle, float(a)), "bcm_axpyz")

    def axpbypz(self, z, x, y, a, b, mask=None):
        if mask is not None:
            L.check(self.lib.bcm_axpbypz_mask(z.handle, x.handle, y.handle, float(a), float(b), mask.handle),
                    "bcm_axpbypz_mask")
            return
        L.check(self.lib.bcm_axpbypz(z.handle, x.handle, y.handle, float(a), float(b)), "bcm_axpbypz")

    def copy_mask(self, y, x, mask):
        """copy_mask (blas.f90:969-992): interior cells where mask is 1."""
        L.check(self.lib.bcm_copy_mask(y.handle, x.handle, mask.handle), "bcm_copy_mask")

    def setup_mask(self, mask):
        """setup_mask (blas.f90:700-751): 1 on interior cells, 0 on ghosts."""
        L.check(self.lib.bcm_setup_mask(mask.handle), "bcm_setup_mask")

    def fill(self, x, a):
        L.check(self.lib.bcm_fill(x.handle, float(a)), "bcm_fill")

    def copy(self, y, x):
        L.check(self.lib.bcm_copy(y.handle, x.handle), "bcm_copy")

    def poisson_assemble(self, A, b, rhs, bvalues):
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        L.check(self.lib.bcm_poisson_assemble(_handles(A) if A is not None else None, b.handle, rhs.handle,
                                              L.dptr(v), len(v)), "bcm_poisson_assemble")

    # -- solvers ----------------------------------------------------------------
    def jacobi_solve(self, x, A, b, x0, omega=1.0, count_max=100, epsilon=1e-8, bvalues=(), mask=None):
        """ILS3D::Jacobi; with a mask ILS3D::Jacobi_Mask."""
        p = L.KrylovParams(float(omega), 0, int(count_max), float(epsilon))
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        cnt = ctypes.c_int()
        res = ctypes.c_double()
        vp = L.dptr(v) if len(v) else None
        if mask is not None:
            rc = self.lib.bcm_jacobi_solve_mask(x.handle, _handles(A), b.handle, x0.handle, mask.handle,
                                                ctypes.byref(p), vp, len(v), ctypes.byref(cnt), ctypes.byref(res))
        else:
            rc = self.lib.bcm_jacobi_solve(x.handle, _handles(A), b.handle, x0.handle, ctypes.byref(p), vp, len(v),
                                           ctypes.byref(cnt), ctypes.byref(res))
        if rc not in (L.OK, L.ERR_NOCONV):
            L.check(rc, "bcm_jacobi_solve")
        return rc, cnt.value, res.value

    def bicgstab(self, x, A, b, work, omega=0.8, count_pre=4, count_max=100, epsilon=1e-8, bvalues=(), mask=None):
        """ILS3D::BiCGSTAB (with a mask ILS3D::BiCGSTAB_Mask); `work` = 9 fields
        {r, r0, p, p_, q_, s, s_, t_, x0}."""
        p = L.KrylovParams(float(omega), int(count_pre), int(count_max), float(epsilon))
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        cnt = ctypes.c_int()
        res = ctypes.c_double()
        vp = L.dptr(v) if len(v) else None
        if mask is not None:
            rc = self.lib.bcm_bicgstab_mask(x.handle, _handles(A), b.handle, _handles(work), mask.handle,
                                            ctypes.byref(p), vp, len(v), ctypes.byref(cnt), ctypes.byref(res))
        else:
            rc = self.lib.bcm_bicgstab(x.handle, _handles(A), b.handle, _handles(work), ctypes.byref(p), vp, len(v),
                                       ctypes.byref(cnt), ctypes.byref(res))
        if rc not in (L.OK, L.ERR_NOCONV):
            L.check(rc, "bcm_bicgstab")
        return rc, cnt.value, res.value

    def solve_fp64(self, x, rhs, bvalues, tol=1e-10, max_iter=2000, count_pre=2, omega=0.8):
        p = L.IRParams(float(tol), float(tol), int(max_iter), 1, int(count_pre), float(omega), 0, 0)
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        it = ctypes.c_int()
        rr = ctypes.c_double()
        rc = self.lib.bcm_solve_fp64(x.handle, rhs.handle, L.dptr(v), len(v), ctypes.byref(p), ctypes.byref(it),
                                     ctypes.byref(rr))
        if rc not in (L.OK, L.ERR_NOCONV):
            L.check(rc, "bcm_solve_fp64")
        return rc, it.value, rr.value

    def solve(self, x, rhs, bvalues, precision="mixed", tol=1e-10, max_iter=5000, count_pre=2, omega=0.8,
              refresh_max=0, refresh_delta=0.0, check_every=8, use_graph=True, seg_tol=None, restart_every=0,
              stall_window=0, stagnation=0):
        """Fused device-resident BiCGSTAB on every segment (bcm_solve).
        Returns one dict per segment: status, iters, refreshes, relres."""
        prec = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED, "fp32": L.PREC_FP32}[precision]
        p = L.SolveParams(prec, float(tol), int(max_iter), int(count_pre), float(omega), int(refresh_max),
                          float(refresh_delta), int(check_every), int(bool(use_graph)), int(restart_every),
                          int(stall_window), int(stagnation))
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        st = (L.SolveStats * self.nseg)()
        tl = None
        if seg_tol is not None:
            tl = np.ascontiguousarray(seg_tol, dtype=np.float64)
            if len(tl) != self.nseg:
                raise ValueError("seg_tol needs one tolerance per segment")
        L.check(self.lib.bcm_solve(x.handle, rhs.handle, L.dptr(v), len(v), ctypes.byref(p), L.dptr(tl), st),
                "bcm_solve")
        names = {L.SOLVE_CONVERGED: "converged", L.SOLVE_NOCONV: "noconv", L.SOLVE_FAILED: "failed"}
        return [{"status": names.get(e.status, e.status), "iters": e.iters, "refreshes": e.refreshes,
                 "relres": e.relres} for e in st]

    def solve_mixed(self, x, rhs, bvalues, tol=1e-10, inner_tol=1e-4, inner_max=500, outer_max=20, count_pre=2,
                    omega=0.8, fp32_only=False, fp32_strict=False):
        """FP32 inner BiCGSTAB inside FP64 iterative refinement (bcm_solve_mixed).
        fp32_strict: stop when an FP32 correction does not lower the FP64
        residual, instead of solving the remaining corrections in FP64."""
        p = L.IRParams(float(tol), float(inner_tol), int(inner_max), int(outer_max), int(count_pre), float(omega),
                       int(bool(fp32_only)), int(bool(fp32_strict)))
        v = np.ascontiguousarray(bvalues, dtype=np.float64)
        outer = ctypes.c_int()
        inner = ctypes.c_int()
        rr = ctypes.c_double()
        rc = self.lib.bcm_solve_mixed(x.handle, rhs.handle, L.dptr(v), len(v), ctypes.byref(p), ctypes.byref(outer),
                                      ctypes.byref(inner), ctypes.byref(rr))
        if rc not in (L.OK, L.ERR_NOCONV):
            L.check(rc, "bcm_solve_mixed")
        return rc, outer.value, inner.value, rr.value


def partition(nblocks, nprocs):
    """Rank of every block under the reference's Partition (Utils/include/
    Partition.h:46-59): contiguous ranges, the first nblocks % nprocs
    ranks one block longer."""
    m, r = divmod(int(nblocks), int(nprocs))
    sizes = [m + 1 if i < r else m for i in range(nprocs)]
    return np.repeat(np.arange(nprocs, dtype=np.int32), sizes)


class CommPlan:
    """One rank's per-peer buffers (bcm_comm_plan; the reference's
    CommBuffer over BlockManager's face lists, src/CommBuffer.cpp:55-94,
    src/BlockManager.cpp:471-711).  `peers`: list of (rank, first face,
    number of faces, byte offset, bytes); `faces`: the bcm_comm_face array
    for bcm_comm_pack / bcm_comm_unpack; `total_bytes`: the buffer size."""

    def __init__(self, blocks, n, vc, dtype, block_rank, rank, direction, axis=-1, level_diff=L.COMM_ANY_LEVEL):
        lib = L.load()
        blocks = np.ascontiguousarray(blocks)
        br = np.ascontiguousarray(block_rank, dtype=np.int32)
        if len(br) != len(blocks):
            raise ValueError("block_rank needs one entry per block")
        dt = L.F64 if np.dtype(dtype) == np.float64 else L.F32
        npeers, nfaces, total = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        args = (blocks.ctypes.data, len(blocks), int(n), int(vc), dt, br.ctypes.data, int(rank), int(direction),
                int(axis), int(level_diff))
        L.check(lib.bcm_comm_plan(*args, None, 0, ctypes.byref(npeers), None, 0, ctypes.byref(nfaces),
                                  ctypes.byref(total)), "bcm_comm_plan")
        peers = (L.CommPeer * max(1, npeers.value))()
        self.faces = (L.CommFace * max(1, nfaces.value))()
        L.check(lib.bcm_comm_plan(*args, peers, npeers.value, ctypes.byref(npeers), self.faces, nfaces.value,
                                  ctypes.byref(nfaces), ctypes.byref(total)), "bcm_comm_plan")
        self.nfaces = nfaces.value
        self.total_bytes = total.value
        self.peers = [(p.rank, p.first, p.nfaces, p.offset, p.bytes) for p in peers[:npeers.value]]


class CommList:
    """A face list uploaded once (bcm_comm_list_create) for repeated packs
    and unpacks on one mesh; pack / unpack enqueue on the mesh's stream and
    return without waiting (Mesh.sync before reading a buffer)."""

    def __init__(self, mesh, faces, nfaces):
        self.lib = mesh.lib
        self.mesh = mesh  # the list must not outlive its mesh
        h = ctypes.c_void_p()
        L.check(self.lib.bcm_comm_list_create(mesh.handle, faces, int(nfaces), ctypes.byref(h)),
                "bcm_comm_list_create")
        self.handle = h

    def pack(self, field, buffer, nbytes):
        L.check(self.lib.bcm_comm_pack_list(field.handle, self.handle, buffer, int(nbytes)), "bcm_comm_pack_list")

    def unpack(self, field, buffer, nbytes):
        L.check(self.lib.bcm_comm_unpack_list(field.handle, self.handle, buffer, int(nbytes)),
                "bcm_comm_unpack_list")

    def close(self):
        if self.handle:
            self.lib.bcm_comm_list_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server: queued Poisson solves admitted by an HBM budget and
    run with continuous batching (bcm_server_*)."""

    PRECISIONS = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED, "fp32": L.PREC_FP32}

    def __init__(self, hbm_budget, max_active=8, n=16, max_iter=20000, count_pre=2, omega=0.8, check_every=8,
                 batch_blocks=0, window_blocks=0, max_wait=0.0):
        """batch_blocks > 0 merges requests of at most that many blocks into
        batched groups of at most that many blocks (0: one request per group).
        window_blocks > 0 launches the smallest groups' iterations first,
        within that many blocks in flight (max_wait > 0: a group waits at
        most that many seconds); 0: every group as soon as it is ready."""
        self.lib = L.load()
        p = L.ServerParams(float(hbm_budget), int(max_active), int(n), int(max_iter), int(count_pre), float(omega),
                           int(check_every), int(batch_blocks), int(window_blocks), float(max_wait))
        h = ctypes.c_void_p()
        L.check(self.lib.bcm_server_create(ctypes.byref(p), ctypes.byref(h)), "bcm_server_create")
        self.handle = h
        self.n = int(n)
        self._keep = {}  # tag -> host arrays the server reads or writes later

    def close(self):
        if self.handle:
            self.lib.bcm_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self):
        return self.lib.bcm_server_clock(self.handle)

    def submit(self, blocks, tol, precision="mixed", rhs=None, seed=0, tag=0, want_x=False, restart=None):
        """Queue a request; `tag` must not name a request still pending (the
        server reads rhs and writes x of the request it names).  restart =
        (dir, label, step[, rank]): start from the iterate in that restart
        file (bcm_server_submit_restart) instead of zero."""
        if int(tag) in self._keep:
            raise L.BCMError(L.ERR_ARG, f"bcm_server_submit (tag {int(tag)} already pending)")
        blocks = np.ascontiguousarray(blocks, dtype=L.BLOCK_DTYPE)
        nx = self.n + 2
        r = None if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64).reshape(len(blocks), nx, nx, nx)
        x = np.zeros((len(blocks), nx, nx, nx)) if want_x else None
        rp = None if r is None else r.ctypes.data
        xp = None if x is None else x.ctypes.data
        if restart is None:
            L.check(self.lib.bcm_server_submit(self.handle, blocks.ctypes.data, len(blocks), rp, int(seed),
                                               float(tol), self.PRECISIONS[precision], xp, int(tag)),
                    "bcm_server_submit")
        else:
            d, label, step = restart[0], restart[1], restart[2]
            rank = restart[3] if len(restart) > 3 else 0
            L.check(self.lib.bcm_server_submit_restart(self.handle, blocks.ctypes.data, len(blocks), rp, int(seed),
                                                       float(tol), self.PRECISIONS[precision], str(d).encode(),
                                                       str(label).encode(), int(rank), int(step), xp, int(tag)),
                    "bcm_server_submit_restart")
        self._keep[int(tag)] = (blocks, r, x)

    def checkpoint(self, tag, dir, label, step, rank=0):
        """Write the running request `tag`'s current iterate as a restart
        file (Dump2 format); returns the iterations it has done."""
        it = ctypes.c_int()
        L.check(self.lib.bcm_server_checkpoint(self.handle, int(tag), str(dir).encode(), str(label).encode(),
                                               int(rank), int(step), ctypes.byref(it)), "bcm_server_checkpoint")
        return it.value

    def progress(self, seconds):
        a, q = ctypes.c_int(), ctypes.c_int()
        L.check(self.lib.bcm_server_progress(self.handle, float(seconds), ctypes.byref(a), ctypes.byref(q)),
                "bcm_server_progress")
        return a.value, q.value

    def poll(self):
        """Finished requests: dicts with tag, status, iters, relres, times and the solution (if asked for)."""
        buf = (L.ServerResult * 64)()
        out = []
        while True:
            k = self.lib.bcm_server_poll(self.handle, buf, 64)
            if k < 0:
                L.check(-k, "bcm_server_poll")
            for e in buf[:k]:
                kept = self._keep.pop(int(e.tag), (None, None, None))
                out.append({"tag": e.tag, "status": e.status, "iters": e.iters, "relres": e.relres,
                            "t_submit": e.t_submit, "t_admit": e.t_admit, "t_done": e.t_done, "nblocks": e.nblocks,
                            "precision": e.precision, "x": kept[2]})
            if k < 64:
                return out

    def stats(self):
        b, a, q = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        L.check(self.lib.bcm_server_stats(self.handle, ctypes.byref(b), ctypes.byref(a), ctypes.byref(q)),
                "bcm_server_stats")
        return {"bytes_in_use": b.value, "active": a.value, "queued": q.value}
