/* Per-block kernels of examples/SolverTest under their Fortran names and
 * by-reference convention, on device-resident block arrays (libbcm_hip.so).
 *
 * These replace the Fortran objects SolverTest links (blas.f90, sup.f90,
 * bca.f90, bcx.f90, bcax.f90, comm.f90) for a build with
 * -D_REAL_IS_DOUBLE_ (examples/make_setting.*: real = double).  The
 * driver's own headers (blas.h, sup.h, bca.h, bcx.h, bcax.h, comm.h)
 * declare the same symbols and stay as they are; this header is the
 * library's statement of them.
 *
 * Conventions, as blas.f90 views its arguments:
 *  - an array argument is one block's (sz(1)+2g) x (sz(2)+2g) x (sz(3)+2g)
 *    doubles in device memory, x fastest: x(1-g:sz(1)+g, 1-g:sz(2)+g,
 *    1-g:sz(3)+g); integer arrays (masks) likewise;
 *  - sz points at 3 ints, g at one; scalar arguments and results are host
 *    pointers;
 *  - each call returns when its results are in place, and runs on a HIP
 *    stream of the calling host thread, so that OpenMP block loops run
 *    their calls concurrently;
 *  - a bad argument (sizes, too few ghost layers for the operation, an
 *    array that is not inside one device allocation of the current device)
 *    prints "bcm: <name>: <reason>" and exits the process, as the
 *    reference's Exit() does.
 * Output and input arrays of a stencil must be different arrays, as
 * Fortran's argument rules require.
 *
 * Not provided: fill_2_ (blas.h:27) and sup_copy_from_buffer_c2f_ /
 * sup_copy_from_buffer_f2c_ (sup.h:69-80), which blas.f90 and sup.f90 do
 * not define; comm_band_cells_ / comm_band_cells_nb_ (comm.h:30-49), which
 * exchange bands between MPI ranks.
 */
#ifndef BCM_FORTRAN_H
#define BCM_FORTRAN_H

#ifdef __cplusplus
extern "C" {
#endif

/* blas.h:23-66 -> blas.f90:18-450 */
void fill_(double* x, double* a, int* sz, int* g);                   /* whole array, ghosts included */
void fill_vf3d_(double* x, double* a, int* sz, int* g, int* ne);     /* ne arrays back to back */
void copy_(double* y, double* x, int* sz, int* g);                   /* whole array */
void copy_integer_(int* y, int* x, int* sz, int* g);                 /* whole array */
void add_(double* C, double* A, double* B, int* sz, int* g);         /* C = A + B (interior) */
void triad_(double* C, double* A, double* B, double* d, int* sz, int* g); /* C = A + d B */
void scal_(double* y, double* a, int* sz, int* g);                   /* y = a y */
void axpy_(double* y, double* x, double* a, int* sz, int* g);        /* y = y + a x */
void xpay_(double* y, double* x, double* a, int* sz, int* g);        /* y = x + a y */
void axpyz_(double* z, double* x, double* y, double* a, int* sz, int* g); /* z = a x + y */
void axpbypz_(double* z, double* x, double* y, double* a, double* b, int* sz, int* g); /* z = a x + b y + z */
void dot_(double* xy, double* y, double* x, int* sz, int* g);        /* interior sum x y */
void dotx2_(double* xy, double* xz, double* x, double* y, double* z, int* sz, int* g);

/* blas.h:68-97 -> blas.f90:452-698 (g >= 1) */
void jacobi_smoother_(double* x1, double* x0, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab,
                      double* At, double* b, double* omega, int* sz, int* g);
/* blas.f90:504-554 takes one array A(0:6, i, j, k) of interleaved
 * coefficients p, w, e, s, n, b, t; blas.h:74-79 declares seven separate
 * arrays, which does not match the routine, so this follows blas.f90. */
void jacobi_smoother2_(double* x1, double* x0, double* A, double* b, double* omega, int* sz, int* g);
void calc_ax_(double* Ax, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab, double* At,
              double* x, int* sz, int* g);
void calc_r_(double* r, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab, double* At, double* x,
             double* b, int* sz, int* g);
void calc_r2_(double* rr, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab, double* At,
              double* x, double* b, int* sz, int* g); /* *rr = interior sum r^2 */

/* blas.h:100-126 -> blas.f90:700-1030.  The masked forms visit the flat
 * range from the first to the last interior cell and act where
 * mask == 1. */
void setup_mask_(double* m, int* mask, int* sz, int* g);
void copy_mask_(double* y, double* x, int* mask, int* sz, int* g);
void dot_mask_(double* xy, double* y, double* x, int* mask, int* sz, int* g);
void calc_ax_mask_(double* Ax, double* Ap, double* Aw, double* Ae, double* As, double* An, double* Ab, double* At,
                   double* x, int* mask, int* sz, int* g);
void jacobi_smoother_mask_(double* x1, double* x0, double* Ap, double* Aw, double* Ae, double* As, double* An,
                           double* Ab, double* At, double* b, int* mask, double* omega, int* sz, int* g);

/* comm.h:23-29 -> comm.f90 (one rank: *total = *a); dot_all_ is dot_
 * summed over ranks. */
void allreduce_(double* total, double* a);
void allreduce_max_(double* total, double* a);
void allreduce_min_(double* total, double* a);
void dot_all_(double* xy, double* y, double* x, int* sz, int* g);

/* sup.h:22-67 -> sup.f90:35-395: a box of sz_c cells from Fortran index
 * i1_src of one block's array (or from a buffer of sz_c values, index
 * i + sz_c(1) (j + sz_c(2) k)) to Fortran index i1_dst of another's (or
 * to a buffer).  _c2f interpolates a coarse array at fine positions
 * (sup_get_intp_index, sup.f90:18-33); _f2c averages 8 fine cells. */
void sup_copy_from_neighbor_(double* data_dst, int* i1_dst, double* data_src, int* i1_src, int* sz_c, int* sz, int* g);
void sup_copy_from_neighbor_c2f_(double* data_dst, int* i1_dst, double* data_src, int* i1_src, int* sz_c, int* sz,
                                 int* g);
void sup_copy_from_neighbor_f2c_(double* data_dst, int* i1_dst, double* data_src, int* i1_src, int* sz_c, int* sz,
                                 int* g);
void sup_copy_to_buffer_(double* buffer, double* data_src, int* i1_src, int* sz_c, int* sz, int* g);
void sup_copy_to_buffer_c2f_(double* buffer, double* data_src, int* i1_src, int* sz_c, int* sz, int* g);
void sup_copy_to_buffer_f2c_(double* buffer, double* data_src, int* i1_src, int* sz_c, int* sz, int* g);
void sup_copy_from_buffer_(double* data_dst, int* i1_dst, double* buffer, int* sz_c, int* sz, int* g);

/* bcx.h -> bcx.f90:18-652: two ghost layers of one face (g >= 2).
 * x1 = X_P, x3 = X_M, x2 = Y_P, x4 = Y_M, x5 = Z_P, x6 = Z_M.  _d: mirror
 * about the Dirichlet value; _d_f: the value itself (a plus face also sets
 * its boundary cell); _n: a Neumann step; _p: periodic copy along the
 * axis (xc unused). */
void bc_x1_d_(double* x, double* xc, int* sz, int* g);
void bc_x3_d_(double* x, double* xc, int* sz, int* g);
void bc_x2_d_(double* x, double* xc, int* sz, int* g);
void bc_x4_d_(double* x, double* xc, int* sz, int* g);
void bc_x5_d_(double* x, double* xc, int* sz, int* g);
void bc_x6_d_(double* x, double* xc, int* sz, int* g);
void bc_x1_d_f_(double* x, double* xc, int* sz, int* g);
void bc_x3_d_f_(double* x, double* xc, int* sz, int* g);
void bc_x2_d_f_(double* x, double* xc, int* sz, int* g);
void bc_x4_d_f_(double* x, double* xc, int* sz, int* g);
void bc_x5_d_f_(double* x, double* xc, int* sz, int* g);
void bc_x6_d_f_(double* x, double* xc, int* sz, int* g);
void bc_x1_n_(double* x, double* xc, int* sz, int* g);
void bc_x3_n_(double* x, double* xc, int* sz, int* g);
void bc_x2_n_(double* x, double* xc, int* sz, int* g);
void bc_x4_n_(double* x, double* xc, int* sz, int* g);
void bc_x5_n_(double* x, double* xc, int* sz, int* g);
void bc_x6_n_(double* x, double* xc, int* sz, int* g);
void bc_x1_p_(double* x, double* xc, int* sz, int* g);
void bc_x2_p_(double* x, double* xc, int* sz, int* g);
void bc_x5_p_(double* x, double* xc, int* sz, int* g);

/* bca.h -> bca.f90:12-394: a face's condition folded into the
 * coefficients of its boundary cells.  _d (Dirichlet): Ap -= Anb,
 * b -= 2 Anb xc, Anb = 0; _n (Neumann): Ap += Anb, b += Anb xc on a minus face and
 * b -= Anb xc on a plus face, Anb = 0. */
void bc_aw_d_(double* Ap, double* Aw, double* b, double* xc, int* sz, int* g);
void bc_ae_d_(double* Ap, double* Ae, double* b, double* xc, int* sz, int* g);
void bc_as_d_(double* Ap, double* As, double* b, double* xc, int* sz, int* g);
void bc_an_d_(double* Ap, double* An, double* b, double* xc, int* sz, int* g);
void bc_ab_d_(double* Ap, double* Ab, double* b, double* xc, int* sz, int* g);
void bc_at_d_(double* Ap, double* At, double* b, double* xc, int* sz, int* g);
void bc_aw_n_(double* Ap, double* Aw, double* b, double* xc, int* sz, int* g);
void bc_ae_n_(double* Ap, double* Ae, double* b, double* xc, int* sz, int* g);
void bc_as_n_(double* Ap, double* As, double* b, double* xc, int* sz, int* g);
void bc_an_n_(double* Ap, double* An, double* b, double* xc, int* sz, int* g);
void bc_ab_n_(double* Ap, double* Ab, double* b, double* xc, int* sz, int* g);
void bc_at_n_(double* Ap, double* At, double* b, double* xc, int* sz, int* g);

/* bcax.h -> bcax.f90:86-280: the Poiseuille inflow on the X_M face;
 * center, org, bsize, csize point at 3 host doubles each (bsize unused,
 * as in bcax.f90). */
void bc_x3_poiseuille_u_(double* x, double* xc, int* sz, int* g, double* center, double* org, double* bsize,
                         double* csize);
void bc_aw_poiseuille_u_(double* Ap, double* Aw, double* b, double* xc, int* sz, int* g, double* center, double* org,
                         double* bsize, double* csize);
void bc_x3_poiseuille_p_(double* x, double* xc, int* sz, int* g, double* center, double* org, double* bsize,
                         double* csize);
void bc_aw_poiseuille_p_(double* Ap, double* Aw, double* b, double* xc, int* sz, int* g, double* center, double* org,
                         double* bsize, double* csize);

#ifdef __cplusplus
}
#endif

#endif
