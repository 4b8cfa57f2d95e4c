/*
 * bcm_hip.h -- C ABI of the MI355X-native BCM block data path.
 *
 * One GPU holds every leaf block of a mesh in HBM.  Fields are stored
 * block-major, each block as a ghosted (n+2vc)^3 array in the reference's
 * Index3DS order (x fastest), i.e. the contiguous layout of
 * include/BlockManager.h:253-283 with block b at offset b*(n+2vc)^3.
 *
 * Every entry point returns an int status (BCM_OK == 0) unless stated.
 * No torch types appear here: plain pointers, sizes and opaque handles.
 * All device work of a mesh is issued on the mesh's own HIP stream (the
 * default priority); bcm_mesh_sync() waits for it.
 *
 * Citations are relative to the BCMTools 1.0.5 source tree.
 */
#ifndef BCM_HIP_H
#define BCM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Face / subface numbering of include/BCMTools.h:48,51. */
enum { BCM_X_M = 0, BCM_X_P, BCM_Y_M, BCM_Y_P, BCM_Z_M, BCM_Z_P, BCM_NUM_FACE };
enum { BCM_SF_00 = 0, BCM_SF_01, BCM_SF_10, BCM_SF_11, BCM_NUM_SUBFACE };

/* Boundary types of Utils/include/BoundaryInfo.h:30-35. */
enum { BCM_BC_INNER = 0, BCM_BC_DIRICHLET = 1, BCM_BC_NEUMANN = 2, BCM_BC_PERIODIC = 3 };

/* Status codes.  The reference calls Exit() (include/BCMTools.h:36-37);
 * this library returns instead. */
enum {
  BCM_OK = 0,
  BCM_ERR_ARG = 1,      /* bad argument, size mismatch, unsupported shape */
  BCM_ERR_HIP = 2,      /* a HIP runtime call failed */
  BCM_ERR_NOMEM = 3,    /* device or host allocation failed */
  BCM_ERR_STATE = 4,    /* handle used in the wrong state */
  BCM_ERR_BALANCE = 5,  /* 2:1 constraint broken (Utils/src/BCMOctree.cpp:462-467) */
  BCM_ERR_NAN = 6,      /* NaN/Inf residual (examples/SolverTest/ILS3D.h:53-59) */
  BCM_ERR_NOCONV = 7,   /* iteration limit reached before the tolerance */
  BCM_ERR_BUSY = 8,     /* admission refused: HBM budget or slot table full */
  BCM_ERR_IO = 9        /* a file could not be opened, read or written */
};

enum { BCM_F64 = 0, BCM_F32 = 1 };

/* ------------------------------------------------------------------ */
/* Block table.  One record per leaf block, in block-ID order.         */
/* Mirrors BlockBase (include/BlockBase.h:68-128), NeighborInfo        */
/* (include/NeighborInfo.h:36-228) and BoundaryInfo                    */
/* (Utils/include/BoundaryInfo.h).  nb_id[f][0] is the neighbour for   */
/* level difference 0 / -1; nb_id[f][0..3] the four finer neighbours   */
/* for level difference +1; -1 means no neighbour.                     */
/* ------------------------------------------------------------------ */
typedef struct bcm_block_info {
  uint64_t pedigree;          /* x16|y16|z16|root12|level4, Utils/include/Pedigree.h:42-73 */
  int32_t level;
  int32_t root_id;
  double origin[3];
  double block_size[3];
  int32_t nb_level_diff[6];   /* neighbour level minus my level: -1, 0, +1 */
  int32_t nb_id[6][4];
  int32_t nb_subface[6];      /* my position on the coarser neighbour's face (diff -1) */
  int32_t bc_type[6];         /* BCM_BC_* */
  int32_t bc_id[6];           /* index into the boundary-value table, -1 if none */
} bcm_block_info;

/* ------------------------------------------------------------------ */
/* Host octree builder: BCMOctree + dividers + BlockFactory            */
/* (Utils/src/BCMOctree.cpp:23-478, Utils/include/SimpleDivider.h,     */
/* Utils/include/BlockFactory.h:89-96) plus an explicit 2:1 balancer,  */
/* which the reference lacks.                                           */
/* ------------------------------------------------------------------ */
/* BCM_DIV_USER: the caller's divider, called once per node (see
 * bcm_tree_build_user).  BCM_ORDER_RANDOM: the leaves in a pseudo-random
 * order, the same on every platform (the reference's BCMOctree::RANDOM
 * shuffles with std::random_shuffle, Utils/src/BCMOctree.cpp:400-404,
 * whose order is platform-dependent). */
enum { BCM_DIV_FLAT = 0, BCM_DIV_SIMPLE = 1, BCM_DIV_SHELL = 2, BCM_DIV_USER = 3 };
enum { BCM_ORDER_Z = 0, BCM_ORDER_HILBERT = 1, BCM_ORDER_RANDOM = 2 };
/* Boundary setters: all outer faces Dirichlet id 0
 * (examples/sor2/BoundaryConditionSetter.h), or the sor1 'x'/'y'/'z'
 * setter (examples/sor1/BoundaryConditionSetter.h): outer faces periodic
 * except the two faces normal to the axis, Dirichlet ids 0 and 1.
 * BCM_BCSET_PER_FACE: the outer faces on domain face f get type
 * face_bc_type[f] (BCM_BC_DIRICHLET or BCM_BC_NEUMANN) and boundary-value
 * id f, as SolverTest's LocalScalar3D takes boundaryType[6] and
 * boundaryValue[6] per domain face (examples/SolverTest/BlockScalar3D.h:67-105). */
enum {
  BCM_BCSET_ALL_DIRICHLET = 0,
  BCM_BCSET_SOR1_X = 1,
  BCM_BCSET_SOR1_Y = 2,
  BCM_BCSET_SOR1_Z = 3,
  BCM_BCSET_PER_FACE = 4
};

#define BCM_MAX_SHELLS 8
typedef struct bcm_tree_params {
  int32_t root_n[3];          /* RootGrid(nx, ny, nz) */
  int32_t periodic[3];        /* RootGrid::setPeriodic{X,Y,Z} */
  double root_origin[3];      /* BlockFactory rootOrigin */
  double root_length;         /* BlockFactory rootLength */
  int32_t divider;            /* BCM_DIV_* */
  int32_t min_level;
  int32_t max_level;
  int32_t n_shells;           /* BCM_DIV_SHELL: refine a node whose box meets any shell */
  double shells[BCM_MAX_SHELLS][5]; /* cx, cy, cz, r_in, r_out in physical coordinates */
  int32_t ordering;           /* BCM_ORDER_* */
  int32_t balance;            /* 1: split leaves until face neighbours differ by <= 1 level */
  int32_t bc_setter;          /* BCM_BCSET_* */
  /* BCM_DIV_SHELL only. shell_mode 0: split while the node's box meets a
   * shell. shell_mode 1: as 0 above the last split; the last split (to
   * max_level) needs the node centre within shell_margin node edges of a
   * shell, which refines a band about one node thick instead of every box
   * the shell touches. */
  int32_t shell_mode;
  double shell_margin;
  int32_t face_bc_type[6];    /* BCM_BCSET_PER_FACE: BCM_BC_* per domain face X_M..Z_P */
} bcm_tree_params;

typedef struct bcm_tree* bcm_tree_t;

/* Replaces `new BCMOctree(rootGrid, divider, ordering)` + the
 * BlockFactory::makeBlock loop of examples/sor1/main.cpp:88-123. */
int bcm_tree_build(const bcm_tree_params* params, bcm_tree_t* out);
/* Divider::operator() (Utils/include/Divider.h:40-46): the node kinds of
 * Divider::NodeType, in its order. */
enum { BCM_NODE_BRANCH = 0, BCM_NODE_LEAF_ACTIVE = 1, BCM_NODE_LEAF_INACTIVE = 2 };
typedef int (*bcm_divider_fn)(void* user, uint64_t pedigree);
/* `new BCMOctree(rootGrid, divider, ordering)` with the caller's divider
 * (Utils/src/BCMOctree.cpp:23-48, 111-124): params->divider must be
 * BCM_DIV_USER; divide(user, pedigree) is called for every node, parents
 * before children, and returns a BCM_NODE_* kind. */
int bcm_tree_build_user(const bcm_tree_params* params, bcm_divider_fn divide, void* user, bcm_tree_t* out);
/* Host threads of bcm_tree_build (0, the default: one per hardware thread,
 * at most 64).  The reference builds its octree on one thread
 * (Utils/src/BCMOctree.cpp:23-48); here the levels of the tree, the 2:1
 * balance checks and the block records are split over threads, and the
 * block table is the same for every thread count.  A caller's divider
 * (bcm_tree_build_user) is still called on one thread, parents before
 * children. */
int bcm_tree_set_threads(int nthreads);
/* BCMOctree::getNumLeafNode (Utils/include/BCMOctree.h:111). Returns the count (>=0) or -BCM_ERR_ARG. */
int bcm_tree_num_blocks(bcm_tree_t tree);
/* Copies the block table (one record per leaf in block-ID order). */
int bcm_tree_get_blocks(bcm_tree_t tree, bcm_block_info* out, int capacity);
/* The destroy calls accept NULL (no-op, BCM_OK), like free(). */
int bcm_tree_destroy(bcm_tree_t tree);

/* ------------------------------------------------------------------ */
/* Device mesh and fields                                               */
/* ------------------------------------------------------------------ */
typedef struct bcm_mesh* bcm_mesh_t;
typedef struct bcm_field* bcm_field_t;

/* BlockManager::registerBlock/endRegisterBlock + prepareForVCUpdate
 * (src/BlockManager.cpp:28-114, 471-648): uploads the block table and
 * the same-level / F2C / C2F / Dirichlet face-task lists. n = cells per
 * block edge (even, as include/BlockBase.h:75-78 demands), vc = ghost
 * width (n/2 >= vc, include/DataClass.h:53-56). */
int bcm_mesh_create(const bcm_block_info* blocks, int nblocks, int n, int vc, bcm_mesh_t* out);
int bcm_mesh_destroy(bcm_mesh_t mesh);
/* counts[0..3] = same-level, F2C, C2F and Dirichlet/Neumann face tasks. */
int bcm_mesh_task_counts(bcm_mesh_t mesh, int counts[4]);
int bcm_mesh_sync(bcm_mesh_t mesh);
/* Per-kernel HIP-event timing on the mesh stream (off by default).
 * BCM_K_SOR_HALO: the halo-gathering sweeps between a bcm_sor_iterate
 * call's first sweep and its last; BCM_K_SOR: every other SOR sweep.
 * BCM_K_JACOBI_HALO: the same for bcm_jacobi_poisson_iterate's
 * halo-gathering steps; BCM_K_JACOBI: every other Jacobi sweep. */
enum {
  BCM_K_GHOST = 0, BCM_K_DIRICHLET, BCM_K_SOR, BCM_K_JACOBI, BCM_K_SPMV, BCM_K_RESID,
  BCM_K_DOT, BCM_K_BLAS1, BCM_K_REDUCE, BCM_K_ASSEMBLE, BCM_K_SOR_HALO, BCM_K_JACOBI_HALO,
  BCM_K_COUNT
};
int bcm_mesh_set_timing(bcm_mesh_t mesh, int enable);
/* Sum of event-measured durations (ms) and launch count since the last reset. */
int bcm_mesh_kernel_time(bcm_mesh_t mesh, int kernel, double* total_ms, long long* launches);
int bcm_mesh_reset_timing(bcm_mesh_t mesh);

/* Scalar3D<T>(size, vc) allocation on every block
 * (BlockManager::setDataClass<Scalar3D<T>,Scalar3DUpdater<T>,T>,
 * include/BlockManager.h:253-283); zero-filled, ghosts included. */
int bcm_field_create(bcm_mesh_t mesh, int dtype, bcm_field_t* out);
int bcm_field_destroy(bcm_field_t field);
/* Host <-> device copies of the whole field (nblocks*(n+2vc)^3 values). */
int bcm_field_upload(bcm_field_t field, const void* host, size_t bytes);
/* Page-locked host memory, zero-filled (no counterpart in the reference,
 * whose fields live in ordinary host memory).  Copies between a field and
 * such memory run at the full PCIe rate; from ordinary memory they are
 * staged by the runtime.  The C++ facade allocates the host arena of every
 * class it mirrors on the device with it.  bcm_host_free accepts NULL. */
int bcm_host_alloc(size_t bytes, void** out);
int bcm_host_free(void* ptr);
int bcm_field_download(bcm_field_t field, void* host, size_t bytes);
/* The same for blocks first_block .. first_block+nblocks-1 only (one
 * segment of a batch, a block range of one mesh); bytes must match. */
int bcm_field_upload_blocks(bcm_field_t field, int first_block, int nblocks, const void* host, size_t bytes);
int bcm_field_download_blocks(bcm_field_t field, int first_block, int nblocks, void* host, size_t bytes);

/* Restart files of SolverTest's LocalScalar3D<real>
 * (examples/SolverTest/LocalScalar3D.cpp:84-357), written from and read
 * into device memory through pinned staging buffers.  The file element
 * is the field's type (the reference's `real`).
 *   Dump2 / Load2 (:226-324): <dir>/dump-<label>-<rank %05d>-<step %010d>.bin,
 *     int32 {nx, ny, nz, vc, 1, nblocks}, then every block's (n+2vc)^3 array.
 *   Dump / Load (:84-222): one file per block,
 *     <dir>/dump-<label>-<rank %05d>-<id %05d>-<step %010d>.bin,
 *     int32 {nx, ny, nz, vc, 1}, then the block's array.
 * The reference writes to ./BIN/ (dir = "./BIN"), creating it.  It imposes
 * the field's boundary conditions before a dump and after a load
 * (LocalScalar3D::ImposeBoundaryCondition); here the caller does that with
 * bcm_ghost_fill and bcm_impose_bc.  A load accepts files of the field's
 * block edge, or of half of it: each fine interior cell then takes its
 * coarse cell (sf3d_copy_x2, examples/SolverTest/bsf3d.f90:18-46) and the
 * ghosts keep their values.  Any other header returns BCM_ERR_ARG, where
 * the reference calls Exit(0); a missing or short file BCM_ERR_IO. */
int bcm_field_dump2(bcm_field_t field, const char* dir, const char* label, int rank, int step);
int bcm_field_load2(bcm_field_t field, const char* dir, const char* label, int rank, int step);
int bcm_field_dump(bcm_field_t field, const char* dir, const char* label, int rank, int step);
int bcm_field_load(bcm_field_t field, const char* dir, const char* label, int rank, int step);

/* VTK output of SolverTest's VtkWriter::writeScalar
 * (examples/SolverTest/VtkWriter.h:733-889) with printVTIC (:75-200): one
 * ImageData file per block, <dir>/<step %010d>/data-<name>-<rank %05d>-
 * <id %05d>-<step %010d>.vti, holding the block's interior as appended raw
 * Float32 cell data (origin and spacing printed in scientific notation with
 * 16 digits), and, on rank 0, the index <dir>/data-<name>-<step %010d>.vthb
 * (vtkHierarchicalBoxDataSet: difflevel + 1 refinement ratios of 2, one
 * DataSet per block with its level, dataset = start_id + id and its
 * amr_box in cells of its level, from the block origin in root lengths:
 * (origin - root_origin) / root_length).  The reference writes to ./VTK
 * (dir = "./VTK").  The interior is packed to FP32 on the device. */
int bcm_field_write_vtk(bcm_field_t field, const char* dir, const char* name, int step, int rank, int start_id,
                        int difflevel, const double root_origin[3], double root_length);
void* bcm_field_device_ptr(bcm_field_t field);
size_t bcm_field_bytes(bcm_field_t field);

/* ------------------------------------------------------------------ */
/* Ghost ("virtual cell") fill                                           */
/* ------------------------------------------------------------------ */
/* kind bits: same-level copy, F2C restriction, C2F prolongation.       */
#define BCM_GF_FLAT 1
#define BCM_GF_F2C 2
#define BCM_GF_C2F 4
#define BCM_GF_ALL 7
/* axis bits (none set = all three axes) */
#define BCM_GF_X 8
#define BCM_GF_Y 16
#define BCM_GF_Z 32
/* BlockManager::updateVC / updateVC_{X,Y,Z} / updateVC_{X,Y,Z}_{Flat,F2C,C2F}
 * (include/BlockManager.h:311-440) -> VCUpdater::copyFromNeighbor
 * (include/VCUpdater.h:65-82) -> Scalar3DUpdater<T>::copyFromNeighbor,
 * copyFromNeighborF2C_0, copyFromNeighborC2F_0
 * (include/Scalar3DUpdater.h:321-347, 519-708). Face ghosts only. */
int bcm_ghost_fill(bcm_field_t field, int mask);

/* VCUpdater::copyFromNeighbor(face, subface) for one block
 * (include/VCUpdater.h:65-82, include/Scalar3DUpdater.h:321-347, 519-708):
 * the ghosts of `face` of `block` from its neighbour -- same level, the
 * finer neighbour on `subface` (F2C), or the coarser one (C2F).
 * BCM_ERR_ARG when that face has no such neighbour. */
int bcm_ghost_fill_face(bcm_field_t field, int block, int face, int subface);

/* Face buffers of an exchange between GPUs (Scalar3DUpdater::
 * copyToCommBuffer, copyToCommBufferF2C, copyToCommBufferC2F and
 * copyFromCommBuffer, copyFromCommBufferF2C, copyFromCommBufferC2F,
 * include/Scalar3DUpdater.h:377-515, 711-901).  One GPU holds every block
 * here, so the exchange never needs them; they are the pack and unpack
 * steps a multi-GPU exchange would run around its transfers, with the
 * reference's buffer layout: a face's box is enumerated x fastest, element
 * m = i + w (j + h k) for a box of w x h x d cells.
 *   pack, BCM_COMM_SAME: the block's interior slab of depth vc at `face`
 *     (w x h x d = vc x n x n for X faces, n x vc x n for Y, n x n x vc for Z);
 *   pack, BCM_COMM_F2C: the fine sender's restriction toward its coarser
 *     neighbour across `face` (a quarter box: vc x n/2 x n/2 for X faces);
 *   pack, BCM_COMM_C2F: the coarse sender's prolongation for the finer
 *     neighbour on quarter `subface` of `face` (vc x n x n for X faces);
 *   unpack, BCM_COMM_SAME or BCM_COMM_C2F: the block's ghost slab at `face`;
 *   unpack, BCM_COMM_F2C: the quarter `subface` of the ghost slab at `face`.
 * offset: the face's first element in `buffer` (elements of the field's
 * type).  `buffer` is device memory of `bytes` bytes; every face must fit. */
enum { BCM_COMM_SAME = 0, BCM_COMM_F2C = 1, BCM_COMM_C2F = 2 };
typedef struct bcm_comm_face {
  int32_t block;
  int32_t face;     /* BCM_X_M .. BCM_Z_P, the block's face toward its neighbour */
  int32_t subface;  /* 0..3: quarter of `face` (pack C2F, unpack F2C); otherwise ignored */
  int32_t kind;     /* BCM_COMM_* */
  int64_t offset;
} bcm_comm_face;
/* Elements one face of `kind` takes in a buffer (n^2 vc, or n^2 vc / 4 for F2C). */
int64_t bcm_comm_face_size(bcm_field_t field, int kind);
int bcm_comm_pack(bcm_field_t field, const bcm_comm_face* faces, int nfaces, void* buffer, size_t bytes);
int bcm_comm_unpack(bcm_field_t field, const bcm_comm_face* faces, int nfaces, const void* buffer, size_t bytes);
/* A face list uploaded to the device once, for exchanges repeated over
 * the same faces (the setup prepareForVCUpdate does once,
 * src/BlockManager.cpp:63-114; each later exchange only packs and unpacks).
 * bcm_comm_pack_list / bcm_comm_unpack_list enqueue the kernel on the
 * mesh's stream and return: no allocation, no host copy, no wait (use
 * bcm_mesh_sync before reading the buffer on the host).  The list belongs
 * to one mesh; fields of other meshes are rejected.  Destroying a list
 * waits for the mesh's stream; destroy every list before its mesh. */
typedef struct bcm_comm_list* bcm_comm_list_t;
int bcm_comm_list_create(bcm_mesh_t mesh, const bcm_comm_face* faces, int nfaces, bcm_comm_list_t* out);
int bcm_comm_list_destroy(bcm_comm_list_t list);
int bcm_comm_pack_list(bcm_field_t field, bcm_comm_list_t list, void* buffer, size_t bytes);
int bcm_comm_unpack_list(bcm_field_t field, bcm_comm_list_t list, const void* buffer, size_t bytes);
/* One rank's per-peer buffers for a partition of a mesh's blocks over
 * ranks (block_rank[b] >= 0 for every block): the layout of the
 * reference's CommBuffer (src/CommBuffer.cpp:55-94, one buffer per peer,
 * each face's slot padded to 8 bytes) over the face lists of
 * BlockManager::setFaceLists / setSeparateFaceLists /
 * setSeparateLevelDiffFaceLists (src/BlockManager.cpp:471-650) as
 * setSend/RecvBufferPointers (:671-711) fill it.  Replaces what
 * prepareForVCUpdate (:63-114) builds for the faces between ranks.
 *   direction   BCM_COMM_SEND: the faces `rank` packs for each peer, in the
 *               order it visits them (block, face, subface); BCM_COMM_RECV:
 *               the faces it unpacks, in each sender's order
 *               (RecvFaceComp, include/BlockManager.h:134-148)
 *   axis        -1: all faces (AtOnce); 0..2: the SeparateXYZ list
 *   level_diff  BCM_COMM_ANY_LEVEL, or -1 / 0 / +1: the
 *               SeparateXYZ_SeparateLevelDiff list of that level difference
 *               (neighbour minus own level, as seen by `rank`'s block;
 *               updateVC_*_F2C receives on +1 and sends on -1, C2F the
 *               reverse, include/BlockManager.h:562-592)
 * The peers come in ascending rank order, each with a contiguous run of
 * faces and a byte range of one buffer of *total_bytes bytes; a face's
 * `offset` is in elements of `dtype` (BCM_F64 / BCM_F32), ready for
 * bcm_comm_pack / bcm_comm_unpack on a field of that mesh.  With
 * peers == faces == NULL only *npeers, *nfaces and *total_bytes are
 * written; otherwise both arrays must hold max_peers / max_faces entries
 * at least. */
enum { BCM_COMM_SEND = 0, BCM_COMM_RECV = 1 };
enum { BCM_COMM_ANY_LEVEL = 2 };
typedef struct bcm_comm_peer {
  int32_t rank;     /* the peer */
  int32_t nfaces;   /* its faces: faces[first .. first + nfaces) */
  int64_t first;
  int64_t offset;   /* byte offset of the peer's buffer */
  int64_t bytes;    /* the peer's message size */
} bcm_comm_peer;
int bcm_comm_plan(const bcm_block_info* blocks, int32_t nblocks, int32_t n, int32_t vc, int32_t dtype,
                  const int32_t* block_rank, int32_t rank, int32_t direction, int32_t axis, int32_t level_diff,
                  bcm_comm_peer* peers, int32_t max_peers, int32_t* npeers, bcm_comm_face* faces,
                  int64_t max_faces, int64_t* nfaces, int64_t* total_bytes);
/* The same with `buffer` in host memory (staged through a device buffer):
 * what the C++ facade's Scalar3DUpdater uses for a driver's send and
 * receive buffers. */
int bcm_comm_pack_host(bcm_field_t field, const bcm_comm_face* faces, int nfaces, void* buffer, size_t bytes);
int bcm_comm_unpack_host(bcm_field_t field, const bcm_comm_face* faces, int nfaces, const void* buffer,
                         size_t bytes);

/* Vector3D<T> (include/Vector3D.h:36-110): three components per cell,
 * interleaved (component l of cell (i,j,k) at 3 Index3DS(i,j,k) + l, the
 * reference's Index3DV), nblocks*3*(n+2vc)^3 values, block-major.  The
 * reference's Vector3DUpdater is a stub (include/Vector3DUpdater.h:35-45);
 * here the exchange applies the scalar rules (copy, F2C, C2F) to each
 * component. */
typedef struct bcm_vfield* bcm_vfield_t;
int bcm_vfield_create(bcm_mesh_t mesh, int dtype, bcm_vfield_t* out);
int bcm_vfield_destroy(bcm_vfield_t field);
int bcm_vfield_upload(bcm_vfield_t field, const void* host, size_t bytes);
int bcm_vfield_download(bcm_vfield_t field, void* host, size_t bytes);
int bcm_vfield_upload_blocks(bcm_vfield_t field, int first_block, int nblocks, const void* host, size_t bytes);
int bcm_vfield_download_blocks(bcm_vfield_t field, int first_block, int nblocks, void* host, size_t bytes);
/* kind and axis bits as bcm_ghost_fill (BCM_GF_BLOCKWISE not accepted). */
int bcm_vfield_ghost_fill(bcm_vfield_t field, int mask);
int bcm_vfield_ghost_fill_face(bcm_vfield_t field, int block, int face, int subface);

/* VCUpdateMethod (include/BlockManager.h:39-46). */
#define BCM_VC_AT_ONCE 0
#define BCM_VC_SEPARATE_XYZ 1
#define BCM_VC_SEPARATE_XYZ_SEPARATE_LEVEL_DIFF 2
/* LocalScalar3D::UpdateVirtualCells (examples/SolverTest/LocalScalar3D.h:
 * 198-223): the ghost exchange in the method's order -- all at once; X, Y,
 * then Z; or F2C on X, Y, Z, then same-level, then C2F.  Every exchange
 * reads interiors only, so the three give the same bits. */
int bcm_update_virtual_cells(bcm_field_t field, int method);
/* LocalScalar3D::ImposeBoundaryCondition (LocalScalar3D.h:225-257):
 * bcm_update_virtual_cells, then bcm_impose_bc with the face values. */
int bcm_impose_boundary_condition(bcm_field_t field, int method, const double* values, int nvalues);

/* Dirichlet mirror g = 2v - u0 on every outer face whose bc_type is
 * DIRICHLET, v = values[bc_id]: Solver::setBoundaryCondition /
 * setDirichletBoundaryInBlock (examples/sor1/Solver.cpp:202-294). */
int bcm_dirichlet_mirror(bcm_field_t field, const double* values, int nvalues);

/* Outer-face conditions of SolverTest's BlockScalar3D::ImposeBoundaryCondition
 * (examples/SolverTest/BlockScalar3D.h:199-206, bcx.f90:18-376) on every
 * outer DIRICHLET or NEUMANN face, v = values[bc_id], u the adjacent
 * interior cell: Dirichlet g1 = 2v - u, g2 = v; Neumann g1 = u -/+ xc,
 * g2 = u -/+ 2 xc on a minus/plus face, xc = v * cellSize.x
 * (BlockScalar3D.h:98-100).  g2 (the second ghost layer) only when vc >= 2.
 * FP64 arithmetic, rounded once for FP32 fields. */
int bcm_impose_bc(bcm_field_t field, const double* values, int nvalues);

/* Per-field outer-face conditions of SolverTest's LocalScalar3D, whose
 * fields each carry one boundary type and value per domain face
 * (examples/SolverTest/LocalScalar3D.h:36-100, BlockScalar3D.h:76-214).
 * types[f] is BCM_BC_DIRICHLET, BCM_BC_NEUMANN, BCM_BC_PERIODIC or
 * BCM_BC_INNER (nothing) for domain face f = X_M..Z_P; values[f] its
 * value, a Neumann value scaled by each block's cellSize.x
 * (BlockScalar3D.h:98-100).  They apply to every block face without a
 * neighbour, in X_M..Z_P order within a block.
 * bcm_impose_face_bc (BlockScalar3D::ImposeBoundaryCondition(), bcx.f90):
 * Dirichlet g1 = 2v - u, g2 = v; Neumann g1 = u -/+ xc, g2 = u -/+ 2 xc on a
 * minus/plus face; periodic (plus faces: bc_x1_p, bc_x2_p, bc_x5_p,
 * bcx.f90:561-652) copies the block's own opposite layers across the whole
 * face, ghost rows included.  g2 and the second periodic layer are written
 * when vc >= 2.
 * bcm_fold_face_bc (BlockScalar3D::ImposeBoundaryCondition(Ap..b),
 * bca.f90:12-394): Dirichlet Ap -= Anb, b -= Anb v 2; Neumann Ap += Anb,
 * b += Anb xc on a minus face, b -= Anb xc on a plus face; Anb = 0;
 * periodic faces unchanged.  A[7] = {Ap, Aw, Ae, As, An, Ab, At}.
 * The Poiseuille inflow (bcax.f90) is available per block through
 * include/bcm_fortran.h. */
int bcm_impose_face_bc(bcm_field_t field, const int32_t types[6], const double values[6]);
int bcm_fold_face_bc(const bcm_field_t A[7], bcm_field_t b, const int32_t types[6], const double values[6]);

/* ------------------------------------------------------------------ */
/* Red-black SOR (examples/sor1/Solver.cpp:132-153,                     */
/* examples/sor2/Solver.cpp:136-160).  Both colours, ghosts frozen.      */
/* source == NULL: sor1 form  f = c0 f + cx(fW+fE) + cy(fS+fN) + cz(fB+fT)  */
/* source != NULL: sor2 form  fNew = c(cx(..)+cy(..)+cz(..) - s);          */
/*                            f = w fNew + (1-w) f                         */
/* ------------------------------------------------------------------ */
int bcm_rbsor_sweep(bcm_field_t f, bcm_field_t source, double omega);
/* One sor1/sor2 outer iteration, `iters` times: SOR on every block, the
 * in-block Dirichlet BC, ghost fill (AtOnce), Dirichlet BC
 * (examples/sor2/Solver.cpp:77-133 with nLoopInner == 1). */
int bcm_sor_iterate(bcm_field_t f, bcm_field_t source, double omega,
                    const double* bvalues, int nbvalues, int iters);
/* bcm_sor_iterate with nLoopInner = `inner` >= 1 (examples/sor1/Solver.cpp:
 * 102-105, examples/sor2/Solver.cpp:106-109): per outer iteration, `inner`
 * times SOR + in-block Dirichlet BC on every block over the ghosts of the
 * last exchange, then ghost fill and Dirichlet BC.  inner == 1 is
 * bcm_sor_iterate; for inner > 1 the HALO schedule is replaced by
 * FUSED_TASKS (the halo sweeps read the neighbours' latest iterate). */
int bcm_sor_iterate_inner(bcm_field_t f, bcm_field_t source, double omega,
                          const double* bvalues, int nbvalues, int iters, int inner);
/* Scheduling of bcm_sor_iterate's outer iteration (all give the same bits):
 *   AUTO          halo-gathering sweeps where the mesh allows, else FUSED_TASKS
 *   FUSED_TASKS   one-pass red-black sweep per block (2x2 cell columns per
 *                 thread, z-planes through LDS) + per-face-task ghost fill
 *   FUSED_GATHER  the same sweep + block-wise ghost gather
 *   HALO          ghost frame built inside the sweep from the previous
 *                 iterate, no separate fill; the iterates and the source
 *                 are held compact (n^3 per block, mesh workspace) between
 *                 the call's first sweep, which reads f's stored ghosts,
 *                 and its last, which writes f (2x1 cell patches per
 *                 thread; vc = 1, n in {16, 32}, every face connected or
 *                 Dirichlet)
 *   COLOUR_TASKS  two one-colour passes (one thread per cell) + task fill */
enum { BCM_SOR_AUTO = 0, BCM_SOR_FUSED_TASKS, BCM_SOR_FUSED_GATHER, BCM_SOR_HALO, BCM_SOR_COLOUR_TASKS,
       BCM_SOR_NUM_VARIANTS };
int bcm_sor_set_variant(bcm_mesh_t mesh, int variant);
/* Returns the mesh's variant, or -BCM_ERR_ARG for a NULL mesh. */
int bcm_sor_get_variant(bcm_mesh_t mesh);
/* Autotuner: times every variant the mesh supports for `iters` outer
 * iterations on a scratch copy of f (f is not modified), keeps the
 * fastest for the mesh and reports it with the time per iteration of
 * each variant (ms_per_iter[BCM_SOR_NUM_VARIANTS]; unsupported: -1). */
int bcm_sor_autotune(bcm_field_t f, bcm_field_t source, double omega, const double* bvalues, int nbvalues,
                     int iters, int* best, double* ms_per_iter);
/* Relative residual of the sor2 system lap_h(f) = s:
 * ||s - lap_h f||_2 / ||s||_2 over interior cells (FP64 reduction). */
int bcm_sor_residual(bcm_field_t f, bcm_field_t source, double* relres);
/* No counterpart in the reference.  bcm_sor_residual for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same kernels and
 * the same result bits, with the partial sums in a fixed device array
 * that every GPU holds once for the life of the process (4 MiB), instead
 * of the mesh's coefficient table, which bcm_sor_residual allocates on
 * first use.  Calls no hipMalloc/hipFree.  Meshes of at most 268,435,456
 * interior cells; larger ones return BCM_ERR_ARG.  Calls that overlap on
 * one GPU (from several host threads) share the array and race. */
int bcm_sor_residual_noalloc(bcm_field_t f, bcm_field_t source, double* relres);

/* ------------------------------------------------------------------ */
/* SolverTest per-block kernels (examples/SolverTest/blas.h:23-126),   */
/* applied to every block.  A[7] = {Ap, Aw, Ae, As, An, Ab, At}.        */
/* ------------------------------------------------------------------ */
/* jacobi_smoother_ (blas.f90:452-502): x1 = x0 - w (x0 - r/Ap). */
int bcm_jacobi7(bcm_field_t x1, bcm_field_t x0, const bcm_field_t A[7], bcm_field_t b, double omega);
/* calc_ax_ (blas.f90:556-600) */
int bcm_calc_ax(bcm_field_t ax, const bcm_field_t A[7], bcm_field_t x);
/* calc_r_ (blas.f90:602-648) */
int bcm_calc_r(bcm_field_t r, const bcm_field_t A[7], bcm_field_t x, bcm_field_t b);
/* calc_r2_ + CalcR2 (blas.f90:650-698, ILS3D.h:168-217): sum of r^2, FP64 out. */
int bcm_calc_r2(const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, double* rr);
/* dot_ + ILS3D::DOT (blas.f90:379-412, ILS3D.h:272-311). Deterministic tree. */
int bcm_dot(bcm_field_t x, bcm_field_t y, double* xy);
/* axpy y+=a x; xpay y=x+a y; axpyz z=a x+y; axpbypz z=a x+b y+z (interior);
 * fill/copy include ghosts (blas.f90:18-377). */
int bcm_axpy(bcm_field_t y, bcm_field_t x, double a);
int bcm_xpay(bcm_field_t y, bcm_field_t x, double a);
int bcm_axpyz(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a);
int bcm_axpbypz(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a, double b);
int bcm_fill(bcm_field_t x, double a);
int bcm_copy(bcm_field_t y, bcm_field_t x);

/* Masked forms (blas.f90:753-1113: jacobi_smoother_mask, calc_ax_mask,
 * calc_r_mask, calc_r2_mask, dot_mask, axpy_mask, xpay_mask, axpyz_mask,
 * axpbypz_mask, copy_mask): only interior cells whose mask value is 1 are
 * written or summed; the others keep their values.  The reference's mask
 * is a LocalScalar3D<int> set to 1 inside and 0 in the ghosts by
 * setup_mask (blas.f90:700-751); here the mask is a field of the
 * operands' type holding 0/1.  The Fortran walks the linear range from the
 * first to the last interior cell, ghosts between rows included; those
 * ghosts carry mask 0 from setup_mask, so visiting interior cells only is
 * the same. */
/* setup_mask (blas.f90:700-751, called at examples/SolverTest/Solver.cpp:
 * 1022): every cell of `mask` to 1 inside the block, 0 in the ghosts. */
int bcm_setup_mask(bcm_field_t mask);
int bcm_jacobi7_mask(bcm_field_t x1, bcm_field_t x0, const bcm_field_t A[7], bcm_field_t b, bcm_field_t mask,
                     double omega);
int bcm_calc_ax_mask(bcm_field_t ax, const bcm_field_t A[7], bcm_field_t x, bcm_field_t mask);
int bcm_calc_r_mask(bcm_field_t r, const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, bcm_field_t mask);
int bcm_calc_r2_mask(const bcm_field_t A[7], bcm_field_t x, bcm_field_t b, bcm_field_t mask, double* rr);
int bcm_dot_mask(bcm_field_t x, bcm_field_t y, bcm_field_t mask, double* xy);
int bcm_axpy_mask(bcm_field_t y, bcm_field_t x, double a, bcm_field_t mask);
int bcm_xpay_mask(bcm_field_t y, bcm_field_t x, double a, bcm_field_t mask);
int bcm_axpyz_mask(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a, bcm_field_t mask);
int bcm_axpbypz_mask(bcm_field_t z, bcm_field_t x, bcm_field_t y, double a, double b, bcm_field_t mask);
int bcm_copy_mask(bcm_field_t y, bcm_field_t x, bcm_field_t mask);

/* Poisson operator A = lap_h (reference sign, bcut.f90:2848-2865 with no
 * cut cells: Aw..At = 1/h^2, Ap = -(l0+l1)-(l2+l3)-(l4+l5)) with the
 * outer-face folds in X_M..Z_P order (BlockScalar3D.h:207-214):
 * Dirichlet bca.f90:12-202 (Ap -= Anb; b -= Anb v 2; Anb = 0), Neumann
 * bca.f90:204-394 (Ap += Anb; b += Anb xc on a minus face, b -= Anb xc on
 * a plus face, xc = v * cellSize.x; Anb = 0).  Writes A[7] (nullable) and
 * b = rhs folded. */
int bcm_poisson_assemble(const bcm_field_t A[7], bcm_field_t b, bcm_field_t rhs,
                         const double* bvalues, int nbvalues);

/* Matrix-free forms of jacobi_smoother_, calc_ax_ and calc_r_ for the
 * Poisson operator above: the coefficients are recomputed per cell in the
 * same operand order as the assembled arrays, so results are bit-identical
 * to the stored-coefficient kernels on bcm_poisson_assemble's output while
 * reading 3 arrays instead of 10 (b must be the folded right-hand side). */
int bcm_jacobi7_poisson(bcm_field_t x1, bcm_field_t x0, bcm_field_t b, double omega);
/* nsweeps steps of ILS3D::Jacobi_Smoother on the Poisson operator
 * (ILS3D.h:65-126 without the boundary step: the operator carries the
 * folded conditions): x <- jacobi_smoother(x) with b, then the ghost
 * exchange (updateVC), in one call.  The result equals that many
 * bcm_jacobi7_poisson + bcm_ghost_fill(BCM_GF_ALL) pairs bit for bit on
 * every cell they write; cells the exchange never writes (edges, corners,
 * faces without a neighbour) keep x's values.  A mesh of vc = 1, n = 8 or
 * 16, whose blocks can all be resident on the GPU at once runs it as one
 * launch with every block in LDS.  A mesh of vc = 1, n = 16 or 32, each of
 * whose faces is connected on every subface or not at all, runs one launch
 * per step on compact copies of x and b, building the ghosts each step
 * needs from the previous step's interiors (nsweeps >= 2; the reads of
 * unconnected faces' ghosts come from x, unchanged until the last step).
 * Any other mesh runs two launches per step.  x and b FP64. */
int bcm_jacobi_poisson_iterate(bcm_field_t x, bcm_field_t b, double omega, int nsweeps);
int bcm_calc_ax_poisson(bcm_field_t ax, bcm_field_t x);
int bcm_calc_r_poisson(bcm_field_t r, bcm_field_t x, bcm_field_t b);

/* ------------------------------------------------------------------ */
/* Solvers                                                               */
/* ------------------------------------------------------------------ */
typedef struct bcm_krylov_params {
  double omega;               /* Jacobi weight (ILS3D Jacobi_Smoother) */
  int32_t count_pre;          /* Jacobi preconditioner sweeps */
  int32_t count_max;          /* iteration limit */
  double epsilon;             /* relative-residual tolerance ||r||/||b|| */
} bcm_krylov_params;

/* ILS3D::Jacobi (ILS3D.h:478-529) on device: sweeps with ghost fill +
 * Dirichlet mirror between sweeps, FP64 residual each sweep.  x0 is work. */
int bcm_jacobi_solve(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, bcm_field_t x0,
                     const bcm_krylov_params* p, const double* bvalues, int nbvalues,
                     int* count, double* residual);
/* ILS3D::BiCGSTAB (ILS3D.h:663-844) on device, Jacobi-preconditioned,
 * in the precision of x (all fields must share it).  work[9] =
 * {r, r0, p, p_, q_, s, s_, t_, x0}. */
int bcm_bicgstab(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, const bcm_field_t work[9],
                 const bcm_krylov_params* p, const double* bvalues, int nbvalues,
                 int* count, double* residual);
/* ILS3D::Jacobi_Mask (ILS3D.h:846-899): the sweeps are masked, the
 * residual norm is not.  ILS3D::BiCGSTAB_Mask (ILS3D.h:901-1097): the dot
 * products, the preconditioner sweeps and A p_, A s_ are masked; the
 * initial residual and the vector updates are not. */
int bcm_jacobi_solve_mask(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, bcm_field_t x0, bcm_field_t mask,
                          const bcm_krylov_params* p, const double* bvalues, int nbvalues,
                          int* count, double* residual);
int bcm_bicgstab_mask(bcm_field_t x, const bcm_field_t A[7], bcm_field_t b, const bcm_field_t work[9],
                      bcm_field_t mask, const bcm_krylov_params* p, const double* bvalues, int nbvalues,
                      int* count, double* residual);

/* Mixed precision: FP64 outer iterative refinement around an FP32
 * Jacobi-preconditioned BiCGSTAB.  Only the FP64 residual r = b - A u,
 * the FP64 update u += d and the FP64 norms run in FP64.  x, b, rhs in
 * FP64; A is the Poisson operator assembled internally in both
 * precisions.  Reports outer/inner iteration counts and the final FP64
 * relative residual ||b - A x|| / ||b||. */
typedef struct bcm_ir_params {
  double tol;                 /* FP64 relative residual target (e.g. 1e-10) */
  double inner_tol;           /* FP32 inner relative tolerance (e.g. 1e-4) */
  int32_t inner_max;          /* inner BiCGSTAB iteration limit */
  int32_t outer_max;          /* refinement step limit */
  int32_t count_pre;          /* Jacobi preconditioner sweeps */
  double omega;               /* Jacobi weight */
  int32_t fp32_only;          /* 1: no refinement, plain FP32 solve to tol */
  int32_t fp32_strict;        /* 1: stop (BCM_ERR_NOCONV, best x so far) when an FP32
                                 correction does not lower the FP64 residual;
                                 0: solve the remaining corrections in FP64 */
} bcm_ir_params;
int bcm_solve_mixed(bcm_field_t x, bcm_field_t rhs, const double* bvalues, int nbvalues,
                    const bcm_ir_params* p, int* outer, int* inner_total, double* relres);
/* Plain FP64 BiCGSTAB on the assembled Poisson operator (the per-mesh
 * FP64 GPU path the mixed solver is compared against). */
int bcm_solve_fp64(bcm_field_t x, bcm_field_t rhs, const double* bvalues, int nbvalues,
                   const bcm_ir_params* p, int* iters, double* relres);

/* ------------------------------------------------------------------ */
/* Batched meshes and the fused device-resident BiCGSTAB                */
/* ------------------------------------------------------------------ */
/* Several independent meshes in one device mesh: the block tables are
 * concatenated (segment s holds counts[s] blocks whose neighbour ids are
 * local to the segment; they are shifted here), so every kernel covers all
 * meshes in one launch while reductions and solver scalars stay per
 * segment.  The reference runs one BlockManager per process
 * (include/BlockManager.h:153-156); each segment is one such mesh.
 * Boundary values index one table shared by all segments. */
int bcm_mesh_create_batch(const bcm_block_info* blocks, const int* counts, int nseg, int n, int vc,
                          bcm_mesh_t* out);
int bcm_mesh_num_segments(bcm_mesh_t mesh);
int bcm_mesh_segment(bcm_mesh_t mesh, int seg, int* first_block, int* nblocks);

/* Ghost-fill method bit for bcm_ghost_fill, with the complete mask only
 * (BCM_GF_ALL, all axes): one work-group per destination block gathers
 * its six faces from the neighbours' interiors.  Same values as the
 * per-face tasks of updateVC (include/Scalar3DUpdater.h:190-266, 321-347,
 * 519-708); n in {4, 8, 16, 32}, vc = 1. */
#define BCM_GF_BLOCKWISE 64

/* ILS3D::BiCGSTAB (examples/SolverTest/ILS3D.h:663-844) with the Jacobi
 * preconditioner (ILS3D.h:452-475) on the folded Poisson operator
 * (bcut.f90:2848-2866, bca.f90:12-202), run entirely on the device: all
 * scalars stay in per-segment device records, the dot products are fused
 * into the kernels that produce their operands and blocks of
 * check_every iterations replay as one hipGraph.  Every segment of the
 * mesh is solved at once; a segment leaves the iteration when its true
 * FP64 relative residual ||b - A x|| / ||b|| reaches its tolerance.
 *   FP64:  the iteration in FP64.
 *   MIXED: FP32 vectors (r, r0, p, q, s, t), FP32 Jacobi sweeps and
 *          vector updates; FP64 for the two SpMV inputs (the
 *          preconditioned p_ and s_), the SpMV arithmetic, every dot
 *          product and the solution correction dx.  FP64 updates
 *          x += dx, r = b - A x whenever the recurrence residual has
 *          dropped by refresh_delta, every refresh_max iterations and when
 *          the recurrence reports convergence (iterative refinement; with
 *          restart_every = 1 each update also restarts the FP32 Krylov
 *          space).
 *   FP32:  as MIXED with FP64 updates only on reported convergence.
 * x: FP64 field holding the initial guess, overwritten with the solution
 * (ghosts filled); rhs: FP64 field whose interior is the right-hand side
 * of A x = rhs before the Dirichlet fold (A = lap_h as assembled by
 * bcm_poisson_assemble).  n in {8, 16, 32}, vc = 1.
 * seg_tol (nullable): per-segment tolerances overriding p->tol.
 * stats[nseg]: outcome per segment. */
enum { BCM_PREC_FP64 = 0, BCM_PREC_MIXED = 1, BCM_PREC_FP32 = 2 };
enum { BCM_SOLVE_ACTIVE = 0, BCM_SOLVE_CONVERGED = 1, BCM_SOLVE_NOCONV = 2, BCM_SOLVE_FAILED = 3 };
typedef struct bcm_solve_params {
  int32_t precision;          /* BCM_PREC_* */
  double tol;                 /* true FP64 relative residual target */
  int32_t max_iter;           /* BiCGSTAB iterations per segment */
  int32_t count_pre;          /* Jacobi preconditioner sweeps (0..16) */
  double omega;               /* Jacobi weight */
  int32_t refresh_max;        /* MIXED: reliable update at least this often (0: 1000) */
  double refresh_delta;       /* MIXED: update when ||r|| fell by this factor (0: 1e-3) */
  int32_t check_every;        /* iterations per replay between host checks (0: 8) */
  int32_t use_graph;          /* 1: capture the iteration block in a hipGraph */
  int32_t restart_every;      /* MIXED/FP32: also restart the Krylov space (r0 = r, p = r)
                                 at every k-th reliable update; 0: every update (k = 1);
                                 < 0: only on breakdown or a failed convergence check */
  int32_t stall_window;       /* MIXED/FP32: FP64 update when the recurrence residual set no
                                 new minimum for this many iterations (0: 40, < 0: off) */
  int32_t stagnation;         /* end a segment as BCM_SOLVE_NOCONV when this many consecutive
                                 true-residual checks did not halve the best ||b - A x||:
                                 the tolerance is below what the arithmetic attains
                                 (0: 8 for FP64, 30 for MIXED/FP32; < 0: off, run to
                                 max_iter) */
} bcm_solve_params;
typedef struct bcm_solve_stats {
  int32_t status;             /* BCM_SOLVE_CONVERGED / NOCONV / FAILED */
  int32_t iters;              /* BiCGSTAB iterations */
  int32_t refreshes;          /* reliable updates after the initial residual */
  int32_t pad;
  double relres;              /* true FP64 relative residual at the last update */
} bcm_solve_stats;
int bcm_solve(bcm_field_t x, bcm_field_t rhs, const double* bvalues, int nbvalues,
              const bcm_solve_params* p, const double* seg_tol, bcm_solve_stats* stats);
/* bcm_solve prints a progress line to stderr at most every `seconds`
 * (segments done, iterations) while it runs; 0 (default) keeps it quiet. */
int bcm_mesh_set_progress(bcm_mesh_t mesh, double seconds);

/* ------------------------------------------------------------------ */
/* Request server (continuous batching of independent Poisson solves)  */
/* ------------------------------------------------------------------ */
/* No counterpart in the reference (one BlockManager per process,
 * include/BlockManager.h:153-156).  Requests (a block table, a right-hand
 * side, a tolerance and a precision) queue in arrival order and are
 * admitted first-fit while their device bytes fit in hbm_budget and fewer
 * than max_active groups run.  A group owns a device mesh, a HIP stream and
 * a fused-solver state (bcm_solve's BiCGSTAB, hipGraph replayed), and
 * bcm_server_progress keeps every group's stream fed without blocking, so
 * groups overlap on the device.  batch_blocks = 0: every request is a group
 * of its own.  batch_blocks > 0: a request of at most batch_blocks blocks
 * takes, at admission, the later queued requests of the same precision
 * that fit with it (arrival order, group total <= batch_blocks blocks, HBM
 * budget) into one batched mesh (bcm_mesh_create_batch), one segment per
 * request; the group's results are reported when its last segment
 * finishes.  Memory is released as soon as a group finishes.  A request's
 * result does not depend on the other requests or on its group.
 * Single-threaded: one host thread calls every bcm_server_* function of a
 * server. */
typedef struct bcm_server* bcm_server_t;
typedef struct bcm_server_params {
  double hbm_budget;          /* device bytes all admitted requests may hold */
  int32_t max_active;         /* requests on the device at once */
  int32_t n;                  /* cells per block edge of every request: 8, 16 or 32 */
  int32_t max_iter;           /* BiCGSTAB iteration limit per request */
  int32_t count_pre;          /* Jacobi preconditioner sweeps */
  double omega;               /* Jacobi weight */
  int32_t check_every;        /* iterations per graph replay (0: 8) */
  int32_t batch_blocks;       /* merge small requests into groups of at most this many blocks (0: off) */
  /* Launch order of the groups' blocks of iterations.  window_blocks = 0:
   * every group's next block as soon as its previous one is done.
   * window_blocks > 0: smallest groups first -- a ready group launches
   * while the blocks of the groups with work in flight, its own included,
   * stay within window_blocks, or when nothing is in flight, or once it
   * has waited max_wait seconds (max_wait <= 0: no such limit); small
   * requests then do not queue behind the kernels of large ones. */
  int32_t window_blocks;
  double max_wait;
} bcm_server_params;
typedef struct bcm_server_result {
  int64_t tag;
  int32_t status;             /* BCM_SOLVE_* or -(BCM_ERR_*) (-BCM_ERR_BUSY: exceeds the budget) */
  int32_t iters;
  double relres;              /* true FP64 relative residual */
  double t_submit, t_admit, t_done;  /* seconds on the server clock */
  int32_t nblocks;
  int32_t precision;
} bcm_server_result;
int bcm_server_create(const bcm_server_params* params, bcm_server_t* out);
int bcm_server_destroy(bcm_server_t server);
/* Seconds since the server was created. */
double bcm_server_clock(bcm_server_t server);
/* Queue a request.  tag: the caller's name for it, unique among the
 * requests queued or running (a tag already pending returns BCM_ERR_ARG;
 * it may be reused once its result has been produced).  blocks: its
 * block table (copied); rhs: host FP64
 * ghosted field (nblocks*(n+2)^3 values, read when the request is
 * admitted, so it must stay valid until then) or NULL for the synthetic
 * right-hand side sin(pi x) sin(pi y) sin(pi z) + 0.01 N(0,1) generated on
 * the device from `seed`; x_out: host buffer of the same size for the
 * solution (NULL: not returned); all outer faces Dirichlet u = 0. */
int bcm_server_submit(bcm_server_t server, const bcm_block_info* blocks, int nblocks, const double* rhs,
                      uint64_t seed, double tol, int precision, double* x_out, int64_t tag);
/* Checkpoint of a running request in SolverTest's restart format
 * (LocalScalar3D::Dump2, examples/SolverTest/LocalScalar3D.cpp:227-281):
 * waits for the iterations already queued for the request's group, then
 * writes the request's current iterate -- x plus the pending FP64
 * correction on the interior, the ghost cells of connected faces gathered
 * from it as the end of a solve leaves them (Dump2 imposes the boundary
 * conditions before it writes) -- to
 * dir/dump-<label>-<rank %05d>-<step %010d>.bin, header int32
 * {n, n, n, 1, 1, nblocks}, the request's blocks in its block-ID order.
 * *iters (may be NULL): iterations done so far.  The request keeps
 * running.  A queued request has no iterate yet: BCM_ERR_STATE; no
 * pending request with this tag: BCM_ERR_ARG. */
int bcm_server_checkpoint(bcm_server_t server, int64_t tag, const char* dir, const char* label, int rank, int step,
                          int* iters);
/* bcm_server_submit starting from an initial guess read, when the request
 * is admitted, from the restart file dir/dump-<label>-<rank>-<step>.bin
 * (written by bcm_server_checkpoint or bcm_field_dump2; Load2's rules,
 * LocalScalar3D.cpp:283-357, a file of half-edge blocks prolonged by
 * sf3d_copy_x2).  Resuming a checkpointed request on a new server
 * continues its solve from the checkpointed iterate with a fresh Krylov
 * space.  The header is checked at submission: BCM_ERR_IO for a missing
 * or short file, BCM_ERR_ARG for another block edge, ghost width or block
 * count. */
int bcm_server_submit_restart(bcm_server_t server, const bcm_block_info* blocks, int nblocks, const double* rhs,
                              uint64_t seed, double tol, int precision, const char* dir, const char* label,
                              int rank, int step, double* x_out, int64_t tag);
/* Drive the admitted requests for up to max_seconds (returns early when
 * nothing is left); finished requests move to the result list. */
int bcm_server_progress(bcm_server_t server, double max_seconds, int* n_active, int* n_queued);
/* Move up to capacity finished results into out; returns how many, or
 * -BCM_ERR_ARG for a NULL server, a negative capacity, or a NULL out with
 * capacity > 0. */
int bcm_server_poll(bcm_server_t server, bcm_server_result* out, int capacity);
int bcm_server_stats(bcm_server_t server, double* bytes_in_use, int* n_active, int* n_queued);

#ifdef __cplusplus
}
#endif

#endif /* BCM_HIP_H */
