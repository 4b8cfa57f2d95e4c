"""Restatement of SolverTest's restart files (numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's cpu_baseline leg as
the checker of the device dump/load entry points; the product never
imports it.

LocalScalar3D<real>::Dump2 / Load2 (examples/SolverTest/LocalScalar3D.cpp:
226-324): one file per rank, named dump-<label>-<rank %05d>-<step %010d>.bin,
holding int32 {nx, ny, nz, vc, ne = 1, nb} and then, for every block in
block-ID order, its whole ghosted (nx+2vc)(ny+2vc)(nz+2vc) array of `real`.
Dump / Load (:84-222): one file per block, dump-<label>-<rank>-<id %05d>-
<step>.bin, int32 {nx, ny, nz, vc, ne = 1} and that block's array.
A load of a file whose blocks have half the edge applies sf3d_copy_x2
(examples/SolverTest/bsf3d.f90:18-46): fine interior cell (i, j, k), 1-based,
takes coarse cell ((i+1)/2, (j+1)/2, (k+1)/2); fine ghosts are not written.
"""
import os

import numpy as np


def name2(directory, label, rank, step):
    return os.path.join(directory, f"dump-{label}-{rank:05d}-{step:010d}.bin")


def name1(directory, label, rank, block_id, step):
    return os.path.join(directory, f"dump-{label}-{rank:05d}-{block_id:05d}-{step:010d}.bin")


def dump2(data, n, vc, path):
    """data: (nb, n+2vc, n+2vc, n+2vc) array of the field type."""
    with open(path, "wb") as fp:
        fp.write(np.array([n, n, n, vc, 1, data.shape[0]], dtype=np.int32).tobytes())
        fp.write(np.ascontiguousarray(data).tobytes())


def dump1(data, n, vc, directory, label, rank, step):
    for b in range(data.shape[0]):
        with open(name1(directory, label, rank, b, step), "wb") as fp:
            fp.write(np.array([n, n, n, vc, 1], dtype=np.int32).tobytes())
            fp.write(np.ascontiguousarray(data[b]).tobytes())


def copy_x2(fine, coarse, n, vc):
    """sf3d_copy_x2 on one block, written with 1-based loops as the Fortran."""
    h = n // 2
    for k in range(1, n + 1):
        for j in range(1, n + 1):
            for i in range(1, n + 1):
                fine[k - 1 + vc, j - 1 + vc, i - 1 + vc] = coarse[(k + 1) // 2 - 1 + vc, (j + 1) // 2 - 1 + vc,
                                                                  (i + 1) // 2 - 1 + vc]
    assert coarse.shape[-1] == h + 2 * vc


def load2(into, n, vc, path):
    """Returns the updated copy of `into` (the field before the load)."""
    out = into.copy()
    with open(path, "rb") as fp:
        hdr = np.frombuffer(fp.read(24), dtype=np.int32)
        nx, nv, ne, nb = int(hdr[0]), int(hdr[3]), int(hdr[4]), int(hdr[5])
        if nv != vc or ne != 1 or nb != into.shape[0] or not (nx == n or 2 * nx == n):
            raise ValueError("header does not match the field")
        c = nx + 2 * nv
        body = np.frombuffer(fp.read(), dtype=into.dtype)
    body = body.reshape(nb, c, c, c)
    if nx == n:
        out[...] = body
    else:
        for b in range(nb):
            copy_x2(out[b], body[b], n, vc)
    return out
