"""MI355X-native BCM (Building-Cube Method) block data path.

Leaf blocks of an octree mesh live in HBM; hand-written HIP kernels do the
ghost fill (same-level copy, fine-to-coarse restriction, coarse-to-fine
prolongation), red-black SOR, 7-point Jacobi and the BiCGSTAB building
blocks.  The kernels sit behind the C ABI of include/bcm_hip.h
(libbcm_hip.so); this package is the host-side mirror of the reference's
BlockManager / Scalar3D / Scalar3DUpdater API over that ABI.
"""
from . import _lib
from ._lib import BCMError, BLOCK_DTYPE
from .core import CommList, CommPlan, Field, Mesh, Server, Tree, partition

__all__ = ["BCMError", "BLOCK_DTYPE", "CommList", "CommPlan", "Field", "Mesh", "Tree", "build", "partition"]


def build(quiet=True):
    """Compile libbcm_hip.so for gfx950 in-tree (make in bcmtools_amd/csrc)."""
    import os
    import subprocess
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
    subprocess.run(["make", "-j8"] + (["-s"] if quiet else []) + ["-C", here], check=True)
