"""Synthetic meshes and right-hand sides of the benchmark configurations.

Level convention: levels count from the unit cube.  A root grid of
2x2x2 roots of length 0.5 therefore starts at level 1, and "refined to
level L" means octree depth L - 1 inside each root.

Right-hand side: f = sin(pi x) sin(pi y) sin(pi z) at cell centres
origin + (i + 0.5) h (the cell-centre convention of
examples/sor2/Solver.cpp:393-400) plus sigma * N(0, 1) noise with
sigma = 0.01.  The noise comes from a counter-based generator keyed by
(seed, block pedigree, cell index), so a cell's value does not depend on
block ordering.  Operator sign: the reference's A = lap_h
(examples/SolverTest/bcut.f90:2857-2866) with b = -f, i.e. -lap(u) = f.
"""
import math

import numpy as np

from . import _lib as L
from .core import Tree

SIGMA = 0.01
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _uniform(h):
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal_noise(seed, pedigrees, n):
    """Standard normal per (block, interior cell), shape (nb, n, n, n)."""
    with np.errstate(over="ignore"):
        ped = np.asarray(pedigrees, dtype=np.uint64)
        cell = np.arange(n * n * n, dtype=np.uint64)
        key = _splitmix(np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + _splitmix(ped))[:, None] + cell[None, :]
        h1 = _splitmix(key)
        h2 = _splitmix(h1 ^ np.uint64(0x632BE59BD9B4E019))
    u1 = 1.0 - _uniform(h1)  # (0, 1]
    u2 = _uniform(h2)
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
    return z.reshape(len(ped), n, n, n)


def rhs_field(blocks, n, vc, seed=0, sigma=SIGMA, freq=1.0, dtype=np.float64, chunk=2048):
    """Ghosted field with f = sin(a x) sin(a y) sin(a z) + sigma N(0,1) inside, a = freq*pi; zero ghosts."""
    nb = len(blocks)
    nx0 = n + 2 * vc
    out = np.zeros((nb, nx0, nx0, nx0), dtype=dtype)
    a = freq * math.pi
    t = (np.arange(n) + 0.5)
    for b0 in range(0, nb, chunk):
        blk = blocks[b0:b0 + chunk]
        org = blk["origin"]
        h = blk["block_size"] / n
        x = org[:, 0, None] + t[None, :] * h[:, 0, None]
        y = org[:, 1, None] + t[None, :] * h[:, 1, None]
        z = org[:, 2, None] + t[None, :] * h[:, 2, None]
        sx, sy, sz = np.sin(a * x), np.sin(a * y), np.sin(a * z)
        f = sz[:, :, None, None] * sy[:, None, :, None] * sx[:, None, None, :]
        if sigma:
            f = f + sigma * normal_noise(seed, blk["pedigree"], n)
        out[b0:b0 + chunk, vc:vc + n, vc:vc + n, vc:vc + n] = f
    return out


def rhs_field_fast(blocks, n, vc, seed=0, sigma=SIGMA, dtype=np.float64):
    """As rhs_field with the noise drawn from numpy's PCG64 stream seeded per
    mesh (block-order dependent, ~10x faster): the batched and served
    workloads generate millions of blocks."""
    nb = len(blocks)
    nx0 = n + 2 * vc
    out = np.zeros((nb, nx0, nx0, nx0), dtype=dtype)
    t = (np.arange(n) + 0.5)
    org = blocks["origin"]
    h = blocks["block_size"] / n
    x = org[:, 0, None] + t[None, :] * h[:, 0, None]
    y = org[:, 1, None] + t[None, :] * h[:, 1, None]
    z = org[:, 2, None] + t[None, :] * h[:, 2, None]
    sx, sy, sz = np.sin(math.pi * x), np.sin(math.pi * y), np.sin(math.pi * z)
    f = sz[:, :, None, None] * sy[:, None, :, None] * sx[:, None, None, :]
    if sigma:
        f += sigma * np.random.default_rng(seed).standard_normal(f.shape)
    out[:, vc:vc + n, vc:vc + n, vc:vc + n] = f
    return out


def config1_tree():
    """Uniform level-2 octree on the unit cube: 4x4x4 = 64 blocks, Z order."""
    return Tree(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=2, ordering=L.ORDER_Z)


def shell_tree(level, r_in=0.2, r_out=0.3, center=(0.5, 0.5, 0.5), margin=0.0, ordering=L.ORDER_Z):
    """RootGrid(2,2,2) of length 0.5, refined where a block meets the shell
    r_in <= |x - c| <= r_out, up to `level` (counted from the unit cube);
    the last split only where the node centre lies within `margin` node
    edges of the shell.  2:1 balanced, all outer faces Dirichlet."""
    return Tree(root_n=(2, 2, 2), root_length=0.5, divider=L.DIV_SHELL, min_level=0, max_level=level - 1,
                shells=[tuple(center) + (r_in, r_out)], ordering=ordering, balance=1, shell_mode=1,
                shell_margin=margin)


def config2_tree(level=5):
    """Shell 0.2 <= r <= 0.3 refined to level 5 with 32^3 blocks: 3,816 blocks (125 M cells)."""
    return shell_tree(level)


def config3_tree(level=7):
    """Mid-surface |x - c| = 0.25 refined to level 7 with 16^3 blocks: 26,888 blocks (110 M cells)."""
    return shell_tree(level, r_in=0.25, r_out=0.25, margin=0.25)


def random_tree(seed, max_level=None, nspheres=None):
    """One mesh of the batched workload: max level uniform in {2..6}, 1-3
    spheres of radius U(0.05, 0.3) refined at their surfaces, 2:1 balanced."""
    rng = np.random.default_rng(seed)
    if max_level is None:
        max_level = int(rng.integers(2, 7))
    if nspheres is None:
        nspheres = int(rng.integers(1, 4))
    shells = []
    for _ in range(nspheres):
        c = rng.uniform(0.2, 0.8, size=3)
        r = float(rng.uniform(0.05, 0.3))
        shells.append((c[0], c[1], c[2], r, r))
    return Tree(root_n=(2, 2, 2), root_length=0.5, divider=L.DIV_SHELL, min_level=0, max_level=max_level - 1,
                shells=shells, ordering=L.ORDER_Z, balance=1, shell_mode=1, shell_margin=0.25)


def serving_tree(seed, target_blocks):
    """A request mesh of the served workload: 1-3 random sphere surfaces
    (as random_tree), refined to the max level (2..8) whose block count is
    closest to target_blocks in ratio."""
    rng = np.random.default_rng(seed)
    nspheres = int(rng.integers(1, 4))
    shells = []
    for _ in range(nspheres):
        c = rng.uniform(0.2, 0.8, size=3)
        r = float(rng.uniform(0.05, 0.3))
        shells.append((c[0], c[1], c[2], r, r))
    best = None
    for level in range(2, 9):
        t = Tree(root_n=(2, 2, 2), root_length=0.5, divider=L.DIV_SHELL, min_level=0, max_level=level - 1,
                 shells=shells, ordering=L.ORDER_Z, balance=1, shell_mode=1, shell_margin=0.25)
        if best is None or abs(math.log(len(t.blocks) / target_blocks)) < abs(math.log(len(best.blocks) / target_blocks)):
            best = t
        if len(t.blocks) >= target_blocks:
            break
    return best


def interior(a, vc):
    """View of the interior cells of a ghosted field array."""
    e = a.shape[-1] - vc
    return a[..., vc:e, vc:e, vc:e]


def face_ghost_mask(n, vc):
    """Boolean (nx0,)*3 mask of interior cells plus face ghosts (no edges/corners)."""
    nx0 = n + 2 * vc
    idx = np.arange(nx0)
    inside = (idx >= vc) & (idx < vc + n)
    ix, iy, iz = np.meshgrid(inside, inside, inside, indexing="ij")
    cnt = ix.astype(int) + iy.astype(int) + iz.astype(int)
    return cnt >= 2
