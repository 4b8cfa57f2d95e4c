"""Host octree builder vs the reference's own block tables (tests/golden)."""
import numpy as np
import pytest

import bcmtools_amd as B
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W

from conftest import load_golden

DIV = {"flat": L.DIV_FLAT, "simple": L.DIV_SIMPLE}
ORD = {"Z": L.ORDER_Z, "Hilbert": L.ORDER_HILBERT}
BCS = {"dir": L.BCSET_ALL_DIRICHLET, "x": L.BCSET_SOR1_X, "y": L.BCSET_SOR1_Y, "z": L.BCSET_SOR1_Z}

MESH_FIXTURES = ["mesh_flat2_z", "mesh_simple13_z", "mesh_simple03_hilbert_x", "mesh_flat1_hilbert_roots_y",
                 "mesh_simple24_hilbert_z"]


def tree_from_args(args):
    a = str(args).split()
    div, lmin, lmax, order = a[0], int(a[1]), int(a[2]), a[3]
    rn = tuple(int(v) for v in a[4:7])
    bc = a[7]
    per = (0, 0, 0)
    if bc in ("x", "y", "z"):
        ax = "xyz".index(bc)
        per = tuple(0 if i == ax else 1 for i in range(3))
    return B.Tree(root_n=rn, periodic=per, divider=DIV[div], min_level=lmin, max_level=lmax, ordering=ORD[order],
                  balance=0, bc_setter=BCS[bc])


@pytest.mark.parametrize("name", MESH_FIXTURES)
def test_block_table_matches_reference(name):
    g = load_golden(name)
    t = tree_from_args(g["args"])
    ref = g["blocks"]
    assert len(t.blocks) == len(ref)
    for field in ref.dtype.names:
        np.testing.assert_array_equal(t.blocks[field], ref[field], err_msg=field)


def test_ghost_fixture_meshes_match_reference_for_simple_divider():
    g = load_golden("ghost_periodic_x_d")
    t = tree_from_args(g["args"])
    assert t.blocks.tobytes() == g["blocks"].tobytes()


def _neighbors_consistent(blocks):
    """Every face relation has its mirror image on the neighbour."""
    rev = [1, 0, 3, 2, 5, 4]
    for b, bi in enumerate(blocks):
        for f in range(6):
            ld = bi["nb_level_diff"][f]
            ids = bi["nb_id"][f]
            if ld == 0 and ids[0] >= 0:
                nb = blocks[ids[0]]
                assert nb["nb_level_diff"][rev[f]] == 0 and nb["nb_id"][rev[f]][0] == b
            elif ld == 1:
                for s in range(4):
                    if ids[s] < 0:
                        continue
                    nb = blocks[ids[s]]
                    assert nb["nb_level_diff"][rev[f]] == -1 and nb["nb_id"][rev[f]][0] == b
                    assert nb["nb_subface"][rev[f]] == s
                    assert nb["level"] == bi["level"] + 1
            elif ld == -1:
                nb = blocks[ids[0]]
                assert nb["level"] == bi["level"] - 1
                assert nb["nb_level_diff"][rev[f]] == 1
                assert nb["nb_id"][rev[f]][bi["nb_subface"][f]] == b


@pytest.mark.parametrize("maker", [lambda: W.config2_tree(4), lambda: W.shell_tree(5, 0.25, 0.25, margin=0.25),
                                   lambda: W.random_tree(3), lambda: W.random_tree(11)])
def test_shell_trees_are_balanced_and_symmetric(maker):
    t = maker()
    blocks = t.blocks
    # 2:1 balance: every face is a neighbour at -1/0/+1 or an outer boundary.
    for bi in blocks:
        for f in range(6):
            has = bi["nb_id"][f][0] >= 0 or (bi["nb_level_diff"][f] == 1 and (bi["nb_id"][f] >= 0).any())
            assert has or bi["bc_type"][f] == L.BC_DIRICHLET
    _neighbors_consistent(blocks)
    # Leaves tile the domain exactly.
    vol = np.prod(blocks["block_size"], axis=1).sum()
    assert vol == pytest.approx(1.0, rel=1e-12)


def test_config_block_counts():
    assert len(W.config1_tree()) == 64
    assert len(W.config2_tree()) == 3816
    assert len(W.config3_tree()) == 26888


def test_unbalanced_tree_without_balancer_is_reported():
    # A deep refinement next to a coarse root breaks the 2:1 relation; without
    # the balancer the builder reports it as the reference aborts
    # (Utils/src/BCMOctree.cpp:462-467); with the balancer the mesh is valid.
    kw = dict(root_n=(2, 1, 1), root_length=0.5, divider=L.DIV_SHELL, max_level=3,
              shells=[(0.49, 0.25, 0.25, 0.0, 0.01)])
    with pytest.raises(B.BCMError) as e:
        B.Tree(balance=0, **kw)
    assert e.value.code == L.ERR_BALANCE
    t1 = B.Tree(balance=1, **kw)
    _neighbors_consistent(t1.blocks)


@pytest.mark.parametrize("bad", [dict(root_n=(0, 1, 1)), dict(max_level=16), dict(min_level=3, max_level=2),
                                 dict(divider=9), dict(divider=L.DIV_SHELL, shells=()), dict(root_length=0.0)])
def test_tree_rejects_bad_parameters(bad):
    kw = dict(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=1)
    kw.update(bad)
    with pytest.raises(B.BCMError) as e:
        B.Tree(**kw)
    assert e.value.code == L.ERR_ARG


def test_rhs_noise_independent_of_block_order():
    t = W.config1_tree()
    a = W.rhs_field(t.blocks, 4, 1, seed=7)
    perm = np.random.default_rng(0).permutation(len(t.blocks))
    b = W.rhs_field(t.blocks[perm], 4, 1, seed=7)
    np.testing.assert_array_equal(a[perm], b)
    noise = W.normal_noise(1, t.blocks["pedigree"], 8).ravel()
    assert abs(noise.mean()) < 0.02 and abs(noise.std() - 1.0) < 0.02


@pytest.mark.parametrize("name,make,digest,count", [
    ("shell5", lambda: W.shell_tree(5), "3037731f804aa7db", 3816),
    ("shell7", lambda: W.shell_tree(7), "e13c1c73b6e59c91", 181000),
    ("random11", lambda: W.random_tree(11, max_level=6), "76d6b4b034dec832", 1289),
    ("simple_periodic", lambda: B.Tree(root_n=(3, 2, 2), periodic=(1, 0, 1), divider=L.DIV_SIMPLE, min_level=1,
                                     max_level=4, ordering=L.ORDER_HILBERT), "f894ae14512d7e9a", 16336),
])
def test_parallel_tree_build_gives_the_serial_block_table(name, make, digest, count):
    """bcm_tree_build splits the levels, the 2:1 balance checks and the
    block records over host threads.  The digests were recorded from the
    single-threaded depth-first builder (one balance pass over every leaf
    until none fails); every thread count gives the same table byte for
    byte."""
    import hashlib
    lib = L.load()
    try:
        for threads in (1, 3, 8):
            assert lib.bcm_tree_set_threads(threads) == L.OK
            t = make()
            assert len(t.blocks) == count, (name, threads)
            assert hashlib.sha256(t.blocks.tobytes()).hexdigest()[:16] == digest, (name, threads)
    finally:
        lib.bcm_tree_set_threads(0)
    assert lib.bcm_tree_set_threads(-1) == L.ERR_ARG
