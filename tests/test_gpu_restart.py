"""Restart files written from and read into device memory.

bcm_field_dump2 / dump produce the same bytes as the restatement of
LocalScalar3D::Dump2 / Dump (oracle/restart.py); load2 / load of files the
restatement wrote reproduce the data bit for bit, including the half-edge
load through sf3d_copy_x2; foreign headers and missing files are errors.
"""
import os
import sys

import numpy as np
import pytest

from bcmtools_amd import Mesh
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W
from oracle import restart as R

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dump2_bytes_and_load2_round_trip(tmp_path, dtype):
    t = W.shell_tree(3)
    n, vc = 16, 1
    case(f"dump2/load2 {np.dtype(dtype).name} nb={len(t.blocks)}")
    rng = np.random.default_rng(2)
    data = rng.standard_normal((len(t.blocks), n + 2, n + 2, n + 2)).astype(dtype)
    mesh = Mesh(t.blocks, n, vc)
    f = mesh.field(dtype, data)
    d = str(tmp_path / "BIN")
    f.dump2(7, "p", rank=0, directory=d)
    got = open(R.name2(d, "p", 0, 7), "rb").read()
    ref_path = str(tmp_path / "ref.bin")
    R.dump2(data, n, vc, ref_path)
    assert got == open(ref_path, "rb").read()
    g = mesh.field(dtype)
    g.load2(7, "p", rank=0, directory=d)
    np.testing.assert_array_equal(g.download(), data)


def test_load2_half_edge_file(tmp_path):
    t = W.shell_tree(2)
    n, vc = 16, 1
    nb = len(t.blocks)
    case(f"load2 half edge nb={nb}")
    rng = np.random.default_rng(3)
    coarse = rng.standard_normal((nb, n // 2 + 2, n // 2 + 2, n // 2 + 2))
    d = tmp_path / "BIN"
    d.mkdir()
    R.dump2(coarse, n // 2, vc, R.name2(str(d), "t", 0, 1))
    before = rng.standard_normal((nb, n + 2, n + 2, n + 2))
    mesh = Mesh(t.blocks, n, vc)
    f = mesh.field(np.float64, before)
    f.load2(1, "t", directory=str(d))
    ref = R.load2(before, n, vc, R.name2(str(d), "t", 0, 1))
    np.testing.assert_array_equal(f.download(), ref)


def test_dump_per_block_files(tmp_path):
    t = W.Tree(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=1)
    n, vc = 8, 2
    case("dump/load per block")
    rng = np.random.default_rng(4)
    data = rng.standard_normal((len(t.blocks), n + 4, n + 4, n + 4))
    mesh = Mesh(t.blocks, n, vc)
    f = mesh.field(np.float64, data)
    d = str(tmp_path / "BIN")
    f.dump(3, "ux", directory=d)
    rd = tmp_path / "ref"
    rd.mkdir()
    R.dump1(data, n, vc, str(rd), "ux", 0, 3)
    for b in range(len(t.blocks)):
        a = open(R.name1(d, "ux", 0, b, 3), "rb").read()
        assert a == open(R.name1(str(rd), "ux", 0, b, 3), "rb").read()
    g = mesh.field(np.float64)
    g.load(3, "ux", directory=str(rd))
    np.testing.assert_array_equal(g.download(), data)


def test_restart_errors(tmp_path):
    t = W.Tree(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=1)
    case("restart errors")
    mesh = Mesh(t.blocks, 8, 1)
    f = mesh.field(np.float64)
    d = tmp_path / "BIN"
    d.mkdir()
    with pytest.raises(L.BCMError) as e:
        f.load2(1, "none", directory=str(d))
    assert e.value.code == L.ERR_IO
    R.dump2(np.zeros((len(t.blocks) + 1, 10, 10, 10)), 8, 1, R.name2(str(d), "x", 0, 1))
    with pytest.raises(L.BCMError) as e:
        f.load2(1, "x", directory=str(d))
    assert e.value.code == L.ERR_ARG
