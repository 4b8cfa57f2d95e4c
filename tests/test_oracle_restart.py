"""Pin the restart-file restatement (oracle/restart.py).

The reference's Dump2/Load2 write raw memory with a fixed int32 header
(examples/SolverTest/LocalScalar3D.cpp:226-324): the header, the file name
and the round trip are checked directly; sf3d_copy_x2's 1-based index map
is checked against the injection it must equal (every 2x2x2 group of fine
cells holds its coarse parent), written with numpy.repeat.
"""
import os

import numpy as np
import pytest

from oracle import restart as R


def test_names_follow_the_reference():
    assert os.path.basename(R.name2("BIN", "p", 3, 120)) == "dump-p-00003-0000000120.bin"
    assert os.path.basename(R.name1("BIN", "ux", 0, 17, 5)) == "dump-ux-00000-00017-0000000005.bin"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dump2_layout_and_round_trip(tmp_path, dtype):
    n, vc, nb = 4, 1, 3
    rng = np.random.default_rng(1)
    data = rng.standard_normal((nb, n + 2, n + 2, n + 2)).astype(dtype)
    path = str(tmp_path / "f.bin")
    R.dump2(data, n, vc, path)
    raw = open(path, "rb").read()
    assert np.frombuffer(raw[:24], dtype=np.int32).tolist() == [n, n, n, vc, 1, nb]
    assert len(raw) == 24 + data.nbytes
    back = R.load2(np.zeros_like(data), n, vc, path)
    np.testing.assert_array_equal(back, data)


@pytest.mark.parametrize("n,vc", [(8, 1), (8, 2), (4, 1)])
def test_copy_x2_is_the_injection(n, vc):
    rng = np.random.default_rng(n + vc)
    h = n // 2
    coarse = rng.standard_normal((h + 2 * vc,) * 3)
    fine = rng.standard_normal((n + 2 * vc,) * 3)
    before = fine.copy()
    R.copy_x2(fine, coarse, n, vc)
    ci = coarse[vc:vc + h, vc:vc + h, vc:vc + h]
    inj = np.repeat(np.repeat(np.repeat(ci, 2, axis=0), 2, axis=1), 2, axis=2)
    np.testing.assert_array_equal(fine[vc:vc + n, vc:vc + n, vc:vc + n], inj)
    mask = np.ones(fine.shape, bool)
    mask[vc:vc + n, vc:vc + n, vc:vc + n] = False
    np.testing.assert_array_equal(fine[mask], before[mask])


def test_load2_rejects_a_foreign_header(tmp_path):
    data = np.zeros((2, 6, 6, 6))
    path = str(tmp_path / "f.bin")
    R.dump2(data, 4, 1, path)
    with pytest.raises(ValueError):
        R.load2(np.zeros((3, 6, 6, 6)), 4, 1, path)
    with pytest.raises(ValueError):
        R.load2(np.zeros((2, 8, 8, 8)), 6, 1, path)
