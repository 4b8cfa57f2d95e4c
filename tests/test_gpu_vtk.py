"""bcm_field_write_vtk writes the same bytes as the restatement of
VtkWriter::writeScalar / printVTIC (oracle/vtk.py), FP64 and FP32 fields,
on an adaptive multi-root mesh."""
import sys

import numpy as np
import pytest

from bcmtools_amd import Mesh
from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W
from oracle import vtk as V

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_write_vtk_bytes(tmp_path, dtype):
    t = W.Tree(root_n=(2, 2, 2), root_length=0.5, divider=L.DIV_SHELL, min_level=0, max_level=2,
               shells=[(0.5, 0.5, 0.5, 0.25, 0.25)], balance=1)
    blocks, n, vc = t.blocks, 8, 1
    print(f"[gpu case] write_vtk {np.dtype(dtype).name} nb={len(blocks)}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(7)
    data = rng.standard_normal((len(blocks), n + 2, n + 2, n + 2)).astype(dtype)
    mesh = Mesh(blocks, n, vc)
    f = mesh.field(dtype, data)
    d = str(tmp_path / "VTK")
    difflevel = int(blocks["level"].max() - blocks["level"].min())
    f.write_vtk(5, "p", difflevel, root_origin=(0.0, 0.0, 0.0), root_length=0.5, directory=d)
    for b, bi in enumerate(blocks):
        got = open(V.vti_name(d, "p", 0, b, 5), "rb").read()
        ref = V.vti_bytes(W.interior(data[b], vc), n, bi["origin"], bi["block_size"][0] / n, "p")
        assert got == ref, b
    got = open(V.vthb_name(d, "p", 5)).read()
    assert got == V.vthb_text(blocks, n, "p", 5, difflevel, (0.0, 0.0, 0.0), 0.5)
