"""Pin the VTK restatement (oracle/vtk.py) against the reference writer's
own bytes (tests/golden/vtk_sphere.npz, written by the compiled
examples/SolverTest/VtkWriter.h), and by what the files must hold:
the .vti XML parses, its extents and spacing describe the block, and the
appended raw data (after "_" and an int32 byte count) decodes to the
block's interior in float32; the .vthb index parses and each amr_box is the
block's cell range at its own level, which for this mesh is known in
closed form (block index along each axis times the block edge)."""
import xml.etree.ElementTree as ET

import numpy as np

from bcmtools_amd import _lib as L
from bcmtools_amd import workloads as W
from oracle import vtk as V


def test_vti_layout():
    n = 4
    rng = np.random.default_rng(0)
    interior = rng.standard_normal((n, n, n))
    raw = V.vti_bytes(interior, n, (0.25, 0.5, 0.0), 0.0625, "p")
    mark = b'<AppendedData encoding="raw">\n_'
    cut = raw.index(mark) + len(mark)
    head, rest = raw[:cut - 1], raw[cut:]
    root = ET.fromstring(head.decode() + "</AppendedData></VTKFile>")
    img = root.find("ImageData")
    assert img.get("WholeExtent") == f"0 {n} 0 {n} 0 {n}"
    assert [float(v) for v in img.get("Origin").split()] == [0.25, 0.5, 0.0]
    assert [float(v) for v in img.get("Spacing").split()] == [0.0625] * 3
    assert img.find("Piece/CellData/DataArray").get("Name") == "p"
    nbytes = int(np.frombuffer(rest[:4], dtype=np.int32)[0])
    assert nbytes == 4 * n ** 3
    data = np.frombuffer(rest[4:4 + nbytes], dtype=np.float32).reshape(n, n, n)
    np.testing.assert_array_equal(data, interior.astype(np.float32))
    assert rest[4 + nbytes:] == b"\n</AppendedData>\n</VTKFile>\n"


def test_vthb_boxes_in_closed_form():
    t = W.Tree(root_n=(1, 1, 1), divider=L.DIV_FLAT, max_level=2)
    n = 8
    text = V.vthb_text(t.blocks, n, "p", 3, 0, (0.0, 0.0, 0.0), 1.0)
    root = ET.fromstring(text)
    sets = root.findall("vtkHierarchicalBoxDataSet/DataSet")
    assert len(sets) == len(t.blocks)
    for i, (ds, bi) in enumerate(zip(sets, t.blocks)):
        assert int(ds.get("group")) == 2 and int(ds.get("dataset")) == i
        box = [float(v) for v in ds.get("amr_box").split()]
        ijk = np.rint(bi["origin"] / 0.25).astype(int)
        for a in range(3):
            assert box[2 * a] == ijk[a] * n and box[2 * a + 1] == ijk[a] * n + n - 1
        assert ds.get("file") == f"./0000000003/data-p-00000-{i:05d}-0000000003.vti"


def test_restatement_reproduces_the_reference_writer():
    """tests/golden/vtk_sphere: examples/SolverTest/VtkWriter.h's own
    writeScalar, compiled from the reference (oracle/ref), on a 344-block
    adaptive mesh with two root cubes; the restatement must give the same
    bytes for the .vthb index and for every kept block's .vti."""
    import os
    from conftest import GOLDEN
    z = np.load(os.path.join(GOLDEN, "vtk_sphere.npz"), allow_pickle=False)
    blocks = z["blocks"].view(L.BLOCK_DTYPE).copy()
    n, vc, step, difflevel = int(z["n"]), int(z["vc"]), int(z["step"]), int(z["difflevel"])
    text = V.vthb_text(blocks, n, "p", step, difflevel, (0.0, 0.0, 0.0), 1.0)
    assert text.encode() == z["vthb"].tobytes()
    for b, fin in zip(z["kept"], z["fin_kept"]):
        bi = blocks[int(b)]
        interior = fin[vc:vc + n, vc:vc + n, vc:vc + n]
        raw = V.vti_bytes(interior, n, bi["origin"], bi["block_size"][0] / n, "p")
        assert raw == z[f"vti_{int(b)}"].tobytes(), int(b)
