"""Restatement of SolverTest's VTK output (numpy + text).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of
bcm_field_write_vtk; the product never imports it.

VtkWriter::writeScalar (examples/SolverTest/VtkWriter.h:733-889) writes, per
block, printVTIC's ImageData file (:75-200): an XML head whose Origin and
Spacing are printed in scientific notation with 16 digits, the marker "_",
an int32 byte count, the block's interior as float32 (x fastest) and a
closing tail; then, on rank 0, a vtkHierarchicalBoxDataSet index listing
every block with its level, dataset id and amr_box, printed with the
stream's default (6 significant digits, %g-like) formatting.
"""
import os

import numpy as np


def vti_bytes(interior, n, origin, dx, label):
    head = ('<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">\n'
            f'<ImageData WholeExtent="0 {n} 0 {n} 0 {n}" '
            f'Origin="{origin[0]:.16e} {origin[1]:.16e} {origin[2]:.16e}" '
            f'Spacing="{dx:.16e} {dx:.16e} {dx:.16e}">\n'
            f'<Piece Extent="0 {n} 0 {n} 0 {n}">\n'
            '<PointData>\n</PointData>\n<CellData>\n'
            f'<DataArray type="Float32" Name="{label}" format="appended" offset="0"/>\n'
            '</CellData>\n<Coordinates>\n</Coordinates>\n</Piece>\n</ImageData>\n'
            '<AppendedData encoding="raw">\n_')
    data = np.ascontiguousarray(interior, dtype=np.float32)
    return (head.encode() + np.int32(data.nbytes).tobytes() + data.tobytes()
            + b"\n</AppendedData>\n</VTKFile>\n")


def _g(x):
    """C++ ostream default formatting of a double."""
    return f"{x:g}"


def vthb_text(blocks, n, name, step, difflevel, root_origin, root_length, rank=0, start_id=0):
    lines = ['<VTKFile type="vtkHierarchicalBoxDataSet" version="0.1">', "<vtkHierarchicalBoxDataSet>"]
    for lv in range(difflevel + 1):
        lines.append(f'<RefinementRatio level="{lv}" refinement="2"/>')
    lines.append("")
    for i, bi in enumerate(blocks):
        level = int(bi["level"])
        lx = float(n * (1 << level))
        box = []
        for a in range(3):
            o2 = (float(bi["origin"][a]) - root_origin[a]) / root_length
            b0 = o2 * lx
            box += [b0, b0 + n - 1]
        f = f"./{step:010d}/data-{name}-{rank:05d}-{i:05d}-{step:010d}.vti"
        lines.append(f'<DataSet group="{level}" dataset="{i + start_id}" amr_box="'
                     + " ".join(_g(v) for v in box) + f'" file="{f}"/>')
    lines += ["</vtkHierarchicalBoxDataSet>", "</VTKFile>", ""]
    return "\n".join(lines)


def vti_name(directory, name, rank, block_id, step):
    return os.path.join(directory, f"{step:010d}", f"data-{name}-{rank:05d}-{block_id:05d}-{step:010d}.vti")


def vthb_name(directory, name, step):
    return os.path.join(directory, f"data-{name}-{step:010d}.vthb")
