"""VTU/PVTU writer (include/msfem_vtu.hpp) on the CPU: the files parse as VTK
XML, point coordinates and fields are the fine grid's, every quad/hexahedron
has positive orientation in VTK's vertex order, and the master record lists
its pieces."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "test_vtu")


def arrays(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {"npts": int(piece.get("NumberOfPoints")), "ncells": int(piece.get("NumberOfCells"))}
    out["points"] = np.array(piece.find("Points/DataArray").text.split(), float).reshape(-1, 3)
    for da in piece.findall("Cells/DataArray"):
        out[da.get("Name")] = np.array(da.text.split(), np.int64)
    for da in piece.findall("PointData/DataArray"):
        out["field:" + da.get("Name")] = np.array(da.text.split(), float)
    return out


@pytest.fixture(scope="module")
def outdir(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("vtu")
    subprocess.run([EXE, str(d)], check=True, capture_output=True)
    return d


def test_vtu_2d(outdir):
    a = arrays(os.path.join(outdir, "cell2d.vtu"))
    assert a["npts"] == 16 and a["ncells"] == 9
    p = a["points"]
    assert np.allclose(p[:, 0].min(), 0.5) and np.allclose(p[:, 0].max(), 1.0) and np.allclose(p[:, 1].max(), 0.25)
    assert np.allclose(a["field:solution"], p[:, 0] + 10 * p[:, 1])
    assert np.array_equal(a["field:twice"], a["field:solution"])
    conn = a["connectivity"].reshape(-1, 4)
    assert np.array_equal(a["offsets"], 4 * np.arange(1, 10)) and (a["types"] == 9).all()
    for q in conn:  # counter-clockwise: positive shoelace area, equal to the fine cell's
        x, y = p[q, 0], p[q, 1]
        area = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
        assert area == pytest.approx((0.5 / 3) * (0.25 / 3))


def test_vtu_3d(outdir):
    a = arrays(os.path.join(outdir, "cell3d.vtu"))
    assert a["npts"] == 27 and a["ncells"] == 8 and (a["types"] == 12).all()
    p = a["points"]
    assert np.allclose(a["field:solution"], p[:, 2])
    for h in a["connectivity"].reshape(-1, 8):  # right-handed: (v1-v0, v3-v0, v4-v0) positive
        v = p[h]
        assert np.linalg.det(np.stack([v[1] - v[0], v[3] - v[0], v[4] - v[0]])) == pytest.approx(0.125)


def test_vtu_quad(outdir):
    """General quadrilateral: points on the bilinear image of the grid, positive
    fine cells tiling the cell (areas sum to its shoelace area), field = x."""
    a = arrays(os.path.join(outdir, "quad2d.vtu"))
    assert a["npts"] == 25 and a["ncells"] == 16
    p = a["points"]
    quad = np.array([[0.1, 0.2], [0.9, 0.1], [0.25, 0.8], [0.7, 1.05]])
    assert np.allclose(p[0, :2], quad[0]) and np.allclose(p[4, :2], quad[1])
    assert np.allclose(p[20, :2], quad[2]) and np.allclose(p[24, :2], quad[3])
    assert np.allclose(p[12, :2], quad.mean(axis=0))
    assert np.allclose(a["field:solution"], p[:, 0])
    total = 0.0
    for q in a["connectivity"].reshape(-1, 4):
        x, y = p[q, 0], p[q, 1]
        area = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
        assert area > 0
        total += area
    c = quad[[0, 1, 3, 2]]
    assert total == pytest.approx(0.5 * np.sum(c[:, 0] * np.roll(c[:, 1], -1) - np.roll(c[:, 0], -1) * c[:, 1]))


def test_pvtu_master(outdir):
    root = ET.parse(os.path.join(outdir, "master.pvtu")).getroot()
    assert root.get("type") == "PUnstructuredGrid"
    assert [p.get("Source") for p in root.iter("Piece")] == ["cell2d.vtu", "cell3d.vtu"]
    assert [d.get("Name") for d in root.find("PUnstructuredGrid/PPointData")] == ["solution"]
