"""The C++ facade (include/msfem_basis.hpp) driving the whole reference pipeline
(DiffusionProblemMultiscale<2>(n_refine, n_refine_local).run(), source/main.cxx:37-39)
on the reference default problem, checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import msfem
import oracle as O

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(__file__), "cpp", "test_facade")


def test_facade_reference_pipeline(tmp_path):
    print("[gpu-case] facade n_refine=2 n_refine_local=5", flush=True)
    out = tmp_path / "facade.bin"
    p = subprocess.run([EXE, "2", "5", str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    nu = int(np.frombuffer(raw[:8], np.int64)[0])
    u = np.frombuffer(raw[8:8 + 8 * nu], np.float64)
    nc_cells = int(np.frombuffer(raw[8 + 8 * nu:16 + 8 * nu], np.int64)[0])
    rest = np.frombuffer(raw[16 + 8 * nu:], np.float64).reshape(nc_cells, 20)
    Ae, Fe = rest[:, :16].reshape(-1, 4, 4), rest[:, 16:]
    # oracle: same cells, reference coefficient (PI_D) at the quadrature points, f = 2,
    # Dirichlet g on x=0,y=0 and Neumann h on x=1,y=1
    nc, nf = 4, 32
    box = msfem.unit_boxes(2, nc)
    n_fine = np.full(nc * nc, nf, np.int32)
    coef = msfem.coef_reference(2, n_fine, box)
    r = O.batch_build(2, box, n_fine, O.COEF_TENSOR_QP, coef, O.RHS_CONST, 2.0)
    assert np.abs(Ae - r["Ae"]).max() / np.abs(r["Ae"]).max() < 1e-10
    assert np.abs(Fe - r["Fe"]).max() / np.abs(r["Fe"]).max() < 1e-10
    Kc, Fc, cm, cv = O.coarse_assemble(2, nc, r["Ae"], r["Fe"], 1)
    ur = O.coarse_solve(2, nc, Kc, Fc, cm, cv)
    assert np.abs(u - ur).max() / np.abs(ur).max() < 1e-10


def test_facade_fine_output(tmp_path):
    """output_global_fine(): one VTU per cell with the fine MsFEM solution and a
    PVTU master record naming them; at the coarse vertices the fine solution
    equals the coarse solution (phi_i(v_j) = delta_ij)."""
    import xml.etree.ElementTree as ET
    print("[gpu-case] facade fine output n_refine=1 n_refine_local=4", flush=True)
    out, vdir = tmp_path / "facade.bin", tmp_path / "vtu"
    vdir.mkdir()
    p = subprocess.run([EXE, "1", "4", str(out), str(vdir)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    nu = int(np.frombuffer(raw[:8], np.int64)[0])
    u = np.frombuffer(raw[8:8 + 8 * nu], np.float64).reshape(3, 3)  # coarse nodes, x fastest
    master = ET.parse(vdir / "solution-ms_fine-2d.pvtu").getroot()
    pieces = [pc.get("Source") for pc in master.iter("Piece")]
    assert len(pieces) == 4
    for src in pieces:
        piece = ET.parse(vdir / src).getroot().find("UnstructuredGrid/Piece")
        assert int(piece.get("NumberOfPoints")) == 289 and int(piece.get("NumberOfCells")) == 256
        pts = np.array(piece.find("Points/DataArray").text.split(), float).reshape(-1, 3)
        sol = np.array(piece.find("PointData/DataArray").text.split(), float)
        for k in (0, 16, 272, 288):  # the cell's corners
            X, Y = int(round(pts[k, 0] * 2)), int(round(pts[k, 1] * 2))
            assert abs(sol[k] - u[Y, X]) <= 1e-12 * max(1.0, abs(u[Y, X]))


def test_facade_general_cells(tmp_path):
    """set_mesh_distortion(0.2, seed 1): every coarse cell a general quadrilateral
    (DiffusionProblemBasis from its corner points); element matrices, loads and
    the coarse solution against the oracle's general-cell path, and the fine
    output's corner points at the moved vertices."""
    import xml.etree.ElementTree as ET
    print("[gpu-case] facade general cells n_refine=2 n_refine_local=5 distortion=0.2", flush=True)
    out, vdir = tmp_path / "facade.bin", tmp_path / "vtu"
    vdir.mkdir()
    p = subprocess.run([EXE, "2", "5", str(out), str(vdir), "0.2"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    nu = int(np.frombuffer(raw[:8], np.int64)[0])
    u = np.frombuffer(raw[8:8 + 8 * nu], np.float64)
    nc_cells = int(np.frombuffer(raw[8 + 8 * nu:16 + 8 * nu], np.int64)[0])
    rest = np.frombuffer(raw[16 + 8 * nu:], np.float64).reshape(nc_cells, 20)
    Ae, Fe = rest[:, :16].reshape(-1, 4, 4), rest[:, 16:]
    nc, nf = 4, 32
    quad = msfem.perturbed_quads(nc, 0.2, 1)
    n_fine = np.full(nc * nc, nf, np.int32)
    coef = msfem.coef_reference_quad(n_fine, quad)
    r = O.batch_build_quad(quad, n_fine, O.COEF_TENSOR_QP, coef, O.RHS_CONST, 2.0)
    assert np.abs(Ae - r["Ae"]).max() / np.abs(r["Ae"]).max() < 1e-10
    assert np.abs(Fe - r["Fe"]).max() / np.abs(r["Fe"]).max() < 1e-10
    Kc, Fc, cm, cv = O.coarse_assemble(2, nc, r["Ae"], r["Fe"], 1)
    ur = O.coarse_solve(2, nc, Kc, Fc, cm, cv)
    assert np.abs(u - ur).max() / np.abs(ur).max() < 1e-10
    master = ET.parse(vdir / "solution-ms_fine-2d.pvtu").getroot()
    pieces = [pc.get("Source") for pc in master.iter("Piece")]
    assert len(pieces) == nc * nc
    for c, src in enumerate(pieces):  # cells in id order
        pts = np.array(ET.parse(vdir / src).getroot().find("UnstructuredGrid/Piece/Points/DataArray").text.split(),
                       float).reshape(-1, 3)
        np1 = nf + 1
        assert np.allclose(pts[[0, nf, nf * np1, np1 * np1 - 1], :2], quad[c], atol=1e-15)


@pytest.mark.parametrize("nr,nrl", [(1, 3), (2, 2)])
def test_facade_reference_pipeline_3d(tmp_path, nr, nrl):
    """DiffusionProblemMultiscale<3> (instantiated by the reference,
    source/base/diffusion_problem_ms.inst.cc:16) on its default problem:
    MatrixCoeff<3> at the QGauss(2) points, f = 2, Dirichlet g on x=0, y=0,
    z=0 and Neumann h on x=1, y=1, z=1; element matrices, loads and the coarse
    solution against the oracle (1e-10)."""
    print(f"[gpu-case] facade 3D n_refine={nr} n_refine_local={nrl}", flush=True)
    out = tmp_path / "facade3.bin"
    p = subprocess.run([EXE, "--dim=3", str(nr), str(nrl), str(out)], capture_output=True, text=True, timeout=200)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    nu = int(np.frombuffer(raw[:8], np.int64)[0])
    u = np.frombuffer(raw[8:8 + 8 * nu], np.float64)
    nc_cells = int(np.frombuffer(raw[8 + 8 * nu:16 + 8 * nu], np.int64)[0])
    rest = np.frombuffer(raw[16 + 8 * nu:], np.float64).reshape(nc_cells, 72)
    Ae, Fe = rest[:, :64].reshape(-1, 8, 8), rest[:, 64:]
    nc, nf = 1 << nr, 1 << nrl
    box = msfem.unit_boxes(3, nc)
    n_fine = np.full(nc ** 3, nf, np.int32)
    coef = msfem.coef_reference(3, n_fine, box)
    r = O.batch_build(3, box, n_fine, O.COEF_TENSOR_QP, coef, O.RHS_CONST, 2.0)
    assert np.abs(Ae - r["Ae"]).max() / np.abs(r["Ae"]).max() < 1e-10
    assert np.abs(Fe - r["Fe"]).max() / np.abs(r["Fe"]).max() < 1e-10
    Kc, Fc, cm, cv = O.coarse_assemble(3, nc, r["Ae"], r["Fe"], 1)
    ur = O.coarse_solve(3, nc, Kc, Fc, cm, cv)
    assert np.abs(u - ur).max() / np.abs(ur).max() < 1e-10


def test_facade_general_hexahedra(tmp_path):
    """DiffusionProblemMultiscale<3> with set_mesh_distortion(0.15, seed 1): every
    coarse cell a general hexahedron (DiffusionProblemBasis<3> from its eight
    corner points); element matrices, loads and the coarse solution against the
    oracle's hexahedron path, the fine output's corner points at the moved
    vertices and the coarse output's points at the moved mesh."""
    import xml.etree.ElementTree as ET
    print("[gpu-case] facade general hexahedra n_refine=2 n_refine_local=3 distortion=0.15", flush=True)
    out, vdir = tmp_path / "facade3h.bin", tmp_path / "vtu"
    vdir.mkdir()
    p = subprocess.run([EXE, "--dim=3", "2", "3", str(out), str(vdir), "0.15"], capture_output=True, text=True,
                       timeout=200)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    nu = int(np.frombuffer(raw[:8], np.int64)[0])
    u = np.frombuffer(raw[8:8 + 8 * nu], np.float64)
    nc_cells = int(np.frombuffer(raw[8 + 8 * nu:16 + 8 * nu], np.int64)[0])
    rest = np.frombuffer(raw[16 + 8 * nu:], np.float64).reshape(nc_cells, 72)
    Ae, Fe = rest[:, :64].reshape(-1, 8, 8), rest[:, 64:]
    nc, nf = 4, 8
    hexa = msfem.perturbed_hexes(nc, 0.15, 1)
    n_fine = np.full(nc ** 3, nf, np.int32)
    coef = msfem.coef_reference_hex(n_fine, hexa)
    r = O.batch_build_hex(hexa, n_fine, O.COEF_TENSOR_QP, coef, O.RHS_CONST, 2.0)
    assert np.abs(Ae - r["Ae"]).max() / np.abs(r["Ae"]).max() < 1e-10
    assert np.abs(Fe - r["Fe"]).max() / np.abs(r["Fe"]).max() < 1e-10
    Kc, Fc, cm, cv = O.coarse_assemble(3, nc, r["Ae"], r["Fe"], 1)
    ur = O.coarse_solve(3, nc, Kc, Fc, cm, cv)
    assert np.abs(u - ur).max() / np.abs(ur).max() < 1e-10
    master = ET.parse(vdir / "solution-ms_fine-3d.pvtu").getroot()
    pieces = [pc.get("Source") for pc in master.iter("Piece")]
    assert len(pieces) == nc ** 3
    np1 = nf + 1
    corners = [x + np1 * (y + np1 * z) for z in (0, nf) for y in (0, nf) for x in (0, nf)]
    for c, src in enumerate(pieces):  # cells in id order
        pts = np.array(ET.parse(vdir / src).getroot().find("UnstructuredGrid/Piece/Points/DataArray").text.split(),
                       float).reshape(-1, 3)
        assert np.allclose(pts[corners], hexa[c], atol=1e-15)
    coarse = ET.parse(vdir / "solution-ms_coarse-3d_refinements-2.0000.vtu").getroot()
    pts = np.array(coarse.find("UnstructuredGrid/Piece/Points/DataArray").text.split(), float).reshape(-1, 3)
    # coarse node (i, j, k) is vertex 0 of cell (i, j, k) where that cell exists
    for k in range(nc):
        for j in range(nc):
            for i in range(nc):
                assert np.allclose(pts[i + (nc + 1) * (j + (nc + 1) * k)], hexa[i + nc * (j + nc * k)][0], atol=1e-15)


@pytest.mark.parametrize("distortion", [None, "0.2"])
def test_facade_coarse_output(tmp_path, distortion):
    """output_global_coarse() (diffusion_problem_ms.tpp:330-380): the coarse
    solution "u" on the coarse mesh (moved vertices for general cells) in
    solution-ms_coarse-2d_refinements-2.0000.vtu, named by the .pvtu record."""
    import xml.etree.ElementTree as ET
    print(f"[gpu-case] facade coarse output distortion={distortion}", flush=True)
    out, vdir = tmp_path / "facade.bin", tmp_path / "vtu"
    vdir.mkdir()
    args = [EXE, "2", "3", str(out), str(vdir)] + ([distortion] if distortion else [])
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    raw = out.read_bytes()
    nu = int(np.frombuffer(raw[:8], np.int64)[0])
    u = np.frombuffer(raw[8:8 + 8 * nu], np.float64)
    master = ET.parse(vdir / "solution-ms_coarse-2d_refinements-2.pvtu").getroot()
    assert [pc.get("Source") for pc in master.iter("Piece")] == ["solution-ms_coarse-2d_refinements-2.0000.vtu"]
    piece = ET.parse(vdir / "solution-ms_coarse-2d_refinements-2.0000.vtu").getroot().find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfPoints")) == 25 and int(piece.get("NumberOfCells")) == 16
    pts = np.array(piece.find("Points/DataArray").text.split(), float).reshape(-1, 3)
    vals = np.array(piece.find("PointData/DataArray").text.split(), float)
    assert np.array_equal(vals, u)
    X, Y = np.arange(25) % 5 / 4, np.arange(25) // 5 / 4
    if distortion is None:
        assert np.allclose(pts[:, 0], X, atol=0) and np.allclose(pts[:, 1], Y, atol=0)
    else:
        quad = msfem.perturbed_quads(4, 0.2, 1)
        for c in range(16):  # every cell's vertices are grid points of the written mesh
            cx, cy = c % 4, c // 4
            for v in range(4):
                node = (cx + (v & 1)) + 5 * (cy + (v >> 1))
                assert np.allclose(pts[node, :2], quad[c, v], atol=1e-15)
