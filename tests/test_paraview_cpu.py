"""The Paraview writer (grains3dgpu_amd/paraview.py) against the reference's
drawing rules (numberOfPoints/numberOfCells/writePoints/writeConnection
_PARAVIEW of Sphere, Box, Rectangle, Cylinder, Cone, Superquadric) and its
file layout (ParaviewPostProcessingWriter.cpp:10-359).

The reference holds no Paraview output to compare with, so the drawings are
pinned by geometry: every point lies on the body's surface, the cells do not
overlap and leave no hole (their volumes sum to the volume of the convex hull
of the points, which for the cylinder and the cone is also a closed form),
and the counts follow the reference's formulas.  The text layout is pinned
on a one-box, one-wall file written out by hand from the reference's
stream statements."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest
from scipy.spatial import ConvexHull

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import paraview as PV
from grains3dgpu_amd import scene as S

SHAPES = {
    "sphere": (A.SPHERE, (0.7,)),
    "box": (A.BOX, (0.3, 0.4, 0.5)),
    "cylinder": (A.CYLINDER, (0.35, 0.8)),
    "cone": (A.CONE, (0.35, 0.8)),
    "ellipsoid": (A.SUPERQUADRIC, (0.3, 0.5, 0.4, 2.0, 2.0)),
    "superquadric": (A.SUPERQUADRIC, (0.3, 0.5, 0.4, 6.0, 3.0)),
}
# numberOfPoints_PARAVIEW / numberOfCells_PARAVIEW of the reference
COUNTS = {"sphere": (4 * 8 * 15 + 3, 8 * 8 * 8), "box": (8, 1), "cylinder": (66, 32), "cone": (34, 32),
          "ellipsoid": (16 * 15 + 3, 256), "superquadric": (16 * 15 + 3, 256)}


def cell_volume(P, ids):
    return ConvexHull(P[list(ids)]).volume


@pytest.mark.parametrize("name", list(SHAPES))
def test_glyph_counts_and_surface(name):
    kind, dims = SHAPES[name]
    pts, cells = PV.glyph(kind, dims)
    P = np.array(pts)
    assert (len(P), len(cells)) == COUNTS[name]
    assert all(0 <= i < len(P) for _, ids in cells for i in ids)
    if name == "sphere":
        np.testing.assert_allclose(np.linalg.norm(P[:-1], axis=1), dims[0], rtol=1e-15)
        np.testing.assert_array_equal(P[-1], 0.0)
    elif name == "cylinder":
        r, h = dims
        np.testing.assert_allclose(np.hypot(P[:64, 0], P[:64, 2]), r, rtol=1e-15)
        np.testing.assert_array_equal(P[:32, 1], -h / 2)
        np.testing.assert_array_equal(P[32:64, 1], h / 2)
    elif name == "cone":
        r, h = dims
        np.testing.assert_allclose(np.hypot(P[:32, 0], P[:32, 2]), r, rtol=1e-15)
        np.testing.assert_array_equal(P[:33, 1], -h / 4)
        np.testing.assert_array_equal(P[33], [0, 3 * (h / 4), 0])
    elif kind == A.SUPERQUADRIC:
        a, b, c, n1, n2 = dims
        Q = P[1:]
        f = (np.abs(Q[:, 0] / a) ** n2 + np.abs(Q[:, 1] / b) ** n2) ** (n1 / n2) + np.abs(Q[:, 2] / c) ** n1
        np.testing.assert_allclose(f, 1.0, rtol=1e-12)


@pytest.mark.parametrize("name", ["sphere", "box", "cylinder", "cone", "ellipsoid"])
def test_cells_tile_the_hull(name):
    """Cells cover the convex hull of the points once: no gap, no overlap."""
    kind, dims = SHAPES[name]
    pts, cells = PV.glyph(kind, dims)
    P = np.array(pts)
    vol = sum(cell_volume(P, ids) for _, ids in cells)
    hull = ConvexHull(P).volume
    assert abs(vol - hull) < 1e-12 * hull
    n = 32
    if name == "cylinder":
        r, h = dims
        assert abs(vol - 0.5 * n * r * r * np.sin(2 * np.pi / n) * h) < 1e-12 * vol
    if name == "cone":
        r, h = dims
        assert abs(vol - 0.5 * n * r * r * np.sin(2 * np.pi / n) * h / 3) < 1e-12 * vol
    if name == "box":
        assert abs(vol - 8 * 0.3 * 0.4 * 0.5) < 1e-14


def test_blocky_superquadric_cells_do_not_overlap():
    """For a blocky superquadric the drawn surface need not be convex; the
    cells from the centre still cover the star-shaped solid once: their
    volumes sum to the volume the surface encloses (divergence theorem over
    the surface quads and triangles)."""
    kind, dims = SHAPES["superquadric"]
    pts, cells = PV.glyph(kind, dims)
    P = np.array(pts)
    vol = sum(cell_volume(P, ids) for _, ids in cells)
    enclosed = 0.0
    for t, ids in cells:
        face = [i for i in ids if i != 0]  # the cell's face on the surface
        for k in range(1, len(face) - 1):
            enclosed += abs(np.linalg.det(P[[face[0], face[k], face[k + 1]]])) / 6
    assert abs(vol - enclosed) < 1e-12 * vol
    from grains3dgpu_amd.shapes import superquadric_volume
    assert 0.9 < vol / superquadric_volume(*dims) < 1.0


# "|" marks the space the reference leaves after each list entry
EXPECTED_OBSTACLES = """<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" >
<UnstructuredGrid>
<Piece NumberOfPoints="4" NumberOfCells="1">
<Points>
<DataArray type="Float32" NumberOfComponents="3" format="ascii">
-0.5 -0.25 0
1.5 -0.25 0
-0.5 0.25 0
1.5 0.25 0
</DataArray>
</Points>
<Cells>
<DataArray type="Int32" Name="connectivity" format="ascii">
0 1 2 3|
</DataArray>
<DataArray type="Int32" Name="offsets" format="ascii">
4|
</DataArray>
<DataArray type="Int32" Name="types" format="ascii">
8|
</DataArray>
</Cells>
<CellData Scalars="Indicator">
<DataArray type="Float32" Name="Indicator" format="ascii">
0|
</DataArray>
</CellData>
</Piece>
</UnstructuredGrid>
</VTKFile>
""".replace("|\n", " \n")

EXPECTED_PARTICLES_HEAD = """<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" >
<UnstructuredGrid>
<Piece NumberOfPoints="8" NumberOfCells="1">
<Points>
<DataArray type="Float32" NumberOfComponents="3" format="ascii">
0.7 1.6 2.5
1.3 1.6 2.5
1.3 1.6 3.5
0.7 1.6 3.5
0.7 2.4 2.5
1.3 2.4 2.5
1.3 2.4 3.5
0.7 2.4 3.5
</DataArray>
</Points>
<Cells>
<DataArray type="Int32" Name="connectivity" format="ascii">
0 1 2 3 4 5 6 7|
</DataArray>
<DataArray type="Int32" Name="offsets" format="ascii">
8|
</DataArray>
<DataArray type="Int32" Name="types" format="ascii">
12|
</DataArray>
</Cells>
<CellData Scalars="NormU,NormOm,CoordNumb">
<DataArray type="Float32" Name="NormU" format="ascii">
5|
</DataArray>
<DataArray type="Float32" Name="NormOm" format="ascii">
1.23457e-05|
</DataArray>
<DataArray type="Float32" Name="CoordNumb" format="ascii">
0|
</DataArray>
</CellData>
</Piece>
</UnstructuredGrid>
</VTKFile>
""".replace("|\n", " \n")


def one_box_scene():
    shapes = [dict(kind=A.BOX, dims=(0.3, 0.4, 0.5), density=1.0),
              dict(kind=A.RECTANGLE, dims=(1.0, 0.25), density=0.0)]
    return S.Scene(shapes=shapes, tables=np.array([S.hooke_row(1e5, 0.8, 0, 0.3)]), n_materials=1,
                   sims=[dict(domain_min=(0, 0, 0), domain_max=(4, 4, 4))],
                   particle_shape=np.zeros(1, np.uint32), pos=np.array([[1.0, 2.0, 3.0]]),
                   quat=np.array([[0, 0, 0, 1.0]]), vel=np.zeros((1, 3)), omega=np.zeros((1, 3)),
                   obstacle_shape=np.array([1], np.uint32),
                   obstacle_tr=np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1, 0.5, 0.0, 0.0]]))


def test_file_layout_and_collections(tmp_path):
    sc = one_box_scene()
    d = str(tmp_path)
    (tmp_path / "run_old.vtu").write_text("x")
    (tmp_path / "run.txt").write_text("kept")
    (tmp_path / "other_Particles.pvd").write_text("kept")
    w = PV.ParaviewWriter(d, "run")
    w.start()
    assert sorted(os.listdir(d)) == ["other_Particles.pvd", "run.txt"]
    vel, om = np.array([[3.0, 4.0, 0.0]]), np.array([[0.0, 1.2345678e-5, 0.0]])
    w.write(0.0, sc, sc.pos, sc.quat, vel, om)
    w.write(2.5e-4, sc, sc.pos, sc.quat, vel, om)
    w.end()
    assert (tmp_path / "run_Obstacles_T0.vtu").read_text() == EXPECTED_OBSTACLES
    assert (tmp_path / "run_Particles_T1.vtu").read_text() == EXPECTED_PARTICLES_HEAD
    pvd = (tmp_path / "run_Particles.pvd").read_text()
    assert pvd == ('<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n'
                   '<Collection>\n<DataSet timestep="0" group="" part="0" file="run_Particles_T0.vtu"/>\n'
                   '<DataSet timestep="0.00025" group="" part="0" file="run_Particles_T1.vtu"/>\n'
                   '</Collection>\n</VTKFile>\n')
    # every file is well-formed XML
    for f in os.listdir(d):
        if f.startswith("run_"):
            ET.parse(os.path.join(d, f))


def test_many_bodies_offsets_and_transforms(tmp_path):
    """A mixed set of rotated bodies: offsets end at the connectivity length,
    ids index the points, and each body's points are its glyph moved by its
    transform."""
    rng = np.random.default_rng(3)
    keys = list(SHAPES)
    kinds = [SHAPES[keys[i % len(keys)]][0] for i in range(12)]
    dims = [SHAPES[keys[i % len(keys)]][1] for i in range(12)]
    tr = S.transform12(rng.uniform(-1, 1, (12, 3)), S.random_quaternions(rng, 12))
    path = str(tmp_path / "m.vtu")
    PV.write_vtu(path, kinds, dims, tr, [("Indicator", np.zeros(12))])
    root = ET.parse(path).getroot()
    arrays = {a.get("Name"): a.text.split() for a in root.iter("DataArray")}
    conn = np.array(arrays["connectivity"], int)
    offs = np.array(arrays["offsets"], int)
    pts = np.array(next(a for a in root.iter("DataArray") if a.get("NumberOfComponents")).text.split(),
                   float).reshape(-1, 3)
    assert offs[-1] == len(conn) and np.all(np.diff(offs) > 0)
    assert conn.max() == len(pts) - 1
    first = 0
    for k, dm, t in zip(kinds, dims, tr):
        g = np.array(PV.glyph(k, dm)[0])
        R = t[:9].reshape(3, 3)
        np.testing.assert_allclose(pts[first:first + len(g)], g @ R.T + t[9:], rtol=1e-5, atol=1e-6)
        first += len(g)
    assert first == len(pts)
