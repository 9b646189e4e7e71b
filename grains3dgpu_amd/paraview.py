"""Paraview output of a run: the ParaviewPostProcessingWriter of the
reference (PostProcessing/src/ParaviewPostProcessingWriter.cpp:10-359).

Each save writes <root>_Particles_T<k>.vtu and <root>_Obstacles_T<k>.vtu
(VTK XML unstructured grids in ASCII) and rewrites <root>_Particles.pvd and
<root>_Obstacles.pvd, the collections of every save so far.  Every body is
drawn as the reference draws it (the numberOfPoints / numberOfCells /
writePoints / writeConnection_PARAVIEW members of each Convex):

  Sphere        8 points per quarter meridian: 15 latitude rings of 32
                points, the two poles and the centre; pyramids between
                rings, tetrahedra at the poles (Sphere.cpp:4,135-295)
  Box           one hexahedron (Box.cpp:155-219)
  Rectangle     one quad, written as the pixel cell type 8 (Rectangle.cpp:142-196)
  Cylinder      32 wedges around the axis (Cylinder.cpp:4,154-259)
  Cone          32 tetrahedra from the base rim, base centre and apex
                (Cone.cpp:5,167-247)
  Superquadric  16 x 16 cells from the centre: tetrahedra at the poles,
                pyramids between the 15 rings (Superquadric.cpp:5,229-428)

Numbers are written as C++ streams write them by default (six significant
digits, the %g form); integer lists end with a space before the newline.
Cell data: particles carry NormU = |v|, NormOm = |omega| and CoordNumb = 0
(the reference writes 0), obstacles Indicator = 0, each repeated once per
cell of the body."""
import math
import os
import re

import numpy as np

from . import _abi as A

VTK_TETRA, VTK_PIXEL, VTK_HEXAHEDRON, VTK_WEDGE, VTK_PYRAMID = 10, 8, 12, 13, 14
SPHERE_Q = 8      # points per quarter meridian of the sphere
ROUND_N = 32      # points around the cylinder and the cone
SQ_N = 16         # latitude and longitude steps of the superquadric


def _g(x):
    """A number as `std::ostream << x` writes it (precision 6, %g)."""
    return "%g" % x


# ---- body-frame glyphs: points (P, 3) and cells [(type, local ids)] ------------

def _sphere(r):
    a = math.pi / (2 * SPHERE_Q)
    per = 4 * SPHERE_Q
    pts = []
    for k in range(2 * SPHERE_Q - 1):
        az = -math.pi / 2 + (k + 1) * a
        lr, z = r * math.cos(az), r * math.sin(az)
        pts += [(lr * math.cos(i * a), lr * math.sin(i * a), z) for i in range(per)]
    pts += [(0.0, 0.0, -r), (0.0, 0.0, r), (0.0, 0.0, 0.0)]
    bottom, top, gc = per * (2 * SPHERE_Q - 1), per * (2 * SPHERE_Q - 1) + 1, per * (2 * SPHERE_Q - 1) + 2
    cells = []
    for k in range(2 * SPHERE_Q - 2):
        for i in range(per):
            j = (i + 1) % per
            cells.append((VTK_PYRAMID, (k * per + i, k * per + j, (k + 1) * per + j, (k + 1) * per + i, gc)))
    for i in range(per):
        cells.append((VTK_TETRA, (i, (i + 1) % per, bottom, gc)))
    last = (2 * SPHERE_Q - 2) * per
    for i in range(per):
        cells.append((VTK_TETRA, (last + i, last + (i + 1) % per, top, gc)))
    return pts, cells


def _box(e):
    ex, ey, ez = e
    pts = [(-ex, -ey, -ez), (ex, -ey, -ez), (ex, -ey, ez), (-ex, -ey, ez),
           (-ex, ey, -ez), (ex, ey, -ez), (ex, ey, ez), (-ex, ey, ez)]
    return pts, [(VTK_HEXAHEDRON, tuple(range(8)))]


def _rectangle(lx, ly):
    return [(-lx, -ly, 0.0), (lx, -ly, 0.0), (-lx, ly, 0.0), (lx, ly, 0.0)], [(VTK_PIXEL, (0, 1, 2, 3))]


def _ring(r, y, n):
    dt = 2 * math.pi / n
    return [(r * math.cos(i * dt), y, r * math.sin(i * dt)) for i in range(n)]


def _cylinder(r, h):
    hh = h / 2
    n = ROUND_N
    pts = _ring(r, -hh, n) + _ring(r, hh, n) + [(0.0, -hh, 0.0), (0.0, hh, 0.0)]
    return pts, [(VTK_WEDGE, (i, (i + 1) % n, 2 * n, n + i, n + (i + 1) % n, 2 * n + 1)) for i in range(n)]


def _cone(r, h):
    q = h / 4
    n = ROUND_N
    pts = _ring(r, -q, n) + [(0.0, -q, 0.0), (0.0, 3 * q, 0.0)]
    return pts, [(VTK_TETRA, (i, (i + 1) % n, n, n + 1)) for i in range(n)]


def _spow(c, e):
    """sign(c) |c|^e, 0 at 0 (Superquadric.cpp:270-300)."""
    if c == 0.0:
        return 0.0
    return -math.pow(-c, e) if c < 0.0 else math.pow(c, e)


def _superquadric(a, b, c, n1, n2):
    e1, e2 = 2.0 / n1, 2.0 / n2
    n = SQ_N
    dth, dph = math.pi / n, 2 * math.pi / n
    pts = [(0.0, 0.0, 0.0), (0.0, 0.0, c), (0.0, 0.0, -c)]
    for i in range(1, n):
        ct, st = math.cos(i * dth), math.sin(i * dth)
        ce, se = _spow(ct, e1), math.pow(st, e1)
        for j in range(n):
            cp, sp = math.cos(j * dph), math.sin(j * dph)
            x = 0.0 if cp == 0.0 else (-a * se * math.pow(-cp, e2) if cp < 0.0 else a * se * math.pow(cp, e2))
            y = 0.0 if sp == 0.0 else (-b * se * math.pow(-sp, e2) if sp < 0.0 else b * se * math.pow(sp, e2))
            pts.append((x, y, c * ce))
    cells = [(VTK_TETRA, (0, 1, 3 + j, 3 + (j + 1) % n)) for j in range(n)]
    for i in range(1, n - 1):
        row = 3 + (i - 1) * n
        for j in range(n):
            p = row + j
            if j < n - 1:
                cells.append((VTK_PYRAMID, (0, p, p + 1, p + n + 1, p + n)))
            else:
                # wrap of the ring: (0, last, first of the ring, first of the next ring, last of the next)
                cells.append((VTK_PYRAMID, (0, p, p - (n - 1), p + 1, p + n)))
    lastrow = 3 + n * (n - 2)
    cells += [(VTK_TETRA, (0, 2, lastrow + j, lastrow + (j + 1) % n)) for j in range(n)]
    return pts, cells


def glyph(kind, dims):
    """Body-frame points and cells of one shape record (kind, dims)."""
    d = [float(x) for x in dims]
    if kind == A.SPHERE:
        return _sphere(d[0])
    if kind == A.BOX:
        return _box(d[:3])
    if kind == A.RECTANGLE:
        return _rectangle(d[0], d[1])
    if kind == A.CYLINDER:
        return _cylinder(d[0], d[1])
    if kind == A.CONE:
        return _cone(d[0], d[1])
    if kind == A.SUPERQUADRIC:
        return _superquadric(*d[:5])
    raise ValueError(f"no Paraview drawing for shape kind {kind} (the reference has none)")


def _place(pts, tr):
    """Transform3::operator() (basis row . p + origin) of every point."""
    P = np.asarray(pts, float)
    R, o = np.asarray(tr[:9], float).reshape(3, 3), np.asarray(tr[9:], float)
    return np.stack([R[k, 0] * P[:, 0] + R[k, 1] * P[:, 1] + R[k, 2] * P[:, 2] + o[k] for k in range(3)], 1)


def _grid(kinds, dims, transforms):
    """Points, connectivity, offsets, types and the cell count of each body."""
    cache, points, conn, offs, types, ncell = {}, [], [], [], [], []
    first, off = 0, 0
    for k, dm, tr in zip(kinds, dims, transforms):
        key = (int(k), tuple(float(x) for x in dm))
        if key not in cache:
            cache[key] = glyph(*key)
        pts, cells = cache[key]
        points.append(_place(pts, tr))
        for t, ids in cells:
            conn += [first + i for i in ids]
            off += len(ids)
            offs.append(off)
            types.append(t)
        ncell.append(len(cells))
        first += len(pts)
    P = np.concatenate(points) if points else np.zeros((0, 3))
    return P, conn, offs, types, ncell


def _ints(v):
    return "".join(f"{x} " for x in v)


def write_vtu(path, kinds, dims, transforms, cell_data):
    """One unstructured grid: `cell_data` is [(name, per-body values)]."""
    P, conn, offs, types, ncell = _grid(kinds, dims, transforms)
    head = 'format="ascii">'
    L = ['<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" >', "<UnstructuredGrid>",
         f'<Piece NumberOfPoints="{len(P)}" NumberOfCells="{sum(ncell)}">', "<Points>",
         f'<DataArray type="Float32" NumberOfComponents="3" {head}']
    L += [f"{_g(x)} {_g(y)} {_g(z)}" for x, y, z in P]
    L += ["</DataArray>", "</Points>", "<Cells>",
          f'<DataArray type="Int32" Name="connectivity" {head}', _ints(conn), "</DataArray>",
          f'<DataArray type="Int32" Name="offsets" {head}', _ints(offs), "</DataArray>",
          f'<DataArray type="Int32" Name="types" {head}', _ints(types), "</DataArray>", "</Cells>",
          f'<CellData Scalars="{",".join(n for n, _ in cell_data)}">']
    for name, vals in cell_data:
        L += [f'<DataArray type="Float32" Name="{name}" {head}',
              "".join(f"{_g(v)} " * c for v, c in zip(vals, ncell)), "</DataArray>"]
    L += ["</CellData>", "</Piece>", "</UnstructuredGrid>", "</VTKFile>"]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


class ParaviewWriter:
    """The writer's life cycle: start() clears the directory's earlier
    <root>_*.pvd/.vtu/.pvtu/.vtp/.pvtp files and opens the collections;
    write() adds one save; end() does nothing (the collections are rewritten
    complete at every save)."""
    HEAD = '<?xml version="1.0"?>\n<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n<Collection>\n'

    def __init__(self, directory, root):
        self.directory, self.root = directory, root
        self.cycle = 0
        self.pvd = {"Obstacles": self.HEAD, "Particles": self.HEAD}

    def start(self):
        os.makedirs(self.directory, exist_ok=True)
        pat = re.compile("^" + re.escape(self.root) + r"_.*\.(pvd|vtu|pvtu|vtp|pvtp)$")
        for name in os.listdir(self.directory):
            if pat.match(name):
                os.remove(os.path.join(self.directory, name))
        self.cycle = 0
        self.pvd = {"Obstacles": self.HEAD, "Particles": self.HEAD}

    def _collection(self, part, t, name):
        self.pvd[part] += f'<DataSet timestep="{_g(t)}" group="" part="0" file="{name}"/>\n'
        with open(os.path.join(self.directory, f"{self.root}_{part}.pvd"), "w") as f:
            f.write(self.pvd[part] + "</Collection>\n</VTKFile>\n")

    def write(self, t, scene, pos, quat, vel, omega):
        """One save at time t of the particles' state (pos, quat, vel, omega
        as the engine returns them) and the scene's obstacles."""
        from .scene import transform12
        shapes = scene.shapes
        kind = lambda s: int(s["kind"])
        dim = lambda s: tuple(s["dims"])
        osh = np.asarray(scene.obstacle_shape, int)
        otr = np.asarray(scene.obstacle_tr, float).reshape(-1, 12)
        name = f"{self.root}_Obstacles_T{self.cycle}.vtu"
        self._collection("Obstacles", t, name)
        write_vtu(os.path.join(self.directory, name), [kind(shapes[s]) for s in osh], [dim(shapes[s]) for s in osh],
                  otr, [("Indicator", np.zeros(len(osh)))])
        psh = np.asarray(scene.particle_shape, int)
        ptr = transform12(np.asarray(pos, float), np.asarray(quat, float))
        name = f"{self.root}_Particles_T{self.cycle}.vtu"
        self._collection("Particles", t, name)
        write_vtu(os.path.join(self.directory, name), [kind(shapes[s]) for s in psh], [dim(shapes[s]) for s in psh],
                  ptr, [("NormU", np.linalg.norm(vel, axis=1)), ("NormOm", np.linalg.norm(omega, axis=1)),
                        ("CoordNumb", np.zeros(len(psh)))])
        self.cycle += 1

    def end(self):
        pass
