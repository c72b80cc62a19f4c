"""Fluent ASCII ``.cas`` meshes of the reference driver, and the map between
its CellArray order and the library's structured index n = j*nx + i.

The reference reads its mesh with ``Ark/Lib/MeshConstructFunc.h:27-358``
driven by ``Ark/Cpp/Mesh/MeshConstruct.cpp:69-123`` and numbers cells in file
order (CellArray); ``Ark/Cpp/Mesh/MeshCartesian.cpp:369-408`` then places each
cell in CarCellArray[i+1][j+1] by its centre.  ``read_cas`` recovers that
placement, so a driver can hand its CellArray-ordered data to the library (and
the flow-field writer its order back).  ``write_cas`` writes a uniform
Cartesian mesh the reader (and the reference) accepts:

* section 10 (nodes): one header ``(10 (0 1 N 0 2))`` then a body with decimal
  ``x y`` pairs (``MeshConstructFunc.h:59-82``);
* section 12 (cells): header with zone 0 allocates, a second header with a
  non-zero zone and element type 3 marks every cell a quad
  (``MeshConstructFunc.h:27-54,83-112``);
* section 13 (faces): one allocating header, then one body per zone with
  ``n0 n1 c_left c_right`` in hex (face_type 2, ``MeshConstructFunc.h:181-275``);
  the left cell is the owner and the face normal is (dy, -dx)/|.|
  (``Ark/Cpp/Mesh/Face_2D.cpp:156-161``), so node order fixes the owner's
  outward normal;
* section 18: periodic face pairs (``MeshConstructFunc.h:330-358``).

Zones: interior faces must be zone 3 (``Ark/Cpp/Mesh/MeshCheck.cpp:151-157``);
each boundary side gets its own zone id (left 4, right 5, bottom 6, top 7).
bc_type 2 = interior, 12/8 = periodic pair, 3 = wall.
"""
import re

import numpy as np

BC_INTERIOR, BC_WALL, BC_PERIODIC, BC_PERIODIC_SHADOW = 2, 3, 12, 8


def write_cas(path, nx, ny, walls_y=False, lx=None, ly=None, cell_order=None):
    """Uniform nx*ny mesh, periodic in x, periodic or walls in y.  cell_order:
    optional permutation; the k-th cell of the file is structured cell
    cell_order[k] (default: structured order)."""
    lx = float(nx) if lx is None else float(lx)
    ly = float(ny) if ly is None else float(ly)
    dx, dy = lx / nx, ly / ny
    if cell_order is None:
        file_index = np.arange(nx * ny)
    else:
        order = np.asarray(cell_order, dtype=np.int64)
        assert sorted(order.tolist()) == list(range(nx * ny)), "cell_order must be a permutation"
        file_index = np.empty(nx * ny, dtype=np.int64)
        file_index[order] = np.arange(nx * ny)

    def node(a, b):
        return b * (nx + 1) + a + 1

    def cell(i, j):
        return int(file_index[j * nx + i]) + 1

    zones = []  # (zone, bc_type, faces[(n0, n1, cl, cr)])
    interior = []
    for j in range(ny):
        for i in range(nx - 1):
            # owner (i,j) on the left, normal +x: node0 below node1
            interior.append((node(i + 1, j), node(i + 1, j + 1), cell(i, j), cell(i + 1, j)))
    for j in range(ny - 1):
        for i in range(nx):
            # owner (i,j) below, normal +y: node0 right of node1
            interior.append((node(i + 1, j + 1), node(i, j + 1), cell(i, j), cell(i, j + 1)))
    zones.append((3, BC_INTERIOR, interior))
    left = [(node(0, j + 1), node(0, j), cell(0, j), 0) for j in range(ny)]
    right = [(node(nx, j), node(nx, j + 1), cell(nx - 1, j), 0) for j in range(ny)]
    bottom = [(node(i, 0), node(i + 1, 0), cell(i, 0), 0) for i in range(nx)]
    top = [(node(i + 1, ny), node(i, ny), cell(i, ny - 1), 0) for i in range(nx)]
    zones.append((4, BC_PERIODIC, left))
    zones.append((5, BC_PERIODIC_SHADOW, right))
    if walls_y:
        zones.append((6, BC_WALL, bottom))
        zones.append((7, BC_WALL, top))
    else:
        zones.append((6, BC_PERIODIC, bottom))
        zones.append((7, BC_PERIODIC_SHADOW, top))

    nfaces = sum(len(z[2]) for z in zones)
    nnodes = (nx + 1) * (ny + 1)
    ncells = nx * ny
    out = []
    out.append('(0 "mpdugks mesh")')
    out.append("(2 2)")
    out.append("(10 (0 1 %x 0 2))" % nnodes)
    out.append("(10 (1 1 %x 1 2)(" % nnodes)
    for b in range(ny + 1):
        for a in range(nx + 1):
            out.append("%.12e %.12e" % (a * dx, b * dy))
    out.append("))")
    out.append("(12 (0 1 %x 0))" % ncells)
    out.append("(12 (2 1 %x 1 3))" % ncells)
    out.append("(13 (0 1 %x 0))" % nfaces)
    first = 1
    face_ids = {}
    for zone, bc, faces in zones:
        last = first + len(faces) - 1
        out.append("(13 (%x %x %x %x 2)(" % (zone, first, last, bc))
        for f in faces:
            out.append("%x %x %x %x" % f)
        out.append("))")
        face_ids[zone] = (first, last)
        first = last + 1
    pairs = []
    (lf, _), (rf, _) = face_ids[4], face_ids[5]
    for j in range(ny):
        pairs.append((lf + j, rf + j))
    if not walls_y:
        (bf, _), (tf, _) = face_ids[6], face_ids[7]
        for i in range(nx):
            pairs.append((bf + i, tf + i))
    out.append("(18 (1 %x 4 5)(" % len(pairs))
    for p in pairs:
        out.append("%x %x" % p)
    out.append("))")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


_HEAD = re.compile(r"^\((\d+)\s+\(([0-9a-fA-F\s]+)\)(\()?\s*$")


def read_cas(path):
    """Parse a uniform Cartesian Fluent mesh.  Returns a dict with nx, ny, dx,
    dy, `order` (int64: order[k] = structured index of file cell k, i.e. of
    CellArray[k]), the face zones {zone: bc_type} and walls_y (the y sides are
    walls)."""
    nodes, faces, zones = None, [], {}
    with open(path) as fh:
        lines = fh.read().split("\n")
    k = 0
    while k < len(lines):
        m = _HEAD.match(lines[k].strip())
        k += 1
        if not m:
            continue
        sec, fields, body = int(m.group(1)), m.group(2).split(), m.group(3)
        if not body:
            continue
        rows = []
        while k < len(lines) and lines[k].strip() != "))":
            rows.append(lines[k].split())
            k += 1
        k += 1
        if sec == 10:
            nodes = np.array([[float(a), float(b)] for a, b in rows])
        elif sec == 13:
            zone, bc = int(fields[0], 16), int(fields[3], 16)
            zones[zone] = bc
            faces += [(int(r[0], 16), int(r[1], 16), int(r[2], 16), int(r[3], 16)) for r in rows]
    if nodes is None or not faces:
        raise ValueError("%s: no node or face section" % path)
    f = np.array(faces, dtype=np.int64)
    ncell = int(max(f[:, 2].max(), f[:, 3].max()))
    # cell centre = mean of the end nodes of its faces (a rectangle's four faces
    # list each corner twice)
    acc = np.zeros((ncell + 1, 2))
    cnt = np.zeros(ncell + 1)
    for side in (2, 3):
        c = f[:, side]
        for e in (0, 1):
            np.add.at(acc, c, nodes[f[:, e] - 1])
            np.add.at(cnt, c, 1)
    centre = acc[1:] / cnt[1:, None]
    xs, ys = np.unique(nodes[:, 0]), np.unique(nodes[:, 1])
    nx, ny = len(xs) - 1, len(ys) - 1
    if nx * ny != ncell:
        raise ValueError("%s: not a uniform Cartesian mesh (%d cells, %dx%d node lines)" % (path, ncell, nx, ny))
    dx, dy = (xs[-1] - xs[0]) / nx, (ys[-1] - ys[0]) / ny
    i = np.floor((centre[:, 0] - xs[0]) / dx).astype(np.int64)
    j = np.floor((centre[:, 1] - ys[0]) / dy).astype(np.int64)
    order = j * nx + i
    if len(np.unique(order)) != ncell:
        raise ValueError("%s: cells do not map one to one onto the grid" % path)
    walls = any(bc == BC_WALL for bc in zones.values())
    return {"nx": nx, "ny": ny, "dx": float(dx), "dy": float(dy), "order": order, "zones": zones,
            "walls_y": walls}
