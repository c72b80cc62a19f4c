"""Fluent ASCII ``.cas`` writer for uniform Cartesian meshes.

Test infrastructure only: it produces the mesh file that the reference solver
reads at start-up, so that the reference can be run as the checker.

The reader this targets is ``Ark/Lib/MeshConstructFunc.h:27-358`` driven by
``Ark/Cpp/Mesh/MeshConstruct.cpp:69-123``:

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
import sys


def write_cas(path, nx, ny, walls_y=False, lx=None, ly=None):
    lx = float(nx) if lx is None else float(lx)
    ly = float(ny) if ly is None else float(ly)
    dx, dy = lx / nx, ly / ny

    def node(a, b):
        return b * (nx + 1) + a + 1

    def cell(i, j):
        return j * nx + i + 1

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
    zones.append((3, 2, interior))
    left = [(node(0, j + 1), node(0, j), cell(0, j), 0) for j in range(ny)]
    right = [(node(nx, j), node(nx, j + 1), cell(nx - 1, j), 0) for j in range(ny)]
    bottom = [(node(i, 0), node(i + 1, 0), cell(i, 0), 0) for i in range(nx)]
    top = [(node(i + 1, ny), node(i, ny), cell(i, ny - 1), 0) for i in range(nx)]
    zones.append((4, 12, left))
    zones.append((5, 8, right))
    if walls_y:
        zones.append((6, 3, bottom))
        zones.append((7, 3, top))
    else:
        zones.append((6, 12, bottom))
        zones.append((7, 8, top))

    nfaces = sum(len(z[2]) for z in zones)
    nnodes = (nx + 1) * (ny + 1)
    ncells = nx * ny
    out = []
    out.append('(0 "mpdugks test mesh")')
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


if __name__ == "__main__":
    # usage: casmesh.py OUT NX NY [walls]
    write_cas(sys.argv[1], int(sys.argv[2]), int(sys.argv[3]),
              walls_y=(len(sys.argv) > 4 and sys.argv[4] == "walls"))
