"""Writers of small Gmsh 4.1 ASCII meshes for the reader tests: structured
boxes of hexahedra or Kuhn tetrahedra in the vertex and cell order of
fvh_mesh_box (tile 0), boxes of prisms, and a mixed mesh of hexahedra, prisms
and pyramids.
Boundary elements are the unmatched local faces, one physical surface per box
side ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"), written in the order of
the faces' first appearance; the volume's physical name is "flow"."""
import numpy as np

from oracle.meshproc import HEXAEDER, PRISM, PYRAMID, TETRAEDER, TRIANGLE, local_faces

GMSH_OF = {TETRAEDER: 4, HEXAEDER: 5, PRISM: 6, PYRAMID: 7}
KUHN = [(0, 1, 2, 6), (0, 1, 5, 6), (0, 3, 2, 6), (0, 3, 7, 6), (0, 4, 5, 6), (0, 4, 7, 6)]
SIDES = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]


def _lattice(n, length):
    nn = [a + 1 for a in n]
    X = np.zeros((nn[0] * nn[1] * nn[2], 3))
    for k in range(nn[2]):
        for j in range(nn[1]):
            for i in range(nn[0]):
                X[(k * nn[1] + j) * nn[0] + i] = [length[0] * i / n[0], length[1] * j / n[1],
                                                  length[2] * k / n[2]]
    vid = lambda i, j, k: (k * nn[1] + j) * nn[0] + i
    return X, vid


def _hex_nodes(vid, i, j, k):
    return [vid(i, j, k), vid(i + 1, j, k), vid(i + 1, j + 1, k), vid(i, j + 1, k),
            vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i + 1, j + 1, k + 1), vid(i, j + 1, k + 1)]


def box_cells(n, length=(1.0, 1.0, 1.0), tets=False):
    X, vid = _lattice(n, length)
    cells = []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                h = _hex_nodes(vid, i, j, k)
                if tets:
                    cells += [(TETRAEDER, [h[q] for q in t]) for t in KUHN]
                else:
                    cells.append((HEXAEDER, h))
    return X, cells


def prism_cells(n, length=(1.0, 1.0, 1.0)):
    """Every hexahedron split into two prisms along its (0, 2) diagonal (cells
    of 5 faces only)."""
    X, vid = _lattice(n, length)
    cells = []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                h = _hex_nodes(vid, i, j, k)
                cells.append((PRISM, [h[0], h[1], h[2], h[4], h[5], h[6]]))
                cells.append((PRISM, [h[0], h[2], h[3], h[4], h[6], h[7]]))
    return X, cells


def mixed_cells(n=(4, 3, 3), length=(1.0, 1.0, 1.0)):
    """Hexahedra, with the column i = j = 0 split into two prisms per hexahedron
    and the hexahedron (2, 1, 1) split into six pyramids around its centre."""
    X, vid = _lattice(n, length)
    X = list(map(list, X))
    cells = []
    hexf = [(0, 1, 2, 3), (4, 5, 6, 7), (0, 4, 7, 3), (0, 4, 5, 1), (1, 5, 6, 2), (3, 7, 6, 2)]
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                h = _hex_nodes(vid, i, j, k)
                if i == 0 and j == 0:
                    cells.append((PRISM, [h[0], h[1], h[2], h[4], h[5], h[6]]))
                    cells.append((PRISM, [h[0], h[2], h[3], h[4], h[6], h[7]]))
                elif (i, j, k) == (2, 1, 1):
                    c = len(X)
                    X.append(list(np.mean(np.array(X)[h], axis=0)))
                    for f in hexf:
                        cells.append((PYRAMID, [h[q] for q in f] + [c]))
                else:
                    cells.append((HEXAEDER, h))
    return np.array(X), cells


def write_msh(path, X, cells, length=(1.0, 1.0, 1.0), jitter_seed=None):
    """Write the cells and their boundary faces as a Gmsh 4.1 ASCII file."""
    X = np.array(X, dtype=np.float64)
    if jitter_seed is not None:
        # move interior vertices a little (keeps every cell valid)
        rng = np.random.default_rng(jitter_seed)
        lo, hi = X.min(0), X.max(0)
        inner = np.all((X > lo + 1e-12) & (X < hi - 1e-12), axis=1)
        X[inner] += 0.02 * rng.uniform(-1, 1, (int(inner.sum()), 3)) * np.asarray(length)
    count = {}
    order = []
    for t, v in cells:
        for ft, fv in local_faces(t, v):
            key = (ft, tuple(sorted(fv)))
            if key not in count:
                order.append((key, ft, fv))
            count[key] = count.get(key, 0) + 1
    lo, hi = X.min(0), X.max(0)
    bnd = {s: [] for s in range(6)}
    seq = []
    for key, ft, fv in order:
        if count[key] != 1:
            continue
        xc = X[fv].mean(0)
        side = None
        for a in range(3):
            if abs(xc[a] - lo[a]) < 1e-12:
                side = 2 * a
            elif abs(xc[a] - hi[a]) < 1e-12:
                side = 2 * a + 1
        assert side is not None
        seq.append((side, ft, fv))
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", "7", '3 1 "flow"']
    out += ['2 %d "%s"' % (s + 2, SIDES[s]) for s in range(6)]
    out += ["$EndPhysicalNames", "$Entities", "0 0 6 1"]
    for s in range(6):
        out.append("%d 0 0 0 1 1 1 1 %d 0" % (s + 1, s + 2))
    out.append("1 0 0 0 1 1 1 1 1 0")
    out += ["$EndEntities", "$Nodes", "1 %d 1 %d" % (len(X), len(X)),
            "3 1 0 %d" % len(X)]
    out += [str(i + 1) for i in range(len(X))]
    out += ["%.17g %.17g %.17g" % tuple(x) for x in X]
    out += ["$EndNodes"]
    # element blocks: boundary runs (entity, type) in face order, then cells
    blocks = []
    for side, ft, fv in seq:
        g = 2 if ft == TRIANGLE else 3
        if blocks and blocks[-1][0] == (2, side + 1, g):
            blocks[-1][1].append(fv)
        else:
            blocks.append(((2, side + 1, g), [fv]))
    for t, v in cells:
        g = GMSH_OF[t]
        if blocks and blocks[-1][0] == (3, 1, g):
            blocks[-1][1].append(v)
        else:
            blocks.append(((3, 1, g), [v]))
    ne = sum(len(b[1]) for b in blocks)
    out += ["$Elements", "%d %d 1 %d" % (len(blocks), ne, ne)]
    tag = 1
    for (d, ent, g), els in blocks:
        out.append("%d %d %d %d" % (d, ent, g, len(els)))
        for v in els:
            out.append(" ".join([str(tag)] + [str(x + 1) for x in v]))
            tag += 1
    out += ["$EndElements"]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return X
