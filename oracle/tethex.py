"""CPU oracle (test infrastructure only): the reference's tet -> 4-hex split with its own node
numbering and vertex orders, Tethex Mesh::convert_3D (src/Tethex.cpp:965-1102), restated
sequentially in the reference's loop order for small meshes.

* edges: IncidenceMatrix (src/Tethex.cpp:374-425) numbers an edge by its position in the
  lower-triangle rows (larger vertex) and sorted columns (smaller vertex): edges sorted by
  (max, min); a cell's local edges run over vertex pairs i, j with v_i > v_j
  (Mesh::edge_numeration, :1232-1270);
* faces: numbered in order of first encounter over (tet, local edge, opposite vertex)
  (Mesh::face_numeration, :1272-1340);
* new nodes: edge midpoints, face and tet centroids, each the running sum of the element's
  vertices divided by their number (Mesh::set_new_vertices, :897-919), after the original
  vertices in the order edges, faces, tets;
* hexes: per tet and tet vertex v (hex 4t + k), [v, e0, f(e0,e1), e1, e2, f(e0,e2), centre,
  f(e1,e2)] with e0..e2 the tet's edges through v in local edge order
  (Mesh::convert_tetrahedra, :1038-1100), front and back faces swapped when the deal.II-ordered
  cell has negative measure (change_vertices_order, :1565-1590; the sign is taken from the
  Jacobian at the cell centre, which has the sign of the cell measure for these cells);
* boundary triangles: three quadrangles [v, e0, face, e1] per triangle, vertices 1 and 3
  swapped when the x-y measure of the deal.II-ordered quadrangle (cell_measure_2D, :1592-1599)
  is negative (Mesh::convert_triangles, :1104-1169).
"""
from __future__ import annotations

import numpy as np

_LEX_OF_UCD = [0, 1, 4, 5, 3, 2, 7, 6]        # lex[k] = ucd[_LEX_OF_UCD[k]]


def _local_edges(v):
    """(i, j) pairs of a cell's local edges in edge_numeration order (v[i] > v[j])."""
    n = len(v)
    return [(i, j) for i in range(n) for j in range(n) if v[i] > v[j]]


def _center_det_lex(x):
    """Jacobian determinant at the centre of a lexicographic hex (8 x 3)."""
    d0 = 0.25 * ((x[1] - x[0]) + (x[3] - x[2]) + (x[5] - x[4]) + (x[7] - x[6]))
    d1 = 0.25 * ((x[2] - x[0]) + (x[3] - x[1]) + (x[6] - x[4]) + (x[7] - x[5]))
    d2 = 0.25 * ((x[4] - x[0]) + (x[5] - x[1]) + (x[6] - x[2]) + (x[7] - x[3]))
    return float(np.dot(d0, np.cross(d1, d2)))


def _measure_2d(xy):
    """cell_measure_2D (src/Tethex.cpp:1592-1599) of 4 points in deal.II order."""
    x = [p[0] for p in xy]
    y = [p[1] for p in xy]
    return (-x[1] * y[0] + x[1] * y[3] + y[0] * x[2] + x[0] * y[1] - x[0] * y[2] - y[1] * x[3]
            - x[2] * y[3] + x[3] * y[2]) / 2.


def split(nodes, tets, tris=None):
    """Returns (all_nodes (n, 3), node_kind (1 vertex, 2 edge, 3 face, 4 tet), hexes (4T, 8) in
    femocs/UCD order, quads (3 F, 4) or None, n_edges, n_faces)."""
    nodes = np.asarray(nodes, np.float64)
    tets = np.asarray(tets, np.int64)
    nv = len(nodes)
    # edge numbers: sorted by (max, min)
    keys = set()
    for t in tets:
        for i, j in _local_edges(t):
            keys.add((int(t[i]), int(t[j])))
    edge_list = sorted(keys)                      # (larger, smaller)
    edge_id = {k: n for n, k in enumerate(edge_list)}
    E = len(edge_list)
    tet_edges = [[edge_id[(int(t[i]), int(t[j]))] for i, j in _local_edges(t)] for t in tets]
    # faces by first encounter; (edge, opposite vertex) -> face
    ev_face = {}
    faces = []
    for t, te in zip(tets, tet_edges):
        for e in te:
            a, b = edge_list[e][1], edge_list[e][0]      # (min, max)
            for w in t:
                w = int(w)
                if w == a or w == b or (e, w) in ev_face:
                    continue
                f = len(faces)
                ev_face[(e, w)] = f
                ev_face[(edge_id[(max(a, w), min(a, w))], b)] = f
                ev_face[(edge_id[(max(b, w), min(b, w))], a)] = f
                faces.append(sorted((a, b, w)))
    F = len(faces)
    T = len(tets)
    out = np.zeros((nv + E + F + T, 3))
    out[:nv] = nodes
    for n, (hi, lo) in enumerate(edge_list):
        for d in range(3):
            out[nv + n, d] = (0.0 + nodes[lo, d] + nodes[hi, d]) / 2
    for n, fv in enumerate(faces):
        for d in range(3):
            out[nv + E + n, d] = (0.0 + nodes[fv[0], d] + nodes[fv[1], d] + nodes[fv[2], d]) / 3
    for n, t in enumerate(tets):
        for d in range(3):
            out[nv + E + F + n, d] = (0.0 + nodes[t[0], d] + nodes[t[1], d] + nodes[t[2], d] + nodes[t[3], d]) / 4
    kind = np.concatenate([np.full(nv, 1), np.full(E, 2), np.full(F, 3), np.full(T, 4)]).astype(np.int8)

    def face_of(e1, e2):
        (h1, l1), (h2, l2) = edge_list[e1], edge_list[e2]
        common = ({h1, l1} & {h2, l2}).pop()
        o1 = l1 if h1 == common else h1
        o2 = l2 if h2 == common else h2
        return ev_face[(edge_id[(max(o1, o2), min(o1, o2))], common)]

    hexes = np.zeros((4 * T, 8), np.int64)
    for n, (t, te) in enumerate(zip(tets, tet_edges)):
        for k in range(4):
            v = int(t[k])
            s = [e for e in te if v in edge_list[e]]
            h = [v, nv + s[0], nv + E + face_of(s[0], s[1]), nv + s[1], nv + s[2],
                 nv + E + face_of(s[0], s[2]), nv + E + F + n, nv + E + face_of(s[1], s[2])]
            if _center_det_lex(out[[h[_LEX_OF_UCD[m]] for m in range(8)]]) < 0:
                h = h[4:] + h[:4]
            hexes[4 * n + k] = h
    quads = None
    if tris is not None:
        tris = np.asarray(tris, np.int64)
        quads = np.zeros((3 * len(tris), 4), np.int64)
        for n, tr in enumerate(tris):
            te = [edge_id[(int(tr[i]), int(tr[j]))] for i, j in _local_edges(tr)]
            for k in range(3):
                v = int(tr[k])
                s = [e for e in te if v in edge_list[e]]
                q = [v, nv + s[0], nv + E + face_of(s[0], s[1]), nv + s[1]]
                if _measure_2d([out[q[0]], out[q[1]], out[q[3]], out[q[2]]]) < 0:
                    q[1], q[3] = q[3], q[1]
                quads[3 * n + k] = q
    return out, kind, hexes.astype(np.int32), None if quads is None else quads.astype(np.int32), E, F
