"""Q2 oracle (test infrastructure, never part of the product path): the Laplace solve of
PoissonSolver with the reference's element degree raised to 2, i.e. a recompile of
include/DealSolver.h:130-131 with shape_degree = 2: FE_Q(2) on the same hexes, QGauss(3)
(quadrature_degree = shape_degree + 1) for cells and QGauss<2>(3) for boundary faces, the
trilinear (MappingQ1) geometry of src/DealSolver.cpp:286,345,392.  The reference holds no Q2
results; this restatement is pinned by the textbook Q2 matrices (tensor products of the 1D
stiffness and mass on [0, 1]) and by patch tests (tests/test_q2_oracle.py).

Nodes of a cell are numbered lexicographically on the 3 x 3 x 3 grid (i + 3 j + 9 k, node
coordinate 0, 1/2, 1 along each axis); the vertex (a, b, c) of the lexicographic Q1 cell is
node 2a + 6b + 18c.  Global dofs: the Q1 dofs (mesh vertices, the import's numbering) first,
then the edges, the faces and the cell interiors, each set in order of first appearance over
the cells.  Dof numbering is internal; results are compared at support points.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

G3 = np.array([0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)])   # QGauss(3) on [0, 1]
W3 = np.array([5.0, 8.0, 5.0]) / 18.0


def basis_1d(x):
    """Q2 Lagrange basis on the nodes 0, 1/2, 1 and its derivative, at the points x."""
    x = np.asarray(x, np.float64)
    L = np.stack([(1 - x) * (1 - 2 * x), 4 * x * (1 - x), x * (2 * x - 1)], -1)
    dL = np.stack([4 * x - 3, 4 - 8 * x, 4 * x - 1], -1)
    return L, dL


# lexicographic node index of the Q1 vertex (a, b, c)
VERTEX_NODE = np.array([2 * a + 6 * b + 18 * c for c in (0, 1) for b in (0, 1) for a in (0, 1)])


def _node_tuple(n):
    return n % 3, (n // 3) % 3, n // 9


def cell_dofs(cells_lex: np.ndarray, n_vertices: int):
    """(cells27 (H, 27), n_dofs): Q2 dofs of the hexes given by their lexicographic Q1 vertices."""
    H = len(cells_lex)
    c27 = -np.ones((H, 27), np.int64)
    c27[:, VERTEX_NODE] = cells_lex
    nxt = n_vertices
    for kind in (1, 2, 3):          # number of non-vertex coordinates: edge, face, interior
        nodes = [n for n in range(27) if sum(1 for t in _node_tuple(n) if t == 1) == kind]
        keys = []
        for n in nodes:
            t = _node_tuple(n)
            # the vertices spanned by the entity: coordinate 1 -> both 0 and 2
            opts = [(0, 1) if v == 1 else (v // 2,) for v in t]
            vs = [cells_lex[:, a + 2 * b + 4 * c] for a in opts[0] for b in opts[1] for c in opts[2]]
            keys.append(np.sort(np.stack(vs, 1), axis=1))
        key = np.stack(keys, 1).reshape(H * len(nodes), -1)      # cell-major, then node order
        _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
        order = np.argsort(first, kind="stable")                 # first appearance
        rank = np.empty(len(order), np.int64)
        rank[order] = np.arange(len(order))
        c27[:, nodes] = (nxt + rank[inv.reshape(-1)]).reshape(H, len(nodes))
        nxt += len(order)
    return c27, nxt


def _geometry(xe, xi):
    """Jacobian (3x3, columns d x / d xi_a) of the trilinear map of cell vertices xe (8, 3,
    lexicographic) at reference point xi."""
    J = np.zeros((3, 3))
    for n in range(8):
        a, b, c = n & 1, (n >> 1) & 1, (n >> 2) & 1
        N = [xi[0] if a else 1 - xi[0], xi[1] if b else 1 - xi[1], xi[2] if c else 1 - xi[2]]
        d = [1 if a else -1, 1 if b else -1, 1 if c else -1]
        g = np.array([d[0] * N[1] * N[2], N[0] * d[1] * N[2], N[0] * N[1] * d[2]])
        J += np.outer(xe[n], g)
    return J


def element_matrix(xe: np.ndarray) -> np.ndarray:
    """27 x 27 stiffness sum_q JxW grad phi_i . grad phi_j, QGauss(3), trilinear geometry."""
    K = np.zeros((27, 27))
    L, dL = basis_1d(G3)
    for qk in range(3):
        for qj in range(3):
            for qi in range(3):
                xi = (G3[qi], G3[qj], G3[qk])
                J = _geometry(xe, xi)
                det = np.linalg.det(J)
                Jit = np.linalg.inv(J).T
                gref = np.zeros((27, 3))
                for n in range(27):
                    i, j, k = _node_tuple(n)
                    gref[n] = [dL[qi, i] * L[qj, j] * L[qk, k], L[qi, i] * dL[qj, j] * L[qk, k],
                               L[qi, i] * L[qj, j] * dL[qk, k]]
                g = gref @ Jit.T
                K += W3[qi] * W3[qj] * W3[qk] * det * (g @ g.T)
    return K


# the 9 nodes of deal.II face f (coordinate f // 2 at its side f % 2), in lexicographic order
FACE_NODES = [[n for n in range(27) if _node_tuple(n)[f // 2] == 2 * (f % 2)] for f in range(6)]


def face_load(xe, face, value):
    """sum_q phi_i g |dx/ds x dx/dt| w over face `face` (QGauss<2>(3)) for its 9 nodes."""
    d = face // 2
    axes = [a for a in range(3) if a != d]
    L, dL = basis_1d(G3)
    out = np.zeros(9)
    nodes = FACE_NODES[face]
    for qt in range(3):
        for qs in range(3):
            xi = [0.0, 0.0, 0.0]
            xi[d] = float(face % 2)
            xi[axes[0]], xi[axes[1]] = G3[qs], G3[qt]
            J = _geometry(xe, xi)
            area = np.linalg.norm(np.cross(J[:, axes[0]], J[:, axes[1]]))
            for m, n in enumerate(nodes):
                t = _node_tuple(n)
                phi = L[qs, t[axes[0]]] * L[qt, t[axes[1]]]
                out[m] += W3[qs] * W3[qt] * area * phi * value
    return out


def support_points(x_vertices, cells_lex, c27, n_dofs):
    """Coordinates of every Q2 dof: vertices, then means of the entity's vertices."""
    P = np.zeros((n_dofs, 3))
    for n in range(27):
        i, j, k = _node_tuple(n)
        opts = [(0, 1) if v == 1 else (v // 2,) for v in (i, j, k)]
        vs = [cells_lex[:, a + 2 * b + 4 * c] for a in opts[0] for b in opts[1] for c in opts[2]]
        P[c27[:, n]] = np.mean(x_vertices[np.stack(vs, 1)], axis=1)
    return P


def laplace(om, applied_field):
    """Q2 PoissonSolver::assemble + solve (laplace mode) on the oracle's Q1 import `om`
    (oracle.Mesh with marks top 8, surface 2, sides 4): phi = 0 on every dof of the id-2
    faces, Neumann applied_field on id 8, direct sparse solve.  Returns (c27, support points,
    solution)."""
    cells = om.cells
    c27, n = cell_dofs(cells, om.n_dofs)
    rows, cols, vals = [], [], []
    for c in range(len(cells)):
        K = element_matrix(om.x[cells[c]])
        rows.append(np.repeat(c27[c], 27))
        cols.append(np.tile(c27[c], 27))
        vals.append(K.reshape(-1))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    b = np.zeros(n)
    mask = np.zeros(n, bool)
    for f in range(len(om.bface_id)):
        c, face, bid = om.bface_cell[f], om.bface_face[f], om.bface_id[f]
        dofs = c27[c, FACE_NODES[face]]
        if bid == 8:
            b[dofs] += face_load(om.x[cells[c]], face, applied_field)
        elif bid == 2:
            mask[dofs] = True
    free = ~mask
    x = np.zeros(n)
    x[free] = spla.spsolve(A[free][:, free].tocsc(), b[free])
    return c27, support_points(om.x, cells, c27, n), x
