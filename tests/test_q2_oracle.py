"""Pins the Q2 oracle (oracle/q2.py): the unit-cube element matrix equals the textbook tensor
product of the 1D Q2 stiffness and mass; symmetry, K 1 = 0 and u^T K u = |a|^2 vol for linear u
on distorted cells; the dof numbering shares every edge, face and vertex dof between
neighbours; patch tests: a linear field on a trilinear (non-affine) mesh and a harmonic
quadratic on an affine mesh are reproduced exactly from their boundary values."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import q2

K1 = np.array([[7.0, -8.0, 1.0], [-8.0, 16.0, -8.0], [1.0, -8.0, 7.0]]) / 3.0
M1 = np.array([[4.0, 2.0, -1.0], [2.0, 16.0, 2.0], [-1.0, 2.0, 4.0]]) / 30.0

UNIT = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], float)


def test_unit_cube_matches_tensor_product():
    K = q2.element_matrix(UNIT)
    ref = np.kron(M1, np.kron(M1, K1)) + np.kron(M1, np.kron(K1, M1)) + np.kron(K1, np.kron(M1, M1))
    assert np.abs(K - ref).max() <= 1e-13


def test_distorted_cell_invariants():
    rng = np.random.default_rng(3)
    xe = UNIT * [2.0, 1.0, 1.5] + rng.uniform(-0.15, 0.15, UNIT.shape)
    K = q2.element_matrix(xe)
    assert np.abs(K - K.T).max() <= 1e-13 * np.abs(K).max()
    assert np.abs(K @ np.ones(27)).max() <= 1e-12 * np.abs(K).max()
    c27 = np.arange(27)[None]
    P = q2.support_points(xe, np.arange(8)[None], c27, 27)
    a = np.array([0.3, -1.2, 0.7])
    u = P @ a
    # volume of the trilinear cell by QGauss(3) (exact for it)
    vol = 0.0
    for qk in range(3):
        for qj in range(3):
            for qi in range(3):
                vol += q2.W3[qi] * q2.W3[qj] * q2.W3[qk] * np.linalg.det(q2._geometry(xe, (q2.G3[qi], q2.G3[qj], q2.G3[qk])))
    assert abs(u @ K @ u - (a @ a) * vol) <= 1e-12 * (a @ a) * vol


def _grid(n, distort=0.0, affine=None, seed=0):
    """n^3 hexes of [0,1]^3 (lexicographic cells), interior vertices jittered by distort * h."""
    g = np.linspace(0.0, 1.0, n + 1)
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    idx = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k
    cells = np.array([[idx(i + a, j + b, k + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
                      for i in range(n) for j in range(n) for k in range(n)])
    if distort:
        inner = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
        X[inner] += np.random.default_rng(seed).uniform(-distort, distort, (inner.sum(), 3)) / n
    if affine is not None:
        X = X @ affine.T
    return X, cells


def test_numbering_shares_entities():
    X, cells = _grid(3)
    c27, n = q2.cell_dofs(cells, len(X))
    assert n == 7 ** 3                      # (2n + 1)^3 support points of a structured grid
    P = q2.support_points(X, cells, c27, n)
    assert len(np.unique(np.round(P, 12), axis=0)) == n
    assert np.array_equal(c27[:, q2.VERTEX_NODE], cells)


def _patch(X, cells, f, boundary):
    c27, n = q2.cell_dofs(cells, len(X))
    P = q2.support_points(X, cells, c27, n)
    rows, cols, vals = [], [], []
    for c in range(len(cells)):
        K = q2.element_matrix(X[cells[c]])
        rows.append(np.repeat(c27[c], 27))
        cols.append(np.tile(c27[c], 27))
        vals.append(K.reshape(-1))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    bd = boundary(P)
    u = np.zeros(n)
    u[bd] = f(P[bd])
    free = ~bd
    rhs = -(A[free][:, bd] @ u[bd])
    u[free] = spla.spsolve(A[free][:, free].tocsc(), rhs)
    return u, f(P)


def test_linear_patch_on_trilinear_mesh():
    X, cells = _grid(4, distort=0.25, seed=5)
    a = np.array([1.0, -2.0, 0.5])
    box = lambda P: np.any((P < 1e-12) | (P > 1 - 1e-12), axis=1)
    u, ref = _patch(X, cells, lambda P: P @ a + 0.25, box)
    assert np.abs(u - ref).max() <= 1e-11


def test_harmonic_quadratic_on_affine_mesh():
    T = np.array([[1.0, 0.2, 0.0], [0.0, 0.9, 0.1], [0.1, 0.0, 1.1]])
    X, cells = _grid(3, affine=T)
    Ti = np.linalg.inv(T)
    box = lambda P: np.any(((P @ Ti.T) < 1e-12) | ((P @ Ti.T) > 1 - 1e-12), axis=1)
    f = lambda P: P[:, 0] ** 2 - P[:, 1] ** 2 + 0.5 * P[:, 0] * P[:, 2]
    u, ref = _patch(X, cells, f, box)
    assert np.abs(u - ref).max() <= 1e-11


def test_face_load_closed_form():
    """On a unit-cube face the load of dof (s, t) is g * w_s * w_t with the integrals 1/6,
    2/3, 1/6 of the 1D Q2 basis; on a sheared face the total is g * area."""
    w = np.array([1.0, 4.0, 1.0]) / 6.0
    for face in range(6):
        f = q2.face_load(UNIT, face, 2.5)
        axes = [a for a in range(3) if a != face // 2]
        ref = [2.5 * w[q2._node_tuple(n)[axes[0]]] * w[q2._node_tuple(n)[axes[1]]] for n in q2.FACE_NODES[face]]
        assert np.abs(f - ref).max() <= 1e-14
    xe = UNIT.copy()
    xe[[4, 5, 6, 7], 0] += 0.3            # shear the top face (face 5) sideways: area unchanged
    assert abs(q2.face_load(xe, 5, 1.5).sum() - 1.5) <= 1e-14
