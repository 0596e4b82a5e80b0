"""Synthetic nanotip meshes in the layout the reference's mesher hands to its solvers.

The reference builds its finite-element mesh from atoms: surface atoms are coarsened,
tetgen makes a Delaunay tetrahedral mesh of surface + bulk + vacuum generator points
(src/TetgenMesh.cpp:211-274), the tets are marked vacuum/bulk, and every tet is split into
four hexahedra around its vertices (src/Tethex.cpp:965-1102).  New hex nodes are the edge
midpoints, face centroids and tet centroids, appended after the tet vertices in that order
(src/Tethex.cpp:897-918, 994-1011).  A hex carries the marker +(1+v) in vacuum and -(1+v)
in the bulk, v being the tet vertex it surrounds (src/TetgenMesh.cpp:353-385).

tetgen and the coarseners are not part of this package: the generator below makes its own
graded point cloud (coarsened FCC surface atoms plus octree-graded interior points) and
triangulates it with scipy's Delaunay (Qhull).  The tet→hex split and every array layout
follow the reference so that the solvers see the same kind of input:

* ``nodes``      (n, 3) float64: tet vertices, then edge midpoints, face centroids, tet centroids
* ``node_kind``  (n,) int8: 1 tet node, 2 edge centroid, 3 face centroid, 4 tet centroid
                 (include/Globals.h:51-54)
* ``tets``       (T, 4) int32, ``tet_marker`` (T,) 3 = vacuum, 1 = bulk (include/Globals.h:37-40)
* ``tet_nbr``    (T, 4) int32: neighbour opposite to local vertex k, -1 on the hull (tetgen's
                 neighbour-list convention, read at src/TetgenCells.cpp:605-612)
* ``hexes``      (4T, 8) int32 in the femocs/UCD node order that DealSolver::import_mesh
                 receives (src/TetgenCells.cpp:673-701), hex 4t+k surrounds vertex k of tet t
* ``hex_marker`` (4T,) int32

UCD orientation: deal.II converts UCD cells to its lexicographic numbering with the
permutation {0,1,5,4,2,3,7,6} (the same permutation the reference applies in
LinearHexahedra::shape_funs_dealii, src/InterpolatorCells.cpp:1355-1358).  A cell is valid
for deal.II when the lexicographic Jacobian is positive, i.e. det(x1-x0, x4-x0, x3-x0) > 0
in UCD order.  The split below emits every hex in that orientation.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

TYPES_BULK = 1
TYPES_VACUUM = 3
KIND_TETNODE, KIND_EDGE, KIND_FACE, KIND_TET = 1, 2, 3, 4

# UCD (femocs) -> lexicographic (deal.II) vertex permutation: lex[k] = ucd[UCD2LEX[k]]
UCD_TO_LEX = np.array([0, 1, 4, 5, 3, 2, 7, 6], dtype=np.int64)


@dataclass
class TipGeometry:
    """Cu nanotip: cylinder of radius ``radius`` and height ``height`` with a hemispherical
    cap, standing on a square slab of side ``slab_width`` and thickness ``slab_thickness``.
    The vacuum box reaches ``top_factor`` x the apex height.  Lengths in Angstrom."""
    radius: float = 30.0
    height: float = 150.0
    slab_width: float = 300.0
    slab_thickness: float = 30.0
    top_factor: float = 10.0
    lattice: float = 3.61          # Cu FCC lattice constant
    jitter: float = 0.05           # uniform +-jitter of atom coordinates
    h_min: float = 7.0             # mesh spacing at the tip surface
    growth: float = 0.5            # spacing growth per Angstrom of distance from the tip
    h_max: float = 400.0
    seed: int = 12345

    @property
    def apex(self) -> float:
        return self.height + self.radius

    @property
    def z_top(self) -> float:
        return self.top_factor * self.apex

    def box(self):
        w = 0.5 * self.slab_width
        return np.array([-w, -w, -self.slab_thickness]), np.array([w, w, self.z_top])


def _sd_tip(g: TipGeometry, p: np.ndarray) -> np.ndarray:
    rho = np.hypot(p[:, 0], p[:, 1])
    dz = p[:, 2] - g.height
    cap = np.sqrt(rho * rho + np.maximum(dz, 0.0) ** 2) - g.radius
    return np.where(dz > 0, cap, rho - g.radius)


def signed_distance(g: TipGeometry, p: np.ndarray) -> np.ndarray:
    """Signed distance (negative inside the metal) of the union tip + slab."""
    w = 0.5 * g.slab_width
    sd_slab = np.maximum.reduce([p[:, 2], -g.slab_thickness - p[:, 2],
                                 np.abs(p[:, 0]) - w, np.abs(p[:, 1]) - w])
    return np.minimum(_sd_tip(g, p), sd_slab)


def spacing(g: TipGeometry, p: np.ndarray) -> np.ndarray:
    d = np.abs(_sd_tip(g, p))
    return np.clip(g.h_min + g.growth * d, g.h_min, g.h_max)


# ----------------------------------------------------------------------------------------
# atoms
# ----------------------------------------------------------------------------------------
_FCC_BASIS = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
_FCC_NN = np.array([[sx * 0.5, sy * 0.5, 0] for sx in (-1, 1) for sy in (-1, 1)]
                   + [[sx * 0.5, 0, sz * 0.5] for sx in (-1, 1) for sz in (-1, 1)]
                   + [[0, sy * 0.5, sz * 0.5] for sy in (-1, 1) for sz in (-1, 1)])


def _lattice_in_box(a: float, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    i0 = np.floor(lo / a).astype(int) - 1
    i1 = np.ceil(hi / a).astype(int) + 1
    gi = [np.arange(i0[d], i1[d] + 1) for d in range(3)]
    cells = np.stack(np.meshgrid(*gi, indexing="ij"), -1).reshape(-1, 1, 3)
    pts = ((cells + _FCC_BASIS[None]) * a).reshape(-1, 3)
    m = np.all((pts >= lo) & (pts <= hi), axis=1)
    return pts[m]


def surface_atoms(g: TipGeometry, rng: np.random.Generator) -> np.ndarray:
    """FCC Cu atoms inside tip + slab that have at least one empty nearest-neighbour site,
    with +-jitter applied.  Returns (A, 3) float64."""
    a = g.lattice
    w = 0.5 * g.slab_width
    # only lattice sites within two lattice constants of the surface are generated
    slab = _lattice_in_box(a, np.array([-w, -w, -2.0 * a]), np.array([w, w, 0.0]))
    r = g.radius + a
    tip = _lattice_in_box(a, np.array([-r, -r, 0.0]), np.array([r, r, g.apex]))
    pts = np.unique(np.round(np.concatenate([slab, tip]) / (0.5 * a)).astype(np.int64), axis=0)
    pts = pts.astype(np.float64) * (0.5 * a)
    inside = signed_distance(g, pts) <= 1e-9
    pts = pts[inside]
    # a site is on the surface if one of its 12 neighbours lies outside the metal; the slab
    # bottom and sides are not free surfaces
    nb = pts[:, None, :] + a * _FCC_NN[None]
    nb_out = signed_distance(g, nb.reshape(-1, 3)).reshape(-1, 12) > 1e-9
    nb_z = nb[:, :, 2]
    free = nb_out & (nb_z > -g.slab_thickness) & (np.abs(nb[:, :, 0]) <= w) & (np.abs(nb[:, :, 1]) <= w)
    pts = pts[free.any(axis=1)]
    pts = pts + rng.uniform(-g.jitter, g.jitter, size=pts.shape)
    order = np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))
    return pts[order]


def _coarsen(g: TipGeometry, pts: np.ndarray) -> np.ndarray:
    """Keep one point per graded cell: cell size = spacing quantised to powers of two."""
    h = spacing(g, pts)
    level = np.floor(np.log2(h / g.h_min) + 1e-9).astype(np.int64)
    size = g.h_min * (2.0 ** level)
    key = np.floor(pts / size[:, None]).astype(np.int64)
    keys = np.concatenate([level[:, None], key], axis=1)
    _, first = np.unique(keys, axis=0, return_index=True)
    return pts[np.sort(first)]


# ----------------------------------------------------------------------------------------
# graded octree points for bulk and vacuum
# ----------------------------------------------------------------------------------------
def _octree_leaves(g: TipGeometry):
    lo, hi = g.box()
    ext = hi - lo
    root = min(g.h_max, ext.min())
    n = np.maximum(1, np.round(ext / root)).astype(int)
    size0 = ext / n
    grid = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3)
    centers = lo + (grid + 0.5) * size0
    sizes = np.tile(size0, (len(centers), 1))
    leaves_c, leaves_s = [], []
    for _ in range(40):
        rad = 0.5 * np.linalg.norm(sizes, axis=1)
        near = np.maximum(np.abs(_sd_tip(g, centers)) - rad, 0.0)
        target = np.clip(g.h_min + g.growth * near, g.h_min, g.h_max)
        split = sizes.max(axis=1) > 1.5 * target
        leaves_c.append(centers[~split])
        leaves_s.append(sizes[~split])
        if not split.any():
            break
        c, s = centers[split], sizes[split] * 0.5
        offs = np.array([[i, j, k] for i in (-0.5, 0.5) for j in (-0.5, 0.5) for k in (-0.5, 0.5)])
        centers = (c[:, None, :] + offs[None] * s[:, None, :]).reshape(-1, 3)
        sizes = np.repeat(s, 8, axis=0)
    return np.concatenate(leaves_c), np.concatenate(leaves_s)


def generator_points(g: TipGeometry, atoms: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    lo, hi = g.box()
    c, s = _octree_leaves(g)
    hs = s.max(axis=1)
    sd = signed_distance(g, c)
    keep = np.abs(sd) > 0.75 * hs
    interior = c[keep] + rng.uniform(-0.15, 0.15, size=(keep.sum(), 3)) * s[keep]
    # box boundary: project leaves touching a face onto it (edges and corners follow)
    bpts = []
    for d in range(3):
        for side, val in ((0, lo[d]), (1, hi[d])):
            touch = np.abs((c[:, d] - 0.5 * s[:, d] * (1 if side == 0 else -1)) - val) < 1e-6 * (1 + abs(val))
            # tangential jitter keeps the face points in general position (a regular grid
            # of them makes cospherical quadruples and flat tets inside the box)
            p = c[touch] + rng.uniform(-0.15, 0.15, size=(touch.sum(), 3)) * s[touch]
            p[:, d] = val
            bpts.append(p)
    bp = np.concatenate(bpts)
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    surf = _coarsen(g, atoms)
    # keep surface generators out of the box faces (slab rim atoms sit on the box sides)
    surf = surf[np.all((surf > lo + 0.25 * g.h_min) & (surf < hi - 0.25 * g.h_min), axis=1)]
    # interior points must not crowd the box faces either
    inner = interior[np.all((interior > lo + 0.3 * g.h_min) & (interior < hi - 0.3 * g.h_min), axis=1)]
    pts = np.round(np.concatenate([surf, inner, bp, corners]), 9)
    # exact duplicates (box edges are touched from two faces) are merged; the first copy is
    # kept in place, so the surface generators stay in front
    _, first = np.unique(pts, axis=0, return_index=True)
    first = np.sort(first)
    return pts[first], int(np.sum(first < len(surf)))


# ----------------------------------------------------------------------------------------
# tetrahedral mesh and the tet -> hex split
# ----------------------------------------------------------------------------------------
def _tet_volumes(x: np.ndarray, tets: np.ndarray) -> np.ndarray:
    a, b, c, d = (x[tets[:, k]] for k in range(4))
    return np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a)) / 6.0


def _face_neighbours(tets: np.ndarray) -> np.ndarray:
    """nbr[t, k] = tet sharing the face opposite local vertex k, or -1."""
    T = len(tets)
    opp = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    faces = np.sort(tets[:, opp], axis=2).reshape(-1, 3)          # (4T, 3)
    owner = np.repeat(np.arange(T), 4)
    local = np.tile(np.arange(4), T)
    key = (faces[:, 0].astype(np.int64) * (1 << 42) + faces[:, 1].astype(np.int64) * (1 << 21)
           + faces[:, 2].astype(np.int64))
    order = np.argsort(key, kind="stable")
    ks = key[order]
    same = ks[1:] == ks[:-1]
    nbr = -np.ones((T, 4), dtype=np.int32)
    i = np.nonzero(same)[0]
    a, b = order[i], order[i + 1]
    nbr[owner[a], local[a]] = owner[b]
    nbr[owner[b], local[b]] = owner[a]
    return nbr


@dataclass
class TipMesh:
    geometry: TipGeometry
    nodes: np.ndarray
    node_kind: np.ndarray
    n_tetnodes: int
    n_surface_nodes: int
    tets: np.ndarray
    tet_marker: np.ndarray
    tet_nbr: np.ndarray
    hexes: np.ndarray
    hex_marker: np.ndarray
    atoms: np.ndarray = field(default=None)

    # --- exports in the form DealSolver::import_mesh receives them
    def vacuum_cells(self) -> np.ndarray:
        """UCD cells of the vacuum hexes (marker > 0), in hex order
        (Hexahedra::export_vacuum, src/TetgenCells.cpp:673-686)."""
        return self.hexes[self.hex_marker > 0]

    def bulk_cells(self) -> np.ndarray:
        """UCD cells of the bulk hexes (marker < 0), src/TetgenCells.cpp:688-701."""
        return self.hexes[self.hex_marker < 0]

    def surface_triangles(self):
        """Faces between a vacuum and a bulk tet (the surface triangles the mesher exports):
        (triangles (n, 3) as the vacuum tet orders the face, tri2tet (n, 2) = (vacuum tet,
        bulk tet)), in vacuum-tet order."""
        tets, nbr, mk = self.tets, self.tet_nbr, self.tet_marker
        opp = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
        t, k = np.nonzero((mk[:, None] == TYPES_VACUUM) & (nbr >= 0))
        other = nbr[t, k]
        sel = mk[other] != TYPES_VACUUM
        t, k, other = t[sel], k[sel], other[sel]
        tris = tets[t[:, None], opp[k]]
        return tris.astype(np.int32), np.stack([t, other], 1).astype(np.int32)

    def surface_quads(self, tris):
        """The 3 quadrangles each face splits into (quad 3f+k around face vertex k): (vertex k,
        midpoint of edge (k, k+1), face centroid, midpoint of edge (k, k+2)), as hex-mesh nodes."""
        nv = self.n_tetnodes
        _, ev, _, fv = _tet_entities(nv, self.tets)
        ekey = ev[:, 0] * nv + ev[:, 1]
        fkey = (fv[:, 0] * nv + fv[:, 1]) * nv + fv[:, 2]
        tris = np.asarray(tris, np.int64)
        fs = np.sort(tris, 1)
        face = len(ev) + nv + np.searchsorted(fkey, (fs[:, 0] * nv + fs[:, 1]) * nv + fs[:, 2])

        def edge(a, b):
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            return nv + np.searchsorted(ekey, lo * nv + hi)

        q = np.empty((len(tris), 3, 4), np.int64)
        for k in range(3):
            a, b, c = tris[:, k], tris[:, (k + 1) % 3], tris[:, (k + 2) % 3]
            q[:, k] = np.stack([a, edge(a, b), face, edge(a, c)], 1)
        return q.reshape(-1, 4).astype(np.int32)

    @property
    def n_tets(self) -> int:
        return len(self.tets)

    @property
    def n_hexes(self) -> int:
        return len(self.hexes)


_PAIRS = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]])


def _tet_entities(nv: int, tets: np.ndarray):
    """Unique edges and faces of a tet mesh in sorted-key order: (tet_edges [T,6], edge vertices
    [E,2], tet_faces [T,4] (face opposite local vertex k), face vertices [F,3])."""
    T = len(tets)
    e = np.sort(tets[:, _PAIRS], axis=2).reshape(-1, 2).astype(np.int64)
    ekey = e[:, 0] * nv + e[:, 1]
    uek, einv = np.unique(ekey, return_inverse=True)
    opp = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    f = np.sort(tets[:, opp], axis=2).reshape(-1, 3).astype(np.int64)
    fkey = (f[:, 0] * nv + f[:, 1]) * nv + f[:, 2]
    ufk, finv = np.unique(fkey, return_inverse=True)
    ev = np.stack([uek // nv, uek % nv], 1)
    fv = np.stack([ufk // (nv * nv), (ufk // nv) % nv, ufk % nv], 1)
    return einv.reshape(T, 6), ev, finv.reshape(T, 4), fv


def hex_node_weights(n_tetnodes: int, tets: np.ndarray):
    """Sparse map from tet-vertex positions to all hex nodes (vertices, edge midpoints, face and
    tet centroids, in the order tet_to_hex emits them): moving tet vertices by D moves the hex
    nodes by W @ D, as Tethex set_new_vertices places them (src/Tethex.cpp:965-1102)."""
    from scipy.sparse import csr_matrix

    _, ev, _, fv = _tet_entities(n_tetnodes, tets)
    T, E, F = len(tets), len(ev), len(fv)
    rows = np.concatenate([np.arange(n_tetnodes), np.repeat(n_tetnodes + np.arange(E), 2),
                           np.repeat(n_tetnodes + E + np.arange(F), 3),
                           np.repeat(n_tetnodes + E + F + np.arange(T), 4)])
    cols = np.concatenate([np.arange(n_tetnodes), ev.ravel(), fv.ravel(), np.asarray(tets).ravel()])
    vals = np.concatenate([np.ones(n_tetnodes), np.full(2 * E, 0.5), np.full(3 * F, 1.0 / 3.0),
                           np.full(4 * T, 0.25)])
    n = n_tetnodes + E + F + T
    return csr_matrix((vals, (rows, cols)), shape=(n, n_tetnodes))


def tet_to_hex(nodes: np.ndarray, tets: np.ndarray):
    """Split every tet into four hexes around its vertices (Tethex convert_3D,
    src/Tethex.cpp:965-1102).  Returns (all_nodes, node_kind, hexes_ucd)."""
    T = len(tets)
    nv = len(nodes)
    pairs = _PAIRS
    tet_edges, ev, tet_faces, fv = _tet_entities(nv, tets)
    E, F = len(ev), len(fv)
    # new nodes: sum of vertices divided by the count, as Tethex set_new_vertices
    emid = (nodes[ev[:, 0]] + nodes[ev[:, 1]]) / 2.0
    fcen = (nodes[fv[:, 0]] + nodes[fv[:, 1]] + nodes[fv[:, 2]]) / 3.0
    tcen = (nodes[tets[:, 0]] + nodes[tets[:, 1]] + nodes[tets[:, 2]] + nodes[tets[:, 3]]) / 4.0
    all_nodes = np.concatenate([nodes, emid, fcen, tcen])
    kind = np.concatenate([np.full(nv, KIND_TETNODE), np.full(E, KIND_EDGE),
                           np.full(F, KIND_FACE), np.full(T, KIND_TET)]).astype(np.int8)
    e0, f0, c0 = nv, nv + E, nv + E + F
    # local edge index for vertex pair (i, j)
    eidx = -np.ones((4, 4), dtype=np.int64)
    for n, (i, j) in enumerate(pairs):
        eidx[i, j] = eidx[j, i] = n
    hexes = np.empty((T, 4, 8), dtype=np.int64)
    for k in range(4):
        j1, j2, j3 = [j for j in range(4) if j != k]
        ea = e0 + tet_edges[:, eidx[k, j1]]
        eb = e0 + tet_edges[:, eidx[k, j2]]
        ec = e0 + tet_edges[:, eidx[k, j3]]
        fab = f0 + tet_faces[:, j3]   # face with k, j1, j2 is opposite j3
        fac = f0 + tet_faces[:, j2]
        fbc = f0 + tet_faces[:, j1]
        cc = c0 + np.arange(T)
        v = tets[:, k].astype(np.int64)
        h = np.stack([v, ea, fab, eb, ec, fac, cc, fbc], 1)
        # orientation: lexicographic order (u, w, v) must be right handed, i.e.
        # det(x1-x0, x4-x0, x3-x0) > 0; otherwise swap the roles of the a and b edges
        x = all_nodes
        d = np.einsum("ij,ij->i", x[h[:, 1]] - x[h[:, 0]],
                      np.cross(x[h[:, 4]] - x[h[:, 0]], x[h[:, 3]] - x[h[:, 0]]))
        flip = d < 0
        hf = np.stack([v, eb, fab, ea, ec, fbc, cc, fac], 1)
        h[flip] = hf[flip]
        hexes[:, k, :] = h
    return all_nodes, kind, hexes.reshape(-1, 8).astype(np.int32)


def make_tip_mesh(g: TipGeometry | None = None, with_atoms: bool = True) -> TipMesh:
    from scipy.spatial import Delaunay

    g = g or TipGeometry()
    rng = np.random.default_rng(g.seed)
    atoms = surface_atoms(g, rng)
    pts, n_surf = generator_points(g, atoms, rng)
    tri = Delaunay(pts, qhull_options="Qbb Qc Qz Q12")
    tets = tri.simplices.astype(np.int64)
    vol = _tet_volumes(pts, tets)
    # drop the flat tets Qhull leaves on the planar box faces (all four vertices on one
    # face); interior slivers are kept, removing them would punch holes into the mesh
    lo, hi = g.box()
    tp = pts[tets]
    flat = np.zeros(len(tets), bool)
    for d in range(3):
        flat |= np.all(tp[:, :, d] == lo[d], axis=1) | np.all(tp[:, :, d] == hi[d], axis=1)
    keep = ~flat
    if np.any(np.abs(vol[keep]) == 0.0):
        raise RuntimeError("degenerate interior tetrahedron")
    tets = tets[keep]
    vol = vol[keep]
    neg = vol < 0
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    # keep only vertices that are used, in their order (tet vertices come first)
    used = np.zeros(len(pts), bool)
    used[tets.ravel()] = True
    remap = -np.ones(len(pts), np.int64)
    remap[used] = np.arange(used.sum())
    pts = pts[used]
    tets = remap[tets]
    cen = pts[tets].mean(axis=1)
    tet_marker = np.where(signed_distance(g, cen) < 0, TYPES_BULK, TYPES_VACUUM).astype(np.int32)
    nbr = _face_neighbours(tets)
    nodes, kind, hexes = tet_to_hex(pts, tets)
    # hex marker: +(1 + tet node) in vacuum, -(1 + tet node) in bulk
    vnode = tets.reshape(-1)
    hm = np.where(np.repeat(tet_marker, 4) == TYPES_VACUUM, 1 + vnode, -1 - vnode).astype(np.int32)
    return TipMesh(geometry=g, nodes=nodes, node_kind=kind, n_tetnodes=len(pts),
                   n_surface_nodes=int(used[:n_surf].sum()), tets=tets.astype(np.int32),
                   tet_marker=tet_marker, tet_nbr=nbr, hexes=hexes, hex_marker=hm,
                   atoms=atoms if with_atoms else None)


def tip_r3nm_geometry(**kw) -> TipGeometry:
    """Nanotip r = 3 nm, h = 15 nm on a 30 x 30 x 3 nm slab, top at 10 x apex height."""
    return TipGeometry(radius=30.0, height=150.0, slab_width=300.0, slab_thickness=30.0, **kw)


def tip_r10nm_geometry(**kw) -> TipGeometry:
    """Nanotip r = 10 nm, h = 60 nm on a 120 x 120 nm slab, top at 10 x apex height."""
    return TipGeometry(radius=100.0, height=600.0, slab_width=1200.0, slab_thickness=60.0, **kw)


def batched_tip_geometry(i: int) -> TipGeometry:
    """Tip i of the batched workload of 64 tips: radius U[2, 12] nm, aspect ratio U[3, 10],
    slab 12 radii wide, its own lattice jitter seed (65k-2.3M vacuum hexes)."""
    rng = np.random.default_rng(4242 + 7919 * i)
    r = rng.uniform(20.0, 120.0)
    a = rng.uniform(3.0, 10.0)
    return TipGeometry(radius=r, height=r * a, slab_width=12 * r, slab_thickness=30.0 if r < 60 else 60.0,
                       seed=20000 + i)


def small_geometry(**kw) -> TipGeometry:
    """A small tip for tests that run the CPU oracle in seconds."""
    for k, v in (("h_min", 6.0), ("growth", 0.5), ("radius", 12.0), ("height", 30.0),
                 ("slab_width", 90.0), ("slab_thickness", 12.0), ("top_factor", 4.0)):
        kw.setdefault(k, v)
    return TipGeometry(**kw)
