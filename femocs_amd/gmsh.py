"""Reader of Gmsh 2.2 ASCII tetrahedral meshes, the format the reference's TetgenMesh::read
takes through Tethex (src/TetgenMesh.cpp:276-306).  Tets carry their physical tag as the
region marker (3 = vacuum, 1 = bulk, include/Globals.h:37-40); the tet -> hex split and the
hex markers follow femocs_amd.mesh.
"""
from __future__ import annotations

import numpy as np

from . import mesh as M


def read_msh(path: str):
    """Returns (nodes (n,3), tets (T,4) 0-based, tet physical tags, tris (F,3), tri tags)."""
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    nodes = tets = None
    tet_tag, tris, tri_tag = [], [], []
    tl = []
    while i < len(lines):
        s = lines[i].strip()
        if s == "$Nodes":
            n = int(lines[i + 1])
            arr = np.array([ln.split() for ln in lines[i + 2:i + 2 + n]], dtype=np.float64)
            ids = arr[:, 0].astype(np.int64)
            nodes = np.zeros((ids.max() + 1, 3))
            nodes[ids] = arr[:, 1:4]
            node_ids = ids
            i += 2 + n
            continue
        if s == "$Elements":
            n = int(lines[i + 1])
            for ln in lines[i + 2:i + 2 + n]:
                v = [int(x) for x in ln.split()]
                etype, ntags = v[1], v[2]
                tag = v[3] if ntags > 0 else 0
                conn = v[3 + ntags:]
                if etype == 4:
                    tl.append(conn)
                    tet_tag.append(tag)
                elif etype == 2:
                    tris.append(conn)
                    tri_tag.append(tag)
            i += 2 + n
            continue
        i += 1
    # compact node numbering (Gmsh ids are 1-based and may have gaps)
    remap = -np.ones(nodes.shape[0], np.int64)
    remap[node_ids] = np.arange(len(node_ids))
    nodes = nodes[node_ids]
    tets = remap[np.array(tl, dtype=np.int64)]
    tris = remap[np.array(tris, dtype=np.int64)] if tris else np.zeros((0, 3), np.int64)
    return nodes, tets, np.array(tet_tag, np.int32), tris, np.array(tri_tag, np.int32)


def tip_mesh_from_msh(path: str) -> M.TipMesh:
    """Tet mesh from a Gmsh file, split into hexahedra like the reference's mesher."""
    nodes, tets, tag, _, _ = read_msh(path)
    vol = M._tet_volumes(nodes, tets)
    neg = vol < 0
    tets = tets.copy()
    tets[neg] = tets[neg][:, [0, 2, 1, 3]]
    marker = np.where(tag == M.TYPES_VACUUM, M.TYPES_VACUUM, M.TYPES_BULK).astype(np.int32)
    nbr = M._face_neighbours(tets)
    all_nodes, kind, hexes = M.tet_to_hex(nodes, tets)
    vnode = tets.reshape(-1)
    hm = np.where(np.repeat(marker, 4) == M.TYPES_VACUUM, 1 + vnode, -1 - vnode).astype(np.int32)
    return M.TipMesh(geometry=None, nodes=all_nodes, node_kind=kind, n_tetnodes=len(nodes), n_surface_nodes=0,
                     tets=tets.astype(np.int32), tet_marker=marker, tet_nbr=nbr, hexes=hexes, hex_marker=hm)
