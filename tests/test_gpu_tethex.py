"""The reference's tet -> 4-hex split on the device against the oracle's sequential
restatement: nodes, hexes and boundary quadrangles bit for bit on a generated tip and on the
reference's own mesh file; a triangle that is no face of the tets is refused."""
import os

import numpy as np
import pytest

from femocs_amd import fgpu, gmsh
from femocs_amd import mesh as M
from oracle import tethex

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _compare(ctx, nodes, tets, tris):
    ref = tethex.split(nodes, tets, tris)
    got = fgpu.tethex_split(ctx, nodes, tets, tris)
    assert got[4:] == ref[4:]                           # numbers of edges and faces
    assert np.array_equal(got[0], ref[0])               # coordinates, bit for bit
    assert np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2], ref[2])               # hexes in the reference's vertex order
    assert np.array_equal(got[3], ref[3])               # quadrangles
    return got


def test_split_of_a_generated_tip(ctx):
    tip = M.make_tip_mesh(M.small_geometry(), with_atoms=False)
    tris, _ = tip.surface_triangles()
    got = _compare(ctx, tip.nodes[:tip.n_tetnodes], tip.tets, tris)
    assert len(got[2]) == 4 * len(tip.tets)


def test_split_of_the_reference_mesh_file(ctx):
    nodes, tets, tags, tris, _ = gmsh.read_msh(os.path.join(HERE, "golden", "hemicone.msh"))
    used = np.unique(tets)
    remap = -np.ones(len(nodes), np.int64)
    remap[used] = np.arange(len(used))
    nodes, tets = nodes[used], remap[tets]
    tris = remap[tris]
    tris = tris[np.all(tris >= 0, axis=1)]
    _compare(ctx, nodes, tets, tris)


def test_split_refuses_a_foreign_triangle(ctx):
    tip = M.make_tip_mesh(M.small_geometry(), with_atoms=False)
    t = tip.tets[0]
    far = tip.tets[-1]
    bad = np.array([[t[0], t[1], far[3] if far[3] not in t else far[2]]])
    with pytest.raises(fgpu.FgpuError):
        fgpu.tethex_split(ctx, tip.nodes[:tip.n_tetnodes], tip.tets, bad)
    with pytest.raises(fgpu.FgpuError):
        fgpu.tethex_split(ctx, tip.nodes[:tip.n_tetnodes], np.array([[0, 0, 1, 2]]), None)
