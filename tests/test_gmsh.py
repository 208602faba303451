"""The Gmsh 4.1 reader on the reference's own meshes (mesh/carre.msh, carre_hybrid.msh,
rectangle.msh, cube.msh, carre_structure.msh, rectangle_RTM.msh, cube_bis.msh, kept gzipped under
tests/golden/meshes): node tags follow
define_ghost_node's priority (partitions/partitions_utils.py:16-111), checked here against the
geometry of each mesh; cell volumes add up to the polygon/box area; the element layout is
convert_2d/3d_cons_to_array's (partitions_utils.py:113-185)."""
import os

import numpy as np
import pytest

from manapy_amd.gmsh import read_gmsh, domain_from_gmsh

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, 'golden', 'meshes')


def path(name):
    return os.path.join(MESHES, name + '.msh.gz')


def _expected_tags(nd, box):
    """Priority of define_ghost_node on a box: x ends (1, 2) over y ends (3, 4) over z ends (5, 6)."""
    x, y, z = nd[:, 0], nd[:, 1], nd[:, 2]
    t = np.zeros(len(nd))
    (x0, x1), (y0, y1) = box[0], box[1]
    if len(box) == 3:
        z0, z1 = box[2]
        t[np.isclose(z, z0)] = 6
        t[np.isclose(z, z1)] = 5
    t[np.isclose(y, y0)] = 4
    t[np.isclose(y, y1)] = 3
    t[np.isclose(x, x1)] = 2
    t[np.isclose(x, x0)] = 1
    return t


@pytest.mark.parametrize("name,dim,box", [("carre", 2, [(0, 1), (0, 1)]), ("rectangle", 2, None),
                                          ("cube", 3, [(0, 1), (0, 1), (0, 1)])])
def test_boundary_tags_follow_the_geometry(name, dim, box):
    el, nd = read_gmsh(path(name), dim)
    if box is None:
        box = [(nd[:, 0].min(), nd[:, 0].max()), (nd[:, 1].min(), nd[:, 1].max())]
    assert np.array_equal(nd[:, 3], _expected_tags(nd, box))
    assert el.dtype == np.int32 and np.all(el[:, -1] == dim + 1)
    assert el[:, :-1].min() == 0 and el[:, :-1].max() == len(nd) - 1


def test_hybrid_mesh_layout_and_area():
    """carre_hybrid: triangles first, then quadrangles (padding 0), area = the polygon areas
    (shoelace) of the elements, which a Domain's cell volumes reproduce; the outer boundary
    of [0, 1] x [0, 0.5] carries the box tags and the holes' boundaries their own."""
    el, nd = read_gmsh(path('carre_hybrid'), 2)
    assert el.shape[1] == 5
    n = el[:, -1]
    ntri = int(np.sum(n == 3))
    assert ntri > 0 and np.all(n[:ntri] == 3) and np.all(n[ntri:] == 4)
    assert np.all(el[:ntri, 3] == 0)
    area = 0.
    for r in el:
        p = nd[r[:r[-1]], :2]
        area += 0.5 * abs(np.dot(p[:, 0], np.roll(p[:, 1], -1)) - np.dot(p[:, 1], np.roll(p[:, 0], -1)))
    d = domain_from_gmsh(path('carre_hybrid'), 2)
    assert abs(d.cells.volume.sum() - area) < 1e-12
    assert area < 0.5  # the holes are cut out
    x, y, t = nd[:, 0], nd[:, 1], nd[:, 3]
    assert np.all(t[np.isclose(x, 0.)] == 1) and np.all(t[np.isclose(x, 1.)] == 2)


@pytest.mark.parametrize("name,dim,measure", [("carre", 2, 1.0), ("rectangle", 2, None), ("cube", 3, 1.0)])
def test_domain_from_gmsh(name, dim, measure):
    d = domain_from_gmsh(path(name), dim)
    el, nd = read_gmsh(path(name), dim)
    if measure is None:
        measure = np.prod(nd[:, :2].max(axis=0) - nd[:, :2].min(axis=0))
    assert abs(d.cells.volume.sum() - measure) < 1e-12 * measure
    # every boundary face lies on the box and carries a tag 1..2*dim; the normals close each cell
    bf = d.boundaryfaces
    assert set(np.unique(d.faces.name[bf])) <= set(range(1, 2 * dim + 1))
    s = np.zeros((int(d.nbcells), 3))
    for k in range(d.cells.faceid.shape[1] - 1):
        f = d.cells.faceid[:, k]
        ok = k < d.cells.faceid[:, -1]
        sign = np.where(d.faces.cellid[f, 0] == np.arange(int(d.nbcells)), 1., -1.)
        s[ok] += (sign[:, None] * d.faces.normal[f, :3])[ok]
    assert np.abs(s).max() < 1e-12


@pytest.mark.parametrize("name,dim,counts,measure", [
    ("cube_bis", 3, {4: 3196, 5: 81, 8: 729}, 2.0),            # tetrahedra, pyramids, hexahedra
    ("carre_structure", 2, {3: 1058, 4: 192}, 0.4),            # triangles and quadrangles
    ("rectangle_RTM", 2, {3: 5206}, 0.008)])
def test_mixed_element_meshes(name, dim, counts, measure):
    """Meshes with several element kinds: the element counts per kind, the box tags of the
    nodes, and cell volumes that add up to the box (the 3D cells split into tetrahedra by
    Compute_3dcentervolumeOfCell's rules)."""
    el, nd = read_gmsh(path(name), dim)
    kinds, n = np.unique(el[:, -1], return_counts=True)
    assert dict(zip(kinds.tolist(), n.tolist())) == counts
    box = [(nd[:, k].min(), nd[:, k].max()) for k in range(dim)]
    assert np.array_equal(nd[:, 3], _expected_tags(nd, box))
    d = domain_from_gmsh(path(name), dim)
    assert abs(d.cells.volume.sum() - measure) < 1e-12 * measure

