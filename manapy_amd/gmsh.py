"""Gmsh 4.1 ASCII meshes in manapy's partition layout (no meshio needed).

``read_gmsh(path, dim)`` returns ``(elements, nodes)`` as ``MeshPartition`` builds them before
writing its partition file (``partitions/meshpartitioning.py:160-194`` reads the mesh with
meshio, ``:255-256`` writes these arrays):

* ``nodes`` -- float64 ``[N, 4]``: x, y, z and the boundary tag of ``define_ghost_node``
  (``partitions/partitions_utils.py:16-111``), from the physical tags of the boundary elements
  (lines in 2D; quadrangles and triangles in 3D): first tags 1 and 2, then 3 and 4 on nodes not
  already 1 or 2 (both passes element type by element type, types in order of first appearance
  in the file), then any tag on nodes still untagged;
* ``elements`` -- int32 ``[C, nv+1]``: vertex indices (0-based, nodes in file order), the vertex
  count in the last column, padding 0 (``convert_2d_cons_to_array`` / ``convert_3d_cons_to_array``,
  ``partitions_utils.py:113-185``): triangles then quadrangles in 2D, tetrahedra, pyramids then
  hexahedra in 3D, each type in file order.

An element's physical tag is the first physical tag of its entity (what meshio reports as
``gmsh:physical``). Periodic meshes (tags x 11) are not handled. ``.msh.gz`` files are read
through gzip.
"""
import gzip

import numpy as np

# Gmsh element type -> (name, number of nodes)
_TYPES = {1: ('line', 2), 2: ('triangle', 3), 3: ('quad', 4), 4: ('tetra', 4), 5: ('hexahedron', 8),
          7: ('pyramid', 5), 15: ('vertex', 1)}
_CELL_TYPES = {2: ('triangle', 'quad'), 3: ('tetra', 'pyramid', 'hexahedron')}
_FACE_TYPES = {2: ('line',), 3: ('quad', 'triangle')}


def _sections(text):
    out = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        ln = lines[i].strip()
        if ln.startswith('$') and not ln.startswith('$End'):
            name = ln[1:]
            j = i + 1
            while lines[j].strip() != '$End' + name:
                j += 1
            out[name] = lines[i + 1:j]
            i = j
        i += 1
    return out


def _entities(sec):
    """{dim: {entity tag: [physical tags]}} from the $Entities section."""
    head = sec[0].split()
    counts = [int(x) for x in head[:4]]
    phys = {0: {}, 1: {}, 2: {}, 3: {}}
    k = 1
    for dim in range(4):
        for _ in range(counts[dim]):
            v = sec[k].split()
            k += 1
            tag = int(v[0])
            p = 4 if dim == 0 else 7  # point: tag x y z; others: tag and the bounding box
            nph = int(v[p])
            phys[dim][tag] = [int(x) for x in v[p + 1:p + 1 + nph]]
    return phys


def _nodes(sec):
    nb, nn = (int(x) for x in sec[0].split()[:2])
    tags = np.empty(nn, dtype=np.int64)
    xyz = np.empty((nn, 3), dtype=np.float64)
    k, o = 1, 0
    for _ in range(nb):
        dim, etag, para, n = (int(x) for x in sec[k].split())
        k += 1
        if para:
            raise ValueError("parametric node coordinates are not supported")
        for j in range(n):
            tags[o + j] = int(sec[k + j])
        k += n
        for j in range(n):
            xyz[o + j] = [float(x) for x in sec[k + j].split()[:3]]
        k += n
        o += n
    return tags, xyz


def _elements(sec, tag2idx, phys):
    """{type name: (node index array [n, nv], physical tag array [n])}, types in file order."""
    nb = int(sec[0].split()[0])
    blocks = {}
    k = 1
    for _ in range(nb):
        dim, etag, et, n = (int(x) for x in sec[k].split())
        k += 1
        if et not in _TYPES:
            raise ValueError("Gmsh element type %d is not supported" % et)
        name, nv = _TYPES[et]
        rows = np.array([[int(x) for x in sec[k + j].split()[1:1 + nv]] for j in range(n)], dtype=np.int64)
        k += n
        ph = phys[dim].get(etag, [])
        tag = ph[0] if ph else 0
        blocks.setdefault(name, ([], []))
        blocks[name][0].append(tag2idx[rows])
        blocks[name][1].append(np.full(n, tag, dtype=np.int64))
    return {nm: (np.concatenate(a), np.concatenate(b)) for nm, (a, b) in blocks.items()}


def boundary_tags(nbnodes, blocks, dim):
    """define_ghost_node (partitions_utils.py:16-111, non-periodic)."""
    tag = np.zeros(nbnodes, dtype=np.int64)
    faces = [nm for nm in blocks if nm in _FACE_TYPES[dim]]
    for nm in faces:
        conn, ph = blocks[nm]
        for k in range(len(ph)):
            if ph[k] in (1, 2):
                tag[conn[k]] = ph[k]
        for k in range(len(ph)):
            if ph[k] in (3, 4):
                for n in conn[k]:
                    if tag[n] not in (1, 2):
                        tag[n] = ph[k]
    for nm in faces:
        conn, ph = blocks[nm]
        for k in range(len(ph)):
            for n in conn[k]:
                if tag[n] == 0:
                    tag[n] = ph[k]
    return tag


def read_gmsh(path, dim):
    """(elements, nodes) of a Gmsh 4.1 ASCII mesh in the partition layout (module doc)."""
    opener = gzip.open if str(path).endswith('.gz') else open
    with opener(path, 'rt') as fh:
        sec = _sections(fh.read())
    ver = sec['MeshFormat'][0].split()
    if not ver[0].startswith('4.1') or ver[1] != '0':
        raise ValueError("only Gmsh 4.1 ASCII meshes are supported (got %s)" % ' '.join(ver))
    phys = _entities(sec['Entities'])
    tags, xyz = _nodes(sec['Nodes'])
    tag2idx = np.zeros(int(tags.max()) + 1, dtype=np.int64)
    tag2idx[tags] = np.arange(len(tags))
    blocks = _elements(sec['Elements'], tag2idx, phys)
    nodes = np.zeros((len(tags), 4), dtype=np.float64)
    nodes[:, :3] = xyz
    nodes[:, 3] = boundary_tags(len(tags), blocks, dim)
    cells = [blocks[nm][0] for nm in _CELL_TYPES[dim] if nm in blocks]
    if not cells:
        raise ValueError("no %dD cells in %s" % (dim, path))
    width = max(c.shape[1] for c in cells)
    elements = np.zeros((sum(len(c) for c in cells), width + 1), dtype=np.int32)
    o = 0
    for c in cells:
        elements[o:o + len(c), :c.shape[1]] = c
        elements[o:o + len(c), -1] = c.shape[1]
        o += len(c)
    return elements, nodes


def domain_from_gmsh(path, dim):
    """A Domain built from a Gmsh 4.1 mesh."""
    from .ddm import Domain
    return Domain(dim, mesh=read_gmsh(path, dim))
