"""Domain construction: manapy's mesh arrays, built with vectorised NumPy.

``Domain(dim, conf, mesh=(elements, nodes))`` produces the arrays that
``manapy/ddm/domain.py:130-612`` builds with per-element Python loops, with the
same numbering, orderings and floating-point expressions, so every downstream
kernel sees identical inputs:

* face numbering = ``np.unique`` of the per-face node tuples sorted in
  descending order (``domain.py:314-315``), reproduced with an integer key
  whose order is the row-lexicographic order ``np.unique(axis=0)`` uses;
* ``faces.cellid[f, 0]`` = first (lowest-numbered) cell listing the face
  (``ddm_utils2d.py:12-20``);
* ``cells.cellnid`` = vertex neighbours in ascending order
  (``ddm_utils2d.py:55-97``);
* ghost mirror centres and per-node ghost lists in boundary-face order
  (``ddm_utils2d.py:163-223``, ``ddm_utils3d.py:87-157``);
* pseudo-Laplacian node weights and diamond coefficients
  (``ddm_utils2d.py:553-705``, ``ddm_utils3d.py:459-527, 813-981``).

Only the single-subdomain case exists here (one GPU = one subdomain): halo
arrays have the shapes the reference gives them at one MPI rank
(``pure_ddm.py:21-33``, ``create_2d_halo_structure`` ``:277-278``).

In 3D the vertex-neighbour and ghost stencils (needed only by gradients and
the diamond scheme) are built on first access, because at millions of
tetrahedra they dominate construction time and a first-order Euler step never
reads them.
"""
import numpy as np

from .base import Struct, make_get_conf, process_conf

__all__ = ['Domain', 'Container']

I32 = np.int32
F64 = np.float64


class Container(object):
    """Attribute bag for ``cells``/``faces``/``nodes``/``halos``. Attributes
    registered with ``lazy`` are computed on first access."""

    def __init__(self, name):
        self.__dict__['_name'] = name
        self.__dict__['_lazy'] = {}

    def lazy(self, attr, builder):
        self._lazy[attr] = builder

    def __getattr__(self, attr):
        lazy = self.__dict__.get('_lazy', {})
        if attr in lazy:
            builder = lazy.pop(attr)
            builder()
            return self.__dict__[attr]
        raise AttributeError("%s has no attribute %r" % (self.__dict__.get('_name'), attr))

    def keys(self):
        return [k for k in self.__dict__ if not k.startswith('_')] + list(self._lazy)


# ---------------------------------------------------------------------------------------------
# small helpers

def _pad_rows(row, val, nrows, min_width=0, fill=-1):
    """Rows given as (row, value) pairs already sorted by row (and by the
    wanted order inside a row) -> int32 ``[nrows, w+1]`` padded with ``fill``,
    last column = count (manapy's padded-connectivity convention)."""
    counts = np.bincount(row, minlength=nrows).astype(np.int64)
    width = max(int(counts.max()) if nrows else 0, min_width)
    out = np.full((nrows, width + 1), fill, dtype=I32)
    if len(row):
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        pos = np.arange(len(row), dtype=np.int64) - start[row]
        out[row, pos] = val
    out[:, -1] = counts
    return out


def _rows_from_padded(padded):
    """Inverse of ``_pad_rows`` for the valid entries: (row, col, value)."""
    cnt = padded[:, -1]
    width = padded.shape[1] - 1
    mask = np.arange(width)[None, :] < cnt[:, None]
    r, c = np.nonzero(mask)
    return r, c, padded[r, c]


def _face_key(rows, nbnodes):
    """Integer key whose ascending order equals the row-lexicographic order of
    ``rows`` (non-negative ids < nbnodes), as ``np.unique(axis=0)`` sorts."""
    key = np.zeros(len(rows), dtype=np.int64)
    n = np.int64(nbnodes)
    for col in range(rows.shape[1]):
        key = key * n + rows[:, col].astype(np.int64)
    return key


# ---------------------------------------------------------------------------------------------
# geometry kernels (vectorised restatements)

def centers_volumes_2d(nodeid, vertex):
    """``Compute_2dcentervolumeOfCell`` (``ddm/ddm_utils2d.py:433-479``)."""
    C = len(nodeid)
    center = np.zeros((C, 3), dtype=F64)
    volume = np.zeros(C, dtype=F64)
    tri = np.nonzero(nodeid[:, -1] == 3)[0]
    if len(tri):
        s = nodeid[tri]
        x1, y1 = vertex[s[:, 0], 0], vertex[s[:, 0], 1]
        x2, y2 = vertex[s[:, 1], 0], vertex[s[:, 1], 1]
        x3, y3 = vertex[s[:, 2], 0], vertex[s[:, 2], 1]
        center[tri, 0] = 1. / 3 * (x1 + x2 + x3)
        center[tri, 1] = 1. / 3 * (y1 + y2 + y3)
        volume[tri] = (1. / 2) * np.abs((x1 - x2) * (y1 - y3) - (x1 - x3) * (y1 - y2))
    quad = np.nonzero(nodeid[:, -1] == 4)[0]
    if len(quad):
        s = nodeid[quad]
        xs = [vertex[s[:, k], 0] for k in range(4)]
        ys = [vertex[s[:, k], 1] for k in range(4)]
        # split_to_triangle: centre = builtin sum / 4 (sequential), then 4 triangles (v_k, v_k+1, centre)
        cx = (((xs[0] + xs[1]) + xs[2]) + xs[3]) / 4
        cy = (((ys[0] + ys[1]) + ys[2]) + ys[3]) / 4
        center[quad, 0] = cx
        center[quad, 1] = cy
        vol = np.zeros(len(quad), dtype=F64)
        for k in range(4):
            x1, y1 = xs[k], ys[k]
            x2, y2 = xs[(k + 1) % 4], ys[(k + 1) % 4]
            x3, y3 = cx, cy
            vol += (1. / 2) * np.abs((x1 - x2) * (y1 - y3) - (x1 - x3) * (y1 - y2))
        volume[quad] = vol
    return center, volume


def _cross(u, v):
    return np.stack([u[:, 1] * v[:, 2] - u[:, 2] * v[:, 1],
                     u[:, 2] * v[:, 0] - u[:, 0] * v[:, 2],
                     u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]], axis=1)


def _dot3(u, v):
    return u[:, 0] * v[:, 0] + u[:, 1] * v[:, 1] + u[:, 2] * v[:, 2]


def centers_volumes_3d(nodeid, vertex):
    """``Compute_3dcentervolumeOfCell`` (``ddm/ddm_utils3d.py:686-764``)."""
    C = len(nodeid)
    center = np.zeros((C, 3), dtype=F64)
    volume = np.zeros(C, dtype=F64)
    for nv in np.unique(nodeid[:, -1]):
        idx = np.nonzero(nodeid[:, -1] == nv)[0]
        verts = [vertex[nodeid[idx, k], :3] for k in range(nv)]
        acc = np.zeros((len(idx), 3), dtype=F64)
        for v in verts:
            acc = acc + v
        center[idx] = acc / nv
        if nv == 4:
            u = verts[1] - verts[0]
            v = verts[2] - verts[0]
            w = verts[3] - verts[0]
            volume[idx] = np.abs(_dot3(u, _cross(v, w))) / 6.0
        elif nv in (8, 5):
            rules = ([[0, 1, 3, 4], [1, 3, 4, 5], [4, 5, 3, 7], [1, 3, 5, 2], [3, 7, 5, 2], [5, 7, 6, 2]]
                     if nv == 8 else [[0, 1, 2, 4], [0, 2, 3, 4]])
            vol = np.zeros(len(idx), dtype=F64)
            for r in rules:
                u = verts[r[1]] - verts[r[0]]
                v = verts[r[2]] - verts[r[0]]
                w = verts[r[3]] - verts[r[0]]
                vol += np.abs(_dot3(u, _cross(v, w))) / 6.0
            volume[idx] = vol
        else:
            raise ValueError("unsupported 3D cell with %d vertices" % nv)
    return center, volume


# local face tables: create_2dfaces (ddm_utils2d.py:482-505) and create_3dfaces (ddm_utils3d.py:766-810)
_FACES_2D = {3: [[0, 1], [1, 2], [2, 0]], 4: [[0, 1], [1, 2], [2, 3], [3, 0]]}
_FACES_3D = {4: [[0, 1, 2], [2, 3, 0], [0, 1, 3], [3, 1, 2]],
             5: [[0, 1, 4], [1, 2, 4], [2, 3, 4], [0, 3, 4], [0, 1, 2, 3]],
             8: [[0, 1, 2, 3], [0, 1, 4, 5], [1, 2, 5, 6], [2, 3, 6, 7], [0, 3, 4, 7], [4, 5, 6, 7]]}


def build_faces(nodeid, nbnodes, dim):
    """``create_{2,3}dfaces`` + the sort/unique of ``domain.py:294-335``.

    Returns ``faces_nodeid [F, maxfacenid+1]``, ``cells_faceid [C, maxcellfid+1]``,
    ``faces_cellid [F, 2]`` and ``cells_cellfid [C, maxcellfid+1]``."""
    C = len(nodeid)
    table = _FACES_2D if dim == 2 else _FACES_3D
    nvs = nodeid[:, -1]
    nfaces_per_cell = np.zeros(C, dtype=np.int64)
    for nv, faces in table.items():
        nfaces_per_cell[nvs == nv] = len(faces)
    if np.any(nfaces_per_cell == 0):
        raise ValueError("unsupported cell type in a %dD mesh" % dim)
    maxcellfid = int(nfaces_per_cell.max())
    maxfacenid = 2 if dim == 2 else max(len(f) for nv, fl in table.items() if np.any(nvs == nv) for f in fl)
    # p_faces rows in cell-major order (k counter of create_*faces)
    offs = np.concatenate([[0], np.cumsum(nfaces_per_cell)[:-1]])
    total = int(nfaces_per_cell.sum())
    p_faces = np.full((total, maxfacenid + 1), -1, dtype=I32)
    p_faces[:, -1] = 2 if dim == 2 else 3
    cellf = np.full((C, maxcellfid), -1, dtype=np.int64)
    for nv, faces in table.items():
        idx = np.nonzero(nvs == nv)[0]
        if not len(idx):
            continue
        for j, loc in enumerate(faces):
            rows = offs[idx] + j
            for k, v in enumerate(loc):
                p_faces[rows, k] = nodeid[idx, v]
            if dim == 3:
                p_faces[rows, -1] = len(loc)
            cellf[idx, j] = rows
    # domain.py:314 -- sort node ids of each face, descending (-1 padding goes last)
    p_faces[:, :-1] = np.sort(p_faces[:, :-1], axis=1)[:, ::-1]
    # domain.py:315 -- np.unique(axis=0, return_inverse=True)
    if np.all(p_faces[:, :-1] >= 0) and float(nbnodes) ** p_faces[:, :-1].shape[1] < 2.0 ** 62:
        key = _face_key(p_faces[:, :-1], nbnodes)
        _, first, inverse = np.unique(key, return_index=True, return_inverse=True)
        faces_nodeid = p_faces[first]
    else:
        faces_nodeid, inverse = np.unique(p_faces, axis=0, return_inverse=True)
    inverse = np.asarray(inverse).reshape(-1)
    F = len(faces_nodeid)

    # create_cell_faceid (ddm_utils2d.py:22-29)
    faceid = np.full((C, maxcellfid + 1), -1, dtype=I32)
    valid = cellf >= 0
    faceid[:, :-1][valid] = inverse[cellf[valid]]
    faceid[:, -1] = valid.sum(axis=1)

    # create_cellsOfFace (ddm_utils2d.py:12-20): first visitor -> [0], other -> [1]
    cell_of = np.repeat(np.arange(C, dtype=np.int64), maxcellfid)[valid.ravel()]
    face_of = faceid[:, :-1][valid].astype(np.int64)
    order = np.argsort(face_of, kind='stable')
    fs = face_of[order]
    cs = cell_of[order]
    counts = np.bincount(fs, minlength=F)
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    cellid = np.full((F, 2), -1, dtype=I32)
    cellid[:, 0] = cs[start]
    two = counts >= 2
    cellid[two, 1] = cs[start[two] + counts[two] - 1]

    # create_NeighborCellByFace (ddm_utils2d.py:32-53)
    fl = faceid[:, :-1].astype(np.int64)
    flc = np.where(valid, fl, 0)
    c0 = cellid[flc, 0]
    c1 = cellid[flc, 1]
    me = np.arange(C)[:, None]
    other = np.where(c0 == me, c1, c0)
    ok = valid & (c1 != -1)
    r, c = np.nonzero(ok)  # row-major: keeps the faceid order inside a row
    cellfid = _pad_rows(r, other[r, c], C, min_width=maxcellfid)
    return faces_nodeid, faceid, cellid, cellfid, maxcellfid, maxfacenid


def face_info_2d_geometry(cellid, fnodeid, namen, vertex, centerc):
    """``create_info_2dfaces`` (``ddm/ddm_utils2d.py:507-551``): names of
    boundary faces, non-unit outward normals, lengths, centres."""
    F = len(cellid)
    n0 = fnodeid[:, 0]
    n1 = fnodeid[:, 1]
    namef = np.zeros(F, dtype=I32)
    bnd = np.nonzero(cellid[:, 1] == -1)[0]
    a = namen[n0[bnd]]
    b = namen[n1[bnd]]
    name = np.full(len(bnd), 100, dtype=I32)
    undecided = np.ones(len(bnd), dtype=bool)
    sel = undecided & (a == b)
    name[sel] = a[sel]
    undecided &= ~sel
    for tag in (3, 4, 33, 44):
        sel = undecided & (((a == tag) & (b != 0)) | ((a != 0) & (b == tag)))
        name[sel] = tag
        undecided &= ~sel
    namef[bnd] = name

    normx = vertex[n0, 1] - vertex[n1, 1]
    normy = vertex[n1, 0] - vertex[n0, 0]
    centerf = 0.5 * (vertex[n0, 0:3] + vertex[n1, 0:3])
    snorm = centerc[cellid[:, 0]] - centerf
    flip = (snorm[:, 0] * normx + snorm[:, 1] * normy) > 0
    normal = np.zeros((F, 3), dtype=F64)
    normal[:, 0] = np.where(flip, -1 * normx, normx)
    normal[:, 1] = np.where(flip, -1 * normy, normy)
    mesure = np.sqrt(normal[:, 0] * normal[:, 0] + normal[:, 1] * normal[:, 1])
    return namef, normal, mesure, centerf


def _name3d(nv_ids, namen):
    """Boundary-face naming of ``create_info_3dfaces`` (``ddm_utils3d.py:549-593``,
    ``:615-669``) for faces with ``k`` nodes (columns of ``nv_ids``)."""
    k = nv_ids.shape[1]
    nm = namen[nv_ids]
    out = np.full(len(nv_ids), 100 if k == 3 else 200, dtype=I32)
    undecided = np.ones(len(nv_ids), dtype=bool)
    sel = np.all(nm == nm[:, :1], axis=1)
    out[sel] = nm[sel, 0]
    undecided &= ~sel
    for tag in (5, 6, 3, 4, 55, 66, 33, 44):
        hit = np.zeros(len(nv_ids), dtype=bool)
        for j in range(k):
            cond = nm[:, j] == tag
            for l in range(k):
                if l != j:
                    cond &= nm[:, l] != 0
            hit |= cond
        sel = undecided & hit
        out[sel] = tag
        undecided &= ~sel
    return out


def face_info_3d_geometry(cellid, fnodeid, namen, vertex, centerc):
    """``create_info_3dfaces`` (``ddm/ddm_utils3d.py:530-684``)."""
    F = len(cellid)
    nn = fnodeid[:, -1]
    centerf = np.zeros((F, 3), dtype=F64)
    norm = np.zeros((F, 3), dtype=F64)
    mesure = np.zeros(F, dtype=F64)
    namef = np.zeros(F, dtype=I32)
    for k in (3, 4):
        idx = np.nonzero(nn == k)[0]
        if not len(idx):
            continue
        v = [vertex[fnodeid[idx, j], :3] for j in range(k)]
        if k == 3:
            centerf[idx] = 1. / 3 * (v[0] + v[1] + v[2])
            u = v[1] - v[0]
            w = v[2] - v[0]
            nrm = 0.5 * _cross(u, w)
            mesure[idx] = np.sqrt(nrm[:, 0] ** 2 + nrm[:, 1] ** 2 + nrm[:, 2] ** 2)
        else:
            centerf[idx] = 1. / 4 * (v[0] + v[1] + v[2] + v[3])
            u = v[1] - v[0]
            w = v[2] - v[0]
            area1 = 0.5 * _cross(u, w)
            nrm = area1 + area1  # area2 = area1 (reference quirk, ddm_utils3d.py:605)
            mag = np.sqrt(area1[:, 0] ** 2 + area1[:, 1] ** 2 + area1[:, 2] ** 2)
            mesure[idx] = mag + mag
        norm[idx] = nrm
        bnd = idx[cellid[idx, 1] == -1]
        if len(bnd):
            namef[bnd] = _name3d(fnodeid[bnd, :k], namen)
    snorm = centerc[cellid[:, 0]] - centerf
    flip = _dot3(snorm, norm) > 0
    normal = np.where(flip[:, None], -1 * norm, norm)
    u = vertex[fnodeid[:, 1], 0:3] - vertex[fnodeid[:, 0], 0:3]
    tangent = u.copy()
    binormal = 0.5 * _cross(u, normal)
    return namef, normal, mesure, centerf, tangent, binormal


def oriente_3dfacenodeid(fnodeid, normal, vertex):
    """``ddm/ddm_utils3d.py:362-403`` (in place on ``fnodeid``)."""
    tri = np.nonzero(fnodeid[:, -1] == 3)[0]
    if len(tri):
        n1, n2, n3 = fnodeid[tri, 0].copy(), fnodeid[tri, 1].copy(), fnodeid[tri, 2].copy()
        v1 = vertex[n2, :3] - vertex[n1, :3]
        v2 = vertex[n3, :3] - vertex[n1, :3]
        flip = _dot3(_cross(v1, v2), normal[tri]) < 0
        fnodeid[tri[flip], 1] = n3[flip]
        fnodeid[tri[flip], 2] = n2[flip]
    quad = np.nonzero(fnodeid[:, -1] == 4)[0]
    if len(quad):
        def orient(a, b, c):
            pa, pb, pc = vertex[a, :2], vertex[b, :2], vertex[c, :2]
            cz = lambda p, q: p[:, 0] * q[:, 1] - p[:, 1] * q[:, 0]
            return cz(pa, pb) + cz(pb, pc) + cz(pc, pa)
        s = [fnodeid[quad, j].copy() for j in range(4)]
        o123 = orient(s[0], s[1], s[2]) < 0
        o134 = orient(s[0], s[2], s[3]) < 0
        o124 = orient(s[0], s[1], s[3]) < 0
        rows = fnodeid[quad]
        # branch 1: abc clockwise
        b1 = o123 & ~o134
        m = b1 & o124
        rows[m, 2] = s[3][m]; rows[m, 3] = s[2][m]
        m = b1 & ~o124
        rows[m, 0] = s[3][m]; rows[m, 3] = s[0][m]
        # branch 2: abc ccw, acd clockwise
        b2 = ~o123 & o134
        m = b2 & o124
        rows[m, 1] = s[2][m]; rows[m, 2] = s[1][m]
        m = b2 & ~o124
        rows[m, 0] = s[1][m]; rows[m, 1] = s[0][m]
        # branch 3: both ccw
        b3 = ~o123 & ~o134
        rows[b3, 0] = s[2][b3]; rows[b3, 2] = s[0][b3]
        fnodeid[quad] = rows


def normal_faces_of_cell(centerc, centerf, faceid, normal, maxcellfid):
    """``create_NormalFacesOfCell`` (``ddm/ddm_utils2d.py:99-120``)."""
    C = len(faceid)
    nf = np.zeros((C, maxcellfid, 3), dtype=F64)
    for j in range(maxcellfid):
        rows = np.nonzero(faceid[:, -1] > j)[0]
        f = faceid[rows, j]
        d = centerc[rows] - centerf[f]
        inward = _dot3(d, normal[f]) < 0.
        nf[rows, j] = np.where(inward[:, None], normal[f], -1.0 * normal[f])
    return nf


def native_topology(nodeid, nbnodes, dim, want_cellnid):
    """The connectivity from the library's host builder (``mfv_domain_topology_build``, C++):
    the arrays of ``build_faces`` and of the node/vertex-neighbour tables, identical to the
    NumPy construction. None when libmfv.so is not built."""
    import ctypes
    try:
        from . import _lib
        _lib.load()
    except Exception:
        return None
    el = np.ascontiguousarray(nodeid, dtype=I32)
    C = len(el)
    t = _lib.MfvDomainTopology()
    _lib.call('mfv_domain_topology_build', int(dim), C, int(el.shape[1]), el.ctypes.data_as(ctypes.c_void_p),
              int(nbnodes), int(bool(want_cellnid)), ctypes.byref(t))
    try:
        F = int(t.nb_faces)

        def grab(ptr, rows, cols):
            if rows == 0 or cols <= 0:
                return np.zeros((rows, max(cols, 0)), dtype=I32)
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int32)), shape=(rows, cols))
            return a.copy()
        out = {'faces_nodeid': grab(t.faces_nodeid, F, t.max_face_nodes + 1),
               'faceid': grab(t.cells_faceid, C, t.max_cell_faces + 1),
               'cellid': grab(t.faces_cellid, F, 2),
               'cellfid': grab(t.cells_cellfid, C, t.max_cell_faces + 1),
               'node_cellid': grab(t.nodes_cellid, int(nbnodes), t.node_cells_width + 1),
               'cellnid': grab(t.cells_cellnid, C, t.cellnid_width + 1) if t.cellnid_width >= 0 else None,
               'maxcellfid': int(t.max_cell_faces), 'maxfacenid': int(t.max_face_nodes)}
    finally:
        _lib.call('mfv_domain_topology_free', ctypes.byref(t))
    return out


def native_geometry_2d(d):
    """The 2D geometry from the library's host builder (``mfv_domain_geometry_2d_build``, C++
    on host threads): the arrays of the NumPy stages below, identical."""
    import ctypes
    from . import _lib
    nodeid = np.ascontiguousarray(d._cells.nodeid, dtype=I32)
    vertex = np.ascontiguousarray(d._nodes.vertex, dtype=F64)
    fn = np.ascontiguousarray(d._faces.nodeid, dtype=I32)
    faceid = np.ascontiguousarray(d._cells.faceid, dtype=I32)
    cellid = np.ascontiguousarray(d._faces.cellid, dtype=I32)
    ncid = np.ascontiguousarray(d._nodes.cellid, dtype=I32)
    C, N, F = len(nodeid), len(vertex), len(fn)
    if fn.shape[1] != 3:
        raise ValueError("2D face rows have two nodes and the count")
    mf = faceid.shape[1] - 1
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    g = _lib.MfvDomainGeometry2d()
    _lib.call('mfv_domain_geometry_2d_build', C, nodeid.shape[1], ptr(nodeid), N, ptr(vertex), F, ptr(fn),
              faceid.shape[1], ptr(faceid), ptr(cellid), ptr(ncid), ncid.shape[1], ctypes.byref(g))
    try:
        def grab(p, shape, ct=ctypes.c_double):
            n = int(np.prod(shape))
            if n == 0:
                return np.zeros(shape, dtype=np.float64 if ct is ctypes.c_double else I32)
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=shape).copy()
        gw, nw = int(g.ghost_width), int(g.ghostnid_width)
        out = {'center': grab(g.cell_center, (C, 3)), 'volume': grab(g.cell_volume, (C,)),
               'name': grab(g.face_name, (F,), ctypes.c_int32), 'normal': grab(g.face_normal, (F, 3)),
               'mesure': grab(g.face_mesure, (F,)), 'fcenter': grab(g.face_center, (F, 3)),
               'nf': grab(g.cell_nf, (C, mf, 3)), 'ghostcenter': grab(g.face_ghostcenter, (F, 3)),
               'dist_ortho': grab(g.dist_ortho, (F,)), 'airDiamond': grab(g.airDiamond, (F,))}
        for k in ('param1', 'param2', 'param3', 'param4'):
            out[k] = grab(getattr(g, k), (F,))
        for k in ('f_1', 'f_2', 'f_3', 'f_4'):
            out[k] = grab(getattr(g, k), (F, 2))
        for k in ('lambda_x', 'lambda_y', 'R_x', 'R_y'):
            out[k] = grab(getattr(g, k), (N,))
        out['number'] = grab(g.number, (N,), ctypes.c_int32)
        out['node_ghostcenter'] = grab(g.node_ghostcenter, (N, gw, 5))
        out['node_ghostfaceinfo'] = grab(g.node_ghostfaceinfo, (N, gw, 4))
        out['ghostid'] = grab(g.node_ghostid, (N, gw + 1), ctypes.c_int32)
        out['ghostnid'] = grab(g.cell_ghostnid, (C, nw + 1), ctypes.c_int32)
    finally:
        _lib.call('mfv_domain_geometry_2d_free', ctypes.byref(g))
    return out


def native_geometry_3d(d):
    """The 3D geometry from the library's host builder (``mfv_domain_geometry_3d_build``, C++
    on host threads): the arrays of the NumPy stages below, identical, with the face table
    re-oriented as ``oriente_3dfacenodeid`` leaves it."""
    import ctypes
    from . import _lib
    nodeid = np.ascontiguousarray(d._cells.nodeid, dtype=I32)
    vertex = np.ascontiguousarray(d._nodes.vertex, dtype=F64)
    fn = np.ascontiguousarray(d._faces.nodeid, dtype=I32)
    cellid = np.ascontiguousarray(d._faces.cellid, dtype=I32)
    ncid = np.ascontiguousarray(d._nodes.cellid, dtype=I32)
    C, N, F, fw = len(nodeid), len(vertex), len(fn), fn.shape[1]
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    g = _lib.MfvDomainGeometry3d()
    _lib.call('mfv_domain_geometry_3d_build', C, nodeid.shape[1], ptr(nodeid), N, ptr(vertex), F, fw, ptr(fn),
              ptr(cellid), ptr(ncid), ncid.shape[1], ctypes.byref(g))
    try:
        def grab(p, shape, ct=ctypes.c_double):
            n = int(np.prod(shape))
            if n == 0:
                return np.zeros(shape, dtype=np.float64 if ct is ctypes.c_double else I32)
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=shape).copy()
        gw, nw = int(g.ghost_width), int(g.ghostnid_width)
        out = {'center': grab(g.cell_center, (C, 3)), 'volume': grab(g.cell_volume, (C,)),
               'name': grab(g.face_name, (F,), ctypes.c_int32), 'normal': grab(g.face_normal, (F, 3)),
               'mesure': grab(g.face_mesure, (F,)), 'fcenter': grab(g.face_center, (F, 3)),
               'tangent': grab(g.face_tangent, (F, 3)), 'binormal': grab(g.face_binormal, (F, 3)),
               'fnodeid': grab(g.face_nodeid, (F, fw), ctypes.c_int32),
               'ghostcenter': grab(g.face_ghostcenter, (F, 4)), 'dist_ortho': grab(g.dist_ortho, (F,)),
               'airDiamond': grab(g.airDiamond, (F,)), 'f_1': grab(g.f_1, (F, 3)), 'f_2': grab(g.f_2, (F, 3))}
        for k in ('param1', 'param2', 'param3'):
            out[k] = grab(getattr(g, k), (F,))
        for k in ('lambda_x', 'lambda_y', 'lambda_z', 'R_x', 'R_y', 'R_z'):
            out[k] = grab(getattr(g, k), (N,))
        out['param4'] = np.zeros(F)
        out['f_3'], out['f_4'] = np.zeros((F, 3)), np.zeros((F, 3))
        out['number'] = grab(g.number, (N,), ctypes.c_int32)
        out['node_ghostcenter'] = grab(g.node_ghostcenter, (N, gw, 6))
        out['node_ghostfaceinfo'] = grab(g.node_ghostfaceinfo, (N, gw, 6))
        out['ghostid'] = grab(g.node_ghostid, (N, gw + 1), ctypes.c_int32)
        out['ghostnid'] = grab(g.cell_ghostnid, (C, nw + 1), ctypes.c_int32)
    finally:
        _lib.call('mfv_domain_geometry_3d_free', ctypes.byref(g))
    return out


# ---------------------------------------------------------------------------------------------

class Domain(object):
    """Single-subdomain counterpart of ``manapy.ddm.Domain``
    (``ddm/domain.py:36-160``): same attribute names, same arrays.

    ``mesh`` is ``(elements, nodes)`` in the partition layout (see
    ``manapy_amd.meshgen``); ``conf`` accepts the reference's option table
    (``domain.py:70-83``)."""

    _parameters = [
        ('backend', 'str', "hip", False, 'kernel backend: hip (device) only'),
        ('topology', 'str', "auto", False, "connectivity builder: 'native' (C++, libmfv.so), 'numpy', or 'auto' "
                                          "(native when the library is built); the arrays are identical"),
        ('multithreading', 'str', 'single', None, 'accepted for API compatibility'),
        ('signature', 'bool', True, False, 'accepted for API compatibility'),
        ('cache', 'bool', False, False, 'accepted for API compatibility'),
        ('float_precision', 'str', "double", None, 'precision of the float arrays'),
        ('int_precision', 'str', "signed", None, 'precision of the integer arrays'),
    ]

    def __init__(self, dim=None, comm=None, conf=None, mesh=None, **kwargs):
        if conf is None:
            conf = Struct()
        self.conf = process_conf(self._parameters, conf, kwargs)
        get = make_get_conf(self.conf, kwargs)
        if dim is None:
            raise ValueError("dim must be given")
        if mesh is None:
            raise ValueError("mesh=(elements, nodes) must be given")
        if get('float_precision') == "single":
            self.float_precision = 'f4'
        else:
            self.float_precision = 'f8'
        self.int_precision = 'i4'
        self.signature = get('signature')
        self._comm = comm
        self._dim = np.int32(dim)
        self._size = 1
        self._rank = 0
        self._backend_name = get('backend')
        self._backend = None
        self._topology_mode = str(get('topology'))
        if self._topology_mode not in ('auto', 'native', 'numpy'):
            raise ValueError("topology must be 'auto', 'native' or 'numpy'")
        if isinstance(mesh, str):
            from .gmsh import read_gmsh
            elements, nodes = read_gmsh(mesh, dim)
        else:
            elements, nodes = mesh
        self.create_domain(np.ascontiguousarray(elements, dtype=I32),
                           np.ascontiguousarray(nodes, dtype=F64))

    # -- construction, in the stage order of domain.py:146-160 ---------------------------------
    def create_domain(self, elements, vertex):
        self._nodes = Container('nodes')
        self._cells = Container('cells')
        self._faces = Container('faces')
        self._halos = Container('halos')
        self._read_partition(elements, vertex)
        self._define_bounds()
        self._geo = None
        native = self._native() is not None
        if not native:
            self._compute_cells_info()
        self._make_neighbors()
        self._define_eltypes()
        self._create_faces_cons()
        if native:  # the geometry in C++ (the NumPy stages below then only copy it in)
            self._geo = native_geometry_2d(self) if self._dim == 2 else native_geometry_3d(self)
            self._cells.center, self._cells.volume = self._geo['center'], self._geo['volume']
        self._define_face_info()
        self._make_halo_info()
        self._update_boundaries()
        self._update_ghost_info()
        self._prepare_communication()
        self._update_periodic_boudaries()
        if self._dim == 2:
            self._update_haloghost()
            self._compute_diamondCell_info()
            self._compute_orthocenter()
        elif self._geo is not None:  # the C++ builder computed these stages already: copy them in
            self._update_haloghost()
            self._compute_diamondCell_info()
            self._geo = None
        else:
            self._cells.lazy('ghostnid', self._update_haloghost)
            self._nodes.lazy('ghostid', self._update_haloghost)
            for a in ('haloghostid',):
                self._nodes.lazy(a, self._update_haloghost)
            self._cells.lazy('haloghostnid', self._update_haloghost)
            for a in ('R_x', 'R_y', 'R_z', 'lambda_x', 'lambda_y', 'lambda_z', 'number'):
                self._nodes.lazy(a, self._compute_diamondCell_info)
            for a in ('airDiamond', 'param1', 'param2', 'param3', 'param4', 'f_1', 'f_2', 'f_3', 'f_4'):
                self._faces.lazy(a, self._compute_diamondCell_info)

    def _read_partition(self, elements, vertex):
        """``pure_ddm.read_mesh_file`` at one rank (``pure_ddm.py:21-33``)."""
        C = len(elements)
        N = len(vertex)
        self._cells.nodeid = elements
        self._cells.tc = np.arange(C, dtype=I32)
        self._cells.loctoglob = np.arange(C, dtype=I32)
        self._cells.globtoloc = np.arange(C, dtype=I32)
        self._nodes.vertex = vertex
        self._nodes.loctoglob = np.arange(N, dtype=I32)
        self._nodes.nparts = []
        self._halos.halosint = np.zeros((2, 2), dtype=I32)
        self._halos.halosext = np.zeros((2, 2), dtype=I32)
        self._halos.centvol = np.zeros((2, 2), dtype=F64)
        self._halos.neigh = []
        self._nbcells = I32(C)
        self._nbnodes = I32(N)

    def _define_bounds(self):
        v = self._nodes.vertex
        self._bounds = np.array([[v[:, d].min(), v[:, d].max()] for d in range(self._dim)], dtype=F64)

    def _compute_cells_info(self):
        if self._dim == 2:
            c, v = centers_volumes_2d(self._cells.nodeid, self._nodes.vertex)
        else:
            c, v = centers_volumes_3d(self._cells.nodeid, self._nodes.vertex)
        self._cells.center = c
        self._cells.volume = v

    def _native(self):
        """The C++ connectivity (faces, node cells, 2D vertex neighbours), built once."""
        if self._topology_mode == 'numpy':
            return None
        if getattr(self, '_topo', None) is None:
            self._topo = native_topology(self._cells.nodeid, int(self._nbnodes), int(self._dim),
                                         want_cellnid=(int(self._dim) == 2))
            if self._topo is None and self._topology_mode == 'native':
                raise RuntimeError("topology='native' needs libmfv.so (run __graft_entry__.build())")
        return self._topo

    def _make_neighbors(self):
        """``create_node_cellid`` (``ddm_utils2d.py:55-97``): node -> cells
        (ascending) now; cell -> vertex neighbours (ascending) in 2D now, in 3D
        on first access."""
        topo = self._native()
        if topo is not None:
            self._nodes.cellid = topo['node_cellid']
            if self._dim == 2:
                self._cells.cellnid = topo['cellnid']
            else:
                self._cells.lazy('cellnid', self._build_cellnid)
            return
        nodeid = self._cells.nodeid
        C = len(nodeid)
        N = int(self._nbnodes)
        nv = nodeid[:, -1]
        maxnv = nodeid.shape[1] - 1
        mask = np.arange(maxnv)[None, :] < nv[:, None]
        cells_rep = np.repeat(np.arange(C, dtype=np.int64), maxnv).reshape(C, maxnv)[mask]
        nodes_rep = nodeid[:, :maxnv][mask].astype(np.int64)
        order = np.argsort(nodes_rep, kind='stable')
        self._nodes.cellid = _pad_rows(nodes_rep[order], cells_rep[order], N)
        if self._dim == 2:
            self._build_cellnid()
        else:
            self._cells.lazy('cellnid', self._build_cellnid)

    def _build_cellnid(self):
        nodeid = self._cells.nodeid
        C = len(nodeid)
        ncid = self._nodes.cellid
        ncnt = ncid[:, -1].astype(np.int64)
        maxnv = nodeid.shape[1] - 1
        keys = []
        for j in range(maxnv):
            rows = np.nonzero(nodeid[:, -1] > j)[0]
            nod = nodeid[rows, j]
            cnt = ncnt[nod]
            for k in range(ncid.shape[1] - 1):
                sel = cnt > k
                r = rows[sel]
                nb = ncid[nod[sel], k].astype(np.int64)
                keep = nb != r
                keys.append(r[keep] * C + nb[keep])
        key = np.unique(np.concatenate(keys))
        self._cells.cellnid = _pad_rows(key // C, key % C, C)

    def _define_eltypes(self):
        nodeid = self._cells.nodeid
        self._maxcellnodeid = int(nodeid[:, -1].max())
        self._typeOfCells = {}
        if self._dim == 2:
            self._nbOfTriangles = int(np.sum(nodeid[:, -1] == 3))
            self._nbOfQuad = int(np.sum(nodeid[:, -1] == 4))
        else:
            self._nbOfTetra = int(np.sum(nodeid[:, -1] == 4))
            self._nbOfpyra = int(np.sum(nodeid[:, -1] == 5))
            self._nbOfQuad = int(np.sum(nodeid[:, -1] == 8))

    def _create_faces_cons(self):
        topo = self._native()
        if topo is not None:
            fn, faceid, cellid, cellfid = topo['faces_nodeid'], topo['faceid'], topo['cellid'], topo['cellfid']
            maxcellfid, maxfacenid = topo['maxcellfid'], topo['maxfacenid']
            self._topo = None  # the tables now live in the containers
        else:
            fn, faceid, cellid, cellfid, maxcellfid, maxfacenid = build_faces(
                self._cells.nodeid, int(self._nbnodes), int(self._dim))
        self._faces.nodeid = fn
        self._cells.faceid = faceid
        self._faces.cellid = cellid
        self._cells.cellfid = cellfid
        self._maxcellfid = I32(maxcellfid)
        self._maxfacenid = I32(maxfacenid)
        self._nbfaces = I32(len(fn))

    def _define_face_info(self):
        self._nodes.name = self._nodes.vertex[:, 3].astype(I32)
        if self._geo is not None:
            g = self._geo
            self._faces.name = g['name']
            self._faces.normal = g['normal']
            self._faces.mesure = g['mesure']
            self._faces.center = g['fcenter']
            if self._dim == 2:
                self._cells.nf = g['nf']
            else:
                self._faces.tangent = g['tangent']
                self._faces.binormal = g['binormal']
                self._cells.nf = np.zeros((int(self._nbcells), int(self._maxcellfid), 3), dtype=F64)
                self._faces.nodeid[...] = g['fnodeid']
        elif self._dim == 2:
            namef, normal, mesure, centerf = face_info_2d_geometry(
                self._faces.cellid, self._faces.nodeid, self._nodes.name, self._nodes.vertex, self._cells.center)
            self._faces.name = namef
            self._faces.normal = normal
            self._faces.mesure = mesure
            self._faces.center = centerf
            self._cells.nf = normal_faces_of_cell(self._cells.center, centerf, self._cells.faceid, normal,
                                                  int(self._maxcellfid))
        else:
            namef, normal, mesure, centerf, tangent, binormal = face_info_3d_geometry(
                self._faces.cellid, self._faces.nodeid, self._nodes.name, self._nodes.vertex, self._cells.center)
            self._faces.name = namef
            self._faces.normal = normal
            self._faces.mesure = mesure
            self._faces.center = centerf
            self._faces.tangent = tangent
            self._faces.binormal = binormal
            self._cells.nf = np.zeros((int(self._nbcells), int(self._maxcellfid), 3), dtype=F64)
            oriente_3dfacenodeid(self._faces.nodeid, normal, self._nodes.vertex)

    def _make_halo_info(self):
        """``create_{2,3}d_halo_structure`` at one rank: ``oldname`` aliases
        ``name`` (``ddm_utils2d.py:234``)."""
        self._nodes.halonid = np.zeros((int(self._nbnodes), 2), dtype=I32)
        self._faces.halofid = np.zeros(int(self._nbfaces), dtype=I32)
        self._faces.oldname = self._faces.name

    def _update_boundaries(self):
        """``domain.py:395-460``."""
        fname = self._faces.name
        nname = self._nodes.name
        w = lambda a, v: np.nonzero(a == v)[0].astype(I32)
        self._nodes.oldname = self._nodes.vertex[:, 3].astype(I32)
        self._innerfaces = w(fname, 0)
        self._infaces = w(fname, 1)
        self._outfaces = w(fname, 2)
        self._upperfaces = w(fname, 3)
        self._bottomfaces = w(fname, 4)
        self._halofaces = np.asarray([], dtype=I32)
        self._periodicinfaces = w(fname, 11)
        self._periodicoutfaces = w(fname, 22)
        self._periodicupperfaces = w(fname, 33)
        self._periodicbottomfaces = w(fname, 44)
        bf = [self._infaces, self._outfaces, self._bottomfaces, self._upperfaces]
        pf = [self._periodicinfaces, self._periodicoutfaces, self._periodicbottomfaces, self._periodicupperfaces]
        self._innernodes = w(nname, 0)
        self._innodes = w(nname, 1)
        self._outnodes = w(nname, 2)
        self._uppernodes = w(nname, 3)
        self._bottomnodes = w(nname, 4)
        self._halonodes = np.asarray([], dtype=I32)
        bn = [self._innodes, self._outnodes, self._bottomnodes, self._uppernodes]
        pn = [w(nname, 11), w(nname, 22), w(nname, 44), w(nname, 33)]
        if self._dim == 3:
            self._frontfaces = w(fname, 5)
            self._backfaces = w(fname, 6)
            self._periodicfrontfaces = w(fname, 55)
            self._periodicbackfaces = w(fname, 66)
            self._frontnodes = w(nname, 5)
            self._backnodes = w(nname, 6)
            bf += [self._backfaces, self._frontfaces]
            pf += [self._periodicbackfaces, self._periodicfrontfaces]
            bn += [self._backnodes, self._frontnodes]
            pn += [w(nname, 66), w(nname, 55)]
        self._boundaryfaces = np.sort(np.concatenate(bf)).astype(I32)
        self._periodicboundaryfaces = np.sort(np.concatenate(pf)).astype(I32)
        self._boundarynodes = np.sort(np.concatenate(bn)).astype(I32)
        self._periodicboundarynodes = np.sort(np.concatenate(pn)).astype(I32)
        if len(self._periodicboundaryfaces):
            raise NotImplementedError("periodic boundaries are not supported (the reference discards their "
                                      "geometry too, ddm/domain.py:507-516)")

    def _update_ghost_info(self):
        """``face_info_2d`` (``ddm_utils2d.py:163-223``) / ``face_info_3d``
        (``ddm_utils3d.py:87-157``): mirror centres of boundary faces and the
        per-node ghost lists (boundary-face order)."""
        if self._geo is not None:
            g = self._geo
            self._faces.ghostcenter = g['ghostcenter']
            self._nodes.ghostcenter = g['node_ghostcenter']
            self._nodes.ghostfaceinfo = g['node_ghostfaceinfo']
            self._cells.halonid = np.zeros((int(self._nbcells), 2), dtype=I32)
            return
        F = int(self._nbfaces)
        bf = self._boundaryfaces.astype(np.int64)
        cellid = self._faces.cellid
        fn = self._faces.nodeid
        vertex = self._nodes.vertex
        centerc = self._cells.center
        oldname = self._faces.oldname
        if self._dim == 2:
            gc = -1 * np.ones((F, 3), dtype=F64)
            nod1 = fn[bf, 1]
            nod2 = fn[bf, 0]
            x1 = vertex[nod1]
            x2 = vertex[nod2]
            v1 = centerc[cellid[bf, 0]]
            gamma = (((v1[:, 0] - x2[:, 0]) * (x1[:, 0] - x2[:, 0]) + (v1[:, 1] - x2[:, 1]) * (x1[:, 1] - x2[:, 1]))
                     / ((x1[:, 0] - x2[:, 0]) * (x1[:, 0] - x2[:, 0]) + (x1[:, 1] - x2[:, 1]) * (x1[:, 1] - x2[:, 1])))
            kk0 = gamma * x1[:, 0] + (1 - gamma) * x2[:, 0]
            kk1 = gamma * x1[:, 1] + (1 - gamma) * x2[:, 1]
            v20 = 2 * kk0 + (-1 * v1[:, 0])
            v21 = 2 * kk1 + (-1 * v1[:, 1])
            gc[bf, 0] = v20
            gc[bf, 1] = v21
            gc[bf, 2] = gamma
            rows = np.stack([v20, v21, cellid[bf, 0].astype(F64), oldname[bf].astype(F64), bf.astype(F64)], axis=1)
            ncols = 5
            info = np.stack([self._faces.center[bf, 0], self._faces.center[bf, 1],
                             self._faces.normal[bf, 0], self._faces.normal[bf, 1]], axis=1)
            per_face_nodes = [nod1, nod2]
        else:
            gc = -1 * np.ones((F, 4), dtype=F64)
            nrm = self._faces.normal[bf] / self._faces.mesure[bf][:, None]
            v1 = centerc[cellid[bf, 0]]
            u = self._faces.center[bf] - v1
            gamma = _dot3(u, nrm)
            kk = v1 + gamma[:, None] * nrm
            v2 = 2 * kk + (-1 * v1)
            gc[bf, 0:3] = v2
            gc[bf, 3] = gamma
            rows = np.concatenate([v2, cellid[bf, 0][:, None].astype(F64), oldname[bf][:, None].astype(F64),
                                   bf[:, None].astype(F64)], axis=1)
            ncols = 6
            info = np.concatenate([self._faces.center[bf], self._faces.normal[bf]], axis=1)
            per_face_nodes = [fn[bf, 1], fn[bf, 0], fn[bf, 2]]
            quad = fn[bf, -1] == 4
            nod4 = np.where(quad, fn[bf, 3], -1)
            per_face_nodes.append(nod4)
        self._faces.ghostcenter = gc
        # per-node lists, ordered by boundary face then by the node's slot in that face
        N = int(self._nbnodes)
        nb = len(bf)
        slots = len(per_face_nodes)
        node_all = np.concatenate(per_face_nodes).astype(np.int64)
        seq_all = np.concatenate([np.arange(nb, dtype=np.int64) * slots + s for s in range(slots)])
        row_all = np.concatenate([np.arange(nb)] * slots)
        keep = node_all >= 0
        node_all, seq_all, row_all = node_all[keep], seq_all[keep], row_all[keep]
        order = np.lexsort((seq_all, node_all))
        node_s = node_all[order]
        row_s = row_all[order]
        counts = np.bincount(node_s, minlength=N)
        g = max(int(counts.max()) if len(counts) else 0, 1)
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        pos = np.arange(len(node_s)) - start[node_s]
        gcn = -1 * np.ones((N, g, ncols), dtype=F64)
        gcn[node_s, pos] = rows[row_s]
        gfi = -1 * np.ones((N, g, info.shape[1]), dtype=F64)
        gfi[node_s, pos] = info[row_s]
        self._nodes.ghostcenter = gcn
        self._nodes.ghostfaceinfo = gfi
        self._cells.halonid = np.zeros((int(self._nbcells), 2), dtype=I32)

    def _prepare_communication(self):
        """``prepare_comm`` at one rank (``comms/communication.py:139-145``)."""
        self._halos.scount = np.zeros(1, dtype=np.uint32)
        self._halos.rcount = np.zeros(1, dtype=np.uint32)
        self._halos.indsend = np.zeros(1, dtype=I32)
        self._halos.comm_ptr = None
        self._nbhalos = 0

    def _update_periodic_boudaries(self):
        """``domain.py:487-524`` without periodic faces."""
        self._nodes.periodicid = np.zeros((int(self._nbnodes), 2), dtype=I32)
        self._cells.periodicnid = np.zeros((int(self._nbcells), 2), dtype=I32)
        self._cells.periodicfid = np.zeros(int(self._nbcells), dtype=I32)
        self._cells.shift = np.zeros((int(self._nbcells), 3), dtype=F64)
        self._BCs = {"in": ["neumann", 1], "out": ["neumann", 2], "upper": ["neumann", 3], "bottom": ["neumann", 4]}
        if self._dim == 3:
            self._BCs["front"] = ["neumann", 5]
            self._BCs["back"] = ["neumann", 6]

    def _update_haloghost(self):
        """``update_haloghost_info_{2,3}d`` padding at one rank and
        ``create_node_ghostid`` (``ddm_utils2d.py:122-161``)."""
        if 'ghostid' in self._nodes.__dict__:
            return
        for cont, names in ((self._nodes, ('ghostid', 'haloghostid')), (self._cells, ('ghostnid', 'haloghostnid'))):
            for nm in names:
                cont._lazy.pop(nm, None)
        N = int(self._nbnodes)
        C = int(self._nbcells)
        gcn = self._nodes.ghostcenter
        if self._geo is not None:
            self._nodes.ghostid = self._geo['ghostid']
            self._halos.sizehaloghost = 0
            self._nodes.haloghostcenter = -1 * np.ones((N, 1, gcn.shape[2]), dtype=F64)
            self._nodes.haloghostfaceinfo = -1 * np.ones((N, 1, self._nodes.ghostfaceinfo.shape[2]), dtype=F64)
            self._cells.haloghostcenter = np.zeros((0, 3), dtype=F64)
            self._cells.ghostnid = self._geo['ghostnid']
            self._nodes.haloghostid = np.zeros((N, 2), dtype=I32)
            self._cells.haloghostnid = np.zeros((C, 2), dtype=I32)
            return
        fid = gcn[:, :, -1]
        valid = fid != -1
        g = gcn.shape[1]
        ghostid = np.full((N, g + 1), -1, dtype=I32)
        ghostid[:, :g][valid] = fid[valid].astype(I32)
        ghostid[:, -1] = valid.sum(axis=1)
        self._nodes.ghostid = ghostid
        self._halos.sizehaloghost = 0
        self._nodes.haloghostcenter = -1 * np.ones((N, 1, gcn.shape[2]), dtype=F64)
        self._nodes.haloghostfaceinfo = -1 * np.ones((N, 1, self._nodes.ghostfaceinfo.shape[2]), dtype=F64)
        self._cells.haloghostcenter = np.zeros((0, 3), dtype=F64)
        # ghostnid: per cell, the sorted union of its nodes' ghost faces
        nodeid = self._cells.nodeid
        F = np.int64(self._nbfaces)
        keys = []
        for j in range(nodeid.shape[1] - 1):
            rows = np.nonzero(nodeid[:, -1] > j)[0]
            nod = nodeid[rows, j]
            for k in range(g):
                sel = ghostid[nod, -1] > k
                keys.append(rows[sel].astype(np.int64) * F + ghostid[nod[sel], k])
        key = np.unique(np.concatenate(keys)) if keys else np.zeros(0, dtype=np.int64)
        self._cells.ghostnid = _pad_rows(key // F, key % F, C)
        self._nodes.haloghostid = np.zeros((N, 2), dtype=I32)
        self._cells.haloghostnid = np.zeros((C, 2), dtype=I32)

    def _compute_diamondCell_info(self):
        """``variables_{2,3}d`` + ``face_gradient_info_{2,3}d``."""
        if 'R_x' in self._nodes.__dict__:
            return
        for a in ('R_x', 'R_y', 'R_z', 'lambda_x', 'lambda_y', 'lambda_z', 'number'):
            self._nodes._lazy.pop(a, None)
        for a in ('airDiamond', 'param1', 'param2', 'param3', 'param4', 'f_1', 'f_2', 'f_3', 'f_4'):
            self._faces._lazy.pop(a, None)
        if self._geo is not None:
            g = self._geo
            N = int(self._nbnodes)
            self._nodes.lambda_x, self._nodes.lambda_y = g['lambda_x'], g['lambda_y']
            self._nodes.lambda_z = g.get('lambda_z', np.zeros(N))
            self._nodes.R_x, self._nodes.R_y, self._nodes.R_z = g['R_x'], g['R_y'], g.get('R_z', np.zeros(N))
            self._nodes.number = g['number']
            for k in ('airDiamond', 'param1', 'param2', 'param3', 'param4', 'f_1', 'f_2', 'f_3', 'f_4'):
                setattr(self._faces, k, g[k])
        elif self._dim == 2:
            self._variables_2d()
            self._face_gradient_info_2d()
        else:
            self._variables_3d()
            self._face_gradient_info_3d()

    def _node_point_lists(self):
        """(points, node) in the accumulation order of ``variables_2d``: the
        node's cells (ascending), then its ghost mirror centres."""
        ncid = self._nodes.cellid
        gid = self._nodes.ghostid
        return ncid, gid

    def _variables_2d(self):
        """``variables_2d`` (``ddm_utils2d.py:608-705``), accumulation order kept."""
        N = int(self._nbnodes)
        vx = self._nodes.vertex[:, 0]
        vy = self._nodes.vertex[:, 1]
        Ixx = np.zeros(N); Iyy = np.zeros(N); Ixy = np.zeros(N)
        Rx_ = np.zeros(N); Ry_ = np.zeros(N)
        number = np.zeros(N, dtype=I32)
        for arr, pts in ((self._nodes.cellid, self._cells.center), (self._nodes.ghostid, self._faces.ghostcenter)):
            cnt = arr[:, -1]
            for j in range(arr.shape[1] - 1):
                rows = np.nonzero(cnt > j)[0]
                p = pts[arr[rows, j]]
                Rx = p[:, 0] - vx[rows]
                Ry = p[:, 1] - vy[rows]
                Ixx[rows] += Rx * Rx
                Iyy[rows] += Ry * Ry
                Ixy[rows] += Rx * Ry
                Rx_[rows] += Rx
                Ry_[rows] += Ry
                number[rows] += 1
        with np.errstate(divide='ignore', invalid='ignore'):
            D = Ixx * Iyy - Ixy * Ixy
            self._nodes.lambda_x = (Ixy * Ry_ - Iyy * Rx_) / D
            self._nodes.lambda_y = (Ixy * Rx_ - Ixx * Ry_) / D
        self._nodes.lambda_z = np.zeros(N)
        self._nodes.R_x = Rx_
        self._nodes.R_y = Ry_
        self._nodes.R_z = np.zeros(N)
        self._nodes.number = number

    def _face_gradient_info_2d(self):
        """``face_gradient_info_2d`` (``ddm_utils2d.py:553-605``)."""
        cellid = self._faces.cellid
        fn = self._faces.nodeid
        name = self._faces.name
        vertex = self._nodes.vertex
        centerc = self._cells.center
        xy1 = vertex[fn[:, 0], 0:2]
        xy2 = vertex[fn[:, 1], 0:2]
        v1 = centerc[cellid[:, 0], 0:2]
        inner = name == 0
        v2 = self._faces.ghostcenter[:, 0:2].copy()
        v2[inner] = centerc[cellid[inner, 1], 0:2]
        f1 = v1 - xy1
        f2 = xy2 - v1
        f3 = v2 - xy2
        f4 = xy1 - v2
        n1 = self._faces.normal[:, 0]
        n2 = self._faces.normal[:, 1]
        air = 0.5 * ((xy2[:, 0] - xy1[:, 0]) * (v2[:, 1] - v1[:, 1]) + (v1[:, 0] - v2[:, 0]) * (xy2[:, 1] - xy1[:, 1]))
        k = 1. / (2. * air)
        self._faces.airDiamond = air
        self._faces.param1 = k * ((f1[:, 1] + f2[:, 1]) * n1 - (f1[:, 0] + f2[:, 0]) * n2)
        self._faces.param2 = k * ((f2[:, 1] + f3[:, 1]) * n1 - (f2[:, 0] + f3[:, 0]) * n2)
        self._faces.param3 = k * ((f3[:, 1] + f4[:, 1]) * n1 - (f3[:, 0] + f4[:, 0]) * n2)
        self._faces.param4 = k * ((f4[:, 1] + f1[:, 1]) * n1 - (f4[:, 0] + f1[:, 0]) * n2)
        self._faces.f_1 = f1
        self._faces.f_2 = f2
        self._faces.f_3 = f3
        self._faces.f_4 = f4

    def _variables_3d(self):
        """``variables_3d`` (``ddm_utils3d.py:813-981``)."""
        N = int(self._nbnodes)
        v = self._nodes.vertex
        I = {k: np.zeros(N) for k in ('xx', 'yy', 'zz', 'xy', 'xz', 'yz')}
        R = [np.zeros(N), np.zeros(N), np.zeros(N)]
        number = np.zeros(N, dtype=I32)
        for arr, pts in ((self._nodes.cellid, self._cells.center), (self._nodes.ghostid, self._faces.ghostcenter)):
            cnt = arr[:, -1]
            for j in range(arr.shape[1] - 1):
                rows = np.nonzero(cnt > j)[0]
                p = pts[arr[rows, j]]
                d = [p[:, a] - v[rows, a] for a in range(3)]
                I['xx'][rows] += d[0] * d[0]
                I['yy'][rows] += d[1] * d[1]
                I['zz'][rows] += d[2] * d[2]
                I['xy'][rows] += d[0] * d[1]
                I['xz'][rows] += d[0] * d[2]
                I['yz'][rows] += d[1] * d[2]
                for a in range(3):
                    R[a][rows] += d[a]
                number[rows] += 1
        xx, yy, zz, xy, xz, yz = (I[k] for k in ('xx', 'yy', 'zz', 'xy', 'xz', 'yz'))
        Rx, Ry, Rz = R
        with np.errstate(divide='ignore', invalid='ignore'):
            D = xx * yy * zz + 2 * xy * xz * yz - xx * yz * yz - yy * xz * xz - zz * xy * xy
            self._nodes.lambda_x = ((yz * yz - yy * zz) * Rx + (xy * zz - xz * yz) * Ry + (xz * yy - xy * yz) * Rz) / D
            self._nodes.lambda_y = ((xy * zz - xz * yz) * Rx + (xz * xz - xx * zz) * Ry + (yz * xx - xz * xy) * Rz) / D
            self._nodes.lambda_z = ((xz * yy - xy * yz) * Rx + (yz * xx - xz * xy) * Ry + (xy * xy - xx * yy) * Rz) / D
        self._nodes.R_x, self._nodes.R_y, self._nodes.R_z = Rx, Ry, Rz
        self._nodes.number = number

    def _face_gradient_info_3d(self):
        """``face_gradient_info_3d`` (``ddm_utils3d.py:459-527``); ``f_1``/``f_2``
        hold the reference's ``n1``/``n2`` and ``param4``/``f_3``/``f_4`` stay zero."""
        cellid = self._faces.cellid
        fn = self._faces.nodeid
        name = self._faces.name
        vertex = self._nodes.vertex
        centerc = self._cells.center
        i1, i2, i3 = fn[:, 0], fn[:, 1], fn[:, 2]
        i4 = np.where(fn[:, -1] == 4, fn[:, 3] if fn.shape[1] > 4 else i3, i3)
        v1 = centerc[cellid[:, 0]]
        inner = name == 0
        v2 = self._faces.ghostcenter[:, 0:3].copy()
        v2[inner] = centerc[cellid[inner, 1]]
        P = lambda i: vertex[i, 0:3]
        s1 = v2 - P(i2); s2 = P(i4) - P(i2); s3 = v1 - P(i2)
        n1 = (0.5 * _cross(s1, s2)) + (0.5 * _cross(s2, s3))
        s4 = v2 - P(i3); s5 = P(i1) - P(i3); s6 = v1 - P(i3)
        n2 = (0.5 * _cross(s4, s5)) + (0.5 * _cross(s5, s6))
        normal = self._faces.normal
        air = _dot3(normal, v2 - v1)
        self._faces.airDiamond = air
        self._faces.param1 = _dot3(n1, normal) / air
        self._faces.param2 = _dot3(n2, normal) / air
        self._faces.param3 = _dot3(normal, normal) / air
        F = int(self._nbfaces)
        self._faces.param4 = np.zeros(F)
        self._faces.f_1 = n1
        self._faces.f_2 = n2
        self._faces.f_3 = np.zeros((F, 3))
        self._faces.f_4 = np.zeros((F, 3))

    def _compute_orthocenter(self):
        """``dist_ortho_function`` (``ddm_utils2d.py:716-744``), non-unit
        normal kept as in the reference."""
        if self._geo is not None:
            self._faces.dist_ortho = self._geo['dist_ortho']
            self._geo = None  # every stage has taken its arrays
            return
        F = int(self._nbfaces)
        d = np.zeros(F, dtype=F64)
        cellid = self._faces.cellid
        cc = self._cells.center
        cf = self._faces.center
        u = self._faces.normal

        def proj(K, f):
            v = cc[K] - cf[f]
            s = v[:, 0] * u[f, 0] + v[:, 1] * u[f, 1]
            p = cc[K] - s[:, None] * u[f]
            dx = cc[K, 0] - p[:, 0]
            dy = cc[K, 1] - p[:, 1]
            return np.sqrt(dx * dx + dy * dy)

        bf = self._boundaryfaces
        d[bf] = 2 * proj(cellid[bf, 0], bf)
        inn = self._innerfaces
        d[inn] = proj(cellid[inn, 0], inn) + proj(cellid[inn, 1], inn)
        self._faces.dist_ortho = d

    # -- properties (domain.py:850-961) ---------------------------------------------------------
    cells = property(lambda self: self._cells)
    faces = property(lambda self: self._faces)
    nodes = property(lambda self: self._nodes)
    halos = property(lambda self: self._halos)
    dim = property(lambda self: self._dim)
    comm = property(lambda self: self._comm)
    nbnodes = property(lambda self: self._nbnodes)
    nbcells = property(lambda self: self._nbcells)
    nbfaces = property(lambda self: self._nbfaces)
    nbhalos = property(lambda self: self._nbhalos)
    innerfaces = property(lambda self: self._innerfaces)
    infaces = property(lambda self: self._infaces)
    outfaces = property(lambda self: self._outfaces)
    bottomfaces = property(lambda self: self._bottomfaces)
    upperfaces = property(lambda self: self._upperfaces)
    frontfaces = property(lambda self: self._frontfaces)
    backfaces = property(lambda self: self._backfaces)
    halofaces = property(lambda self: self._halofaces)
    innernodes = property(lambda self: self._innernodes)
    innodes = property(lambda self: self._innodes)
    outnodes = property(lambda self: self._outnodes)
    bottomnodes = property(lambda self: self._bottomnodes)
    uppernodes = property(lambda self: self._uppernodes)
    frontnodes = property(lambda self: self._frontnodes)
    backnodes = property(lambda self: self._backnodes)
    halonodes = property(lambda self: self._halonodes)
    boundaryfaces = property(lambda self: self._boundaryfaces)
    boundarynodes = property(lambda self: self._boundarynodes)
    periodicboundaryfaces = property(lambda self: self._periodicboundaryfaces)
    periodicboundarynodes = property(lambda self: self._periodicboundarynodes)
    typeOfCells = property(lambda self: self._typeOfCells)
    bounds = property(lambda self: self._bounds)
    maxcellfid = property(lambda self: int(self._maxcellfid))

    @property
    def device_mesh(self):
        """The GPU copy of the hot-path arrays (built on first use). 2D meshes carry the
        least-squares stencil; 3D meshes upload it on the first gradient call."""
        if getattr(self, '_device_mesh', None) is None:
            from .device import DeviceMesh
            self._device_mesh = DeviceMesh(self, lsq=(int(self._dim) == 2))
        return self._device_mesh

    @property
    def device_diamond(self):
        """The GPU copy of the diamond-scheme geometry (2D, built on first use)."""
        if getattr(self, '_device_diamond', None) is None:
            from .diamond import DeviceDiamond
            self._device_diamond = DeviceDiamond(self.device_mesh, host=getattr(self, '_diamond_host', None))
        return self._device_diamond

    @property
    def backend(self):
        if self._backend is None:
            from .backend import HIPBackend
            self._backend = HIPBackend.for_domain(self)
        return self._backend

    # -- output (ddm/domain.py:624-849): VTU pieces written in the background -------------------
    def _writer(self):
        if getattr(self, '_vtu', None) is None:
            from .vtu import VtuWriter
            self._vtu = VtuWriter(self, 'results')
        return self._vtu

    def save_on_cell(self, dt=0, time=0, niter=0, miter=0, value=None):
        if value is None:
            raise ValueError("value must be given")
        return self._writer().save(miter, ['w'], [value], on_cells=True)

    def save_on_node(self, dt=0, time=0, niter=0, miter=0, value=None):
        if value is None:
            raise ValueError("value must be given")
        return self._writer().save(miter, ['w'], [value], on_cells=False)

    def save_on_cell_multi(self, dt=0, time=0, niter=0, miter=0, variables=None, values=None, file_format="vtu"):
        if file_format != "vtu":
            raise NotImplementedError("only VTU output is written")
        return self._writer().save(miter, list(variables), list(values), on_cells=True)

    def save_on_node_multi(self, dt=0, time=0, niter=0, miter=0, variables=None, values=None, file_format="vtu"):
        if file_format != "vtu":
            raise NotImplementedError("only VTU output is written")
        return self._writer().save(miter, list(variables), list(values), on_cells=False)

    def wait_outputs(self):
        """Join every pending background VTU write."""
        if getattr(self, '_vtu', None) is not None:
            self._vtu.wait()

    def __str__(self):
        return ('\n> dim   :: %d\n> total cells  :: %d\n> total nodes  :: %d\n> total faces  :: %d\n'
                % (self.dim, self.nbcells, self.nbnodes, self.nbfaces))
