"""Device layout of the diamond scheme, 2D and 3D (``mfv_diamond`` in include/mfv.h).

The reference recomputes the pseudo-Laplacian weight of every node incidence inside
each loop that needs it (``ast/functions2d.py:221-225`` in ``centertovertex_2d``,
``:362-364`` and ``:380-382`` in ``get_triplet_2d``)::

    alpha = (1. + lambda_x[n]*(x - x_n) + lambda_y[n]*(y - y_n)) /
            (number[n] + lambda_x[n]*R_x[n] + lambda_y[n]*R_y[n])

The expression depends on geometry only, so it is evaluated once here, elementwise
with the same operation order (hence the same bits), and stored per incidence in CSR
form: ``node_cell`` (``nodes.cellid`` order) and ``node_ghost`` (the valid entries of
``nodes.ghostcenter``, whose centres are ``faces.ghostcenter`` of the listed face). In 3D the
weight has the z terms of ``ast/functions3d.py:265-268`` and ``:532-535``.
"""
import ctypes

import numpy as np

from . import _lib

I32 = np.int32


def _alpha(domain, nodes, x, y, z=None):
    n = domain.nodes
    lx, ly = n.lambda_x[nodes], n.lambda_y[nodes]
    xd = x - n.vertex[nodes, 0]
    yd = y - n.vertex[nodes, 1]
    if z is None:
        den = n.number[nodes] + lx * n.R_x[nodes] + ly * n.R_y[nodes]
        return (1. + lx * xd + ly * yd) / den
    lz = n.lambda_z[nodes]
    zd = z - n.vertex[nodes, 2]
    den = n.number[nodes] + lx * n.R_x[nodes] + ly * n.R_y[nodes] + lz * n.R_z[nodes]
    return (1. + lx * xd + ly * yd + lz * zd) / den


def face_nodes4(domain):
    """[F][4] nodes A, B, C, D of each 3D face (D = C on a triangle, functions3d.py:519-522)."""
    fn = domain.faces.nodeid
    four = fn[:, -1] == 4
    D = np.where(four, fn[:, 3], fn[:, 2]) if fn.shape[1] > 4 else fn[:, 2]
    if fn.shape[1] <= 4 and np.any(four):
        raise ValueError("a 4-node face in a 3-node face table")
    return np.ascontiguousarray(np.column_stack([fn[:, 0], fn[:, 1], fn[:, 2], D]), dtype=I32)


def diamond_layout(domain):
    """Host (numpy) arrays of the diamond descriptor that the mfv_mesh does not hold."""
    dim = int(domain.dim)
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    d = domain
    N = int(d.nbnodes)
    ncid = d.nodes.cellid
    cnt = ncid[:, -1].astype(np.int64)
    ptr = np.zeros(N + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(cnt)
    rows = np.repeat(np.arange(N), cnt)
    mask = np.arange(ncid.shape[1] - 1)[None, :] < cnt[:, None]
    cols = ncid[:, :-1][mask].astype(np.int64)
    cc = d.cells.center
    a_cell = _alpha(d, rows, cc[cols, 0], cc[cols, 1], cc[cols, 2] if dim == 3 else None)

    # ghost centres of a node: (x, y, owner, .., face) in 2D, (x, y, z, owner, .., face) in 3D
    gcn = d.nodes.ghostcenter
    valid = gcn[:, :, -1] != -1
    if np.any(valid[:, 1:] & ~valid[:, :-1]):
        raise ValueError("nodes.ghostcenter has a valid entry after an empty one")
    if np.any(valid != (gcn[:, :, dim] != -1)):
        raise ValueError("nodes.ghostcenter owner and face disagree")
    gcnt = valid.sum(axis=1).astype(np.int64)
    gptr = np.zeros(N + 1, dtype=np.int64)
    gptr[1:] = np.cumsum(gcnt)
    grow = np.repeat(np.arange(N), gcnt)
    gface = gcn[:, :, -1][valid].astype(np.int64)
    gowner = gcn[:, :, dim][valid].astype(np.int64)
    gxyz = gcn[:, :, 0:dim][valid]
    fgc = d.faces.ghostcenter
    if not all(np.array_equal(gxyz[:, k], fgc[gface, k]) for k in range(dim)):
        raise ValueError("nodes.ghostcenter and faces.ghostcenter disagree")
    # centertovertex walks nodes.ghostid: the same faces in the same order
    gid = d.nodes.ghostid
    gmask = np.arange(gid.shape[1] - 1)[None, :] < gid[:, -1][:, None]
    if not (np.array_equal(gid[:, -1].astype(np.int64), gcnt) and np.array_equal(gid[:, :-1][gmask], gface)):
        raise ValueError("nodes.ghostid and nodes.ghostcenter list different faces")
    a_ghost = _alpha(d, grow, gxyz[:, 0], gxyz[:, 1], gxyz[:, 2] if dim == 3 else None)

    fn = d.faces.nodeid
    if dim == 2:
        nodes = np.ascontiguousarray(fn[:, :2], dtype=I32)
        fk = np.concatenate([d.faces.f_1[:, :2], d.faces.f_2[:, :2], d.faces.f_3[:, :2], d.faces.f_4[:, :2]], axis=1)
    else:
        nodes = face_nodes4(d)
        fk = np.concatenate([d.faces.f_1[:, :3], d.faces.f_2[:, :3]], axis=1)
    return {
        'N': N,
        'dim': dim,
        'face_nodes': nodes,
        'face_param': np.ascontiguousarray(np.stack([d.faces.param1, d.faces.param2, d.faces.param3,
                                                     d.faces.param4], axis=1), dtype=np.float64),
        'face_fk': np.ascontiguousarray(fk, dtype=np.float64),
        'face_air': np.ascontiguousarray(d.faces.airDiamond, dtype=np.float64),
        'node_cell_ptr': ptr.astype(I32),
        'node_cell': cols.astype(I32),
        'node_cell_alpha': a_cell.astype(np.float64),
        'node_ghost_ptr': gptr.astype(I32),
        'node_ghost_face': gface.astype(I32),
        'node_ghost_owner': gowner.astype(I32),
        'node_ghost_alpha': a_ghost.astype(np.float64),
    }


def bc_tables(domain, dirichlet_index, dirichletfaces, neumannfaces):
    """Per-variable boundary tables of the diamond kernels (include/mfv.h):
    node_dir[N] (node oldname in BCdirichlet), face_dir_rank[F] (position in the sorted
    dirichletfaces list), face_bc[F] (1 = Neumann list, 2 = Dirichlet list)."""
    N, F = int(domain.nbnodes), int(domain.nbfaces)
    node_dir = np.isin(domain.nodes.oldname, np.asarray(dirichlet_index, dtype=np.int64)).astype(np.uint8)
    rank = np.full(F, -1, dtype=I32)
    dfs = np.asarray(dirichletfaces, dtype=np.int64)
    rank[dfs] = np.arange(len(dfs), dtype=I32)
    face_bc = np.zeros(F, dtype=np.int8)
    face_bc[np.asarray(neumannfaces, dtype=np.int64)] = 1
    face_bc[dfs] = 2
    assert len(node_dir) == N
    return node_dir, rank, face_bc


class DeviceDiamond(object):
    """Device copy of the diamond geometry and its ``mfv_diamond`` descriptor; the
    arrays it shares with the DeviceMesh (cell faces, volumes, normals) are not copied."""

    def __init__(self, dmesh, host=None):
        import torch
        host = host if host is not None else diamond_layout(dmesh.domain)
        self.dmesh = dmesh
        self.N = host['N']
        self.dim = host['dim']
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dmesh.device)
                  for k, v in host.items() if isinstance(v, np.ndarray)}
        d = _lib.MfvDiamond()
        d.nb_cells = dmesh.C
        d.nb_faces = dmesh.F
        d.nb_nodes = self.N
        d.max_cell_faces = dmesh.MF
        d.dim = self.dim
        for k in ('cell_faces', 'cell_nfaces', 'order_sorted', 'cell_volume', 'face_cells', 'face_normal'):
            setattr(d, k, dmesh.t[k].data_ptr())
        for k in ('face_nodes', 'face_param', 'face_fk', 'face_air', 'node_cell_ptr', 'node_cell',
                  'node_cell_alpha', 'node_ghost_ptr', 'node_ghost_face', 'node_ghost_owner', 'node_ghost_alpha'):
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d

    @property
    def ptr(self):
        return ctypes.byref(self.desc)

    def width(self, node_dir, face_dir_rank):
        """Largest number of distinct columns in a row of the assembled matrix (the ELL width
        it needs), from the assembly's first pass."""
        import torch
        from .device import stream_ptr
        C = self.dmesh.C
        sw = torch.zeros(max(1, (C + 63) // 64), dtype=torch.int32, device=self.dmesh.device)
        info = torch.zeros(2, dtype=torch.int32, device=self.dmesh.device)
        _lib.call('mfv_diamond_widths', self.ptr, node_dir.data_ptr(), face_dir_rank.data_ptr(), sw.data_ptr(),
                  info.data_ptr(), stream_ptr())
        ov, wmax = info.tolist()
        if ov:
            raise _lib.MfvError("a diamond matrix row has more distinct columns than the assembly holds")
        return max(1, int(wmax))
