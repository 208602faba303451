"""Device layout of the 2D diamond scheme (``mfv_diamond`` in include/mfv.h).

The reference recomputes the pseudo-Laplacian weight of every node incidence inside
each loop that needs it (``ast/functions2d.py:221-225`` in ``centertovertex_2d``,
``:362-364`` and ``:380-382`` in ``get_triplet_2d``)::

    alpha = (1. + lambda_x[n]*(x - x_n) + lambda_y[n]*(y - y_n)) /
            (number[n] + lambda_x[n]*R_x[n] + lambda_y[n]*R_y[n])

The expression depends on geometry only, so it is evaluated once here, elementwise
with the same operation order (hence the same bits), and stored per incidence in CSR
form: ``node_cell`` (``nodes.cellid`` order) and ``node_ghost`` (the valid entries of
``nodes.ghostcenter``, whose centres are ``faces.ghostcenter`` of the listed face).
"""
import ctypes

import numpy as np

from . import _lib

I32 = np.int32


def _alpha(domain, nodes, x, y):
    n = domain.nodes
    lx, ly = n.lambda_x[nodes], n.lambda_y[nodes]
    xd = x - n.vertex[nodes, 0]
    yd = y - n.vertex[nodes, 1]
    den = n.number[nodes] + lx * n.R_x[nodes] + ly * n.R_y[nodes]
    return (1. + lx * xd + ly * yd) / den


def diamond_layout(domain):
    """Host (numpy) arrays of the diamond descriptor that the mfv_mesh does not hold."""
    if int(domain.dim) != 2:
        raise NotImplementedError("the diamond scheme is implemented for 2D meshes")
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
    a_cell = _alpha(d, rows, cc[cols, 0], cc[cols, 1])

    gcn = d.nodes.ghostcenter
    valid = gcn[:, :, -1] != -1
    gcnt = valid.sum(axis=1).astype(np.int64)
    gptr = np.zeros(N + 1, dtype=np.int64)
    gptr[1:] = np.cumsum(gcnt)
    grow = np.repeat(np.arange(N), gcnt)
    gface = gcn[:, :, -1][valid].astype(np.int64)
    gowner = gcn[:, :, 2][valid].astype(np.int64)
    gxy = gcn[:, :, 0:2][valid]
    fgc = d.faces.ghostcenter
    if not (np.array_equal(gxy[:, 0], fgc[gface, 0]) and np.array_equal(gxy[:, 1], fgc[gface, 1])):
        raise ValueError("nodes.ghostcenter and faces.ghostcenter disagree")
    a_ghost = _alpha(d, grow, gxy[:, 0], gxy[:, 1])

    fn = d.faces.nodeid
    return {
        'N': N,
        'face_nodes': np.ascontiguousarray(fn[:, :2], dtype=I32),
        'face_param': np.ascontiguousarray(np.stack([d.faces.param1, d.faces.param2, d.faces.param3,
                                                     d.faces.param4], axis=1), dtype=np.float64),
        'face_fk': np.ascontiguousarray(np.concatenate([d.faces.f_1[:, :2], d.faces.f_2[:, :2], d.faces.f_3[:, :2],
                                                        d.faces.f_4[:, :2]], axis=1), dtype=np.float64),
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
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dmesh.device)
                  for k, v in host.items() if isinstance(v, np.ndarray)}
        d = _lib.MfvDiamond()
        d.nb_cells = dmesh.C
        d.nb_faces = dmesh.F
        d.nb_nodes = self.N
        d.max_cell_faces = dmesh.MF
        for k in ('cell_faces', 'cell_nfaces', 'order_sorted', 'cell_volume', 'face_cells', 'face_normal'):
            setattr(d, k, dmesh.t[k].data_ptr())
        for k in ('face_nodes', 'face_param', 'face_fk', 'face_air', 'node_cell_ptr', 'node_cell',
                  'node_cell_alpha', 'node_ghost_ptr', 'node_ghost_face', 'node_ghost_owner', 'node_ghost_alpha'):
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d

    @property
    def ptr(self):
        return ctypes.byref(self.desc)
