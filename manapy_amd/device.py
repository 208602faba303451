"""Device residency: GPU layout of a Domain and the mfv_mesh descriptor.

``gpu_layout(domain)`` regroups the reference arrays for coalesced access
(per-cell tables become column-major ``[slot][cell]``) and precomputes the two
residual orders that make the cell-centric gather reproduce the reference's
face-ordered scatters:

* ``order_split`` -- inner faces by increasing id, then boundary faces by
  increasing id (the loop order of ``explicitscheme_convective_2d``,
  ``solvers/advec/fvm_utils.py:52-152``, and ``explicitscheme_convective_SW``,
  ``solvers/shallowater/fvm_utils.py:363-496``);
* ``order_sorted`` -- all faces by increasing id (the single loop of
  ``explicitscheme_euler_3d_rusanov``, ``solvers/euler/fvm_utils3d.py:1101``).

``DeviceMesh`` uploads the layout with PyTorch (device memory and streams
only) and exposes the ``mfv_mesh`` descriptor passed to every C-ABI call.
"""
import ctypes

import numpy as np

from . import _lib

I32 = np.int32


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise _lib.MfvError("no ROCm GPU visible: the manapy_amd kernels run on the GPU only")
    return torch


def gpu_layout(domain, lsq=True):
    """Host-side arrays of the device mesh (numpy)."""
    dim = int(domain.dim)
    C = int(domain.nbcells)
    F = int(domain.nbfaces)
    faceid = domain.cells.faceid
    MF = faceid.shape[1] - 1
    nf = faceid[:, -1].astype(I32)
    fids = faceid[:, :MF]
    name = domain.faces.name.astype(I32)
    bad = np.nonzero((domain.faces.cellid[:, 1] == -1) & ~np.isin(name, [1, 2, 3, 4, 5, 6]))[0]
    if len(bad):
        raise ValueError("%d boundary faces have no boundary tag (name %s): the reference skips such faces "
                         "in every loop; tag the mesh boundary" % (len(bad), np.unique(name[bad])))
    valid = np.arange(MF)[None, :] < nf[:, None]
    big = np.int64(1) << 40
    fid64 = np.where(valid, fids, big).astype(np.int64)
    is_bnd = (name[np.where(valid, fids, 0)] != 0) & valid
    order_split = np.argsort(fid64 + is_bnd.astype(np.int64) * (np.int64(1) << 32), axis=1, kind='stable')
    order_sorted = np.argsort(fid64, axis=1, kind='stable')
    out = {
        'dim': dim, 'C': C, 'F': F, 'MF': MF,
        'face_cells': np.ascontiguousarray(domain.faces.cellid, dtype=I32),
        'face_name': name,
        'face_normal': np.ascontiguousarray(domain.faces.normal[:, :3], dtype=np.float64),
        'face_mesure': np.ascontiguousarray(domain.faces.mesure, dtype=np.float64),
        'face_center': np.ascontiguousarray(domain.faces.center[:, :3], dtype=np.float64),
        'cell_center': np.ascontiguousarray(domain.cells.center[:, :3], dtype=np.float64),
        'cell_volume': np.ascontiguousarray(domain.cells.volume, dtype=np.float64),
        'cell_faces': np.ascontiguousarray(fids.T, dtype=I32),
        'cell_nfaces': nf,
        'order_split': np.ascontiguousarray(order_split.T, dtype=np.uint8),
        'order_sorted': np.ascontiguousarray(order_sorted.T, dtype=np.uint8),
        'bnd_faces': np.ascontiguousarray(domain.boundaryfaces, dtype=I32),
    }
    gc = domain.faces.ghostcenter
    ghost = np.zeros((F, 3), dtype=np.float64)
    ghost[:, :min(3, gc.shape[1]) if dim == 3 else 2] = gc[:, :3 if dim == 3 else 2]
    out['face_ghost'] = ghost
    if lsq:
        cnid = domain.cells.cellnid
        gnid = domain.cells.ghostnid
        K = cnid.shape[1] - 1
        G = gnid.shape[1] - 1
        out['K'] = K
        out['G'] = G
        out['lsq_cells'] = np.ascontiguousarray(cnid[:, :K].T, dtype=I32) if K else np.zeros((1, C), I32)
        out['lsq_ncells'] = np.ascontiguousarray(cnid[:, -1], dtype=I32)
        out['lsq_ghosts'] = np.ascontiguousarray(gnid[:, :G].T, dtype=I32) if G else np.zeros((1, C), I32)
        out['lsq_nghosts'] = np.ascontiguousarray(gnid[:, -1], dtype=I32)
    return out


def dptr(t):
    """Device address of a contiguous CUDA/ROCm tensor (0 for None)."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise _lib.MfvError("expected a contiguous device tensor")
    return t.data_ptr()


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class DeviceMesh(object):
    """Device copy of a Domain's hot-path arrays and its ``mfv_mesh`` descriptor."""

    def __init__(self, domain, lsq=True, device=None, host=None):
        torch = require_gpu()
        self.domain = domain
        self.device = torch.device(device if device is not None else 'cuda')
        host = host if host is not None else gpu_layout(domain, lsq=lsq)
        self.host = host
        self.dim = host['dim']
        self.C = host['C']
        self.F = host['F']
        self.MF = host['MF']
        self.t = {}
        for k, v in host.items():
            if isinstance(v, np.ndarray):
                self.t[k] = torch.from_numpy(np.ascontiguousarray(v)).to(self.device)
        d = _lib.MfvMesh()
        d.dim = self.dim
        d.nb_cells = self.C
        d.nb_faces = self.F
        d.max_cell_faces = self.MF
        d.lsq_width = host.get('K', 0)
        d.ghost_width = host.get('G', 0)
        for k in ('face_cells', 'face_name', 'face_normal', 'face_mesure', 'face_center', 'face_ghost',
                  'cell_center', 'cell_volume', 'cell_faces', 'cell_nfaces', 'order_split', 'order_sorted',
                  'bnd_faces'):
            setattr(d, k, self.t[k].data_ptr())
        if lsq:
            for k in ('lsq_cells', 'lsq_ncells', 'lsq_ghosts', 'lsq_nghosts'):
                setattr(d, k, self.t[k].data_ptr())
        d.nb_bnd = len(host['bnd_faces'])
        self.desc = d
        self.lsq = lsq

    @property
    def ptr(self):
        return ctypes.byref(self.desc)

    def ensure_lsq(self):
        """Upload the least-squares stencil (cellnid / ghostnid) if it is not resident yet."""
        if self.lsq:
            return
        import torch
        host = gpu_layout(self.domain, lsq=True)
        for k in ('lsq_cells', 'lsq_ncells', 'lsq_ghosts', 'lsq_nghosts'):
            self.t[k] = torch.from_numpy(host[k]).to(self.device)
            setattr(self.desc, k, self.t[k].data_ptr())
        self.desc.lsq_width = host['K']
        self.desc.ghost_width = host['G']
        self.lsq = True

    @property
    def dist_ortho(self):
        import torch
        if 'dist_ortho' not in self.t:
            self.t['dist_ortho'] = torch.as_tensor(self.domain.faces.dist_ortho, device=self.device)
        return self.t['dist_ortho']

    def zeros(self, *shape, dtype=None):
        import torch
        return torch.zeros(*shape, dtype=dtype or torch.float64, device=self.device)

    def to_device(self, a):
        import torch
        return torch.as_tensor(np.ascontiguousarray(a), device=self.device)


def sw_packed_layout(domain):
    """Host tables of the packed shallow-water path (see mfv_sw_packed in include/mfv.h)."""
    C = int(domain.nbcells)
    faceid = domain.cells.faceid
    MF = faceid.shape[1] - 1
    if MF not in (3, 4):
        raise ValueError("packed shallow-water path needs 3 or 4 faces per cell")
    nf = faceid[:, -1]
    cellid = domain.faces.cellid
    slot_mask = np.arange(MF)[None, :] < nf[:, None]
    fids = np.where(slot_mask, faceid[:, :MF], 0).astype(np.int64)
    rows = np.repeat(np.arange(C, dtype=np.int64), MF).reshape(C, MF)
    a = cellid[fids, 0].astype(np.int64)
    b = cellid[fids, 1].astype(np.int64)
    side = (b == rows) & (b >= 0)
    other = np.where(side, a, b)
    # slot of face f in the other cell's row
    F = int(domain.nbfaces)
    r_idx, s_idx = np.nonzero(slot_mask)
    key = fids[r_idx, s_idx] * 2 + (cellid[fids[r_idx, s_idx], 1] == r_idx)
    lut = np.full(2 * F, -1, dtype=np.int64)
    lut[key] = s_idx
    wall = (b < 0) | ~slot_mask
    other_side = (~side).astype(np.int64)  # the other cell is cellid[1] when we are cellid[0]
    k = lut[fids * 2 + other_side]
    code = (other << 3) | (k << 1) | side.astype(np.int64)
    code = np.where(wall, -1, code)
    if np.any(code[~wall] < 0) or np.any(k[~wall] < 0):
        raise ValueError("inconsistent face/cell tables")
    nbr = np.ascontiguousarray(code.T, dtype=np.int32)
    cf = domain.faces.center
    cc = domain.cells.center
    cfrec = np.zeros((C, MF, 4))
    cfrec[:, :, 0] = np.where(slot_mask, cf[fids, 0] - cc[:, None, 0], 0.)
    cfrec[:, :, 1] = np.where(slot_mask, cf[fids, 1] - cc[:, None, 1], 0.)
    cfgeo = np.zeros((C, MF, 4))
    cfgeo[:, :, 0] = np.where(slot_mask, domain.faces.normal[fids, 0], 0.)
    cfgeo[:, :, 1] = np.where(slot_mask, domain.faces.normal[fids, 1], 0.)
    cfgeo[:, :, 2] = np.where(slot_mask, domain.faces.mesure[fids], 1.)
    L = gpu_layout(domain, lsq=False)
    osplit = L['order_split'].T.astype(np.uint32)  # [C][MF]
    order = np.zeros(C, dtype=np.uint32)
    for r in range(MF):
        order |= (osplit[:, r] & 3) << (2 * r)
    return {'nbr': nbr, 'cfrec': cfrec, 'cfgeo': cfgeo, 'order': order.astype(np.uint8), 'MF': MF}


class SwPacked(object):
    """Device tables and records of the packed shallow-water path for a DeviceMesh."""

    def __init__(self, dmesh):
        import torch
        host = sw_packed_layout(dmesh.domain)
        C = dmesh.C
        MF = host['MF']
        dev = dmesh.device
        self.cellrec = torch.zeros((C, 8), dtype=torch.float64, device=dev)
        self.facerec = torch.zeros((C, MF, 4), dtype=torch.float64, device=dev)
        self.cfrec = torch.as_tensor(host['cfrec'], device=dev)
        self.cfgeo = torch.as_tensor(host['cfgeo'], device=dev)
        self.nbr = torch.as_tensor(host['nbr'], device=dev)
        self.order = torch.as_tensor(host['order'], device=dev)
        d = _lib.MfvSwPacked()
        d.nb_cells = C
        d.max_cell_faces = MF
        d.cellrec = self.cellrec.data_ptr()
        d.facerec = self.facerec.data_ptr()
        d.cfrec = self.cfrec.data_ptr()
        d.cfgeo = self.cfgeo.data_ptr()
        d.nbr = self.nbr.data_ptr()
        d.order = self.order.data_ptr()
        self.desc = d

    @property
    def ptr(self):
        return ctypes.byref(self.desc)


def sw_tiles_layout(nbr, C, T=256, tile_cell0=None, tile_ncells=None):
    """Tiles of the fused mixed step (mfv_sw_tiles in include/mfv.h) from the neighbour codes
    nbr [MF][C]: by default consecutive runs of T cells; halo lists and block-local codes."""
    nbr = np.asarray(nbr)
    MF = nbr.shape[0]
    if tile_cell0 is None:
        tile_cell0 = np.arange(0, C, T, dtype=np.int64)
        tile_ncells = np.minimum(T, C - tile_cell0)
    tile_cell0 = np.asarray(tile_cell0, dtype=np.int64)
    tile_ncells = np.asarray(tile_ncells, dtype=np.int64)
    if np.any(tile_ncells > T) or np.any(tile_ncells <= 0):
        raise ValueError("tiles must hold 1..T cells")
    nt = len(tile_cell0)
    tile_of = np.full(C, -1, dtype=np.int64)
    for_tile = np.repeat(np.arange(nt), tile_ncells)
    cells = np.concatenate([np.arange(a, a + n) for a, n in zip(tile_cell0, tile_ncells)]) if nt else np.zeros(0, int)
    tile_of[cells] = for_tile
    code = nbr.astype(np.int64)                               # [MF][C]
    valid = code >= 0
    j = np.where(valid, code >> 3, 0)
    ti = np.broadcast_to(tile_of[None, :], code.shape)       # tile of the cell owning the slot
    covered = ti >= 0
    outside = valid & covered & (tile_of[j] != ti)
    # halo of each tile: the distinct outside neighbours, ascending
    key = np.unique(ti[outside] * np.int64(C) + j[outside])
    h_tile = key // C
    h_cell = key % C
    halo_ptr = np.zeros(nt + 1, dtype=np.int64)
    np.add.at(halo_ptr, h_tile + 1, 1)
    halo_ptr = np.cumsum(halo_ptr)
    local = np.where(valid, j - tile_cell0[np.maximum(ti, 0)], 0)
    pos = np.searchsorted(key, ti * np.int64(C) + j)
    local = np.where(outside, tile_ncells[np.maximum(ti, 0)] + (pos - halo_ptr[np.maximum(ti, 0)]), local)
    lnbr = np.where(valid, (local << 3) | (code & 7), -1)
    max_halo = int(np.max(np.diff(halo_ptr))) if nt else 0
    return {'T': int(T), 'nb_tiles': nt, 'max_halo': max_halo, 'MF': MF,
            'tile_cell0': tile_cell0.astype(np.int32), 'tile_ncells': tile_ncells.astype(np.int32),
            'halo_ptr': halo_ptr.astype(np.int32), 'halo_cells': h_cell.astype(np.int32),
            'lnbr': np.ascontiguousarray(lnbr, dtype=np.int32)}


class SwTiles(object):
    """Device copy of a tiles layout and its mfv_sw_tiles descriptor."""

    def __init__(self, host, device):
        import torch
        self.host = host
        self.t = {k: torch.from_numpy(np.ascontiguousarray(host[k])).to(device)
                  for k in ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'lnbr')}
        if host['halo_cells'].size == 0:
            self.t['halo_cells'] = torch.zeros(1, dtype=torch.int32, device=device)
        d = _lib.MfvSwTiles()
        d.nb_tiles, d.tile_size, d.max_halo = host['nb_tiles'], host['T'], host['max_halo']
        for k in ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'lnbr'):
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d

    @property
    def ptr(self):
        return ctypes.byref(self.desc)


class SwMixed(object):
    """Device tables of the mixed-precision shallow-water path (mfv_sw_mixed)."""

    def __init__(self, dmesh, host=None):
        import torch
        host = host if host is not None else sw_packed_layout(dmesh.domain)
        C = dmesh.C
        MF = host['MF']
        dev = dmesh.device
        f32 = torch.float32
        self.state = torch.zeros((C, 4), dtype=torch.float64, device=dev)
        self.state2 = torch.zeros((C, 4), dtype=torch.float64, device=dev)  # fused step's second buffer
        self.nbr_host = host['nbr']
        self.vol = dmesh.t['cell_volume']
        self.facerec = torch.zeros((C, MF, 4), dtype=f32, device=dev)
        self.geoA = torch.zeros((C, MF, 4), dtype=f32, device=dev)
        self.geoB = torch.zeros((C, MF, 2), dtype=f32, device=dev)
        self.nbr = torch.as_tensor(host['nbr'], device=dev)
        self.order = torch.as_tensor(host['order'], device=dev)
        self.partial = torch.zeros(3 * ((C + 255) // 256) + 3, dtype=torch.float64, device=dev)
        self.sums = torch.zeros(3, dtype=torch.float64, device=dev)
        d = _lib.MfvSwMixed()
        d.nb_cells = C
        d.max_cell_faces = MF
        for k in ('state', 'vol', 'facerec', 'geoA', 'geoB', 'nbr', 'order'):
            setattr(d, k, getattr(self, k).data_ptr())
        self.desc = d
        self._tiles = {}

    def tiles(self, T=256):
        """The fused step's tiles of T cells (built on first use)."""
        if T not in self._tiles:
            self._tiles[T] = SwTiles(sw_tiles_layout(self.nbr_host, self.desc.nb_cells, T), self.state.device)
        return self._tiles[T]

    def step_tiled(self, dmesh, grav, cfl, order, ring, step, T=256, stream=None):
        """One fused step: reads the current state buffer, writes the other, swaps them."""
        _lib.call('mfv_sw_step_mixed_tiled', dmesh.ptr, self.ptr, self.tiles(T).ptr, self.state2.data_ptr(),
                  float(grav), float(cfl), int(order), ring.data_ptr(), int(step),
                  stream if stream is not None else stream_ptr())
        self.state, self.state2 = self.state2, self.state
        self.desc.state = self.state.data_ptr()

    @property
    def ptr(self):
        return ctypes.byref(self.desc)
