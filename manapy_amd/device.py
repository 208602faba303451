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
        self.nbr_host = host['nbr']
        self._lsq = dmesh.host
        self._tiles = {}
        self.cellrec2 = None  # second record buffer of the fused tiled step (on first use)
        self._rec_synced = None
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

    def pack(self, dmesh, h, hu, hv, Z, zgrad, order, stream=None):
        """mfv_sw_pack into the current record buffer (marks the second buffer stale)."""
        _lib.call('mfv_sw_pack', dmesh.ptr, self.ptr, h.data_ptr(), hu.data_ptr(), hv.data_ptr(), Z.data_ptr(),
                  zgrad.data_ptr(), int(order), stream if stream is not None else stream_ptr())
        self._rec_synced = None

    def tiles(self, T=128):
        """Tiles of the fused float64 step for blocks of T threads (built on first use)."""
        if T not in self._tiles:
            self._tiles[T] = SwTiles(sw_tiles_native(self.nbr_host, self.desc.nb_cells, T, T, self._lsq),
                                     self.cellrec.device, stage=True)
        return self._tiles[T]

    def lds_bytes(self, T=128):
        h = self.tiles(T).host
        return ((h['max_ext'] + 1) // 2 * 2) * 40 + h['max_items'] * h['MF'] * 24

    def step_tiled(self, dmesh, grav, cfl, order, ring, step, T=128, stream=None, noalloc=False):
        """One fused float64 step: reads the current record buffer, writes the other, swaps
        them. The second buffer takes a copy of the records (static fields included) after
        pack(); the static fields change only there. noalloc: the entry point that allocates
        nothing (mfv_sw_step_packed_tiled_noalloc; same bits)."""
        if self.cellrec2 is None or self._rec_synced is not self.cellrec:
            if self.cellrec2 is None:
                import torch
                self.cellrec2 = torch.empty_like(self.cellrec)
            self.cellrec2.copy_(self.cellrec)
        _lib.call('mfv_sw_step_packed_tiled_noalloc' if noalloc else 'mfv_sw_step_packed_tiled', dmesh.ptr, self.ptr,
                  self.tiles(T).ptr, self.cellrec2.data_ptr(),
                  float(grav), float(cfl), int(order), ring.data_ptr(), int(step),
                  stream if stream is not None else stream_ptr())
        self.cellrec, self.cellrec2 = self.cellrec2, self.cellrec
        self.desc.cellrec = self.cellrec.data_ptr()
        self._rec_synced = self.cellrec


def sw_tiles_native(nbr, C, T, N, lsq, ranges=None):
    """Tiles built by the library's host builder (mfv_sw_tiles_build): the same tables as
    sw_tiles_layout, tiles grown under the block's item budget. ranges: (cell0s, ncells)
    of the instances, default the whole mesh."""
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    nbr = i32(nbr)
    MF = nbr.shape[0]
    K, G = int(lsq['lsq_cells'].shape[0]), int(lsq['lsq_ghosts'].shape[0])
    arrs = [i32(lsq[k]) for k in ('lsq_cells', 'lsq_ncells', 'lsq_ghosts', 'lsq_nghosts', 'face_cells')]
    r0, rn = (np.array([0]), np.array([C])) if ranges is None else ranges
    r0, rn = i32(r0), i32(rn)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    h = _lib.MfvSwTilesHost()
    _lib.call('mfv_sw_tiles_build', int(C), MF, P(nbr), K, P(arrs[0]), P(arrs[1]), G, P(arrs[2]), P(arrs[3]),
              P(arrs[4]), len(r0), P(r0), P(rn), int(T), int(N), ctypes.byref(h))
    try:
        nt, I = h.nb_tiles, h.nb_items

        def take(ptr, n, dt, shape=None):
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,)).copy()
            return a.reshape(shape) if shape is not None else a
        out = {'T': int(T), 'nb_tiles': nt, 'MF': MF, 'max_halo': h.max_halo, 'max_ext': h.max_ext, 'items': I,
               'max_items': h.max_items,
               'tile_cell0': take(h.tile_cell0, nt, np.int32), 'tile_ncells': take(h.tile_ncells, nt, np.int32),
               'tile_nhalo': take(h.tile_nhalo, nt, np.int32), 'ext_ptr': take(h.ext_ptr, nt + 1, np.int32),
               'ext_cells': take(h.ext_cells, max(h.ext_len, 1), np.int32)[:h.ext_len],
               'item_ptr': take(h.item_ptr, nt + 1, np.int32),
               'lnbr': take(h.lnbr, MF * C, np.int32, (MF, C)),
               'lst': take(h.lst, sw_tiles_kp(K) * I, np.uint16, (I, sw_tiles_kp(K))),
               'lgh': take(h.lgh, max(G, 1) * I, np.uint16, (max(G, 1), I)),
               'lfn': take(h.lfn, MF * I, np.int32, (MF, I))}
        out['tile_next'] = (np.diff(out['ext_ptr']) + out['tile_ncells']).astype(np.int32)
    finally:
        _lib.call('mfv_sw_tiles_host_free', ctypes.byref(h))
    return out


def sw_tiles_kp(K):
    """Row stride of the tiles' stencil table (MFV_SW_TILES_KP): K + 1 (the handles and the
    counts entry) rounded up to a multiple of 8, so that a row is whole 16-byte vectors."""
    return (int(K) + 8) & ~7


def sw_tiles_fitted(nbr, C, T=256, tile_cell0=None, tile_ncells=None, lsq=None, N=None):
    """sw_tiles_layout, then tiles whose cells + halo exceed the block of T threads are split
    into runs of 64-cell multiples until every tile is reconstructed in one pass."""
    L = sw_tiles_layout(nbr, C, T, tile_cell0, tile_ncells, lsq, N)
    for _ in range(6):
        items = np.diff(L['item_ptr'])
        big = items > T
        if not big.any():
            break
        c0, nc = L['tile_cell0'].astype(np.int64), L['tile_ncells'].astype(np.int64)
        new0, newn = [], []
        for a, n, b in zip(c0, nc, big):
            if b and n > 64:
                h = max(64, (n // 2 + 63) // 64 * 64)
                new0 += [a, a + h]
                newn += [h, n - h]
            else:
                new0.append(a)
                newn.append(n)
        L = sw_tiles_layout(nbr, C, T, np.array(new0), np.array(newn), lsq)
    return L


def sw_tiles_layout(nbr, C, T=256, tile_cell0=None, tile_ncells=None, lsq=None, N=None):
    """Tiles of the fused mixed step (mfv_sw_tiles in include/mfv.h) from the neighbour codes
    nbr [MF][C] and the least-squares tables ``lsq`` (lsq_cells [K][C], lsq_ncells,
    lsq_ghosts [G][C], lsq_nghosts, face_cells of a gpu_layout): by default consecutive runs
    of N cells (N <= T, default T; T is the block size). A tile's items are its cells then its
    halo (face neighbours outside it);
    its ext list is its items then every other cell their reconstruction reads (stencil
    cells, ghost owners, face neighbours). Tables give positions in the ext list."""
    nbr = np.asarray(nbr)
    MF = nbr.shape[0]
    if tile_cell0 is None:
        N = T if N is None else int(N)
        if N % 64 or not 0 < N <= T:
            raise ValueError("cells per tile must be a multiple of 64 in (0, T]")
        tile_cell0 = np.arange(0, C, N, dtype=np.int64)
        tile_ncells = np.minimum(N, C - tile_cell0)
    tile_cell0 = np.asarray(tile_cell0, dtype=np.int64)
    tile_ncells = np.asarray(tile_ncells, dtype=np.int64)
    if np.any(tile_ncells > T) or np.any(tile_ncells <= 0):
        raise ValueError("tiles must hold 1..T cells")
    nt = len(tile_cell0)
    tile_of = np.full(C, -1, dtype=np.int64)
    cells = np.concatenate([np.arange(a, a + n) for a, n in zip(tile_cell0, tile_ncells)]) if nt else np.zeros(0, int)
    tile_of[cells] = np.repeat(np.arange(nt), tile_ncells)
    code = nbr.astype(np.int64)                               # [MF][C]
    valid = code >= 0
    j = np.where(valid, code >> 3, 0)
    ti = np.broadcast_to(tile_of[None, :], code.shape)
    outside = valid & (ti >= 0) & (tile_of[j] != ti)
    key = np.unique(ti[outside] * np.int64(C) + j[outside])  # halo pairs (tile, cell), ascending
    h_tile, h_cell = key // C, key % C
    nhalo = np.bincount(h_tile, minlength=nt).astype(np.int64)
    hptr = np.concatenate([[0], np.cumsum(nhalo)])
    # block-local codes of the flux gather: positions in the item (= ext) list
    t0 = np.maximum(ti, 0)
    local = np.where(valid, j - tile_cell0[t0], 0)
    pos = np.searchsorted(key, ti * np.int64(C) + j)
    local = np.where(outside, tile_ncells[t0] + (pos - hptr[t0]), local)
    lnbr = np.where(valid, (local << 3) | (code & 7), -1)
    out = {'T': int(T), 'nb_tiles': nt, 'MF': MF,
           'tile_cell0': tile_cell0.astype(np.int32), 'tile_ncells': tile_ncells.astype(np.int32),
           'tile_nhalo': nhalo.astype(np.int32), 'max_halo': int(nhalo.max()) if nt else 0,
           'lnbr': np.ascontiguousarray(lnbr, dtype=np.int32)}
    if lsq is None:
        raise ValueError("the least-squares tables (lsq=gpu_layout(...)) are required")
    # items: tile cells then halo, per tile
    n_items = tile_ncells + nhalo
    item_ptr = np.concatenate([[0], np.cumsum(n_items)])
    I = int(item_ptr[-1])
    item_tile = np.repeat(np.arange(nt), n_items)
    item_cell = np.empty(I, dtype=np.int64)
    first = item_ptr[:-1]
    is_cell = np.zeros(I, dtype=bool)
    for_cells = np.concatenate([np.arange(f, f + n) for f, n in zip(first, tile_ncells)]) if nt else np.zeros(0, int)
    is_cell[for_cells] = True
    item_cell[for_cells] = cells
    item_cell[~is_cell] = h_cell                               # halo in (tile, cell) order
    K = lsq['lsq_cells'].shape[0]
    G = lsq['lsq_ghosts'].shape[0]
    nn = lsq['lsq_ncells'][item_cell]
    ng = lsq['lsq_nghosts'][item_cell]
    sten = lsq['lsq_cells'][:, item_cell].astype(np.int64)     # [K][I]
    sten_ok = np.arange(K)[:, None] < nn[None, :]
    gf = lsq['lsq_ghosts'][:, item_cell].astype(np.int64)      # [G][I] face ids
    g_ok = np.arange(G)[:, None] < ng[None, :]
    gown = np.where(g_ok, lsq['face_cells'][np.where(g_ok, gf, 0), 0], 0).astype(np.int64)
    fn_code = code[:, item_cell]                               # [MF][I]
    fn_ok = fn_code >= 0
    fn = np.where(fn_ok, fn_code >> 3, 0)
    # ext list: items first (tile cells, halo), then every other cell read, ascending
    want_t = np.concatenate([np.broadcast_to(item_tile, sten.shape)[sten_ok],
                             np.broadcast_to(item_tile, gown.shape)[g_ok],
                             np.broadcast_to(item_tile, fn.shape)[fn_ok]])
    want_c = np.concatenate([sten[sten_ok], gown[g_ok], fn[fn_ok]])
    wkey = np.unique(want_t * np.int64(C) + want_c)
    ikey = item_tile * np.int64(C) + item_cell
    order = np.argsort(ikey, kind='stable')
    ikey_sorted = ikey[order]
    in_items = np.isin(wkey, ikey_sorted, assume_unique=False)
    xkey = wkey[~in_items]                                     # extras, ascending per tile
    x_tile = xkey // C
    nextra = np.bincount(x_tile, minlength=nt).astype(np.int64)
    xptr = np.concatenate([[0], np.cumsum(nextra)])
    n_ext = n_items + nextra
    ext_ptr = np.concatenate([[0], np.cumsum(nhalo + nextra)])  # ext_cells holds halo + extras
    ext_cells = np.empty(int(ext_ptr[-1]), dtype=np.int64)
    hpos = ext_ptr[:-1]
    for_halo = np.concatenate([np.arange(a, a + n) for a, n in zip(hpos, nhalo)]) if nt else np.zeros(0, int)
    ext_cells[for_halo] = h_cell
    for_x = np.concatenate([np.arange(a + h, a + h + n) for a, h, n in zip(hpos, nhalo, nextra)]) \
        if nt else np.zeros(0, int)
    ext_cells[for_x] = xkey % C

    def ext_pos(t, c):
        """Position of cell c in tile t's ext list (all queried pairs exist)."""
        t = np.asarray(t, dtype=np.int64)
        c = np.asarray(c, dtype=np.int64)
        k = t * np.int64(C) + c
        r = np.full(k.shape, -1, dtype=np.int64)
        inc = (c >= tile_cell0[t]) & (c < tile_cell0[t] + tile_ncells[t])
        r[inc] = (c - tile_cell0[t])[inc]
        ph = np.searchsorted(key, k)
        ph_ok = (~inc) & (ph < len(key))
        ph_ok[ph_ok] = key[ph[ph_ok]] == k[ph_ok]
        r[ph_ok] = (tile_ncells[t] + ph - hptr[t])[ph_ok]
        px = np.searchsorted(xkey, k)
        px_ok = (r < 0) & (px < len(xkey))
        px_ok[px_ok] = xkey[px[px_ok]] == k[px_ok]
        r[px_ok] = (n_items[t] + px - xptr[t])[px_ok]
        if np.any(r < 0):
            raise AssertionError("ext position lookup failed")
        return r
    it = np.broadcast_to(item_tile, sten.shape)
    lst = np.zeros(sten.shape, dtype=np.int64)
    lst[sten_ok] = ext_pos(it[sten_ok], sten[sten_ok])
    KP = sw_tiles_kp(K)
    lst_rows = np.zeros((I, KP), dtype=np.uint16)  # one row per item, padded, counts last
    lst_rows[:, :K] = lst.T
    lst_rows[:, KP - 1] = nn | (ng << 8)
    lgh = np.zeros(gown.shape, dtype=np.int64)
    itg = np.broadcast_to(item_tile, gown.shape)
    lgh[g_ok] = ext_pos(itg[g_ok], gown[g_ok])
    lfn = np.full(fn.shape, -1, dtype=np.int64)
    itf = np.broadcast_to(item_tile, fn.shape)
    lfn[fn_ok] = (ext_pos(itf[fn_ok], fn[fn_ok]) << 3) | (fn_code[fn_ok] & 7)  # local neighbour codes
    if n_ext.max(initial=0) >= 65535:
        raise ValueError("a tile reads more than 65535 cells")
    out.update({'max_ext': int(n_ext.max()) if nt else 0, 'items': I, 'max_items': int(n_items.max()) if nt else 0,
                'ext_ptr': ext_ptr.astype(np.int32), 'ext_cells': ext_cells.astype(np.int32),
                'item_ptr': item_ptr.astype(np.int32), 'tile_next': n_ext.astype(np.int32),
                'lst': lst_rows,
                'lgh': np.ascontiguousarray(lgh, dtype=np.uint16) if G else np.zeros((1, max(I, 1)), np.uint16),
                'lfn': np.ascontiguousarray(lfn, dtype=np.int32)})
    return out


def face_tiles_native(cell_faces, cell_nfaces, face_cells, T):
    """Face tiles of T consecutive cells (mfv_face_tiles_build): halo, faces touching each tile
    and the per-cell-slot face positions. cell_faces [MF][C], cell_nfaces [C], face_cells [F][2]."""
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    cf, nf, fc = i32(cell_faces), i32(cell_nfaces), i32(face_cells)
    MF, C = cf.shape
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    h = _lib.MfvFaceTilesHost()
    _lib.call('mfv_face_tiles_build', int(C), int(MF), P(cf), P(nf), int(fc.shape[0]), P(fc), int(T),
              ctypes.byref(h))
    try:
        nt = h.nb_tiles

        def take(ptr, n, shape=None):
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_int32)), (max(n, 1),))[:n].copy()
            return a.reshape(shape) if shape is not None else a
        out = {'T': int(T), 'nb_tiles': nt, 'max_items': h.max_items, 'max_faces': h.max_faces, 'MF': MF, 'C': C,
               'tile_cell0': take(h.tile_cell0, nt), 'tile_ncells': take(h.tile_ncells, nt),
               'halo_ptr': take(h.halo_ptr, nt + 1), 'halo_cells': take(h.halo_cells, h.halo_len),
               'face_ptr': take(h.face_ptr, nt + 1), 'face_id': take(h.face_id, h.face_len),
               'face_l': take(h.face_l, h.face_len), 'face_r': take(h.face_r, h.face_len),
               'lface': take(h.lface, MF * C, (MF, C))}
    finally:
        _lib.call('mfv_face_tiles_host_free', ctypes.byref(h))
    return out


class FaceTiles(object):
    """Device copy of a face-tiles layout (mfv_face_tiles) for a DeviceMesh, built on the host
    from its cell/face tables."""

    ARRAYS = ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'face_ptr', 'face_id', 'face_l', 'face_r',
              'lface')

    def __init__(self, dmesh, T=256):
        import torch
        H = dmesh.host
        self.host = face_tiles_native(H['cell_faces'], H['cell_nfaces'], H['face_cells'], T)
        self.t = {}
        d = _lib.MfvFaceTiles()
        d.nb_tiles, d.tile_size = self.host['nb_tiles'], T
        d.max_items, d.max_faces = self.host['max_items'], self.host['max_faces']
        for k in self.ARRAYS:
            a = np.ascontiguousarray(self.host[k])
            self.t[k] = torch.from_numpy(a if a.size else np.zeros(1, dtype=np.int32)).to(dmesh.device)
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d

    @property
    def ptr(self):
        return ctypes.byref(self.desc)

    def lds_bytes(self):
        return 48 * self.host['max_items'] + 40 * self.host['max_faces']


def euler_tet_native(cell_faces, face_cells, face_name, face_normal, face_mesure):
    """Tables of the packed Euler step (mfv_euler_tet_build), per cell and in the order of its
    faces by increasing original id: what lies across each face (``cell_nb``: bit 31 boundary,
    bit 30 the cell is the right side, bits 0-29 the other cell or the boundary name), the
    face's geometry (``cell_geo`` [4][2][C][2]: (nx, ny), (nz, mesure)), and the slot of each
    face in cells.faceid (``cell_rank``, 2 bits per face). cell_faces [4][C], face_cells [F][2],
    face_name [F], face_normal [F][3], face_mesure [F]."""
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    cf, fc, fn = i32(cell_faces), i32(face_cells), i32(face_name)
    nrm, ms = f64(face_normal), f64(face_mesure)
    MF, C = cf.shape
    F = fc.shape[0]
    if MF != 4:
        raise ValueError("the packed Euler step needs tetrahedra (4 faces per cell), the mesh has %d slots" % MF)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    h = _lib.MfvEulerTetHost()
    _lib.call('mfv_euler_tet_build', int(C), int(F), P(cf), P(fc), P(fn), P(nrm), P(ms), ctypes.byref(h))
    try:
        def take(ptr, ctype, n, dtype):
            a = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (max(n, 1),))[:n]
            return np.array(a, dtype=dtype)
        out = {'C': C, 'F': F,
               'cell_nb': take(h.cell_nb, ctypes.c_int32, 4 * C, np.int32).reshape(C, 4),
               'cell_rank': take(h.cell_rank, ctypes.c_uint8, C, np.uint8),
               'cell_geo': take(h.cell_geo, ctypes.c_double, 16 * C, np.float64).reshape(4, 2, C, 2)}
    finally:
        _lib.call('mfv_euler_tet_host_free', ctypes.byref(h))
    return out


class EulerTet(object):
    """Device copy of the packed Euler tables (mfv_euler_tet) of a tetrahedral DeviceMesh."""

    ARRAYS = ('cell_nb', 'cell_rank', 'cell_geo')

    def __init__(self, dmesh):
        import torch
        H = dmesh.host
        self.host = euler_tet_native(H['cell_faces'], H['face_cells'], H['face_name'], H['face_normal'],
                                     H['face_mesure'])
        self.t = {}
        d = _lib.MfvEulerTet()
        d.nb_cells, d.nb_faces = self.host['C'], self.host['F']
        for k in self.ARRAYS:
            a = np.ascontiguousarray(self.host[k])
            self.t[k] = torch.from_numpy(a if a.size else np.zeros(4, dtype=a.dtype)).to(dmesh.device)
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d

    @property
    def ptr(self):
        return ctypes.byref(self.desc)


def euler_tet_tiles_host(H, T):
    """Face tiles of the packed Euler step from a host mesh layout ``H`` (gpu_layout): the
    tiles, halos and face entries of face_tiles_native, each entry with its own geometry
    (normal, mesure) and item pair (left item; right item, or -name on a boundary face), and
    per cell the entry codes of its faces in the residual order. Within a tile, the entries
    whose cells are both tile cells come first (the face-phase step reads the records of halo
    cells from HBM, so the waves of the first entries take them all from LDS)."""
    ft = face_tiles_native(H['cell_faces'], H['cell_nfaces'], H['face_cells'], T)
    fid = ft['face_id'].astype(np.int64)
    name = H['face_name'][fid].astype(np.int64)
    if np.any((name == 0) & (ft['face_r'] < 0)) or np.any(name < 0):
        raise ValueError("an inner face (name 0) needs two cells and names must be >= 0")
    geo = np.ascontiguousarray(np.column_stack([H['face_normal'][fid], H['face_mesure'][fid]]), dtype=np.float64)
    lr = np.ascontiguousarray(np.column_stack([ft['face_l'], np.where(name == 0, ft['face_r'], -name)]),
                              dtype=np.int32)
    order = H['order_sorted'].astype(np.int64)              # [4][C]: slot of the k-th face
    code = np.take_along_axis(ft['lface'], order, axis=0).T.astype(np.int64)  # [C][4]
    fp = ft['face_ptr'].astype(np.int64)
    nt = len(fp) - 1
    et = np.repeat(np.arange(nt), np.diff(fp))
    nct = ft['tile_ncells'].astype(np.int64)[et]
    halo = (lr[:, 0] >= nct) | (lr[:, 1] >= nct)
    perm = np.argsort(2 * et + halo, kind='stable')
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    geo, lr = np.ascontiguousarray(geo[perm]), np.ascontiguousarray(lr[perm])
    C = code.shape[0]
    ncs = ft['tile_ncells'].astype(np.int64)
    ct = np.repeat(np.arange(nt), ncs)
    if len(ct) != C or np.any(ft['tile_cell0'] != np.cumsum(ncs) - ncs):
        raise ValueError("face tiles do not cover the cells in order")
    base = fp[ct][:, None]
    code = ((inv[base + (code >> 1)] - base) << 1) | (code & 1)
    code = np.ascontiguousarray(code, dtype=np.int32)
    out = {k: ft[k] for k in ('T', 'nb_tiles', 'max_items', 'max_faces', 'tile_cell0', 'tile_ncells', 'halo_ptr',
                              'halo_cells', 'face_ptr')}
    out.update({'face_geo': geo, 'face_lr': lr, 'cell_code': code, 'face_id': ft['face_id'][perm]})
    return out


class EulerTetTiles(object):
    """Device copy of the packed Euler step's face tiles (mfv_euler_tet_tiles)."""

    ARRAYS = ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'face_ptr', 'face_geo', 'face_lr', 'cell_code')

    def __init__(self, dmesh, T=256):
        import torch
        if dmesh.MF != 4:
            raise ValueError("the packed Euler step needs tetrahedra")
        self.host = euler_tet_tiles_host(dmesh.host, T)
        self.t = {}
        d = _lib.MfvEulerTetTiles()
        d.nb_tiles, d.tile_size = self.host['nb_tiles'], T
        d.max_items, d.max_faces = self.host['max_items'], self.host['max_faces']
        for k in self.ARRAYS:
            a = np.ascontiguousarray(self.host[k])
            self.t[k] = torch.from_numpy(a if a.size else np.zeros(4, dtype=a.dtype)).to(dmesh.device)
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d

    @property
    def ptr(self):
        return ctypes.byref(self.desc)

    def lds_bytes(self, cells=False):
        """LDS of a block: the face-phase step holds the tile cells' records and the face
        fluxes; the single-phase step (cells, mfv_euler3d_step_cells_tiled) the records of the
        tile cells and of their halo."""
        if cells:
            return 48 * self.host['max_items']
        return 48 * self.host['T'] + 40 * self.host['max_faces']


class Euler32(object):
    """FP32-bulk Euler state on face tiles: float32 records [C][8], the tiles' float32 face
    geometry, and the float64 conservation sums of the last step (mfv_euler3d_step32_tiled)."""

    def __init__(self, dmesh, T=256, et=None, tt=None):
        import torch
        self.m = dmesh
        self.et = et if et is not None else EulerTet(dmesh)
        self.tt = tt if tt is not None else EulerTetTiles(dmesh, T)
        n = int(self.tt.host['face_ptr'][-1])
        self.geo32 = torch.empty((max(n, 1), 4), dtype=torch.float32, device=dmesh.device)
        _lib.call('mfv_euler32_tile_geometry', self.tt.ptr, n, self.geo32.data_ptr(), stream_ptr())
        C = int(dmesh.C)
        self.rec = [torch.zeros((C, 8), dtype=torch.float32, device=dmesh.device) for _ in range(2)]
        self.cons = torch.zeros(5, dtype=torch.float64, device=dmesh.device)
        self.k = 0

    def lds_bytes(self):
        return 24 * self.tt.host['max_items'] + 24 * self.tt.host['max_faces']

    def pack(self, U, stream=None):
        """float64 [6][C] state -> the current float32 records."""
        self.k = 0
        _lib.call('mfv_euler32_pack', self.m.ptr, U.data_ptr(), self.rec[0].data_ptr(),
                  stream if stream is not None else stream_ptr())

    def unpack(self, U, stream=None):
        _lib.call('mfv_euler32_unpack', self.m.ptr, self.rec[self.k % 2].data_ptr(), U.data_ptr(),
                  stream if stream is not None else stream_ptr())

    def step(self, gamma, cfl, quirk, ring, step, sums=True, stream=None):
        a, b = self.rec[self.k % 2], self.rec[(self.k + 1) % 2]
        _lib.call('mfv_euler3d_step32_tiled', self.m.ptr, self.et.ptr, self.tt.ptr, self.geo32.data_ptr(),
                  a.data_ptr(), b.data_ptr(), float(gamma), float(cfl), int(quirk), ring.data_ptr(), int(step),
                  self.cons.data_ptr() if sums else None, stream if stream is not None else stream_ptr())
        self.k += 1


class SwTiles(object):
    """Device copy of a tiles layout and its mfv_sw_tiles descriptor (stage=True copies every
    cell state a tile reads into LDS before reconstructing)."""

    ARRAYS = ('tile_cell0', 'tile_ncells', 'tile_nhalo', 'ext_ptr', 'ext_cells', 'item_ptr', 'lnbr', 'lst', 'lgh',
              'lfn')

    def __init__(self, host, device, stage=False):
        import torch
        self.host = host
        self.t = {}
        for k in self.ARRAYS:
            a = np.ascontiguousarray(host[k])
            if a.dtype == np.uint16:  # torch has no uint16: move the bits as int16
                a = a.view(np.int16)
            if a.size == 0:
                a = np.zeros(1, dtype=a.dtype)
            self.t[k] = torch.from_numpy(a).to(device)
        d = _lib.MfvSwTiles()
        d.nb_tiles, d.tile_size, d.max_halo = host['nb_tiles'], host['T'], host['max_halo']
        d.max_ext, d.nb_items, d.stage_states = host['max_ext'], host['items'], int(bool(stage))
        d.max_items = host['max_items']
        for k in self.ARRAYS:
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d

    def lds_bytes(self):
        MF = self.host['MF']
        rec = self.host['max_items'] * MF * 16
        return rec + (self.host['max_ext'] * 48 if self.desc.stage_states else 0)

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
        self.geoB = torch.zeros((C, MF, 4), dtype=f32, device=dev)
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
        self._layouts = {}
        self._lsq = dmesh.host

    def tiles(self, T=256, stage=True, N=None):
        """The fused step's tiles of N (default T) cells for blocks of T threads (built on first use)."""
        N = T if N is None else N
        if (T, N, stage) not in self._tiles:
            if (T, N) not in self._layouts:
                self._layouts[(T, N)] = sw_tiles_native(self.nbr_host, self.desc.nb_cells, T, N, self._lsq)
            self._tiles[(T, N, stage)] = SwTiles(self._layouts[(T, N)], self.state.device, stage)
        return self._tiles[(T, N, stage)]

    def step_tiled(self, dmesh, grav, cfl, order, ring, step, T=256, stage=True, stream=None, N=None):
        """One fused step: reads the current state buffer, writes the other, swaps them."""
        _lib.call('mfv_sw_step_mixed_tiled', dmesh.ptr, self.ptr, self.tiles(T, stage, N).ptr, self.state2.data_ptr(),
                  float(grav), float(cfl), int(order), ring.data_ptr(), int(step),
                  stream if stream is not None else stream_ptr())
        self.state, self.state2 = self.state2, self.state
        self.desc.state = self.state.data_ptr()

    CONFIGS = (('split', 256, 256), ('tiled', 256, 256), ('staged', 128, 128), ('staged', 256, 256),
               ('staged', 512, 512))

    def step(self, dmesh, grav, cfl, order, ring, step, config=('staged', 256, 256), stream=None):
        """One step with ``config``: ('split', _, _) runs the reconstruction and flux kernels,
        ('tiled', T, N) / ('staged', T, N) the fused step on tiles of N cells, T threads."""
        kind, T, N = config
        s = stream if stream is not None else stream_ptr()
        if kind == 'split':
            _lib.call('mfv_sw_recon_mixed', dmesh.ptr, self.ptr, int(order), s)
            _lib.call('mfv_sw_flux_mixed', dmesh.ptr, self.ptr, float(grav), float(cfl), ring.data_ptr(), int(step), s)
        else:
            self.step_tiled(dmesh, grav, cfl, order, ring, step, T=T, stage=(kind == 'staged'), stream=s, N=N)

    def autotune(self, dmesh, grav, cfl, order, ring, steps=10):
        """Times each configuration on ``steps`` steps (state and ring restored after) and
        returns (best, {config: ms per step})."""
        import torch
        saved = (self.state.clone(), ring.clone())
        st = torch.cuda.current_stream()
        res = {}
        for cfg in self.CONFIGS:
            if cfg[0] != 'split':
                try:
                    if self.tiles(cfg[1], cfg[0] == 'staged', cfg[2]).lds_bytes() > 64 * 1024:
                        continue
                except ValueError:
                    continue
            self.step(dmesh, grav, cfl, order, ring, 0, cfg)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for k in range(steps):
                self.step(dmesh, grav, cfl, order, ring, 1 + k, cfg)
            e1.record(st)
            e1.synchronize()
            res[cfg] = e0.elapsed_time(e1) / steps
            self.state.copy_(saved[0])
            ring.copy_(saved[1])
        return min(res, key=res.get), res

    @property
    def ptr(self):
        return ctypes.byref(self.desc)
