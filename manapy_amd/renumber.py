"""Cell renumbering along a space-filling curve for the device layout.

The gather kernels read every neighbour of a cell. With the mesh generator's or a mesh
file's numbering, neighbours can sit thousands of cells apart, so the cache lines a wave
pulls in for its neighbours are evicted before the neighbouring waves use them. Ordering
cells along a Hilbert curve of their centres keeps a cell's neighbours within a few waves
of it.

``renumbered(domain, perm)`` is a read-only view of a Domain with the cells in the new
order and every cell index remapped; faces keep their ids. Each cell keeps its own
ordered lists (faces in ``cells.faceid`` order, the least-squares stencil in
``cells.cellnid`` order), so every per-cell sum of the kernels is formed from the same
terms in the same order and the results are bit-identical to the original numbering,
permuted.
"""
import numpy as np

from .base import Struct

__all__ = ['hilbert_order_2d', 'hilbert_order_3d', 'morton_order_3d', 'cell_order', 'renumbered',
           'RenumberedDomain']


def _quantize(coords, bits):
    """Integer grid coordinates of the points, one scale for every axis (a square grid cell,
    so the curve's locality does not depend on the domain's aspect ratio)."""
    lo = [float(np.min(c)) for c in coords]
    span = max(float(np.max(c)) - l for c, l in zip(coords, lo))
    span = span if span > 0 else 1.0
    return [np.floor((np.asarray(c, dtype=np.float64) - l) / span * ((1 << bits) - 1) + 0.5).astype(np.int64)
            for c, l in zip(coords, lo)]


def hilbert_order_2d(x, y, bits=16):
    """Indices that sort points by their position along a Hilbert curve of order ``bits``
    (the iterative quadrant-rotation construction of the curve index)."""
    xi, yi = _quantize([x, y], bits)
    d = np.zeros(len(xi), dtype=np.int64)
    s = np.int64(1) << (bits - 1)
    n = np.int64(1) << bits
    while s > 0:
        rx = ((xi & s) > 0).astype(np.int64)
        ry = ((yi & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate the quadrant so the sub-curve is in standard orientation
        flip = (ry == 0)
        refl = flip & (rx == 1)
        xi = np.where(refl, n - 1 - xi, xi)
        yi = np.where(refl, n - 1 - yi, yi)
        xs = np.where(flip, yi, xi)
        ys = np.where(flip, xi, yi)
        xi, yi = xs, ys
        s >>= 1
    return np.argsort(d, kind='stable')


def _spread3(v):
    v = v.astype(np.uint64) & np.uint64(0x1fffff)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1f00000000ffff)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1f0000ff0000ff)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100f00f00f00f00f)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10c30c30c30c30c3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_order_3d(x, y, z, bits=21):
    """Indices that sort points by their Morton (Z-order) key."""
    xi, yi, zi = _quantize([x, y, z], bits)
    key = _spread3(xi) | (_spread3(yi) << np.uint64(1)) | (_spread3(zi) << np.uint64(2))
    return np.argsort(key, kind='stable')


def hilbert_order_3d(x, y, z, bits=16):
    """Indices that sort points by their position along a 3D Hilbert curve of order ``bits``
    (Skilling's axes-to-transpose construction, "Programming the Hilbert curve", AIP Conf.
    Proc. 707, 2004; the transposed index is then read out bit-plane by bit-plane). Unlike
    the Morton order the curve never jumps, so a run of consecutive cells is a more compact
    region: on Kuhn tetrahedra, tiles of 256 cells have 4 % fewer halo cells."""
    X = [a.astype(np.int64) for a in _quantize([x, y, z], bits)]
    n = 3
    Q = 1 << (bits - 1)
    while Q > 1:  # undo the excess work of the reflected Gray code, top bit plane first
        P = Q - 1
        for i in range(n):
            hit = (X[i] & Q) != 0
            inv = X[0] ^ P
            if i == 0:
                X[0] = np.where(hit, inv, X[0])
            else:
                t = (X[0] ^ X[i]) & P
                X[0] = np.where(hit, inv, X[0] ^ t)
                X[i] = np.where(hit, X[i], X[i] ^ t)
        Q >>= 1
    for i in range(1, n):  # Gray encode
        X[i] = X[i] ^ X[i - 1]
    t = np.zeros_like(X[0])
    Q = 1 << (bits - 1)
    while Q > 1:
        t = np.where((X[n - 1] & Q) != 0, t ^ (Q - 1), t)
        Q >>= 1
    key = np.zeros_like(X[0])
    for b in range(bits - 1, -1, -1):
        for i in range(n):
            key = (key << 1) | (((X[i] ^ t) >> b) & 1)
    return np.argsort(key, kind='stable')


def cell_order(domain):
    """New-to-old cell permutation of a Domain: Hilbert order of the cell centres (2D or 3D)."""
    c = domain.cells.center
    if int(domain.dim) == 2:
        return hilbert_order_2d(c[:, 0], c[:, 1])
    return hilbert_order_3d(c[:, 0], c[:, 1], c[:, 2])


def _remap_padded(a, iperm):
    """Rows with the count in the last column: valid entries mapped through iperm."""
    a = np.asarray(a)
    out = a.copy()
    W = a.shape[1] - 1
    valid = np.arange(W)[None, :] < a[:, -1][:, None]
    body = out[:, :W]
    body[valid] = iperm[body[valid]]
    return out


class _RenumberedCells(Struct):
    """Cell tables of a RenumberedDomain. The least-squares stencil tables (cellnid, ghostnid)
    are remapped on first access: first-order paths never build them (on a 6M-tetrahedra mesh
    the base Domain's vertex-neighbour table alone takes most of a minute)."""

    def __getattr__(self, name):
        if name in ('cellnid', 'ghostnid'):
            base, perm, iperm = self.__dict__['_src']
            bc = base.cells
            if name == 'cellnid':
                v = _remap_padded(bc.cellnid[perm], iperm).astype(bc.cellnid.dtype)
            else:
                v = np.ascontiguousarray(bc.ghostnid[perm])
            setattr(self, name, v)
            return v
        raise AttributeError(name)


class RenumberedDomain(object):
    """Domain view with cells in ``perm`` order (perm[new] = old); see the module doc."""

    def __init__(self, domain, perm):
        perm = np.asarray(perm, dtype=np.int64)
        C = int(domain.nbcells)
        if perm.shape != (C,) or not np.array_equal(np.sort(perm), np.arange(C)):
            raise ValueError("perm must be a permutation of the %d cells" % C)
        iperm = np.empty(C, dtype=np.int64)
        iperm[perm] = np.arange(C)
        self.base = domain
        self.perm = perm
        self.iperm = iperm
        self.dim = domain.dim
        self.nbcells = domain.nbcells
        self.nbfaces = domain.nbfaces
        self.nbnodes = domain.nbnodes
        self.maxcellfid = domain.maxcellfid
        self.innerfaces = domain.innerfaces
        self.boundaryfaces = domain.boundaryfaces
        bc = domain.cells
        cells = _RenumberedCells()
        cells.__dict__['_src'] = (domain, perm, iperm)
        cells.faceid = np.ascontiguousarray(bc.faceid[perm])
        cells.center = np.ascontiguousarray(bc.center[perm])
        cells.volume = np.ascontiguousarray(bc.volume[perm])
        cells.nodeid = np.ascontiguousarray(bc.nodeid[perm])
        cells.cellfid = _remap_padded(bc.cellfid[perm], iperm).astype(bc.cellfid.dtype)
        if hasattr(bc, 'nf'):
            cells.nf = np.ascontiguousarray(bc.nf[perm])
        self.cells = cells
        bf = domain.faces
        faces = Struct()
        cid = np.asarray(bf.cellid)
        faces.cellid = np.where(cid >= 0, iperm[np.maximum(cid, 0)], -1).astype(cid.dtype)
        for k in ('name', 'normal', 'mesure', 'center', 'ghostcenter', 'nodeid'):
            setattr(faces, k, getattr(bf, k))
        self.faces = faces
        self.nodes = domain.nodes

    def to_new(self, a):
        """Cell array in the original numbering -> the new numbering (axis 0 or -1)."""
        a = np.asarray(a)
        return np.ascontiguousarray(a[..., self.perm] if a.shape[-1] == len(self.perm) and a.ndim > 1
                                    else a[self.perm])

    def to_old(self, a):
        """Inverse of to_new."""
        a = np.asarray(a)
        return np.ascontiguousarray(a[..., self.iperm] if a.shape[-1] == len(self.perm) and a.ndim > 1
                                    else a[self.iperm])


def renumbered(domain, perm=None):
    """Renumbered view of ``domain`` (space-filling-curve order unless ``perm`` is given)."""
    return RenumberedDomain(domain, cell_order(domain) if perm is None else perm)
