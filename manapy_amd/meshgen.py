"""Seeded synthetic meshes in manapy's partition layout.

A mesh is returned as ``(elements, nodes)``:

* ``elements`` -- int32 ``[C, nv+1]``: vertex ids, last column = vertex count
  (the ``elements`` dataset written by ``partitions/meshpartitioning.py:255``);
* ``nodes`` -- float64 ``[N, 4]``: x, y, z and the boundary tag
  (``partitions/meshpartitioning.py:256``).

Boundary tags follow ``partitions/partitions_utils.py:16-111``: on the unit
box, x = xmin -> 1 ("in"), x = xmax -> 2 ("out"), y = ymax -> 3 ("upper"),
y = ymin -> 4 ("bottom"), and in 3D z = zmax -> 5 ("front"), z = zmin -> 6
("back") (the physical surfaces of ``mesh/cube.geo``). Tags 1/2 win over 3/4,
which win over 5/6, so corner and edge nodes take the x tag first.

Interior nodes are jittered uniformly by ``+/- jitter * h`` per axis from
``numpy.random.default_rng(seed)``; boundary nodes keep their place so the
domain stays a box.
"""
import numpy as np


def _tags_from_coords(x, y, z, box, dim):
    tag = np.zeros(len(x), dtype=np.float64)
    (x0, x1), (y0, y1) = box[0], box[1]
    if dim == 3:
        z0, z1 = box[2]
        tag[np.isclose(z, z0)] = 6
        tag[np.isclose(z, z1)] = 5
    tag[np.isclose(y, y0)] = 4
    tag[np.isclose(y, y1)] = 3
    tag[np.isclose(x, x1)] = 2
    tag[np.isclose(x, x0)] = 1
    return tag


def jittered_triangles(nx, ny, lx=1.0, ly=1.0, jitter=0.1, seed=0):
    """Rectangle ``[0, lx] x [0, ly]`` split into ``nx * ny`` squares, each cut
    into two triangles along the (i, j) -> (i+1, j+1) diagonal: ``2 nx ny``
    counter-clockwise triangles, cells numbered square by square, row-major."""
    hx, hy = lx / nx, ly / ny
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing='xy')
    x = (ii * hx).ravel().astype(np.float64)
    y = (jj * hy).ravel().astype(np.float64)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=(len(x), 2))
    interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)).ravel()
    x[interior] += jit[interior, 0] * hx
    y[interior] += jit[interior, 1] * hy
    # exact boundary coordinates
    x[(ii == nx).ravel()] = lx
    y[(jj == ny).ravel()] = ly
    z = np.zeros_like(x)
    tag = _tags_from_coords(x, y, z, [(0.0, lx), (0.0, ly)], 2)
    nodes = np.stack([x, y, z, tag], axis=1)

    si, sj = np.meshgrid(np.arange(nx), np.arange(ny), indexing='xy')
    si = si.ravel()
    sj = sj.ravel()
    a = sj * (nx + 1) + si
    b = a + 1
    c = a + (nx + 1) + 1
    d = a + (nx + 1)
    elements = np.empty((2 * nx * ny, 4), dtype=np.int32)
    elements[0::2, 0] = a
    elements[0::2, 1] = b
    elements[0::2, 2] = c
    elements[1::2, 0] = a
    elements[1::2, 1] = c
    elements[1::2, 2] = d
    elements[:, 3] = 3
    return elements, nodes


# Kuhn (Freudenthal) split of a cube along its v000 -> v111 diagonal: one tet
# per axis permutation. Every cube uses the same diagonal direction, so the
# split is conforming across cube faces.
_KUHN_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def kuhn_tetrahedra(n, lx=1.0, jitter=0.08, seed=0):
    """Unit cube (edge ``lx``) as ``n^3`` cubes, each split into 6 tets."""
    h = lx / n
    ii, jj, kk = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing='ij')
    # node id = (k * (n+1) + j) * (n+1) + i  (x fastest)
    ii = ii.transpose(2, 1, 0).ravel()
    jj = jj.transpose(2, 1, 0).ravel()
    kk = kk.transpose(2, 1, 0).ravel()
    x = (ii * h).astype(np.float64)
    y = (jj * h).astype(np.float64)
    z = (kk * h).astype(np.float64)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=(len(x), 3))
    interior = (ii > 0) & (ii < n) & (jj > 0) & (jj < n) & (kk > 0) & (kk < n)
    x[interior] += jit[interior, 0] * h
    y[interior] += jit[interior, 1] * h
    z[interior] += jit[interior, 2] * h
    x[ii == n] = lx
    y[jj == n] = lx
    z[kk == n] = lx
    tag = _tags_from_coords(x, y, z, [(0.0, lx)] * 3, 3)
    nodes = np.stack([x, y, z, tag], axis=1)

    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
    ci = ci.transpose(2, 1, 0).ravel()
    cj = cj.transpose(2, 1, 0).ravel()
    ck = ck.transpose(2, 1, 0).ravel()
    stride = np.array([1, n + 1, (n + 1) * (n + 1)], dtype=np.int64)
    base = ci * stride[0] + cj * stride[1] + ck * stride[2]
    elements = np.empty((6 * len(base), 5), dtype=np.int32)
    for t, perm in enumerate(_KUHN_PERMS):
        v = base.copy()
        elements[t::6, 0] = v
        v = v + stride[perm[0]]
        elements[t::6, 1] = v
        v = v + stride[perm[1]]
        elements[t::6, 2] = v
        v = v + stride[perm[2]]
        elements[t::6, 3] = v
    elements[:, 4] = 4
    return elements, nodes


def gaussian_bumps(center, nbumps, rng, box, amp=(0.0, 0.3), width=(2.0, 8.0)):
    """Sum of ``nbumps`` Gaussians ``A exp(-|x - c|^2 / w^2)`` evaluated at the
    rows of ``center`` (``[C, >=2]``); A, w and c are drawn from ``rng``."""
    (x0, x1), (y0, y1) = box
    a = rng.uniform(amp[0], amp[1], size=nbumps)
    w = rng.uniform(width[0], width[1], size=nbumps)
    cx = rng.uniform(x0, x1, size=nbumps)
    cy = rng.uniform(y0, y1, size=nbumps)
    z = np.zeros(len(center), dtype=np.float64)
    for k in range(nbumps):
        r2 = (center[:, 0] - cx[k]) ** 2 + (center[:, 1] - cy[k]) ** 2
        z += a[k] * np.exp(-r2 / (w[k] * w[k]))
    return z
