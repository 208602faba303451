"""Tables of the packed Euler step (mfv_euler_tet_build): faces renumbered by first use along
the cell order, each cell's faces in slot order under the new ids, the residual order equal
to the ascending-original-id order of the device layout, and face records carrying the
reference's normal, mesure, cells and boundary name."""
import ctypes

import numpy as np
import pytest

from conftest import make_tet_domain


def _tables(d):
    from manapy_amd.device import euler_tet_native, gpu_layout
    L = gpu_layout(d, lsq=False)
    return L, euler_tet_native(L['cell_faces'], L['face_cells'], L['face_name'], L['face_normal'],
                               L['face_mesure'])


@pytest.mark.parametrize("n,renumber", [(4, False), (5, True)])
def test_first_use_numbering_and_records(n, renumber):
    d = make_tet_domain(n, 0.08, n)
    if renumber:
        from manapy_amd.renumber import renumbered
        d = renumbered(d)
    L, T = _tables(d)
    C, F = int(d.nbcells), int(d.nbfaces)
    old = L['cell_faces'].T                     # [C][4] original ids, slot order
    new = T['cell_faces']
    # new[i][s] = newid[old[i][s]] for one permutation newid of the faces
    newid = np.full(F, -1)
    newid[old.ravel()] = new.ravel()
    assert np.array_equal(np.sort(newid), np.arange(F))
    assert np.array_equal(newid[old], new)
    # first use: walking the cells in order, each face not seen before takes the next id
    seen, nxt = set(), 0
    for f in new.ravel():
        if f not in seen:
            assert f == nxt
            seen.add(f)
            nxt += 1
    # residual order = the device layout's ascending-original-id order
    ranks = np.stack([(T['cell_rank'] >> (2 * k)) & 3 for k in range(4)], axis=1)
    assert np.array_equal(ranks, L['order_sorted'].T.astype(np.int64))
    assert np.all(np.diff(np.take_along_axis(old, ranks, axis=1), axis=1) > 0)
    # face records in the new order
    perm = np.argsort(newid)                    # new -> old
    assert np.array_equal(T['face_geo'][:, :3], L['face_normal'][perm])
    assert np.array_equal(T['face_geo'][:, 3], L['face_mesure'][perm])
    # what lies across each slot: the other cell (inner) or the name (boundary), and the side
    fc, nm = L['face_cells'][old], L['face_name'][old]        # [C][4](2), [C][4]
    me = np.arange(C)[:, None]
    right = fc[..., 1] == me
    assert np.all(right | (fc[..., 0] == me))
    v = T['cell_nb'].astype(np.int64) & 0xffffffff
    assert np.array_equal((v >> 31) == 1, nm != 0)
    assert np.array_equal(((v >> 30) & 1) == 1, right)
    want = np.where(nm == 0, np.where(right, fc[..., 0], fc[..., 1]), nm)
    assert np.array_equal(v & ((1 << 30) - 1), want)
    assert np.all(nm[fc[..., 1] < 0] >= 1)


def test_builder_rejects_bad_meshes():
    from manapy_amd import _lib
    from manapy_amd.device import euler_tet_native, gpu_layout
    d = make_tet_domain(3, 0.08, 1)
    L = gpu_layout(d, lsq=False)
    args = lambda **kw: [kw.get(k, L[k]) for k in ('cell_faces', 'face_cells', 'face_name', 'face_normal',
                                                   'face_mesure')]
    cf = L['cell_faces'].copy()
    cf[2, 7] = -1                               # a cell with three faces
    with pytest.raises(_lib.MfvError, match="four faces"):
        euler_tet_native(*args(cell_faces=cf))
    fn = L['face_name'].copy()
    b = int(np.nonzero(L['face_cells'][:, 1] < 0)[0][0])
    fn[b] = 0                                   # a boundary face without a name
    with pytest.raises(_lib.MfvError, match="inner face"):
        euler_tet_native(*args(face_name=fn))
    cf = L['cell_faces'].copy()
    cf[1, 4] = cf[0, 4]                         # a face twice in one cell
    with pytest.raises(_lib.MfvError, match="twice"):
        euler_tet_native(*args(cell_faces=cf))
    cf = L['cell_faces'].copy()
    far = int(np.nonzero((L['face_cells'][:, 0] != 4) & (L['face_cells'][:, 1] != 4))[0][0])
    cf[3, 4] = far                              # a face the cell is no side of
    with pytest.raises(_lib.MfvError, match="no side"):
        euler_tet_native(*args(cell_faces=cf))
    with pytest.raises(ValueError, match="tetrahedra"):
        euler_tet_native(np.zeros((3, 4), np.int32), *args()[1:])
