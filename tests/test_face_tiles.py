"""Face tiles of the tiled Euler step (mfv_face_tiles_build): every face touching a tile is
listed once, its sides resolve to the right cells through the tile's item list, the halo is
exactly the outside face neighbours, and each cell slot names its face and side."""
import numpy as np
import pytest

from conftest import make_tri_domain


def _kuhn(n, seed=0):
    from manapy_amd.ddm import Domain
    from manapy_amd.meshgen import kuhn_tetrahedra
    el, nd = kuhn_tetrahedra(n, 1.0, 0.08, seed)
    return Domain(3, mesh=(el, nd))


def _check(L, cell_faces, nfaces, face_cells):
    MF, C = cell_faces.shape
    T = L['T']
    assert L['nb_tiles'] == (C + T - 1) // T
    for t in range(L['nb_tiles']):
        c0, nc = int(L['tile_cell0'][t]), int(L['tile_ncells'][t])
        assert c0 == t * T and nc == min(T, C - c0)
        halo = L['halo_cells'][L['halo_ptr'][t]:L['halo_ptr'][t + 1]]
        items = np.concatenate([np.arange(c0, c0 + nc), halo])
        assert len(items) <= L['max_items'] and np.all(np.diff(halo) > 0)
        faces = L['face_id'][L['face_ptr'][t]:L['face_ptr'][t + 1]]
        assert len(faces) <= L['max_faces'] and np.all(np.diff(faces) > 0)
        want_f, want_h = set(), set()
        for c in range(c0, c0 + nc):
            for s in range(nfaces[c]):
                f = int(cell_faces[s, c])
                want_f.add(f)
                o = face_cells[f, 1] if face_cells[f, 0] == c else face_cells[f, 0]
                if o >= 0 and not (c0 <= o < c0 + nc):
                    want_h.add(int(o))
                code = int(L['lface'][s, c])
                assert faces[code >> 1] == f and (code & 1) == int(face_cells[f, 1] == c)
        assert want_f == set(faces.tolist()) and want_h == set(halo.tolist())
        fl = L['face_l'][L['face_ptr'][t]:L['face_ptr'][t + 1]]
        fr = L['face_r'][L['face_ptr'][t]:L['face_ptr'][t + 1]]
        assert np.array_equal(items[fl], face_cells[faces, 0])
        bnd = face_cells[faces, 1] < 0
        assert np.all(fr[bnd] == -1) and np.array_equal(items[fr[~bnd]], face_cells[faces[~bnd], 1])


@pytest.mark.parametrize("T", [128, 256])
def test_face_tiles_of_a_tetrahedral_mesh(T):
    from manapy_amd.device import face_tiles_native, gpu_layout
    d = _kuhn(7, 3)
    H = gpu_layout(d, lsq=False)
    L = face_tiles_native(H['cell_faces'], H['cell_nfaces'], H['face_cells'], T)
    _check(L, H['cell_faces'], H['cell_nfaces'], H['face_cells'])
    assert L['max_faces'] < 4 * T  # faces shared inside a tile are listed once


def test_face_tiles_of_a_triangle_mesh():
    from manapy_amd.device import face_tiles_native, gpu_layout
    d = make_tri_domain(13, 9, 10., 7., 0.1, 4)
    H = gpu_layout(d, lsq=False)
    L = face_tiles_native(H['cell_faces'], H['cell_nfaces'], H['face_cells'], 128)
    _check(L, H['cell_faces'], H['cell_nfaces'], H['face_cells'])


def test_face_tiles_reject_inconsistent_tables():
    from manapy_amd import _lib
    from manapy_amd.device import face_tiles_native, gpu_layout
    d = make_tri_domain(5, 4, 1., 1., 0.1, 1)
    H = gpu_layout(d, lsq=False)
    bad = H['cell_faces'].copy()
    bad[0, 0] = bad[0, 5]  # a face that does not list cell 0
    with pytest.raises(_lib.MfvError):
        face_tiles_native(bad, H['cell_nfaces'], H['face_cells'], 128)
