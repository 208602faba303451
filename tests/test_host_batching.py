"""Host-side logic of the batched paths (no GPU): the slice-range allocator of the sparse
solve arena and the concatenated layout of shallow-water ensembles."""
import numpy as np

from conftest import make_tri_domain


def test_range_allocator_first_fit_and_coalescing():
    from manapy_amd.krylov import RangeAllocator
    a = RangeAllocator(100)
    x = a.alloc(30)
    y = a.alloc(30)
    z = a.alloc(30)
    assert (x, y, z) == (0, 30, 60)
    assert a.alloc(20) is None and a.alloc(10) == 90
    a.release(30, 30)
    assert a.alloc(40) is None
    a.release(0, 30)
    assert a.largest() == 60 and a.alloc(50) == 0
    a.release(0, 50)
    a.release(60, 30)
    a.release(90, 10)
    assert a.free == [(0, 100)] and a.available() == 100


def test_ensemble_layout_is_block_diagonal_copy_of_single_layouts():
    from manapy_amd.device import gpu_layout, sw_packed_layout
    from manapy_amd.ensemble import ensemble_layout
    doms = [make_tri_domain(7, 5, 1., 1., 0.1, 1), make_tri_domain(11, 9, 2., 1., 0.1, 2),
            make_tri_domain(4, 4, 1., 1., 0.1, 3)]
    E = ensemble_layout(doms)
    assert E['C'] % 64 == 0 and len(E['wave_inst']) == E['C'] // 64
    for i, d in enumerate(doms):
        L, P = gpu_layout(d, lsq=True), sw_packed_layout(d)
        c0, n = int(E['inst_cell0'][i]), int(E['inst_ncells'][i])
        f0, nf = int(E['inst_faces0'][i]), int(E['inst_nfaces'][i])
        assert c0 % 64 == 0 and n == L['C'] and nf == L['F']
        assert np.all(E['wave_inst'][c0 // 64:(c0 + n + 63) // 64] == i)
        fc = E['face_cells'][f0:f0 + nf]
        assert np.array_equal(np.where(fc >= 0, fc - c0, -1), L['face_cells'])
        cf = E['cell_faces'][:L['MF'], c0:c0 + n]
        assert np.array_equal(np.where(cf >= 0, cf - f0, -1), L['cell_faces'])
        K = L['lsq_cells'].shape[0]
        nn = L['lsq_ncells']
        lc = E['lsq_cells'][:K, c0:c0 + n] - c0
        mask = np.arange(K)[:, None] < nn[None, :]
        assert np.array_equal(lc[mask], L['lsq_cells'][mask])
        nb = E['nbr'][:, c0:c0 + n]
        assert np.array_equal(np.where(nb >= 0, nb - (c0 << 3), -1), P['nbr'])
        assert np.array_equal(E['cell_volume'][c0:c0 + n], L['cell_volume'])
        pad = slice(c0 + n, c0 + (n + 63) // 64 * 64)
        assert np.all(E['cell_nfaces'][pad] == 0) and np.all(E['cell_volume'][pad] == 1.)
