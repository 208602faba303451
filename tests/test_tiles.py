"""Tiles of the fused mixed step: every block-local neighbour code names the same neighbour
cell, slot and side as the global code, halos hold exactly the outside face neighbours, and
ensemble tiles never straddle instances."""
import numpy as np
import pytest

from conftest import make_tri_domain


def _ext(L, t):
    """Tile t's ext list as global cell ids: its cells, its halo, then the extra cells."""
    a, n = int(L['tile_cell0'][t]), int(L['tile_ncells'][t])
    return np.concatenate([np.arange(a, a + n), L['ext_cells'][L['ext_ptr'][t]:L['ext_ptr'][t + 1]]])


def _check(L, nbr, lsq, C):
    MF = nbr.shape[0]
    c0, nc, nh, ln = (L[k] for k in ('tile_cell0', 'tile_ncells', 'tile_nhalo', 'lnbr'))
    covered = np.zeros(C, dtype=bool)
    K, G = lsq['lsq_cells'].shape[0], lsq['lsq_ghosts'].shape[0]
    for t in range(L['nb_tiles']):
        a, n, h = int(c0[t]), int(nc[t]), int(nh[t])
        assert 1 <= n <= L['T'] and h <= L['max_halo']
        covered[a:a + n] = True
        ext = _ext(L, t)
        assert len(ext) == L['tile_next'][t] <= L['max_ext'] and len(set(ext.tolist())) == len(ext)
        halo = ext[n:n + h]
        assert np.all(np.diff(halo) > 0) and not np.any((halo >= a) & (halo < a + n))
        want = set()
        for i in range(a, a + n):
            for s in range(MF):
                g, l = int(nbr[s, i]), int(ln[s, i])
                if g < 0:
                    assert l == -1
                    continue
                if not (a <= (g >> 3) < a + n):
                    want.add(g >> 3)
                assert ext[l >> 3] == g >> 3 and (l & 7) == (g & 7) and (l >> 3) < n + h
        assert want == set(halo.tolist())
        # every item's reconstruction inputs resolve through the ext list
        for q in range(n + h):
            r = int(L['item_ptr'][t]) + q
            c = int(ext[q])
            for jj in range(int(lsq['lsq_ncells'][c])):
                assert ext[L['lst'][r, jj]] == lsq['lsq_cells'][jj, c]
            # the row's last entry holds the item's stencil and ghost counts
            assert L['lst'][r, -1] == int(lsq['lsq_ncells'][c]) | int(lsq['lsq_nghosts'][c]) << 8
            for jj in range(int(lsq['lsq_nghosts'][c])):
                f = lsq['lsq_ghosts'][jj, c]
                assert ext[L['lgh'][jj, r]] == lsq['face_cells'][f, 0]
            for s in range(MF):
                g = int(nbr[s, c])
                lf = int(L['lfn'][s, r])
                assert (lf == -1) if g < 0 else (ext[lf >> 3] == g >> 3 and lf & 7 == g & 7)
        assert int(L['item_ptr'][t + 1]) - int(L['item_ptr'][t]) == n + h
    assert covered.all()


@pytest.mark.parametrize("T,N", [(256, 256), (256, 192), (128, 64), (512, 512)])
def test_native_builder_fits_the_block_and_matches_numpy(T, N):
    """The library's host builder grows tiles under the block's item budget; for the tiles it
    chose, every table equals the numpy construction's."""
    from manapy_amd.renumber import renumbered
    from manapy_amd.device import sw_packed_layout, sw_tiles_layout, sw_tiles_native, gpu_layout
    d = renumbered(make_tri_domain(41, 29, 100., 70., 0.1, 5))
    P = sw_packed_layout(d)
    C = int(d.nbcells)
    lsq = gpu_layout(d)
    L = sw_tiles_native(P['nbr'], C, T, N, lsq)
    _check(L, P['nbr'], lsq, C)
    items = np.diff(L['item_ptr'])
    assert np.all((items <= T) | (L['tile_ncells'] == 64))
    assert np.all(L['tile_ncells'] <= N) and np.all(L['tile_cell0'] % 64 == 0)
    R = sw_tiles_layout(P['nbr'], C, T, L['tile_cell0'], L['tile_ncells'], lsq=lsq)
    for k in ('tile_nhalo', 'ext_ptr', 'ext_cells', 'item_ptr', 'lnbr', 'lst', 'lgh', 'lfn'):
        assert np.array_equal(np.asarray(L[k]), np.asarray(R[k])), k
    assert (L['max_ext'], L['max_items'], L['items']) == (R['max_ext'], R['max_items'], R['items'])


@pytest.mark.parametrize("T", [256, 512])
def test_tiles_of_a_renumbered_mesh(T):
    from manapy_amd.renumber import renumbered
    from manapy_amd.device import sw_packed_layout, sw_tiles_layout
    d = renumbered(make_tri_domain(37, 23, 100., 60., 0.1, 2))
    from manapy_amd.device import gpu_layout
    P = sw_packed_layout(d)
    C = int(d.nbcells)
    lsq = gpu_layout(d)
    L = sw_tiles_layout(P['nbr'], C, T, lsq=lsq)
    _check(L, P['nbr'], lsq, C)
    # the Hilbert order keeps halos well below the tile size
    assert L['tile_nhalo'].mean() < 0.4 * T


def test_ensemble_tiles_stay_inside_instances():
    from manapy_amd.ensemble import ensemble_layout, ensemble_tiles
    doms = [make_tri_domain(n, m, 10., 10., 0.1, s) for n, m, s in ((9, 7, 1), (23, 11, 2), (5, 5, 3))]
    E = ensemble_layout(doms)
    L = ensemble_tiles(E, 256)
    _check(L, E['nbr'], E, E['C'])
    inst_end = E['inst_cell0'] + ((E['inst_ncells'] + 63) // 64) * 64
    for a, n in zip(L['tile_cell0'], L['tile_ncells']):
        k = np.searchsorted(E['inst_cell0'], a, side='right') - 1
        assert a >= E['inst_cell0'][k] and a + n <= inst_end[k] and a % 64 == 0 and n % 64 == 0
