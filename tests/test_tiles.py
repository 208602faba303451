"""Tiles of the fused mixed step: every block-local neighbour code names the same neighbour
cell, slot and side as the global code, halos hold exactly the outside face neighbours, and
ensemble tiles never straddle instances."""
import numpy as np
import pytest

from conftest import make_tri_domain


def _check(L, nbr, C):
    MF = nbr.shape[0]
    c0, nc, hp, hc, ln = (L[k] for k in ('tile_cell0', 'tile_ncells', 'halo_ptr', 'halo_cells', 'lnbr'))
    covered = np.zeros(C, dtype=bool)
    for t in range(L['nb_tiles']):
        a, n = int(c0[t]), int(nc[t])
        assert 1 <= n <= L['T']
        covered[a:a + n] = True
        halo = hc[hp[t]:hp[t + 1]]
        assert len(halo) <= L['max_halo']
        assert np.all(np.diff(halo) > 0) and not np.any((halo >= a) & (halo < a + n))
        want = set()
        for i in range(a, a + n):
            for s in range(MF):
                g, l = int(nbr[s, i]), int(ln[s, i])
                if g < 0:
                    assert l == -1
                    continue
                j = g >> 3
                if not (a <= j < a + n):
                    want.add(j)
                loc = l >> 3
                got = a + loc if loc < n else int(halo[loc - n])
                assert got == j and (l & 7) == (g & 7)
        assert want == set(halo.tolist())
    assert covered.all()


@pytest.mark.parametrize("T", [256, 512])
def test_tiles_of_a_renumbered_mesh(T):
    from manapy_amd.renumber import renumbered
    from manapy_amd.device import sw_packed_layout, sw_tiles_layout
    d = renumbered(make_tri_domain(37, 23, 100., 60., 0.1, 2))
    P = sw_packed_layout(d)
    C = int(d.nbcells)
    L = sw_tiles_layout(P['nbr'], C, T)
    _check(L, P['nbr'], C)
    # the Hilbert order keeps halos well below the tile size
    assert np.diff(L['halo_ptr']).mean() < 0.4 * T


def test_ensemble_tiles_stay_inside_instances():
    from manapy_amd.ensemble import ensemble_layout, ensemble_tiles
    doms = [make_tri_domain(n, m, 10., 10., 0.1, s) for n, m, s in ((9, 7, 1), (23, 11, 2), (5, 5, 3))]
    E = ensemble_layout(doms)
    L = ensemble_tiles(E, 256)
    _check(L, E['nbr'], E['C'])
    inst_end = E['inst_cell0'] + ((E['inst_ncells'] + 63) // 64) * 64
    for a, n in zip(L['tile_cell0'], L['tile_ncells']):
        k = np.searchsorted(E['inst_cell0'], a, side='right') - 1
        assert a >= E['inst_cell0'][k] and a + n <= inst_end[k] and a % 64 == 0 and n % 64 == 0
