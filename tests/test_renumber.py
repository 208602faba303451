"""Cell renumbering along a space-filling curve: the renumbered Domain view is a pure
relabelling, so the CPU restatement run on it gives the original results permuted, bit
for bit, and neighbours end up close in index."""
import numpy as np

from conftest import make_tri_domain, make_tet_domain


def test_hilbert_order_is_a_permutation_with_local_neighbours():
    from manapy_amd.renumber import renumbered
    d = make_tri_domain(40, 30, 100., 75., 0.1, 2)
    r = renumbered(d)
    C = int(d.nbcells)
    assert np.array_equal(np.sort(r.perm), np.arange(C))
    assert np.array_equal(r.iperm[r.perm], np.arange(C))

    def far(dom):
        cid = dom.faces.cellid
        inner = cid[:, 1] >= 0
        return np.mean(np.abs(cid[inner, 0] - cid[inner, 1]) > 64)
    # row-major numbering puts the row above 2 nx = 80 cells away (a third of the inner
    # faces); along the curve most neighbours are within one 64-cell wave
    assert far(d) > 0.3 and far(r) < 0.1
    # cell arrays follow the permutation, face ids are unchanged
    assert np.array_equal(r.cells.center, d.cells.center[r.perm])
    assert np.array_equal(r.cells.faceid, d.cells.faceid[r.perm])
    cid = d.faces.cellid
    assert np.array_equal(r.faces.cellid[:, 0], r.iperm[cid[:, 0]])
    b = cid[:, 1] >= 0
    assert np.array_equal(r.faces.cellid[b, 1], r.iperm[cid[b, 1]])
    assert np.all(r.faces.cellid[~b, 1] == -1)


def test_hilbert_3d_is_a_continuous_curve_and_keeps_tiles_compact():
    """On a 2^k grid the 3D Hilbert order visits every point once and moves by one grid step
    at a time (Morton jumps); on Kuhn tetrahedra its 256-cell runs touch fewer outside cells
    than Morton's."""
    from manapy_amd.renumber import hilbert_order_3d, morton_order_3d, renumbered
    for k in (1, 2, 3, 4):
        m = 1 << k
        g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing='ij'), -1).reshape(-1, 3)
        o = hilbert_order_3d(g[:, 0].astype(float), g[:, 1].astype(float), g[:, 2].astype(float), bits=k)
        assert np.array_equal(np.sort(o), np.arange(m ** 3))
        assert np.all(np.abs(np.diff(g[o], axis=0)).sum(axis=1) == 1)
        om = morton_order_3d(g[:, 0].astype(float), g[:, 1].astype(float), g[:, 2].astype(float), bits=k)
        if k > 1:
            assert np.any(np.abs(np.diff(g[om], axis=0)).sum(axis=1) > 1)
    d = make_tet_domain(12, 0.08, 4)
    c = d.cells.center

    def outside(perm, T=256):
        r = renumbered(d, perm)
        cid = r.faces.cellid
        inner = cid[:, 1] >= 0
        a, b = cid[inner, 0], cid[inner, 1]
        return np.mean((a // T) != (b // T))
    assert outside(hilbert_order_3d(c[:, 0], c[:, 1], c[:, 2])) < outside(morton_order_3d(c[:, 0], c[:, 1], c[:, 2]))


def test_renumbered_sw_steps_equal_permuted_original(oracle_lib):
    from manapy_amd.renumber import renumbered
    d = make_tri_domain(18, 15, 100., 100., 0.1, 4)
    r = renumbered(d)
    C = int(d.nbcells)
    x = d.cells.center[:, 0]
    Z = 0.2 * np.exp(-((x - 40.) ** 2) / 200.)
    h0 = np.where(x < 50., 2., 1.)
    hh, uu, vv = h0.copy(), np.zeros(C), np.zeros(C)
    zg = oracle_lib.z_gradient(d, Z)
    t = [oracle_lib.sw_step(d, hh, uu, vv, Z, zg, 9.81, 0.8, 2) for _ in range(8)]
    Zr = r.to_new(Z)
    h2, u2, v2 = r.to_new(h0), np.zeros(C), np.zeros(C)
    zg2 = oracle_lib.z_gradient(r, Zr)
    t2 = [oracle_lib.sw_step(r, h2, u2, v2, Zr, zg2, 9.81, 0.8, 2) for _ in range(8)]
    assert t == t2
    for a, b in ((hh, h2), (uu, u2), (vv, v2)):
        assert np.array_equal(r.to_new(a), b)


def test_renumbered_srnh_and_euler_equal_permuted_original(oracle_lib):
    from manapy_amd.renumber import renumbered
    d = make_tri_domain(12, 10, 1., 1., 0.1, 6)
    r = renumbered(d)
    C = int(d.nbcells)
    x = d.cells.center[:, 0]
    S = np.zeros((5, C))
    S[0] = np.where(x < .5, 2., 1.)
    S[3] = 0.1 * S[0]
    S[4] = 0.05 * np.sin(4 * x)
    S2 = r.to_new(S)
    for _ in range(4):
        oracle_lib.srnh_step(d, S, 9.81, 0.8, 2, quirk=False)
        oracle_lib.srnh_step(r, S2, 9.81, 0.8, 2, quirk=False)
    assert np.array_equal(r.to_new(S), S2)
    t = make_tet_domain(3, 0.08, 1)
    rt = renumbered(t)
    assert int(rt.dim) == 3
    assert np.array_equal(np.sort(rt.perm), np.arange(int(t.nbcells)))
