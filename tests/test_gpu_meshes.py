"""GPU parity on the reference's own Gmsh meshes (mesh/carre.msh: triangles; carre_hybrid.msh:
triangles and quadrangles around two holes; rectangle.msh; rectangle_RTM.msh; carre_structure.msh:
triangles and quadrangles; cube.msh: tetrahedra; cube_bis.msh: tetrahedra, pyramids and
hexahedra), read by the
Gmsh 4.1 reader: advection, the float64 shallow-water and Euler steps, bit for bit against the
oracle; the mixed-precision and FP32-bulk steps bit for bit against their float32 models."""
import os

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

HERE = os.path.dirname(os.path.abspath(__file__))
DOMS = {}


def domain(name, dim):
    from manapy_amd.gmsh import domain_from_gmsh
    if name not in DOMS:
        DOMS[name] = domain_from_gmsh(os.path.join(HERE, 'golden', 'meshes', name + '.msh.gz'), dim)
    return DOMS[name]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


def dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def call(name, *args):
    from manapy_amd import _lib
    from manapy_amd.device import stream_ptr
    _lib.call(name, *args, stream_ptr())


def P(t):
    return t.data_ptr()


@pytest.mark.parametrize("name,dim,order", [("carre", 2, 1), ("carre", 2, 2), ("carre_hybrid", 2, 2),
                                            ("rectangle", 2, 2), ("cube", 3, 1), ("cube", 3, 2),
                                            ("carre_structure", 2, 2), ("rectangle_RTM", 2, 2),
                                            ("cube_bis", 3, 1), ("cube_bis", 3, 2)])
def test_advection_on_reference_meshes(torch, name, dim, order, oracle_lib):
    from manapy_amd.device import DeviceMesh
    d = domain(name, dim)
    m = DeviceMesh(d)
    C, F = int(d.nbcells), int(d.nbfaces)
    c = d.cells.center
    w0 = np.exp(-((c[:, 0] - .3) ** 2 + (c[:, 1] - .25) ** 2) / .01)
    uf, vf = np.full(F, 1.0), np.full(F, 0.5)
    wf = np.full(F, -0.3) if dim == 3 else np.zeros(F)
    uc, vc, wc = np.zeros(C), np.zeros(C), np.zeros(C)
    oracle_lib.facetocell(d, uf, uc)
    oracle_lib.facetocell(d, vf, vc)
    oracle_lib.facetocell(d, wf, wc)
    w_ref = w0.copy()
    dts = [oracle_lib.advection_step(d, w_ref, uc, vc, wc, uf, vf, wf, 0.8, order) for _ in range(50)]
    a, b = dev(torch, w0), dev(torch, np.zeros(C))
    scratch = dev(torch, np.zeros((dim + 1) * C))
    dt0, ring = dev(torch, np.zeros(1)), dev(torch, np.zeros(4))
    tv = [dev(torch, x) for x in (uc, vc, wc, uf, vf, wf)]
    call('mfv_time_step', m.ptr, P(tv[0]), P(tv[1]), P(tv[2]), 0.8, P(dt0), None)
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(50):
        call('mfv_advection_step', m.ptr, P(a), P(b), *[P(x) for x in tv], 0.8, order, P(ring), s, P(scratch))
        a, b = b, a
    assert host(ring)[49 % 3] == dts[-1]
    assert np.array_equal(host(a), w_ref)


def _sw_case(d):
    from manapy_amd.meshgen import gaussian_bumps
    c = d.cells.center
    lo, hi = c[:, :2].min(axis=0), c[:, :2].max(axis=0)
    scale = (hi[0] - lo[0]) / 50.  # the bench's dam break (2 m / 1 m over 100 m), shrunk to the mesh
    Z = gaussian_bumps(c, 4, np.random.default_rng(7), [(lo[0], hi[0]), (lo[1], hi[1])]) * scale / 2.
    h = np.where(c[:, 0] < lo[0] + 0.4 * (hi[0] - lo[0]), 2.0, 1.0) * scale
    return Z, h


@pytest.mark.parametrize("name", ["carre", "carre_hybrid", "rectangle", "carre_structure", "rectangle_RTM"])
def test_sw_float64_and_mixed_on_reference_meshes(torch, name, oracle_lib):
    """The bench's float64 path (packed records, fused tiled step) and the two-kernel packed
    step against 40 oracle steps of the dam break over bumps (walls, also around the hybrid
    mesh's holes), bit for bit; the mixed-precision step against its float32 model."""
    from manapy_amd.device import DeviceMesh, SwPacked, SwMixed
    from manapy_amd.renumber import renumbered
    d = renumbered(domain(name, 2))
    C = int(d.nbcells)
    Z, h0 = _sw_case(d)
    G = 9.81
    zg_ref = oracle_lib.z_gradient(d, Z)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    model = oracle_lib.MixedModel(d, h0, np.zeros(C), np.zeros(C), Z, zg_ref)
    dt_m = oracle_lib.time_step_sw(d, h0, np.zeros(C), np.zeros(C), 0.8, G)
    for _ in range(40):
        dt_last = oracle_lib.sw_step(d, h, hu, hv, Z, zg_ref, G, 0.8, 2)
        dt_m = model.step(G, 0.8, dt_m)
    m = DeviceMesh(d)
    th, tz, tZ, zg = dev(torch, h0), dev(torch, np.zeros(C)), dev(torch, Z), dev(torch, zg_ref)
    dt0 = dev(torch, np.zeros(1))
    call('mfv_time_step_sw', m.ptr, P(th), P(tz), P(tz), 0.8, G, 0., 0., P(dt0), None)
    o = [dev(torch, np.zeros(C)) for _ in range(3)]
    for fused in (True, False):
        pk = SwPacked(m)
        pk.pack(m, th, tz, tz, tZ, zg, 2)
        ring = dev(torch, np.zeros(4))
        call('mfv_dt_ring_init', P(ring), P(dt0))
        for s in range(40):
            if fused:
                pk.step_tiled(m, G, 0.8, 2, ring, s, T=128)
            else:
                call('mfv_sw_recon', m.ptr, pk.ptr, 2)
                call('mfv_sw_flux_packed', m.ptr, pk.ptr, G, 0.8, P(ring), s)
        call('mfv_sw_unpack', m.ptr, pk.ptr, P(o[0]), P(o[1]), P(o[2]))
        for a, b in zip(o, (h, hu, hv)):
            assert np.array_equal(host(a), b), fused
        assert host(ring)[39 % 3] == dt_last
    p = SwMixed(m)
    call('mfv_sw_pack_mixed', m.ptr, p.ptr, P(th), P(tz), P(tz), P(tZ), P(zg), 2)
    ring = dev(torch, np.zeros(4))
    call('mfv_dt_ring_init', P(ring), P(dt0))
    for s in range(40):
        call('mfv_sw_recon_mixed', m.ptr, p.ptr, 2)
        call('mfv_sw_flux_mixed', m.ptr, p.ptr, G, 0.8, P(ring), s)
    assert np.array_equal(host(p.state), model.state)
    assert np.abs(hu).max() > 0.  # the dam has moved


@pytest.mark.parametrize("quirk", [0, 1])
def test_euler_on_the_reference_cube(torch, quirk, oracle_lib):
    """cube.msh (tetrahedra): the fused per-cell step, the packed face-tiled step and the
    FP32-bulk step on the Sod state, 40 steps: float64 paths equal the oracle bit for bit, the
    FP32-bulk step equals its float32 model."""
    from manapy_amd.device import DeviceMesh, EulerTet, EulerTetTiles, Euler32
    from manapy_amd.renumber import renumbered
    from manapy_amd.solvers.euler import initialisation_euler_3d
    d = renumbered(domain('cube', 3))
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = initialisation_euler_3d(d.cells.center)
    U = U0.copy()
    G_ = np.zeros((6, F))
    model = oracle_lib.Euler32Model(d, U0, quirk)
    dt = float(np.min(oracle_lib.euler_dt(d, U0, 0.8, 1.4)))
    dt_m = dt
    for _ in range(40):
        dt_last = oracle_lib.euler_step(d, U, G_, 0.8, 1.4, quirk)
        dt_m = model.step(0.8, dt_m)
    m = DeviceMesh(d, lsq=False)
    tU = dev(torch, U0)
    # per-cell fused step on [6][C]
    S = [dev(torch, U0), dev(torch, np.zeros((6, C)))]
    ring = dev(torch, np.array([dt, 1e6, 1e6, 0.]))
    for s in range(40):
        call('mfv_euler3d_step_fused', m.ptr, P(S[s % 2]), P(S[(s + 1) % 2]), 1.4, 0.8, quirk, P(ring), s)
    assert np.array_equal(host(S[0]), U)
    assert host(ring)[39 % 3] == dt_last
    # packed records on face tiles
    et, tt = EulerTet(m), EulerTetTiles(m, 128)
    R = [torch.zeros((C, 8), dtype=torch.float64, device='cuda') for _ in range(2)]
    call('mfv_euler3d_pack', m.ptr, P(tU), P(R[0]))
    ring = dev(torch, np.array([dt, 1e6, 1e6, 0.]))
    for s in range(40):
        call('mfv_euler3d_step_packed_tiled', m.ptr, et.ptr, tt.ptr, P(R[s % 2]), P(R[(s + 1) % 2]), 1.4, 0.8, quirk,
             P(ring), s)
    back = dev(torch, np.zeros((6, C)))
    call('mfv_euler3d_unpack', m.ptr, P(R[0]), P(back))
    assert np.array_equal(host(back), U)
    # FP32-bulk
    e = Euler32(m, et=et, tt=tt)
    e.pack(tU)
    ring = dev(torch, np.array([dt, 1e6, 1e6, 0.]))
    for s in range(40):
        e.step(1.4, 0.8, quirk, ring, s)
    assert np.array_equal(host(e.rec[e.k % 2]), model.state)
    assert host(ring)[40 % 3] == dt_m
