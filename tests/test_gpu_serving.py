"""Continuous-batching serving of implicit advection-diffusion solves: each request's result
is bit-identical alone and in a mix of overlapping requests, matches an FP64 direct solve of
the same implicit steps, and admission respects the HBM budget."""
import numpy as np
import pytest

from conftest import gpu_available, make_tri_domain

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _request(rid, nx, ny, seed, D, vel, nsteps, arrival=0.):
    from manapy_amd.serving import AdvDiffRequest
    d = make_tri_domain(nx, ny, 1.0, 1.0, 0.1, 900 + seed)
    c = d.cells.center
    u0 = np.exp(-((c[:, 0] - .3) ** 2 + (c[:, 1] - .4) ** 2) / .01)
    return AdvDiffRequest(rid, d, u0, velocity=vel, D=D, nsteps=nsteps, arrival=arrival)


def test_alone_equals_mix():
    from manapy_amd.serving import ServingEngine
    specs = [(0, 25, 19, 1, 3e-3, (1.0, 0.5), 6), (1, 12, 30, 2, 5e-2, (-0.4, 0.9), 9),
             (2, 31, 28, 3, 1e-3, (0.7, -0.7), 5)]
    alone = {}
    for rid, nx, ny, seed, D, vel, n in specs:
        eng = ServingEngine(max_cells=1 << 14, max_inst=4)
        r = _request(rid, nx, ny, seed, D, vel, n)
        eng.run([r])
        alone[rid] = r.result
    eng = ServingEngine(max_cells=1 << 14, max_inst=4, n_iter=3)
    reqs = [_request(rid, nx, ny, seed, D, vel, n, arrival=0.002 * k)
            for k, (rid, nx, ny, seed, D, vel, n) in enumerate(specs)]
    done = eng.run(reqs)
    assert len(done) == 3 and eng.max_active >= 2
    for r in done:
        assert np.array_equal(r.result, alone[r.rid]), r.rid
    # eager solver ticks instead of graph replays: same bits
    eng = ServingEngine(max_cells=1 << 14, max_inst=4, n_iter=5, graph=False)
    reqs = [_request(rid, nx, ny, seed, D, vel, n, arrival=0.001 * k)
            for k, (rid, nx, ny, seed, D, vel, n) in enumerate(specs)]
    for r in eng.run(reqs):
        assert np.array_equal(r.result, alone[r.rid]), r.rid


def test_request_matches_fp64_direct_solve(oracle_lib):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from oracle import diamond as dm
    from manapy_amd.serving import ServingEngine
    r = _request(7, 22, 18, 4, 2e-2, (0.8, 0.3), 4)
    eng = ServingEngine(max_cells=1 << 12, max_inst=2)
    eng.run([r])
    d = r.domain
    C, F = int(d.nbcells), int(d.nbfaces)
    data, row, col = dm.get_triplet_2d(d, [], dm.matrixinnerfaces(d), np.zeros(0, np.int32))
    A = sp.csr_matrix((data, (row, col)), shape=(C, C))
    uf, vf, wf = np.full(F, 0.8), np.full(F, 0.3), np.zeros(F)
    uc, vc, wc = np.zeros(C), np.zeros(C), np.zeros(C)
    oracle_lib.facetocell(d, uf, uc)
    oracle_lib.facetocell(d, vf, vc)
    u = r.u0.copy()
    for _ in range(4):
        dt = oracle_lib.advection_step(d, u, uc, vc, wc, uf, vf, wf, 0.8, 1)
        M = (sp.identity(C) - dt * 2e-2 * A).tocsc()
        u = sla.spsolve(M, u)
    err = np.linalg.norm(r.result - u) / np.linalg.norm(u)
    assert err < 1e-8, err
    assert r.dt == oracle_lib.time_step(d, uc, vc, wc, 0.8)


def test_admission_by_hbm_budget():
    from manapy_amd.serving import ServingEngine
    reqs = [_request(k, 14, 14, 10 + k, 1e-2, (1.0, 0.0), 3, arrival=0.) for k in range(4)]
    eng = ServingEngine(max_cells=1 << 14, max_inst=8)
    one = reqs[0].footprint(eng.arena)
    eng.hbm_budget = int(1.5 * one)
    done = eng.run(reqs)
    assert len(done) == 4 and eng.max_active == 1
    assert all(r.t_done >= r.t_admit >= r.arrival for r in done)


def test_failed_request_is_isolated():
    """A request whose right-hand side holds a NaN fails alone: the other requests of the mix
    finish with the bits they have when served alone."""
    from manapy_amd.serving import ServingEngine, PHASE_DONE, PHASE_FAILED
    specs = [(0, 18, 15, 1, 3e-3, (1.0, 0.5), 4), (1, 14, 20, 2, 5e-2, (-0.4, 0.9), 5),
             (2, 16, 16, 3, 1e-2, (0.7, -0.7), 3)]
    alone = {}
    for rid, nx, ny, seed, D, vel, n in specs:
        if rid == 1:
            continue
        eng = ServingEngine(max_cells=1 << 14, max_inst=4)
        r = _request(rid, nx, ny, seed, D, vel, n)
        eng.run([r])
        alone[rid] = r.result
    eng = ServingEngine(max_cells=1 << 14, max_inst=4, n_iter=3)
    reqs = [_request(rid, nx, ny, seed, D, vel, n, arrival=0.001 * k)
            for k, (rid, nx, ny, seed, D, vel, n) in enumerate(specs)]
    reqs[1].u0[len(reqs[1].u0) // 2] = np.nan
    done = eng.run(reqs)
    assert len(done) == 3
    by = {r.rid: r for r in done}
    assert by[1].phase == PHASE_FAILED and by[1].failed and by[1].result is None and "failed" in by[1].error
    for rid in (0, 2):
        assert by[rid].phase == PHASE_DONE, by[rid].error
        assert np.array_equal(by[rid].result, alone[rid]), rid
    assert not eng.active and eng.in_use == 0


def test_request_that_never_fits_is_rejected():
    """A request above the HBM budget, or larger than the whole arena, is rejected at
    submission; the others are served."""
    from manapy_amd.serving import ServingEngine, PHASE_DONE, PHASE_REJECTED
    small = _request(0, 10, 10, 5, 1e-2, (1.0, 0.0), 2)
    big = _request(1, 40, 40, 6, 1e-2, (1.0, 0.0), 2)
    eng = ServingEngine(max_cells=1 << 14, max_inst=4)
    eng.hbm_budget = int(1.2 * small.footprint(eng.arena))
    done = eng.run([big, small])
    by = {r.rid: r for r in done}
    assert by[1].phase == PHASE_REJECTED and "budget" in by[1].error
    assert by[0].phase == PHASE_DONE
    eng = ServingEngine(max_cells=1 << 10, max_inst=4)
    huge = _request(2, 30, 30, 7, 1e-2, (1.0, 0.0), 2)
    done = eng.run([huge])
    assert done[0].phase == PHASE_REJECTED and "slices" in done[0].error


def test_raw_mesh_requests_are_prepared_on_arrival():
    """A request that carries its raw mesh (elements, nodes) and a callable initial field is
    prepared (Domain, device tables) by the engine's host threads after it arrives, while the
    GPU serves the others; its result is the same as the request built from a ready Domain,
    and its latency includes that preparation."""
    from manapy_amd.meshgen import jittered_triangles
    from manapy_amd.serving import AdvDiffRequest, ServingEngine
    blob = lambda c: np.exp(-((c[:, 0] - .3) ** 2 + (c[:, 1] - .4) ** 2) / .01)
    meshes = [jittered_triangles(n, n, 1.0, 1.0, 0.1, 950 + n) for n in (14, 31, 22)]
    ref = {}
    for k, (el, nd) in enumerate(meshes):
        from manapy_amd.ddm import Domain
        d = Domain(2, mesh=(el, nd))
        r = AdvDiffRequest(k, d, blob(d.cells.center), velocity=(0.6, -0.2), D=4e-3, nsteps=5)
        ServingEngine(max_cells=1 << 13, max_inst=4).run([r])
        ref[k] = r.result
    reqs = [AdvDiffRequest(k, (el, nd), blob, velocity=(0.6, -0.2), D=4e-3, nsteps=5, arrival=0.003 * k)
            for k, (el, nd) in enumerate(meshes)]
    assert not any(r.prepared for r in reqs)
    done = ServingEngine(max_cells=1 << 13, max_inst=4, prep_threads=2).run(reqs)
    assert len(done) == 3
    for r in done:
        assert r.prepared and r.prepare_s > 0
        assert r.t_done - r.arrival >= r.prepare_s
        assert np.array_equal(r.result, ref[r.rid]), r.rid


def test_request_whose_preparation_fails_is_rejected():
    """A raw-mesh request whose Domain cannot be built (an element names a node that does not
    exist) is rejected with the builder's error while the others are served."""
    from manapy_amd.meshgen import jittered_triangles
    from manapy_amd.serving import AdvDiffRequest, ServingEngine, PHASE_DONE, PHASE_REJECTED
    blob = lambda c: np.exp(-((c[:, 0] - .3) ** 2 + (c[:, 1] - .4) ** 2) / .01)
    el, nd = jittered_triangles(12, 12, 1.0, 1.0, 0.1, 961)
    bad_el = el.copy()
    bad_el[0, 0] = len(nd) + 5
    reqs = [AdvDiffRequest(0, (bad_el, nd), blob, velocity=(1.0, 0.0), D=1e-2, nsteps=2, arrival=0.),
            AdvDiffRequest(1, (el, nd), blob, velocity=(1.0, 0.0), D=1e-2, nsteps=2, arrival=0.)]
    eng = ServingEngine(max_cells=1 << 14, max_inst=4)
    done = eng.run(reqs)
    by = {r.rid: r for r in done}
    assert len(done) == 2
    assert by[0].phase == PHASE_REJECTED and "preparation failed" in by[0].error and "out of" in by[0].error
    assert by[1].phase == PHASE_DONE and by[1].result is not None
    assert not eng.active and eng.in_use == 0
