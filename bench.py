"""Benchmark: the explicit finite-volume step of the 2D shallow-water dam break on one GPU.

Workload (default): [0,100]^2 split into 1000 x 1000 x 2 = 2M triangles, interior nodes
jittered by +/-10% of h (seed 0); h = 2 for x < 50 else 1, u = v = 0; bathymetry = 8 seeded
Gaussian bumps (amplitude 0-0.3, width 2-8); second-order MUSCL with least-squares gradients
and the Barth-Jespersen limiter, hydrostatic reconstruction + Rusanov flux, wall boundaries,
CFL 0.8; 500 steps in float64. One step = gradient/limiter kernel + fused flux/update/dt
kernel; dt stays on the device (no host synchronisation inside the timed loop).

    python bench.py [--steps 500] [--warmup 20] [--nx 1000] [--no-cpu]

Prints one JSON line (metric, value = whole-job Mcell-updates/s, roofline of the dominant
kernel from HIP events over the timed region, and a CPU baseline on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s + %peak HBM GB/s at stated FP64 error; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell-update of a whole step (each array the step needs counted
# once at its element width): shallow water second order float64 / float32, Euler first order
# float64 / float32, advection first order float64; a Krylov (BiCGStab) iteration per matrix
# row with float32 ELL values.
ALGO_B_SW64, ALGO_B_MIXED = 356.0, 214.0
ALGO_B_EULER64, ALGO_B_EULER32 = 208.0, 124.0
ALGO_B_ADV64 = 112.0
ALGO_B_KRYLOV32 = 164.0


def measured_traffic(kernel, cells):
    """HBM bytes per launch of ``kernel`` on this workload from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (scaled by the copy calibration), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'hbm_traffic_sw_dam_break_2M.json')
    try:
        with open(path) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None
    k = data.get('kernels', {}).get(kernel)
    if not k or k.get('cells', data.get('cells')) != cells:
        return None
    return int(k['bytes'])


def krylov_traffic():
    """HBM bytes per row-iteration of the BiCGStab tick from the committed counter runs of the
    serving workload, or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'hbm_traffic_sw_dam_break_2M.json')
    try:
        with open(path) as fh:
            return int(json.load(fh)['krylov_row_iteration']['bytes'])
    except (OSError, ValueError, KeyError):
        return None


def log(msg):
    """Progress on stderr (the JSON result line is the only stdout output)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)
G = 9.81
CFL = 0.8


def kernel_bytes_per_cell_packed(MF, K_avg, G_avg):
    """Algorithmic HBM bytes per cell of the packed kernels: every array a kernel needs,
    counted once per cell at its element width (a neighbour's record is another cell's own
    record, already counted)."""
    recon = (8 * 8                # own record: cx, cy, eta, hu, hv (+h, Z, vol in the same line)
             + 4 * K_avg + 4      # cellnid entries + count
             + G_avg * (4 + 8 + 24) + 4  # ghostnid entries, face_cells, ghost centre + count
             + MF * 4 + 4         # neighbour codes + face count
             + MF * 32            # cfrec: rx, ry, z_f
             + MF * 32)           # write facerec
    flux = (8 * 8                 # own record (read) -- h, hu, hv, Z, vol
            + 4 * 8               # write eta, hu, hv, h
            + MF * 32             # own facerec slots (neighbours read theirs)
            + MF * 32             # cfgeo: n0, n1, mesure
            + MF * 4 + 1 + 4)     # neighbour codes, residual order, face count
    return recon, flux


def kernel_bytes_per_cell(F, C, K_avg, G_avg):
    """Algorithmic HBM bytes per cell of each kernel (every array the kernel needs, counted
    once, at its element width; face arrays scaled by F/C)."""
    fpc = F / C
    grad = (3 * 8 + 8            # h, hu, hv, Z
            + 16                 # cell centre (x, y)
            + 4 * K_avg + 4      # cellnid entries + count
            + 4 * G_avg + 4 + 16 * G_avg  # ghostnid entries + count + ghost centres
            + 3 * 4 + 4          # cell faces + count
            + fpc * (8 + 16)     # face cells, face centre
            + 9 * 8)             # write eta/hu/hv gradients and psi
    flux = (3 * 8 + 8            # state, Z
            + 9 * 8 + 2 * 8      # limited gradients, Z gradient
            + 16 + 8             # centre, volume
            + 3 * 4 + 3 + 4      # cell faces, residual order, count
            + fpc * (8 + 16 + 8 + 16)  # face cells, normal, mesure, face centre
            + 3 * 8)             # write new state
    return grad, flux


def build_case(nx, seed=0):
    from manapy_amd.ddm import Domain
    from manapy_amd.meshgen import jittered_triangles, gaussian_bumps
    el, nd = jittered_triangles(nx, nx, 100., 100., 0.1, seed)
    d = Domain(2, mesh=(el, nd))
    c = d.cells.center
    Z = gaussian_bumps(c, 8, np.random.default_rng(seed), [(0., 100.), (0., 100.)])
    h = np.where(c[:, 0] < 50., 2.0, 1.0)
    return d, Z, h


def cpu_baseline(d, Z, h0, block, repeats):
    """The oracle (single-threaded C restatement of the reference loops) on the same mesh and
    initial state: one warm-up step, then ``repeats`` timed blocks of ``block`` steps, each
    continuing from the last. Returns (median Mcell-updates/s over the blocks, the block rates,
    the total number of steps taken, the state (h, hu, hv) after them)."""
    from oracle import oracle
    C = int(d.nbcells)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg = oracle.z_gradient(d, Z)
    oracle.sw_step(d, h, hu, hv, Z, zg, G, CFL, 2)  # warm-up (its step counts for the parity run)
    rates = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for _ in range(block):
            oracle.sw_step(d, h, hu, hv, Z, zg, G, CFL, 2)
        rates.append(C * block / (time.perf_counter() - t0) / 1e6)
    return float(np.median(rates)), rates, 1 + block * repeats, (h, hu, hv)


def cpu_baseline_omp(d, Z, h0, block, repeats):
    """The same step on all host cores: the oracle's code with OpenMP over cells and faces
    (oracle/omp_baseline.c, bit-identical to the serial oracle), timed like cpu_baseline."""
    from oracle import oracle
    try:
        omp = oracle.OmpSwStep(d, Z)
    except OSError as exc:
        log("all-cores CPU baseline unavailable: %s" % exc)
        return None
    C = int(d.nbcells)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    omp.step(h, hu, hv, G, CFL, 2)
    rates = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for _ in range(block):
            omp.step(h, hu, hv, G, CFL, 2)
        rates.append(C * block / (time.perf_counter() - t0) / 1e6)
    return {"value": round(float(np.median(rates)), 3), "unit": "Mcell-updates/s", "cores": omp.threads(),
            "kind": "port", "block_rates": [round(x, 3) for x in rates],
            "sample": "median of %d blocks of %d steps of the same 2M-cell case, OpenMP over cells and faces "
                      "(the oracle's code, bit-identical to the serial oracle), threads = OMP_NUM_THREADS"
                      % (repeats, block)}


def region_step_time(step, first, n):
    """Device time per step of ``n`` un-instrumented steps, from HIP events recorded on the
    launch stream (torch's current stream, the one every step is launched on) around them."""
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for k in range(n):
        step(first + k)
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / n


def rel_errors(a, b, vol):
    """Relative L1 (Variable.norml2 order 1: sum |a - b| |K| / sum |b| |K|) and relative L-inf."""
    return {"rel_L1": float(np.sum(np.abs(a - b) * vol) / np.sum(np.abs(b) * vol)),
            "rel_Linf": float(np.abs(a - b).max() / np.abs(b).max())}


def mixed_bytes_per_cell(MF, K_avg, G_avg):
    """Algorithmic HBM bytes per cell of the two mixed-precision kernels (each array the
    kernel needs counted once at its width; neighbour records are other cells' own)."""
    recon = (32 + 16                    # own state (h, hu, hv, Z), centre (x, y)
             + 4 * K_avg + 4            # cellnid entries + count
             + G_avg * (4 + 4 + 16) + 4  # ghostnid entries, owner, ghost centre + count
             + 4 * MF + 4               # neighbour codes + face count
             + 4 * MF + 16 * MF         # geoA.w (z_f - Z), geoB (rx, ry, |n|, 1/|n|)
             + 16 * MF)                 # write face records
    flux = (32 + 24 + 8                 # state read + write, volume
            + 4 + 1                     # face count, residual order
            + 16 * MF + 16 * MF         # own face records, geoA
            + 4 * MF)                   # neighbour codes
    return recon, flux


def fused_bytes_per_cell(MF, K_avg, G_avg):
    """Algorithmic HBM bytes per cell of the fused float64 tiled step (records of other cells
    are their own; the halo's repeated reconstruction is extra work, not extra bytes)."""
    return (64 + 32                     # own record read, new eta/hu/hv/h written
            + 4 + 1                     # face count, residual order
            + 4 * K_avg + 4             # cellnid entries + count
            + G_avg * (4 + 4 + 16) + 4  # ghostnid entries, owner, ghost centre + count
            + 4 * MF + 4 * MF           # global and block-local neighbour codes
            + 32 * MF + 32 * MF)        # cfrec, cfgeo


def tiled_bytes_per_cell(MF, K_avg, G_avg):
    """Algorithmic HBM bytes per cell of the fused tiled mixed step (the halo's repeated
    reconstruction is extra work, not extra algorithmic bytes)."""
    return (32 + 32 + 16 + 8           # state read + write, centre, volume
            + 4 + 1                     # face count, residual order
            + 4 * K_avg + 4             # cellnid entries + count
            + G_avg * (4 + 4 + 16) + 4  # ghostnid entries, owner, ghost centre + count
            + 4 * MF + 4 * MF           # global and block-local neighbour codes
            + 16 * MF + 16 * MF)        # geoA, geoB


def time_steps(step, warmup, steps):
    """Warm-up steps, then ``steps`` steps timed on the host clock between synchronisations, then
    ``steps`` more between two HIP events on the launch stream. Returns (s per step of the host
    loop, device s per step from the events)."""
    import torch
    for k in range(warmup):
        step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        step(warmup + k)
    torch.cuda.synchronize()
    t_loop = (time.perf_counter() - t0) / steps
    return t_loop, region_step_time(step, warmup + steps, steps)


def cpu_blocks(run_block, cells, total, blocks):
    """Runs ``total`` steps as ``blocks`` timed blocks (the last takes the remainder); returns
    (median Mcell-updates/s, the block rates)."""
    rates = []
    base = total // blocks
    for b in range(blocks):
        n = base if b < blocks - 1 else total - base * (blocks - 1)
        t0 = time.perf_counter()
        run_block(n)
        rates.append(cells * n / (time.perf_counter() - t0) / 1e6)
    return float(np.median(rates)), rates


def run_advection(n, steps, warmup, cpu=True):
    """Advection of a Gaussian blob on n x n x 2 jittered triangles of the unit square (u = 1,
    v = 0.5 on every face, all-Neumann boundaries, order 1, CFL 0.8): the fused device step
    (mfv_advection_step) against the oracle's step, bit for bit, after every timed step."""
    import torch
    from manapy_amd import _lib
    from manapy_amd.ddm import Domain
    from manapy_amd.device import DeviceMesh, stream_ptr
    from manapy_amd.meshgen import jittered_triangles
    el, nd = jittered_triangles(n, n, 1.0, 1.0, 0.1, 0)
    d = Domain(2, mesh=(el, nd))
    C, F = int(d.nbcells), int(d.nbfaces)
    c = d.cells.center
    w0 = np.exp(-((c[:, 0] - .3) ** 2 + (c[:, 1] - .3) ** 2) / .005)
    m = DeviceMesh(d, lsq=False)
    s = stream_ptr()
    dv = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    uf, vf, wf = dv(np.full(F, 1.0)), dv(np.full(F, 0.5)), dv(np.zeros(F))
    uc, vc, wc = dv(np.zeros(C)), dv(np.zeros(C)), dv(np.zeros(C))
    for a, b in ((uf, uc), (vf, vc), (wf, wc)):
        _lib.call('mfv_facetocell', m.ptr, a.data_ptr(), b.data_ptr(), s)
    W = [dv(w0), dv(np.zeros(C))]
    ring, dt0 = dv(np.zeros(4)), dv(np.zeros(1))
    _lib.call('mfv_time_step', m.ptr, uc.data_ptr(), vc.data_ptr(), wc.data_ptr(), CFL, dt0.data_ptr(), None, s)
    _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)

    def step(k, sp=s):
        _lib.call('mfv_advection_step', m.ptr, W[k % 2].data_ptr(), W[(k + 1) % 2].data_ptr(), uc.data_ptr(),
                  vc.data_ptr(), wc.data_ptr(), uf.data_ptr(), vf.data_ptr(), wf.data_ptr(), CFL, 1,
                  ring.data_ptr(), k, None, sp)
    t_loop, t_ev = time_steps(step, warmup, steps)
    nsteps = warmup + 2 * steps
    # the same steps captured in a HIP graph (102 steps: a multiple of 3 for the dt ring and of 2
    # for the state buffers, so each replay continues where the last one ended), replayed 5 times
    graph = None
    if nsteps % 6 == 0:
        try:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for k in range(6):  # first use of the side stream (its dt scratch) outside the capture
                    step(k, side.cuda_stream)
            torch.cuda.current_stream().wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=side):
                for k in range(102):
                    step(k, side.cuda_stream)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                g.replay()
            e1.record()
            e1.synchronize()
            nsteps += 6 + 5 * 102
            graph = {"steps": 510, "ms_per_step": round(e0.elapsed_time(e1) / 510, 5),
                     "value": round(C * 510 / (e0.elapsed_time(e1) * 1e-3) / 1e6, 3)}
        except Exception as exc:  # a capture this build cannot do is reported, not fatal
            graph = {"error": str(exc)[:200]}
    w_gpu = W[nsteps % 2].cpu().numpy()
    # algorithmic bytes of k_adv_step (order 1): w read + write, volume, u_c v_c w_c (dt),
    # faceid + order entries; per face (x F/C): cellid, normal (3 doubles), u_f v_f w_f
    MF = int(d.cells.faceid.shape[1] - 1)
    bpc = (8 + 8 + 8 + 24 + 4 + (4 + 1) * MF) + (8 + 24 + 24) * F / C
    rl = roofline("k_adv_step + k_dt_finish", ALGO_B_ADV64, C, t_ev, bpc, measured_traffic('k_adv_step<2, 1>', C))
    rl["bound"] = "latency (the step's arrays, %.1f MB, stay in the 256 MB Infinity Cache)" % (bpc * C / 1e6)
    out = {"workload": "advection of a Gaussian blob, %dx%dx2 jittered triangles, order 1, CFL 0.8" % (n, n),
           "cells": C, "faces": F, "steps": steps, "dtype": "f64",
           "value": round(C / t_loop / 1e6, 3), "unit": "Mcell-updates/s", "ms_per_step": round(t_loop * 1e3, 5),
           "roofline": rl, "graph": graph}
    if cpu:
        from oracle import oracle as O
        u_, v_, w_ = np.full(F, 1.0), np.full(F, 0.5), np.zeros(F)
        ucc, vcc, wcc = np.zeros(C), np.zeros(C), np.zeros(C)
        for a, b in ((u_, ucc), (v_, vcc), (w_, wcc)):
            O.facetocell(d, a, b)
        w = w0.copy()

        def block(k):
            for _ in range(k):
                O.advection_step(d, w, ucc, vcc, wcc, u_, v_, w_, CFL, 1)
        v, rates = cpu_blocks(block, C, nsteps, 6)
        out["parity"] = {"steps": nsteps, "bitwise_equal_to_oracle": bool(np.array_equal(w, w_gpu)),
                         "max_abs_diff": float(np.abs(w - w_gpu).max())}
        out["cpu_baseline"] = {"value": round(v, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
                               "block_rates": [round(x, 3) for x in rates],
                               "sample": "median of 6 blocks over the %d steps of the parity run, single-threaded "
                                         "C restatement of the reference loops" % nsteps}
    return out


def run_euler(n, steps, warmup, cpu=True, cpu_steps=5, renumber='hilbert'):
    """3D Sod shock tube on n^3 cubes x 6 Kuhn tetrahedra of the unit cube (interior nodes
    jittered by +/-8 % of h, slip walls, first-order Rusanov, CFL 0.8): the fastest of the
    fused device steps (per-cell on [6][C], packed records, face tiles). Parity: the first ``cpu_steps`` steps against the oracle, bit for
    bit, at full size (the same run times the CPU baseline)."""
    import torch
    from manapy_amd import _lib
    from manapy_amd.ddm import Domain
    from manapy_amd.device import DeviceMesh, stream_ptr
    from manapy_amd.meshgen import kuhn_tetrahedra
    from manapy_amd.solvers.euler import initialisation_euler_3d
    t0 = time.perf_counter()
    el, nd = kuhn_tetrahedra(n, 1.0, 0.08, 0)
    d = Domain(3, mesh=(el, nd))
    if renumber in ('hilbert', 'morton'):  # cells along a space-filling curve (a relabelling: same bits)
        from manapy_amd.renumber import hilbert_order_3d, morton_order_3d, renumbered
        c = d.cells.center
        order = hilbert_order_3d if renumber == 'hilbert' else morton_order_3d
        d = renumbered(d, order(c[:, 0], c[:, 1], c[:, 2]))
    t_mesh = time.perf_counter() - t0
    C, F = int(d.nbcells), int(d.nbfaces)
    U0 = initialisation_euler_3d(d.cells.center)
    m = DeviceMesh(d, lsq=False)
    s = stream_ptr()
    dv = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    S = [dv(U0), dv(np.zeros((6, C)))]
    ring, dt0 = dv(np.zeros(4)), dv(np.zeros(1))

    from manapy_amd.device import EulerTet, EulerTetTiles, FaceTiles
    tiles, ptiles, ctiles = {}, {}, {}
    for T in (128, 256, 512):
        if T < 512:
            ft = FaceTiles(m, T)
            if ft.lds_bytes() <= 64 * 1024:
                tiles[T] = ft
        pt = EulerTetTiles(m, T)
        if pt.lds_bytes() <= 64 * 1024:
            ptiles[T] = pt
        if T > 128 and pt.lds_bytes(cells=True) <= 64 * 1024:
            ctiles[T] = pt
    et = EulerTet(m)
    R = [torch.empty((C, 8), dtype=torch.float64, device='cuda') for _ in range(2)]
    path = ['fused']  # 'fused': per-cell step on [6][C]; 'tiled:T': face tiles; 'packed': [C][8] records

    def init():
        S[0].copy_(dv(U0))
        _lib.call('mfv_euler3d_time_step', m.ptr, S[0].data_ptr(), CFL, 1.4, dt0.data_ptr(), None, s)
        _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
        _lib.call('mfv_euler3d_pack', m.ptr, S[0].data_ptr(), R[0].data_ptr(), s)

    def step(k):
        if path[0] == 'fused':
            _lib.call('mfv_euler3d_step_fused', m.ptr, S[k % 2].data_ptr(), S[(k + 1) % 2].data_ptr(), 1.4, CFL, 0,
                      ring.data_ptr(), k, s)
        elif path[0] == 'packed':
            _lib.call('mfv_euler3d_step_packed', m.ptr, et.ptr, R[k % 2].data_ptr(), R[(k + 1) % 2].data_ptr(), 1.4,
                      CFL, 0, ring.data_ptr(), k, s)
        elif path[0].startswith('packed-tiled:'):
            _lib.call('mfv_euler3d_step_packed_tiled', m.ptr, et.ptr, ptiles[int(path[0].split(':')[1])].ptr,
                      R[k % 2].data_ptr(), R[(k + 1) % 2].data_ptr(), 1.4, CFL, 0, ring.data_ptr(), k, s)
        elif path[0].startswith('cells-tiled:'):
            _lib.call('mfv_euler3d_step_cells_tiled', m.ptr, et.ptr, ctiles[int(path[0].split(':')[1])].ptr,
                      R[k % 2].data_ptr(), R[(k + 1) % 2].data_ptr(), 1.4, CFL, 0, ring.data_ptr(), k, s)
        else:
            _lib.call('mfv_euler3d_step_tiled', m.ptr, tiles[int(path[0].split(':')[1])].ptr, S[k % 2].data_ptr(),
                      S[(k + 1) % 2].data_ptr(), 1.4, CFL, 0, ring.data_ptr(), k, s)
    probe = {}
    for cand in (['fused', 'packed'] + ['packed-tiled:%d' % T for T in sorted(ptiles)] +
                 ['cells-tiled:%d' % T for T in sorted(ctiles)] + ['tiled:%d' % T for T in sorted(tiles)]):
        path[0] = cand
        init()
        probe[cand] = time_steps(step, 2, 10)[1] * 1e3
    best = min(probe, key=probe.get)
    path[0] = best
    log("euler: path %s %s" % (best, {k: round(v, 4) for k, v in probe.items()}))
    init()
    t_loop, t_ev = time_steps(step, warmup, steps)
    MF = int(d.cells.faceid.shape[1] - 1)
    if best == 'packed':
        # record read + write (64 B each: six conserved/primitive values, volume, pad), the
        # per-cell across-codes (16 B), face geometry (4 x 32 B) and rank byte
        bpc = 64 + 64 + 16 + 128 + 1
        kname = "k_euler_step_packed"
    elif best.startswith('packed-tiled:'):
        # records read + written, per-cell entry codes (16 B), rank byte; per face (x F/C): one
        # 40-byte entry (geometry, item pair)
        bpc = 64 + 64 + 16 + 1 + 40 * F / C
        kname = "k_euler_step_packed_tiled<%s>" % best.split(':')[1]
    elif best.startswith('cells-tiled:'):
        # the same tables as the face-tiled step (the fluxes stay in registers)
        bpc = 64 + 64 + 16 + 1 + 40 * F / C
        kname = "k_euler_step_cells_tiled<%s>" % best.split(':')[1]
    else:
        # U read + write (6 doubles each), volume, face count; per face slot: order entry +
        # faceid; per face (x F/C): cellid, normal, mesure, name
        bpc = (48 + 48 + 8 + 4 + (1 + 4) * MF) + (8 + 24 + 8 + 4) * F / C
        kname = "k_euler_step_tet" if best == 'fused' else "k_euler_step_tiled<%s>" % best.split(':')[1]
    out = {"workload": "3D Sod shock tube, %d^3 x 6 Kuhn tetrahedra, first-order Rusanov, slip walls, CFL 0.8" % n,
           "cells": C, "faces": F, "steps": steps, "dtype": "f64", "mesh_build_s": round(t_mesh, 1),
           "value": round(C / t_loop / 1e6, 3), "unit": "Mcell-updates/s", "ms_per_step": round(t_loop * 1e3, 5),
           "path": best, "path_ms_per_step": {k: round(v, 4) for k, v in probe.items()}, "cell_order": renumber,
           "roofline": roofline(kname + " + k_dt_finish", ALGO_B_EULER64, C, t_ev, bpc,
                                measured_traffic(kname, C))}
    # FP32-bulk variant on the same tiles: throughput, then 200-step (or --euler-steps) accuracy
    # against the float64 step and the float64 conservation sums
    # tiles whose FP32-bulk LDS (24 B per item and per face entry) fits: 256 and 512 cells
    cand32 = {}
    for T in (256, 512, 128):
        tt = ptiles.get(T) or ctiles.get(T)
        if tt is not None and 24 * tt.host['max_items'] + 24 * tt.host['max_faces'] <= 64 * 1024:
            cand32[T] = tt
    if cand32:
        from manapy_amd.device import Euler32
        e32s = {T: Euler32(m, et=et, tt=tt) for T, tt in cand32.items()}
        cur = {}

        def init32():
            S[0].copy_(dv(U0))
            _lib.call('mfv_euler3d_time_step', m.ptr, S[0].data_ptr(), CFL, 1.4, dt0.data_ptr(), None, s)
            _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
            cur['e'].pack(S[0], s)

        def step32(k):
            cur['e'].step(1.4, CFL, 0, ring, k, sums=False, stream=s)
        probe32 = {}
        for T, e in e32s.items():
            cur['e'] = e
            init32()
            probe32[T] = time_steps(step32, 2, 10)[1] * 1e3
        T32 = min(probe32, key=probe32.get)
        e32 = cur['e'] = e32s[T32]
        log("euler fp32-bulk: tiles %d %s" % (T32, {k: round(v, 4) for k, v in probe32.items()}))
        init32()
        t32_loop, t32_ev = time_steps(step32, warmup, steps)
        # accuracy: both paths from the initial state for ``steps`` steps
        path[0] = best
        init()
        for k in range(steps):
            step(k)
        if best.startswith('packed'):
            _lib.call('mfv_euler3d_unpack', m.ptr, R[steps % 2].data_ptr(), S[steps % 2].data_ptr(), s)
        U64 = S[steps % 2].cpu().numpy()
        init32()
        for k in range(steps):
            e32.step(1.4, CFL, 0, ring, k, sums=(k == steps - 1), stream=s)
        U32t = torch.empty((6, C), dtype=torch.float64, device='cuda')
        e32.unpack(U32t, s)
        U32 = U32t.cpu().numpy()
        cons32 = e32.cons.cpu().numpy()
        vol = d.cells.volume
        cons64 = np.array([np.sum(U64[k] * vol) for k in (0, 2, 3, 4, 5)])
        errs = {nm: rel_errors(U32[k], U64[k], vol) for k, nm in enumerate(('rho', 'P', 'rhou', 'rhov', 'rhow',
                                                                             'rhoE'))}
        b32 = 32 + 32 + 16 + 1 + 24 * F / C
        out["fp32_bulk"] = {
            "value": round(C / t32_loop / 1e6, 3), "unit": "Mcell-updates/s", "ms_per_step": round(t32_loop * 1e3, 5),
            "speedup_vs_fp64": round(t_loop / t32_loop, 3), "tile": T32,
            "tile_ms_per_step": {str(k): round(v, 4) for k, v in probe32.items()},
            "dtype": "f32 records, geometry and face fluxes; f64 gathers, update, dt and conservation sums",
            "roofline": roofline("k_euler_step32_tiled<%d> + k_dt_finish" % T32, ALGO_B_EULER32, C, t32_ev, b32,
                                 measured_traffic("k_euler_step32_tiled<%d>" % T32, C)),
            "accuracy": {"steps": steps, "error_vs_fp64": errs,
                         "max_rel_L1": max(v["rel_L1"] for v in errs.values()), "target_rel_L1": 1e-5,
                         "conservation_sums_f64": {"fields": ["rho", "rhou", "rhov", "rhow", "rhoE"],
                                                   "fp32_bulk": cons32.tolist(), "fp64_run": cons64.tolist(),
                                                   "rel_diff": ((cons32 - cons64) / np.abs(cons64)).tolist()}}}
        log("euler fp32-bulk: %.1f Mcell-updates/s, max rel L1 %.3g" % (out["fp32_bulk"]["value"],
                                                                       out["fp32_bulk"]["accuracy"]["max_rel_L1"]))
    if cpu and cpu_steps:
        from oracle import oracle as O
        path[0] = best
        init()
        for k in range(cpu_steps):
            step(k)
        if best.startswith('packed'):
            _lib.call('mfv_euler3d_unpack', m.ptr, R[cpu_steps % 2].data_ptr(), S[cpu_steps % 2].data_ptr(), s)
        U_gpu = S[cpu_steps % 2].cpu().numpy()
        U = U0.copy()
        G_ = np.zeros((6, F))

        def block(k):
            for _ in range(k):
                O.euler_step(d, U, G_, CFL, 1.4, 0)
        v, rates = cpu_blocks(block, C, cpu_steps, cpu_steps)
        out["parity"] = {"steps": cpu_steps, "bitwise_equal_to_oracle": bool(np.array_equal(U, U_gpu)),
                         "max_abs_diff": float(np.abs(U - U_gpu).max())}
        out["cpu_baseline"] = {"value": round(v, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
                               "block_rates": [round(x, 3) for x in rates],
                               "sample": "median of %d single steps of the same case (the parity run), "
                                         "single-threaded C restatement of the reference loops" % cpu_steps}
    return out


def run_mixed(d, m, Z, h0, zgrad, warmup, steps, mcell_fp64, kernel='auto', tile=256):
    """The same case on the mixed-precision path: throughput and the step's roofline (its
    accuracy against float64 is the separate fixed-length leg ``mixed_accuracy``). kernel=
    'tiled' / 'staged' run the fused one-launch step (states from HBM / staged in LDS), 'split'
    the reconstruction + flux kernels, 'auto' the fastest of them on this mesh."""
    import torch
    from manapy_amd import _lib
    from manapy_amd.device import SwMixed, stream_ptr
    C = int(d.nbcells)
    s = stream_ptr()
    p = SwMixed(m)
    dv = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    th, tz, tZ = dv(h0), dv(np.zeros(C)), dv(Z)
    _lib.call('mfv_sw_pack_mixed', m.ptr, p.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), tZ.data_ptr(),
              zgrad.data_ptr(), 2, s)
    _lib.call('mfv_sw_conservation', m.ptr, p.ptr, p.partial.data_ptr(), p.partial.numel(), p.sums.data_ptr(), s)
    mass0 = float(p.sums[0])
    dt0 = dv(np.zeros(1))
    ring = dv(np.zeros(4))
    _lib.call('mfv_time_step_sw', m.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), CFL, G, 0., 0., dt0.data_ptr(),
              None, s)
    _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
    tune = None
    ncell = tile
    if kernel == 'auto':
        (kernel, tile, ncell), tune = p.autotune(m, G, CFL, 2, ring)
        log("mixed: autotuned %s %d/%d %s" % (kernel, tile, ncell,
                                              {"%s:%d:%d" % c: round(v, 4) for c, v in tune.items()}))

    def step(k):
        if kernel in ('tiled', 'staged'):
            p.step_tiled(m, G, CFL, 2, ring, k, T=tile, stage=(kernel == 'staged'), stream=s, N=ncell)
        else:
            _lib.call('mfv_sw_recon_mixed', m.ptr, p.ptr, 2, s)
            _lib.call('mfv_sw_flux_mixed', m.ptr, p.ptr, G, CFL, ring.data_ptr(), k, s)

    for k in range(warmup):
        step(k)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for k in range(steps):
        step(warmup + k)
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t1
    t_dev = region_step_time(step, warmup + steps, steps)
    _lib.call('mfv_sw_conservation', m.ptr, p.ptr, p.partial.data_ptr(), p.partial.numel(), p.sums.data_ptr(), s)
    mass1 = float(p.sums[0])
    K_avg = float(d.cells.cellnid[:, -1].mean())
    G_avg = float(d.cells.ghostnid[:, -1].mean())
    MF = int(d.cells.faceid.shape[1] - 1)
    mcell = C * steps / t_loop / 1e6
    if kernel in ('tiled', 'staged'):
        kname, bmodel = 'k_sw_step32_tiled + k_dt_finish', tiled_bytes_per_cell(MF, K_avg, G_avg)
        trace_key = "k_sw_step32_tiled:%s:%d:%d" % (kernel, tile, ncell)
    else:
        kname, bmodel = 'k_sw_recon32 + k_sw_flux32', sum(mixed_bytes_per_cell(MF, K_avg, G_avg))
        trace_key = None
    return {
        "value": round(mcell, 3), "unit": "Mcell-updates/s", "ms_per_step": round(t_loop / steps * 1e3, 5),
        "kernel": kernel, "tile": tile if kernel != 'split' else None,
        "autotune_ms_per_step": {"%s:%d:%d" % c: round(v, 4) for c, v in tune.items()} if tune else None,
        "speedup_vs_fp64": round(mcell / mcell_fp64, 3),
        "dtype": "f32 flux/gradient arithmetic and face records, f64 state/gathers/dt/sums",
        "mass_rel_drift": (mass1 - mass0) / mass0,
        "roofline": roofline(kname, ALGO_B_MIXED, C, t_dev, bmodel,
                             measured_traffic(trace_key, C) if trace_key else None),
    }


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """HBM bandwidth of a plain device-to-device copy of ``nbytes`` (bytes read + written per
    second, HIP events around ``reps`` copies after two untimed ones): the measured ceiling the
    roofline fractions can be read against, beside the 8 TB/s of the specification."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    del src, dst
    return 2 * nbytes / t / 1e9


def roofline(kernel, bytes_per_cell, cells, t_step, model_bytes=None, traffic=None):
    """Roofline entry of one step: algorithmic bytes per cell-update (the ALGO_B_* counts above)
    x cells / the device time per step from HIP events over the timed region."""
    ach = bytes_per_cell * cells / t_step / 1e9
    r = {"bound": "hbm", "kernel": kernel, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_cell": bytes_per_cell,
         "avg_launch_us": round(t_step * 1e6, 2)}
    if model_bytes is not None:
        r["bytes_per_cell_kernel_model"] = round(model_bytes, 1)
        r["frac_kernel_model"] = round(model_bytes * cells / t_step / 1e9 / HBM_PEAK_GBS, 4)
    return r


def fp64_sw_run(d, m, Z, h0, zgrad_t, checkpoints, T=256):
    """The float64 fused step from the initial state; (h, hu, hv) on the host after each
    checkpoint's number of steps (device numbering)."""
    import torch
    from manapy_amd import _lib
    from manapy_amd.device import SwPacked, stream_ptr
    C = int(d.nbcells)
    s = stream_ptr()
    dv = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    S0 = dv(np.stack([h0, np.zeros(C), np.zeros(C)]))
    tZ = dv(Z)
    pk = SwPacked(m)
    pk.pack(m, S0[0], S0[1], S0[2], tZ, zgrad_t, 2, s)
    dt0, ring = dv(np.zeros(1)), dv(np.zeros(4))
    _lib.call('mfv_time_step_sw', m.ptr, S0[0].data_ptr(), S0[1].data_ptr(), S0[2].data_ptr(), CFL, G, 0., 0.,
              dt0.data_ptr(), None, s)
    _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
    out = {}
    o = [torch.empty(C, dtype=torch.float64, device='cuda') for _ in range(3)]
    for k in range(max(checkpoints)):
        pk.step_tiled(m, G, CFL, 2, ring, k, T=T, stream=s)
        if k + 1 in checkpoints:
            _lib.call('mfv_sw_unpack', m.ptr, pk.ptr, o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), s)
            out[k + 1] = [x.cpu().numpy() for x in o]
    return out


def mixed_accuracy(d, m, Z, h0, zgrad_t, checkpoints=(25, 100, 500), T=256):
    """Per-field error of the mixed-precision path against the float64 path on the same case
    after each checkpoint's number of steps (independent of --steps), and its mass drift."""
    import torch
    from manapy_amd import _lib
    from manapy_amd.device import SwMixed, stream_ptr
    C = int(d.nbcells)
    s = stream_ptr()
    ref = fp64_sw_run(d, m, Z, h0, zgrad_t, checkpoints, T)
    dv = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    p = SwMixed(m)
    th, tz, tZ = dv(h0), dv(np.zeros(C)), dv(Z)
    _lib.call('mfv_sw_pack_mixed', m.ptr, p.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), tZ.data_ptr(),
              zgrad_t.data_ptr(), 2, s)
    _lib.call('mfv_sw_conservation', m.ptr, p.ptr, p.partial.data_ptr(), p.partial.numel(), p.sums.data_ptr(), s)
    mass0 = float(p.sums[0])
    dt0, ring = dv(np.zeros(1)), dv(np.zeros(4))
    _lib.call('mfv_time_step_sw', m.ptr, th.data_ptr(), tz.data_ptr(), tz.data_ptr(), CFL, G, 0., 0., dt0.data_ptr(),
              None, s)
    _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
    vol = d.cells.volume
    o = [torch.empty(C, dtype=torch.float64, device='cuda') for _ in range(3)]
    res = {}
    worst = 0.
    for k in range(max(checkpoints)):
        p.step_tiled(m, G, CFL, 2, ring, k, T=T, stage=True, stream=s)
        if k + 1 in checkpoints:
            _lib.call('mfv_sw_unpack_mixed', m.ptr, p.ptr, o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), s)
            e = {nm: rel_errors(a.cpu().numpy(), b, vol) for nm, a, b in zip(('h', 'hu', 'hv'), o, ref[k + 1])}
            worst = max(worst, max(v["rel_L1"] for v in e.values()))
            res[str(k + 1)] = e
    _lib.call('mfv_sw_conservation', m.ptr, p.ptr, p.partial.data_ptr(), p.partial.numel(), p.sums.data_ptr(), s)
    mass1 = float(p.sums[0])
    return {"steps": list(checkpoints), "error_vs_fp64": res, "max_rel_L1": worst, "target_rel_L1": 1e-5,
            "meets_target": bool(worst <= 1e-5), "mass_rel_drift": (mass1 - mass0) / mass0}


def run_ensemble(n_inst, steps, seed=0, n_check=None):
    """Batched ensemble: n_inst shallow-water instances (sizes log-uniform in [5k, 200k]
    triangles, aspect ratio 1-4, 1-10 bumps, random dam) advanced together by a captured graph
    with the autotuned launch configuration; error of every instance (or n_check of them)
    against its own float64 run."""
    import torch
    from manapy_amd import _lib
    from manapy_amd.ddm import Domain
    from manapy_amd.device import DeviceMesh, SwPacked, stream_ptr
    from manapy_amd.ensemble import SwEnsemble, random_instance, instance_state
    if steps % 6:
        raise ValueError("ensemble steps must be a multiple of 6 (dt ring period 3, two state buffers)")
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    sizes = np.exp(rng.uniform(np.log(5e3), np.log(2e5), n_inst))
    doms, states = [], []
    for k, c in enumerate(sizes):
        if k % 32 == 0:
            log("ensemble: building instance %d / %d" % (k, n_inst))
        el, nd, lx, ly, r = random_instance(rng, c)
        d = Domain(2, mesh=(el, nd))
        doms.append(d)
        states.append(instance_state(d, lx, ly, r))
    t_build = time.perf_counter() - t0
    log("ensemble: uploading %d instances" % n_inst)
    e = SwEnsemble(doms, states)
    tune = e.autotune(steps=6)
    log("ensemble: launch configuration %s" % (tune['config'],))
    e.step(6)
    torch.cuda.synchronize()
    g = e.graph(steps)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    ev0.record()
    e.run_graph(steps, chunk=steps)
    ev1.record()
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t1
    t_dev = ev0.elapsed_time(ev1) * 1e-3 / steps
    cells = int(sum(int(d.nbcells) for d in doms))
    mcell = cells * steps / t_loop / 1e6
    # float64 reference for a sample of instances (same initial state, same 6 + steps steps)
    errs = []
    pick = np.arange(n_inst) if n_check is None else np.linspace(0, n_inst - 1, n_check).astype(int)
    log("ensemble: float64 runs of %d instances for the error" % len(pick))
    s = stream_ptr()
    for i in pick:
        d = doms[i]
        h0, hu0, hv0, Z = states[i]
        C, F = int(d.nbcells), int(d.nbfaces)
        m = DeviceMesh(d)
        pk = SwPacked(m)
        dv = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
        tZ, zgh, zg = dv(Z), dv(np.zeros(F)), dv(np.zeros((3, C)))
        bf = torch.as_tensor(d.boundaryfaces, device='cuda')
        _lib.call('mfv_ghost_value', m.ptr, 0, tZ.data_ptr(), zgh.data_ptr(), bf.data_ptr(), len(bf), s)
        _lib.call('mfv_cell_gradient', m.ptr, tZ.data_ptr(), zgh.data_ptr(), zg[0].data_ptr(), zg[1].data_ptr(),
                  zg[2].data_ptr(), s)
        th, thu, thv = dv(h0), dv(hu0), dv(hv0)
        _lib.call('mfv_sw_pack', m.ptr, pk.ptr, th.data_ptr(), thu.data_ptr(), thv.data_ptr(), tZ.data_ptr(),
                  zg.data_ptr(), 2, s)
        dt0, ring = dv(np.zeros(1)), dv(np.zeros(4))
        _lib.call('mfv_time_step_sw', m.ptr, th.data_ptr(), thu.data_ptr(), thv.data_ptr(), CFL, G, 0., 0.,
                  dt0.data_ptr(), None, s)
        _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)
        for k in range(6 + steps):
            _lib.call('mfv_sw_recon', m.ptr, pk.ptr, 2, s)
            _lib.call('mfv_sw_flux_packed', m.ptr, pk.ptr, G, CFL, ring.data_ptr(), k, s)
        _lib.call('mfv_sw_unpack', m.ptr, pk.ptr, th.data_ptr(), thu.data_ptr(), thv.data_ptr(), s)
        ref = [x.cpu().numpy() for x in (th, thu, thv)]
        got = e.instance_state(i)
        vol = d.cells.volume
        errs.append({nm: float(np.sum(np.abs(a - b) * vol) / np.sum(np.abs(b) * vol))
                     for nm, a, b in zip(('h', 'hu', 'hv'), got, ref)})
    sums = e.conservation()
    return {
        "value": round(mcell, 3), "unit": "Mcell-updates/s", "instances": n_inst, "cells": cells,
        "steps": steps, "ms_per_step": round(t_loop / steps * 1e3, 4), "graph": "one captured graph of all steps",
        "config": list(tune['config']), "autotune_ms_per_step": tune['ms_per_step'],
        "roofline": roofline("k_sw_step32_tiled (all instances) + k_batch_dt_fold", ALGO_B_MIXED, cells, t_dev,
                             traffic=measured_traffic("k_sw_step32_tiled:batch:%s:%d:%d" % tuple(tune['config']),
                                                      cells)),
        "error_vs_fp64": {"instances_checked": len(pick),
                          "max_rel_L1": {nm: max(x[nm] for x in errs) for nm in ('h', 'hu', 'hv')},
                          "median_rel_L1": {nm: float(np.median([x[nm] for x in errs])) for nm in ('h', 'hu', 'hv')}},
        "max_rel_L1": max(max(x.values()) for x in errs),
        "mass_totals_finite": bool(np.all(np.isfinite(sums))), "build_s": round(t_build, 1),
    }


def run_serving(n_req, seed=1, nsteps=20, n_iter=6, prep_threads=8, n_capacity=64):
    """Implicit advection-diffusion requests (meshes log-uniform in [10k, 1M] cells, D
    log-uniform in [1e-3, 1e-1], 20 steps each, FP64 refinement to 1e-10) served with
    continuous batching. Each request carries its raw mesh (elements, nodes); the engine builds
    its Domain and device tables on host threads from its arrival on, so latency counts that
    preprocessing. Capacity from the first ``n_capacity`` requests back to back, then all the
    requests with Poisson arrivals at 80 % of it; p50 / p99 latency from arrival to result."""
    import torch
    from manapy_amd.meshgen import jittered_triangles
    from manapy_amd.serving import AdvDiffRequest, ServingEngine, latency_stats
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    specs = []
    for k in range(n_req):
        if k % 64 == 0:
            log("serving: generating request mesh %d / %d" % (k, n_req))
        cells = np.exp(rng.uniform(np.log(1e4), np.log(1e6)))
        n = max(8, int(round(np.sqrt(cells / 2.))))
        el, nd = jittered_triangles(n, n, 1.0, 1.0, 0.1, int(rng.integers(1 << 30)))
        x0, y0 = rng.uniform(.2, .5, 2)
        ang = rng.uniform(0, 2 * np.pi)
        D = float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1))))
        specs.append(((el, nd), (x0, y0), (np.cos(ang), np.sin(ang)), D))
    t_gen = time.perf_counter() - t0
    total_cells = sum(len(sp[0][0]) for sp in specs)

    def blob(x0, y0):
        return lambda c: np.exp(-((c[:, 0] - x0) ** 2 + (c[:, 1] - y0) ** 2) / .01)

    def make(sel, arrivals):
        return [AdvDiffRequest(k, specs[k][0], blob(*specs[k][1]), velocity=specs[k][2], D=specs[k][3],
                               nsteps=nsteps, arrival=a) for k, a in zip(sel, arrivals)]

    budget = int(3.0e9)
    kw = dict(max_cells=max(1 << 22, 4 * 1000000), max_inst=64, hbm_budget=budget, n_iter=n_iter,
              prep_threads=prep_threads)
    ncap = min(n_capacity, n_req)
    eng = ServingEngine(**kw)
    log("serving: capacity run (%d requests back to back)" % ncap)
    t1 = time.perf_counter()
    eng.run(make(range(ncap), np.zeros(ncap)))
    makespan = time.perf_counter() - t1
    capacity = ncap / makespan
    rate = 0.8 * capacity
    arrivals = np.cumsum(rng.exponential(1. / rate, n_req))
    log("serving: Poisson run (%d requests at %.2f req/s)" % (n_req, rate))
    eng = ServingEngine(time_ticks=True, **kw)
    t2 = time.perf_counter()
    done = eng.run(make(range(n_req), arrivals))
    span = time.perf_counter() - t2
    lat = latency_stats(done)
    ok = [r for r in done if r.failed is False]
    rows_iters = float(sum(r.C * r.iterations for r in ok))
    t_ticks = eng.tick_seconds()
    ach = ALGO_B_KRYLOV32 * rows_iters / t_ticks / 1e9
    return {
        "requests": n_req, "completed": len(ok), "failed": len(done) - len(ok), "steps_per_request": nsteps,
        "cells_total": total_cells, "capacity_req_per_s": round(capacity, 3),
        "capacity_sample": "first %d requests back to back" % ncap,
        "arrival_rate_req_per_s": round(rate, 3), "throughput_req_per_s": round(n_req / span, 3),
        "throughput_Mcell_steps_per_s": round(total_cells * nsteps / span / 1e6, 3),
        "latency_p50_s": round(lat['p50_s'], 4), "latency_p99_s": round(lat['p99_s'], 4),
        "latency_mean_s": round(lat['mean_s'], 4), "latency_max_s": round(lat['max_s'], 4),
        "latency_includes": "mesh preprocessing (Domain + device tables) from arrival, admission wait, solve",
        "prepare_p50_s": round(lat['prepare_p50_s'], 4), "prepare_p99_s": round(lat['prepare_p99_s'], 4),
        "prep_threads": prep_threads, "max_concurrent": eng.max_active, "hbm_budget_bytes": budget,
        "krylov_iterations_mean": float(np.mean([r.iterations for r in ok])), "solve_tol": 1e-10,
        "request_generation_s": round(t_gen, 1), "solver_tick": "HIP graph replay",
        "bicgstab_iterations_per_tick": n_iter,
        "roofline": {"bound": "hbm", "kernel": "BiCGStab tick (ELL SpMV + vector kernels), all requests",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": krylov_traffic(),
                     "traffic_unit": "HBM bytes per row-iteration (rocprof counters of the serving workload)",
                     "bytes_per_row_iteration": ALGO_B_KRYLOV32,
                     "tick_device_s": round(t_ticks, 3), "row_iterations": rows_iters},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--nx', type=int, default=1000)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-block', type=int, default=4, help='steps per timed block of the CPU baseline')
    ap.add_argument('--cpu-repeats', type=int, default=5, help='timed blocks of the CPU baseline (median)')
    ap.add_argument('--path', choices=('auto', 'fused', 'packed', 'unpacked'), default='auto',
                    help='float64 path: fused (one tiled launch per step, records in LDS), packed (per-cell records '
                         '+ per-face MUSCL values, two kernels), unpacked (SoA kernels), auto (the faster of fused '
                         'and packed on this mesh)')
    ap.add_argument('--no-mixed', action='store_true', help='skip the mixed-precision run of the same case')
    ap.add_argument('--ensemble', type=int, default=256, help='instances of the ensemble run (0: skip)')
    ap.add_argument('--ensemble-steps', type=int, default=204)
    ap.add_argument('--serving', type=int, default=256, help='requests of the serving run (0: skip)')
    ap.add_argument('--serving-iters', type=int, default=6, help='BiCGStab iterations per solver tick')
    ap.add_argument('--advection', type=int, default=200,
                    help='n of the n x n x 2 triangle advection run (0: skip)')
    ap.add_argument('--advection-steps', type=int, default=100)
    ap.add_argument('--euler', type=int, default=100, help='n of the n^3 x 6 tetrahedra Euler run (0: skip)')
    ap.add_argument('--euler-steps', type=int, default=200)
    ap.add_argument('--euler-renumber', choices=('hilbert', 'morton', 'none'), default='hilbert')
    ap.add_argument('--tile', type=int, choices=(128, 256, 512), default=256, help='cells per tile of the fused step')
    ap.add_argument('--mixed-kernel', choices=('auto', 'split', 'tiled', 'staged'), default='auto',
                    help='mixed-precision step: autotuned (default), two kernels, fused, fused with LDS-staged states')
    ap.add_argument('--renumber', choices=('hilbert', 'none'), default='hilbert',
                    help='device cell order: Hilbert curve of the cell centres (default) or the mesh numbering')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='only run a device-to-device copy of 1 GiB five times (calibrates HBM counters)')
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 2 * src.numel() * 8, "read_bytes": src.numel() * 8}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one GPU = one subdomain); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from manapy_amd import _lib
    from manapy_amd.device import DeviceMesh, stream_ptr

    t0 = time.perf_counter()
    d0, Z0, h00 = build_case(args.nx)
    t_mesh = time.perf_counter() - t0
    if args.renumber == 'hilbert':
        from manapy_amd.renumber import renumbered
        d = renumbered(d0)
        Z, h0 = d.to_new(Z0), d.to_new(h00)
    else:
        d, Z, h0 = d0, Z0, h00
    C, F = int(d.nbcells), int(d.nbfaces)
    m = DeviceMesh(d)
    s = stream_ptr()
    dv = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    S = [dv(np.stack([h0, np.zeros(C), np.zeros(C)])), dv(np.zeros((3, C)))]
    tZ = dv(Z)
    zghost = dv(np.zeros(F))
    zgrad = dv(np.zeros((3, C)))
    bf = torch.as_tensor(d.boundaryfaces, device='cuda')
    _lib.call('mfv_ghost_value', m.ptr, 0, tZ.data_ptr(), zghost.data_ptr(), bf.data_ptr(), len(bf), s)
    _lib.call('mfv_cell_gradient', m.ptr, tZ.data_ptr(), zghost.data_ptr(), zgrad[0].data_ptr(),
              zgrad[1].data_ptr(), zgrad[2].data_ptr(), s)
    grad = dv(np.zeros((9, C)))
    dt0 = dv(np.zeros(1))
    ring = dv(np.zeros(4))
    _lib.call('mfv_time_step_sw', m.ptr, S[0][0].data_ptr(), S[0][1].data_ptr(), S[0][2].data_ptr(), CFL, G, 0.,
              0., dt0.data_ptr(), None, s)
    _lib.call('mfv_dt_ring_init', ring.data_ptr(), dt0.data_ptr(), s)

    path, fused_T, path_ms = args.path, None, None
    if path in ('auto', 'fused', 'packed'):
        from manapy_amd.device import SwPacked
        pk = SwPacked(m)
        pk.pack(m, S[0][0], S[0][1], S[0][2], tZ, zgrad, 2, s)
        if path in ('auto', 'fused'):
            cands = [('fused', T) for T in (128, 256) if pk.lds_bytes(T) <= 64 * 1024]
            cands += [('packed', None)] if path == 'auto' else []
            path_ms = {}
            ring0 = ring.clone()
            for kind, T in cands:  # a few steps of each, then the state and ring are restored
                ring.copy_(ring0)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                for k in range(13):
                    if k == 3:
                        e0.record()
                    if kind == 'fused':
                        pk.step_tiled(m, G, CFL, 2, ring, k, T=T, stream=s)
                    else:
                        _lib.call('mfv_sw_recon', m.ptr, pk.ptr, 2, s)
                        _lib.call('mfv_sw_flux_packed', m.ptr, pk.ptr, G, CFL, ring.data_ptr(), k, s)
                e1.record()
                e1.synchronize()
                path_ms["%s:%s" % (kind, T)] = e0.elapsed_time(e1) / 10
                pk.pack(m, S[0][0], S[0][1], S[0][2], tZ, zgrad, 2, s)
            ring.copy_(ring0)
            best = min(path_ms, key=path_ms.get)
            path = best.split(':')[0]
            fused_T = int(best.split(':')[1]) if path == 'fused' else None
            log("fp64: path %s %s" % (best, {k: round(v, 4) for k, v in path_ms.items()}))
        if path == 'fused':
            names = ('k_sw_step_packed_tiled', 'k_dt_finish')

            def launch_a(k):
                pk.step_tiled(m, G, CFL, 2, ring, k, T=fused_T, stream=s)

            def launch_b(k):
                pass
        else:
            names = ('k_sw_recon', 'k_sw_flux_packed')

            def launch_a(k):
                _lib.call('mfv_sw_recon', m.ptr, pk.ptr, 2, s)

            def launch_b(k):
                _lib.call('mfv_sw_flux_packed', m.ptr, pk.ptr, G, CFL, ring.data_ptr(), k, s)
    else:
        names = ('k_sw_gradients', 'k_sw_step')

        def launch_a(k):
            a = S[k % 2]
            _lib.call('mfv_sw_gradients', m.ptr, a[0].data_ptr(), a[1].data_ptr(), a[2].data_ptr(), tZ.data_ptr(),
                      grad.data_ptr(), s)

        def launch_b(k):
            a, b = S[k % 2], S[(k + 1) % 2]
            _lib.call('mfv_sw_step_fused', m.ptr, a.data_ptr(), b.data_ptr(), tZ.data_ptr(), grad.data_ptr(),
                      zgrad.data_ptr(), G, CFL, 2, ring.data_ptr(), k, s)

    def step(k):
        launch_a(k)
        launch_b(k)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()

    # timed loop: no events inside, host-side wall clock bracketed by synchronisations
    t1 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k)
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t1
    # the same number of steps again between two HIP events on the launch stream: the device
    # time per step that the roofline divides by
    t_dev = region_step_time(step, args.warmup + args.steps, args.steps)

    K_avg = float(d.cells.cellnid[:, -1].mean())
    G_avg = float(d.cells.ghostnid[:, -1].mean())
    MF = int(d.cells.faceid.shape[1] - 1)
    if path == 'fused':
        bmodel = fused_bytes_per_cell(MF, K_avg, G_avg)
        trace_key = names[0] + ":%d" % fused_T
        kname = names[0] + " + k_dt_finish"
    elif path == 'packed':
        bmodel = sum(kernel_bytes_per_cell_packed(MF, K_avg, G_avg))
        trace_key, kname = None, " + ".join(names)
    else:
        bmodel = sum(kernel_bytes_per_cell(F, C, K_avg, G_avg))
        trace_key, kname = None, " + ".join(names)
    mcell = C * args.steps / t_loop / 1e6

    out = {
        "metric": METRIC,
        "value": round(mcell, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_loop / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh, dam break over seeded Gaussian bumps)",
        "config": {"workload": "shallow-water dam break, 2D, MUSCL + LSQ + Barth-Jespersen + Rusanov, walls",
                   "cells": C, "faces": F, "nx": args.nx, "cfl": CFL, "order": 2, "precision": "fp64 plain",
                   "path": path, "tile": fused_T, "cell_order": args.renumber},
        "roofline": roofline(kname, ALGO_B_SW64, C, t_dev, bmodel,
                             measured_traffic(trace_key, C) if trace_key else None),
        "fp64_path_ms_per_step": path_ms,
        "mesh_build_s": round(t_mesh, 2),
    }
    log("fp64: %.1f Mcell-updates/s" % mcell)
    bw = copy_bandwidth()
    out["hbm_copy_GBs"] = round(bw, 1)
    out["roofline"]["frac_of_copy"] = round(out["roofline"]["achieved"] / bw, 4)
    log("device copy: %.1f GB/s (the headline step reaches %.3f of it)" % (bw, out["roofline"]["frac_of_copy"]))
    if not args.no_cpu:
        v, rates, n_cpu, cpu_state = cpu_baseline(d0, Z0, h00, args.cpu_block, args.cpu_repeats)
        out["cpu_baseline"] = {"value": round(v, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
                               "sample": "median of %d blocks of %d steps of the same 2M-cell case (after one "
                                         "warm-up step), single-threaded C restatement of the reference loops"
                                         % (args.cpu_repeats, args.cpu_block),
                               "block_rates": [round(x, 3) for x in rates]}
        omp = cpu_baseline_omp(d0, Z0, h00, args.cpu_block, args.cpu_repeats)
        if omp is not None:
            out["cpu_baseline_all_cores"] = omp
        # full-size parity: the float64 GPU path over the same number of steps as the oracle took
        got = fp64_sw_run(d, m, Z, h0, zgrad, (n_cpu,), T=fused_T or 256)[n_cpu]
        if hasattr(d, 'to_old'):
            got = [d.to_old(x) for x in got]
        diff = max(float(np.abs(a - b).max()) for a, b in zip(got, cpu_state))
        out["parity"] = {"steps": n_cpu, "bitwise_equal_to_oracle": bool(all(np.array_equal(a, b)
                                                                              for a, b in zip(got, cpu_state))),
                         "max_abs_diff": diff}
        log("fp64 parity after %d steps: %s" % (n_cpu, out["parity"]["bitwise_equal_to_oracle"]))
    if not args.no_mixed:
        out["mixed"] = run_mixed(d, m, Z, h0, zgrad, args.warmup, args.steps, mcell, args.mixed_kernel, args.tile)
        log("mixed (%s): %.1f Mcell-updates/s" % (out["mixed"]["kernel"], out["mixed"]["value"]))
        out["mixed_accuracy"] = mixed_accuracy(d, m, Z, h0, zgrad)
        log("mixed accuracy: max rel L1 %.3g over %s steps" % (out["mixed_accuracy"]["max_rel_L1"],
                                                              out["mixed_accuracy"]["steps"]))
        out["mixed_split"] = run_mixed(d, m, Z, h0, zgrad, args.warmup, args.steps, mcell, 'split')
        log("mixed (two kernels): %.1f Mcell-updates/s" % out["mixed_split"]["value"])
    if args.ensemble:
        out["ensemble"] = run_ensemble(args.ensemble, args.ensemble_steps)
        log("ensemble: %.1f Mcell-updates/s" % out["ensemble"]["value"])
    if args.serving:
        out["serving"] = run_serving(args.serving, n_iter=args.serving_iters)
        log("serving: p50 %.3f s, p99 %.3f s" % (out["serving"]["latency_p50_s"], out["serving"]["latency_p99_s"]))
    if args.advection:
        out["advection"] = run_advection(args.advection, args.advection_steps, 10, not args.no_cpu)
        log("advection: %.1f Mcell-updates/s" % out["advection"]["value"])
    if args.euler:
        out["euler"] = run_euler(args.euler, args.euler_steps, 5, not args.no_cpu, renumber=args.euler_renumber)
        log("euler: %.1f Mcell-updates/s" % out["euler"]["value"])
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
