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


def cpu_baseline(d, Z, h0, steps, budget_s):
    """The oracle (single-threaded C restatement of the reference loop structure) on the same
    mesh and initial state, timed over a bounded number of steps."""
    from oracle import oracle
    C = int(d.nbcells)
    h, hu, hv = h0.copy(), np.zeros(C), np.zeros(C)
    zg = oracle.z_gradient(d, Z)
    oracle.sw_step(d, h, hu, hv, Z, zg, G, CFL, 2)  # warm-up
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        oracle.sw_step(d, h, hu, hv, Z, zg, G, CFL, 2)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    t = time.perf_counter() - t0
    return C * done / t / 1e6, done, t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--nx', type=int, default=1000)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-steps', type=int, default=20)
    ap.add_argument('--cpu-budget', type=float, default=20.0)
    ap.add_argument('--path', choices=('packed', 'unpacked'), default='packed',
                    help='packed: per-cell records + per-face MUSCL values (default); unpacked: SoA kernels')
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
    d, Z, h0 = build_case(args.nx)
    t_mesh = time.perf_counter() - t0
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

    step_no = [0]
    if args.path == 'packed':
        from manapy_amd.device import SwPacked
        pk = SwPacked(m)
        _lib.call('mfv_sw_pack', m.ptr, pk.ptr, S[0][0].data_ptr(), S[0][1].data_ptr(), S[0][2].data_ptr(),
                  tZ.data_ptr(), zgrad.data_ptr(), 2, s)
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

    def step(ev=None):
        k = step_no[0]
        if ev is not None:
            ev[0].record()
        launch_a(k)
        if ev is not None:
            ev[1].record()
        launch_b(k)
        if ev is not None:
            ev[2].record()
        step_no[0] = k + 1

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed loop: no events inside, host-side wall clock bracketed by synchronisations
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t1

    # per-kernel durations: HIP events on the launch stream over a second timed pass
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    for e in evs:
        step(e)
    torch.cuda.synchronize()
    t_grad = np.array([a.elapsed_time(b) for a, b, _ in evs]) * 1e-3
    t_flux = np.array([b.elapsed_time(c) for _, b, c in evs]) * 1e-3

    K_avg = float(d.cells.cellnid[:, -1].mean())
    G_avg = float(d.cells.ghostnid[:, -1].mean())
    if args.path == 'packed':
        bg, bfx = kernel_bytes_per_cell_packed(int(d.cells.faceid.shape[1] - 1), K_avg, G_avg)
    else:
        bg, bfx = kernel_bytes_per_cell(F, C, K_avg, G_avg)
    ach_flux = bfx * C / t_flux.mean() / 1e9
    ach_grad = bg * C / t_grad.mean() / 1e9
    dom_is_flux = t_flux.mean() >= t_grad.mean()
    ach, bpc, tk = (ach_flux, bfx, t_flux) if dom_is_flux else (ach_grad, bg, t_grad)
    mcell = C * args.steps / t_loop / 1e6
    step_bytes_survey = 356.0  # FP64 B per cell-update of the whole C2 step (algorithmic model)

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
                   "path": args.path},
        "roofline": {"bound": "hbm", "kernel": names[1] if dom_is_flux else names[0],
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell": round(bpc, 1), "avg_launch_us": round(float(tk.mean()) * 1e6, 2)},
        "kernels": {names[0]: {"avg_us": round(float(t_grad.mean()) * 1e6, 2),
                               "bytes_per_cell": round(bg, 1), "GB/s": round(ach_grad, 1)},
                    names[1]: {"avg_us": round(float(t_flux.mean()) * 1e6, 2),
                               "bytes_per_cell": round(bfx, 1), "GB/s": round(ach_flux, 1)}},
        "step_hbm_frac_356B": round(step_bytes_survey * C * args.steps / t_loop / 1e9 / HBM_PEAK_GBS, 4),
        "mesh_build_s": round(t_mesh, 2),
    }
    if not args.no_cpu:
        v, n, t = cpu_baseline(d, Z, h0, args.cpu_steps, args.cpu_budget)
        out["cpu_baseline"] = {"value": round(v, 3), "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
                               "sample": "%d steps of the same 2M-cell case (%.1f s), single-threaded C "
                                         "restatement of the reference loops" % (n, t)}
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
