#!/usr/bin/env python3
"""Throughput of the residual-and-update path on one MI355X.

Workload (default): the bump channel at 192x64x64 hexes x 6 = 4,718,592 tets,
split into 8 blocks (2x2x2) whose interfaces are exchanged on the device;
second-order Roe with Green-Gauss gradients and the Venkatakrishnan limiter
(K=5), M=0.675, 5-stage RK, local CFL 1.5, FP64. A step is one RK iteration
(5 stages, the interface exchange and the residual/force monitors) over the
whole mesh, launched through e3d_iterate; the mesh and state are resident in
HBM before the timed region starts.

metric: Mcell-RK-stage/s = real cells x RK stages x steps / wall time.
roofline: the kernel with the largest share of the timed region (interp:
gradient + limiter + face states, or flux_update: residual + RK update + BCs
on a stage with dissipation), algorithmic bytes per launch / average launch
time from HIP events recorded on the context's stream during the timed region;
"rooflines" carries the same figures for both kernels.
cpu_baseline: the oracle (plain-C restatement, one core) on a bounded sample of
the same workload: one of the 8 blocks (589,824 cells) for 4 RK iterations
(about 10 s).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Compulsory bytes per cell and RK stage of the whole second-order Roe/GG/VK
# stage at this workload (K_A + K_B + K_C + per-iteration work / 5, below).
MODEL_BYTES_PER_CELL_STAGE = 1110.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fused_kernels_bump192x64x64_traffic.json")


def algorithmic_bytes(blocks):
    """Algorithmic bytes per launch of the two stage kernels.

    The compulsory-traffic model of this path: every array element counted once
    per kernel that must read or write it (W 40 B/cell, grad 120, lim 40, W0 40,
    conv+diss 80, dt 8, V 8; per face S 24 + rL, rR 48 + left/right 8 = 80 B).
    interp is the gradient + limiter kernel K_A = 208(N+G) + 80F + 16N;
    flux_update does the residual kernel K_B = 200(N+G) + 80F + 120N and the
    update + BC kernel K_C = 176N + 112G of that model in one launch."""
    N = sum(b["nreal"] for b in blocks)
    G = sum(b["nall"] - b["nreal"] for b in blocks)
    F = sum(b["nface"] for b in blocks)
    ka = 208 * (N + G) + 80 * F + 16 * N
    kb = 200 * (N + G) + 80 * F + 120 * N
    kc = 176 * N + 112 * G
    return {"interp": ka, "flux_update": kb + kc}, N, G, F


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--parts", type=int, nargs=3, default=[2, 2, 2])
    ap.add_argument("--cpu-iters", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, one 1 GiB device-to-device copy: the reference kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (blocks would shard "
              "across GPUs by contiguous ranges, but that path is not built)", file=sys.stderr)
        sys.exit(2)

    import euler3d_amd as E

    t_mesh = time.perf_counter()
    mesh = E.Mesh.bump(a.nx, a.ny, a.nz, 0.1, 4, tuple(a.parts))
    blocks = [mesh.block(b) for b in range(mesh.nblocks)]
    t_mesh = time.perf_counter() - t_mesh
    alg, N, G, F = algorithmic_bytes(blocks)
    params = E.make_params(0.675, 0.0, flux=E.FLUX_ROE, order=2, gradient=E.GRAD_GREEN_GAUSS,
                           limiter=E.LIM_VENKATAKRISHNAN, k=5.0, stages=5, max_iter=10 ** 9, criterion=-1e9)
    ctx = E.Context(0)
    ctx.set_params(params)
    ctx.upload(mesh)
    ctx.init_freestream()
    print(f"[bench] mesh {N} cells, {G} ghosts, {F} faces, {mesh.nblocks} blocks (built in {t_mesh:.1f}s); "
          f"warmup {a.warmup}", file=sys.stderr, flush=True)
    if a.warmup:
        ctx.iterate(a.warmup, history=False)
    ctx.op("sync")
    E.lib().e3d_enable_timing(ctx._h, 1)
    t0 = time.perf_counter()
    hist = ctx.iterate(a.steps)
    ctx.op("sync")
    t1 = time.perf_counter()
    wall = t1 - t0
    stages = 5
    value = N * stages * a.steps / wall / 1e6
    kt = {}
    for name in ("interp", "flux_update0", "flux_update", "flux_update_nodiss", "exchange", "monitors"):
        ms, n = ctx.kernel_time(name)
        kt[name] = {"avg_ms": ms, "launches": n}
    measured = {}
    if os.path.exists(TRAFFIC_FILE) and (a.nx, a.ny, a.nz, tuple(a.parts)) == (192, 64, 64, (2, 2, 2)):
        with open(TRAFFIC_FILE) as f:
            measured = json.load(f)["hbm_bytes_per_launch"]
    rooflines = {}
    for name in ("interp", "flux_update"):
        t = kt[name]
        ach = alg[name] / (t["avg_ms"] * 1e-3) / 1e9 if t["avg_ms"] > 0 else 0.0
        rooflines[name] = {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": measured.get(name),
                           "algorithmic_bytes_per_launch": alg[name], "avg_ms": t["avg_ms"],
                           "share_of_step": t["avg_ms"] * t["launches"] / (wall * 1e3)}
    dominant = max(rooflines, key=lambda k: rooflines[k]["share_of_step"])
    rms_ro = hist[-1]["rms"][0] if hist else float("nan")
    print(f"[bench] {a.steps} steps in {wall:.3f}s; rms_ro {rms_ro:.6e}; kernels "
          + ", ".join(f"{k} {v['avg_ms']:.3f}ms x{v['launches']}" for k, v in kt.items()), file=sys.stderr, flush=True)
    ctx.close()
    if a.calib_copy:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library already loaded
        n = ctypes.c_size_t(1 << 30)
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        ok = hip.hipMalloc(ctypes.byref(src), n) == 0 and hip.hipMalloc(ctypes.byref(dst), n) == 0
        ok = ok and hip.hipMemset(src, 1, n) == 0 and hip.hipMemcpy(dst, src, n, 3) == 0
        ok = ok and hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        print(f"[bench] calibration copy of 1 GiB: {'done' if ok else 'FAILED'}", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        import oracle as O
        b0 = blocks[0]
        S = O.Solver([b0], O.params(0.675, 0.0, flux=0, order=2, gradient=0, limiter=1, k=5.0, stages=5,
                                    max_iter=10 ** 9, criterion=-1e9), interfaces=False)
        S.init_freestream()
        c0 = time.perf_counter()
        S.iterate(a.cpu_iters)
        c1 = time.perf_counter()
        cpu = {"value": b0["nreal"] * stages * a.cpu_iters / (c1 - c0) / 1e6, "unit": "Mcell-RK-stage/s",
               "cores": 1, "kind": "port",
               "sample": f"oracle (plain-C restatement, 1 thread) on block 0 of the same mesh "
                         f"({b0['nreal']} cells), {a.cpu_iters} RK iteration(s) of 5 stages, "
                         f"{c1 - c0:.1f}s"}
        print(f"[bench] cpu baseline {cpu['value']:.3f} Mcell-RK-stage/s", file=sys.stderr, flush=True)

    out = {
        "metric": "Mcell-RK-stage/s (FP64, %HBM roof) + time to 8-order residual drop, warm vs cold",
        "value": value, "unit": "Mcell-RK-stage/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"bump channel {a.nx}x{a.ny}x{a.nz}x6 tets ({N} cells), "
                               f"{mesh.nblocks} blocks on one device, 2nd-order Roe + Green-Gauss + "
                               "Venkatakrishnan K=5, M=0.675, 5-stage RK, CFL 1.5",
                   "cells": N, "ghosts": G, "faces": F, "blocks": mesh.nblocks, "stages": stages},
        "hbm_roof_frac_overall": value / (HBM_PEAK_GBS * 1e9 / MODEL_BYTES_PER_CELL_STAGE / 1e6),
        "roofline": rooflines[dominant],
        "rooflines": rooflines,
        "kernels": kt,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
