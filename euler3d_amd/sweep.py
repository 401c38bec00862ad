"""Batched sweep of bump-channel instances of uneven size, advanced together.

The sweep: ``n`` instances, instance i drawn from ``numpy.random.default_rng(seed0 + i)``:
N_x uniform in 32..128, N_y = N_z = N_x // 3, six tets per hexahedron, bump
thickness h ~ U[0.04, 0.12], Mach ~ U[0.3, 0.8], inlet angle ~ U[-2, 6] deg;
second-order Roe with Green-Gauss gradients and the Venkatakrishnan limiter
(K = 5), 5-stage RK, CFL 1.5, FP64. Every instance runs until its density
residual has dropped ``criterion`` decades (its own stop test) or it reaches
``max_iter``; all of them advance in one batch (``Context.upload_batch`` /
``Context.batch_run``: every RK stage is one launch over all instances, the
iteration replayed as a hipGraph, converged instances skipped).

Every mesh is numbered by the generator in bricks of 4 hexahedra (fixed: the
numbering sets the summation order, so the results do not depend on any
timing). The device memory order, the tile order and the tile shape (cells
per workgroup), which move addresses only, are chosen by the autotuner
(``euler3d_amd.autotune``) per size class and then by a vote weighted by cells.

usage: python -m euler3d_amd.sweep [--instances 64] [--max-iter 1000] [--criterion -8]
                                   [--scale 1.0] [--no-graph] [--no-autotune] [--json out.json]
``--scale`` multiplies N_x (and so N_y, N_z) to run a smaller sweep of the same shape.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import FLUX_ROE, GRAD_GREEN_GAUSS, LIM_VENKATAKRISHNAN, Context, Mesh, make_params

SCHEME = dict(flux=FLUX_ROE, order=2, gradient=GRAD_GREEN_GAUSS, limiter=LIM_VENKATAKRISHNAN, k=5.0, stages=5)


def instances(n=64, seed0=1000, scale=1.0):
    """The sweep's instances: mesh size, bump height and flow conditions, seeded per instance."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        nx = int(rng.integers(32, 129))
        h = float(rng.uniform(0.04, 0.12))
        mach = float(rng.uniform(0.3, 0.8))
        aoa = float(rng.uniform(-2.0, 6.0))
        nx = max(6, int(round(nx * scale)))
        out.append(dict(seed=seed0 + i, nx=nx, ny=max(2, nx // 3), nz=max(2, nx // 3), h=h, mach=mach, aoa_deg=aoa))
    return out


def run(n=64, max_iter=1000, criterion=-8.0, scale=1.0, use_graph=True, autotune=True, seed0=1000, log=None):
    from . import autotune as at
    cases = instances(n, seed0, scale)
    tile_votes, mem_votes, cell_votes = {}, {}, {}
    t0 = time.perf_counter()
    for c in cases:
        sc = at.size_class(c["nx"])
        order, mem, cells = at.tune(sc, sc // 3, sc // 3, SCHEME, verbose=bool(log)) if autotune else (-1, 0, 256)
        c["brick"] = at.DEFAULT_BRICK
        c["cells"] = 6 * c["nx"] * c["ny"] * c["nz"]
        c["tuned"] = dict(tile_order=order, memory_order=mem, tile_cells=cells)
        tile_votes[order] = tile_votes.get(order, 0.0) + c["cells"]
        mem_votes[mem] = mem_votes.get(mem, 0.0) + c["cells"]
        cell_votes[cells] = cell_votes.get(cells, 0.0) + c["cells"]
    t_tune = time.perf_counter() - t0
    # the mesh builder runs outside the interpreter lock: build the instances in parallel
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=8) as pool:
        meshes = list(pool.map(lambda c: Mesh.bump(c["nx"], c["ny"], c["nz"], c["h"], c["brick"], (1, 1, 1)), cases))
    t_mesh = time.perf_counter() - t0 - t_tune
    # one device memory order, tile order and tile shape for the whole batch, by cells
    tile_order = max(tile_votes, key=tile_votes.get)
    memory_order = max(mem_votes, key=mem_votes.get)
    tile_cells = max(cell_votes, key=cell_votes.get)
    ctx = Context(0)
    ctx.set_params(make_params(0.5, 0.0, max_iter=max_iter, criterion=criterion, **SCHEME))
    ctx.set_ordering(memory_order)
    ctx.upload_batch(meshes, [dict(mach=c["mach"], aoa_deg=c["aoa_deg"], cmac=1.0, criterion=criterion,
                                   max_iter=max_iter) for c in cases])
    ctx.set_tile_order(tile_order)
    ctx.set_tile_cells(tile_cells)
    ctx.init_freestream()
    ctx.op("sync")
    if log:
        print(f"[sweep] {n} instances, {sum(c['cells'] for c in cases)} cells, tuned in {t_tune:.1f}s, meshes built "
              f"in {t_mesh:.1f}s, tile order {tile_order}, memory order {memory_order}, tile {tile_cells} cells",
              file=log, flush=True)
    t1 = time.perf_counter()
    done, status = ctx.batch_run(max_iter, use_graph=use_graph)
    wall = time.perf_counter() - t1
    cell_stages = 0
    for c, s in zip(cases, status):
        # a non-finite residual norm stops an instance (the reference would run on to max_iter)
        diverged = not all(math.isfinite(x) for x in s["last"]["rms"])
        c.update(iterations=s["iterations"], stopped=s["stopped"], stop_seconds=s["stop_seconds"],
                 drop=s["last"]["drop"][0], cl=s["last"]["cl"], cd=s["last"]["cd"], diverged=diverged,
                 reached_criterion=bool(s["last"]["drop"][0] <= criterion))
        cell_stages += c["cells"] * SCHEME["stages"] * s["iterations"]
    ctx.close()
    for m in meshes:
        m.close()
    return {"instances": cases, "wall_seconds": wall, "iterations_launched": done,
            "cells": sum(c["cells"] for c in cases), "mcell_rk_stage_per_s": cell_stages / wall / 1e6,
            "reached_criterion": sum(c["reached_criterion"] for c in cases),
            "diverged": [c["seed"] for c in cases if c["diverged"]], "tile_order": tile_order,
            "memory_order": memory_order, "tile_cells": tile_cells,
            "mesh_seconds": t_mesh, "tune_seconds": t_tune, "max_iter": max_iter, "criterion": criterion, "graph": use_graph}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--criterion", type=float, default=-8.0)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--json")
    a = ap.parse_args(argv)
    r = run(a.instances, a.max_iter, a.criterion, a.scale, not a.no_graph, not a.no_autotune, a.seed0, log=sys.stderr)
    print(f"[sweep] {r['cells']} cells in {len(r['instances'])} instances: {r['iterations_launched']} iterations "
          f"in {r['wall_seconds']:.2f}s = {r['mcell_rk_stage_per_s']:.0f} Mcell-RK-stage/s; "
          f"{r['reached_criterion']} reached a {-a.criterion:g}-decade drop; diverged (non-finite residual): "
          f"seeds {r['diverged']}", file=sys.stderr, flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(r, f, indent=1)
    return r


if __name__ == "__main__":
    main()
