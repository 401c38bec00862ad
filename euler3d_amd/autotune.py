"""Autotuner for the stage kernels' tile shape and memory order.

Three knobs change how the fused stage kernels walk memory, and none of them
changes a result bit (every cell keeps its faces in the reference's order and
the residual-norm partial sums are formed per 64 cells whatever the tiling):

* the device memory order of cells and faces (``e3d_set_ordering``: the
  file's numbering, reverse Cuthill-McKee or a Morton curve);
* the workgroup-to-tile order: tile i on workgroup i, one contiguous run of
  tiles per XCD (each XCD has its own L2), or runs of 4096 cells dealt to the
  XCDs in turn (``e3d_set_tile_order``);
* the tile shape: 64, 128 or 256 cells per workgroup (``e3d_set_tile_cells``).

``tune(nx, ny, nz, scheme)`` times every combination on the GPU (a few RK
iterations through ``Context.iterate``) on the mesh as numbered by its
generator (bricks of ``brick**3`` hexahedra, 4 by default) and keeps the
fastest; results are cached in a JSON file keyed by mesh size, numbering and
scheme, so a batch of many instances tunes once per size class. Because the
knobs are addresses only, two tuning passes that pick differently still give
the same results bit for bit.

The cell numbering itself (the brick size) is another numbering of the same
mesh: it changes the summation order and so the result bits. It is searched
only when asked (``numberings=``), never by default.
"""
from __future__ import annotations

import json
import os
import time

from . import (FLUX_ROE, GRAD_GREEN_GAUSS, LIM_VENKATAKRISHNAN, ORDER_MORTON, ORDER_RCM, ORDER_REFERENCE,
               TILES_LINEAR, TILES_XCD_CHUNKS, TILES_XCD_CONTIGUOUS, Context, Mesh, lib, make_params)

TILE_ORDERS = (TILES_XCD_CHUNKS, TILES_XCD_CONTIGUOUS, TILES_LINEAR)
MEMORY_ORDERS = (ORDER_REFERENCE, ORDER_MORTON, ORDER_RCM)
TILE_CELLS = (256, 128, 64)
DEFAULT_BRICK = 4
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def size_class(nx):
    """Instances are tuned per class of streamwise size: 32-63, 64-95, 96-127, 128+."""
    return max(32, min(128, (nx // 32) * 32))


def _key(nx, ny, nz, scheme, brick):
    return (f"{nx}x{ny}x{nz}/b{brick}/f{scheme['flux']}o{scheme['order']}g{scheme['gradient']}"
            f"l{scheme['limiter']}")


def measure(mesh, memory_order, scheme, iters=6, warmup=2):
    """Milliseconds per RK iteration of one mesh under one memory order, for every
    (tile order, tile cells) pair: {(order, cells): ms}."""
    ctx = Context(0)
    ctx.set_params(make_params(0.6, 1.0, max_iter=10 ** 9, criterion=-1e9, **scheme))
    ctx.set_ordering(memory_order)
    ctx.upload(mesh)
    out = {}
    for cells in TILE_CELLS:
        ctx.set_tile_cells(cells)
        for tile_order in TILE_ORDERS:
            ctx.set_tile_order(tile_order)
            ctx.init_freestream()
            ctx.iterate(warmup, history=False)
            ctx.op("sync")
            t0 = time.perf_counter()
            ctx.iterate(iters, history=False)
            ctx.op("sync")
            out[(tile_order, cells)] = (time.perf_counter() - t0) * 1e3 / iters
    ctx.close()
    return out


def tune(nx, ny, nz, scheme=None, cache=DEFAULT_CACHE, verbose=False, brick=DEFAULT_BRICK, numberings=None):
    """Best (tile_order, memory_order, tile_cells) for this mesh size, numbering and
    scheme, measured once and cached. numberings: brick sizes to search as well
    (changes result bits); then the result is (brick, tile_order, memory_order,
    tile_cells)."""
    scheme = scheme or dict(flux=FLUX_ROE, order=2, gradient=GRAD_GREEN_GAUSS, limiter=LIM_VENKATAKRISHNAN, k=5.0,
                            stages=5)
    bricks = tuple(numberings) if numberings else (brick,)
    key = _key(nx, ny, nz, scheme, ",".join(map(str, bricks)))
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table and "tile_cells" in table[key]:
        t = table[key]
    else:
        lib()
        trials = []
        for b in bricks:
            mesh = Mesh.bump(nx, ny, nz, 0.1, b, (1, 1, 1))
            for mem in MEMORY_ORDERS:
                for (order, cells), ms in measure(mesh, mem, scheme).items():
                    trials.append({"brick": b, "tile_order": order, "memory_order": mem, "tile_cells": cells,
                                   "ms_per_iteration": ms})
                    if verbose:
                        print(f"[autotune] {key} brick {b} memory order {mem} tile order {order} tile {cells} "
                              f"cells: {ms:.3f} ms/iteration", flush=True)
            mesh.close()
        best = min(trials, key=lambda t: t["ms_per_iteration"])
        t = table[key] = dict(best, trials=trials)
        if cache:
            try:
                with open(cache, "w") as f:
                    json.dump(table, f, indent=1)
            except OSError:
                pass
    if numberings:
        return t["brick"], t["tile_order"], t["memory_order"], t["tile_cells"]
    return t["tile_order"], t["memory_order"], t["tile_cells"]
