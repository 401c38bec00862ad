"""Autotuner for the stage kernels' memory order.

Three knobs change how the fused stage kernels walk memory:

* the cell numbering of the generated bump-channel mesh: hexahedra grouped in
  bricks of ``brick**3`` (0 = plain lexicographic), which sets how far apart
  in memory the neighbours gathered by a cell are. This is a different
  numbering of the same mesh, so the results are the reference's on that
  numbering (the same physics, that numbering's summation order);
* the device memory order of cells and faces (``e3d_set_ordering``: the
  file's numbering, reverse Cuthill-McKee or a Morton curve), which changes
  addresses only: every result bit stays the same;
* the workgroup-to-tile order: tile i on workgroup i, one contiguous run of
  tiles per XCD (each XCD has its own L2), or runs of 16 tiles dealt to the
  XCDs in turn; addresses only.

``tune(nx, ny, nz, scheme)`` times every combination on the GPU (a few RK
iterations through ``Context.iterate``) and keeps the fastest; results are
cached in a JSON file keyed by mesh size and scheme, so a batch of many
instances tunes once per size class.
"""
from __future__ import annotations

import json
import os
import time

from . import (FLUX_ROE, GRAD_GREEN_GAUSS, LIM_VENKATAKRISHNAN, ORDER_MORTON, ORDER_RCM, ORDER_REFERENCE,
               TILES_LINEAR, TILES_XCD_CHUNKS, TILES_XCD_CONTIGUOUS, Context, Mesh, lib, make_params)

BRICKS = (0, 2, 4, 8)
TILE_ORDERS = (TILES_XCD_CHUNKS, TILES_XCD_CONTIGUOUS, TILES_LINEAR)
MEMORY_ORDERS = (ORDER_REFERENCE, ORDER_MORTON, ORDER_RCM)
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def size_class(nx):
    """Instances are tuned per class of streamwise size: 32-63, 64-95, 96-127, 128+."""
    return max(32, min(128, (nx // 32) * 32))


def _key(nx, ny, nz, scheme):
    return f"{nx}x{ny}x{nz}/f{scheme['flux']}o{scheme['order']}g{scheme['gradient']}l{scheme['limiter']}"


def measure(mesh, tile_orders, memory_order, scheme, iters=6, warmup=2):
    """Milliseconds per RK iteration of one mesh under one memory order, for each tile order."""
    ctx = Context(0)
    ctx.set_params(make_params(0.6, 1.0, max_iter=10 ** 9, criterion=-1e9, **scheme))
    ctx.set_ordering(memory_order)
    ctx.upload(mesh)
    out = []
    for tile_order in tile_orders:
        ctx.set_tile_order(tile_order)
        ctx.init_freestream()
        ctx.iterate(warmup, history=False)
        ctx.op("sync")
        t0 = time.perf_counter()
        ctx.iterate(iters, history=False)
        ctx.op("sync")
        out.append((time.perf_counter() - t0) * 1e3 / iters)
    ctx.close()
    return out


def tune(nx, ny, nz, scheme=None, cache=DEFAULT_CACHE, verbose=False):
    """Best (brick, tile_order, memory_order) for this mesh size and scheme, measured once and cached."""
    scheme = scheme or dict(flux=FLUX_ROE, order=2, gradient=GRAD_GREEN_GAUSS, limiter=LIM_VENKATAKRISHNAN, k=5.0,
                            stages=5)
    key = _key(nx, ny, nz, scheme)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table and "memory_order" in table[key]:
        t = table[key]
        return t["brick"], t["tile_order"], t["memory_order"]
    lib()
    trials = []
    for brick in BRICKS:
        mesh = Mesh.bump(nx, ny, nz, 0.1, brick, (1, 1, 1))
        for mem in MEMORY_ORDERS:
            for order, ms in zip(TILE_ORDERS, measure(mesh, TILE_ORDERS, mem, scheme)):
                trials.append({"brick": brick, "tile_order": order, "memory_order": mem, "ms_per_iteration": ms})
                if verbose:
                    print(f"[autotune] {key} brick {brick} memory order {mem} tile order {order}: {ms:.3f} "
                          f"ms/iteration", flush=True)
        mesh.close()
    best = min(trials, key=lambda t: t["ms_per_iteration"])
    table[key] = dict(best, trials=trials)
    if cache:
        try:
            with open(cache, "w") as f:
                json.dump(table, f, indent=1)
        except OSError:
            pass
    return best["brick"], best["tile_order"], best["memory_order"]
