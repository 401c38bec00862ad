"""Autotuner for the stage kernels' memory order.

Two knobs change how the fused stage kernels walk memory without changing a
single result bit (per-cell sums always run over the cell's faces in the
reference's order):

* the cell ordering of the generated bump-channel mesh: hexahedra grouped in
  bricks of ``brick**3`` (0 = plain lexicographic), which sets how far apart
  in memory the neighbours gathered by a cell are;
* the workgroup-to-tile order: eight contiguous runs of tiles, one per XCD
  (each XCD has its own L2), or tile i on workgroup i.

``tune(nx, ny, nz, scheme)`` times every combination on the GPU (a few RK
iterations through ``Context.iterate`` with HIP-event kernel timing) and keeps
the fastest; results are cached in a JSON file keyed by mesh size and scheme,
so a batch of many instances tunes once per size class.
"""
from __future__ import annotations

import json
import os
import time

from . import (FLUX_ROE, GRAD_GREEN_GAUSS, LIM_VENKATAKRISHNAN, Context, Mesh, lib, make_params)

BRICKS = (0, 2, 4, 8)
TILE_ORDERS = (1, 0)
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def size_class(nx):
    """Instances are tuned per class of streamwise size: 32-63, 64-95, 96-127, 128+."""
    return max(32, min(128, (nx // 32) * 32))


def _key(nx, ny, nz, scheme):
    return f"{nx}x{ny}x{nz}/f{scheme['flux']}o{scheme['order']}g{scheme['gradient']}l{scheme['limiter']}"


def measure(nx, ny, nz, brick, tile_order, scheme, iters=6, warmup=2):
    """Milliseconds per RK iteration of one configuration."""
    mesh = Mesh.bump(nx, ny, nz, 0.1, brick, (1, 1, 1))
    ctx = Context(0)
    ctx.set_params(make_params(0.6, 1.0, max_iter=10 ** 9, criterion=-1e9, **scheme))
    ctx.upload(mesh)
    ctx.set_tile_order(tile_order)
    ctx.init_freestream()
    ctx.iterate(warmup, history=False)
    ctx.op("sync")
    t0 = time.perf_counter()
    ctx.iterate(iters, history=False)
    ctx.op("sync")
    ms = (time.perf_counter() - t0) * 1e3 / iters
    ctx.close()
    mesh.close()
    return ms


def tune(nx, ny, nz, scheme=None, cache=DEFAULT_CACHE, verbose=False):
    """Best (brick, tile_order) for this mesh size and scheme, measured once and cached."""
    scheme = scheme or dict(flux=FLUX_ROE, order=2, gradient=GRAD_GREEN_GAUSS, limiter=LIM_VENKATAKRISHNAN, k=5.0,
                            stages=5)
    key = _key(nx, ny, nz, scheme)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table:
        return table[key]["brick"], table[key]["tile_order"]
    lib()
    trials = []
    for brick in BRICKS:
        for order in TILE_ORDERS:
            ms = measure(nx, ny, nz, brick, order, scheme)
            trials.append({"brick": brick, "tile_order": order, "ms_per_iteration": ms})
            if verbose:
                print(f"[autotune] {key} brick {brick} tile order {order}: {ms:.3f} ms/iteration", flush=True)
    best = min(trials, key=lambda t: t["ms_per_iteration"])
    table[key] = dict(best, trials=trials)
    if cache:
        try:
            with open(cache, "w") as f:
                json.dump(table, f, indent=1)
        except OSError:
            pass
    return best["brick"], best["tile_order"]
