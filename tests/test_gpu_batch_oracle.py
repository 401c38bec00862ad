"""Batched path against the oracle, instance by instance.

* The batched sweep at full size: 64 bump-channel instances of uneven size
  drawn as ``euler3d_amd.sweep.instances`` draws them (instance i seeded with
  1000 + i: N_x ~ U{32..128}, N_y = N_z = N_x // 3, h ~ U[0.04, 0.12],
  M ~ U[0.3, 0.8], inlet angle ~ U[-2, 6] deg; 31.1 M tets together), every
  mesh numbered in bricks of 4 hexahedra (fixed, not the autotuner's choice),
  second-order Roe + Green-Gauss + Venkatakrishnan (K = 5), 5-stage RK. Two
  iterations through e3d_batch_run replaying the hipGraph; every instance's
  state and dissipation must equal the oracle's run of that instance alone
  (values, +0 == -0), its monitor rows within 1e-12 relative (the device sums
  the residual norms per tile, the oracle cell by cell). At this size the
  automatic tile order deals runs of 16 tiles to the XCDs.
* Staggered stops: small instances with their own iteration limits and one
  stop on the drop criterion (PostProcessing.cpp:59-83, per instance), run to
  the end in one batch; each instance's final state, stop iteration and every
  monitor row must match the oracle's run of that instance.

The oracle runs each instance on its own host thread (at most 16)."""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import euler3d_amd as E
import oracle as O
from euler3d_amd.sweep import SCHEME, instances
from helpers import values_equal

pytestmark = pytest.mark.gpu

OPTS = dict(omega=1e-16, **SCHEME)


def _rel_close(a, b, rtol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300)))


def _oracle_run(mesh, c, n, criterion):
    S = O.Solver([mesh.block(b) for b in range(mesh.nblocks)],
                 O.params(c["mach"], c["aoa_deg"], max_iter=c["max_iter"], criterion=criterion, **OPTS))
    S.init_freestream()
    rows = S.iterate(n)
    return [(B.W.copy(), B.diss.reshape(5, B.nall)[:, :B.nreal].reshape(-1).copy()) for B in S.blocks], rows


def _check_instance(ctx, i, oracle_blocks, oracle_rows, device_rows, label):
    for b, (W, d) in zip(ctx.batch_blocks(i), oracle_blocks):
        mine = ctx.get(b, E.ARR_W).reshape(-1)
        same = values_equal(mine, W)
        assert same.all(), f"{label}: {np.sum(~same)} state values differ, first {np.flatnonzero(~same)[:5]}"
        mine = ctx.get(b, E.ARR_DISS).reshape(-1)
        same = values_equal(mine, d)
        assert same.all(), f"{label}: {np.sum(~same)} dissipation values differ"
    assert len(device_rows) == len(oracle_rows), (label, len(device_rows), len(oracle_rows))
    for a, r in zip(device_rows, oracle_rows):
        assert a["iter"] == r["iter"] and a["stop"] == r["stop"], (label, a["iter"], a["stop"], r["stop"])
        assert _rel_close(a["rms"], r["rms"]), (label, a["iter"], a["rms"], r["rms"])
        assert _rel_close([a["cl"], a["cd"], a["cmx"], a["cmy"], a["cmz"]],
                          [r["cl"], r["cd"], r["cmx"], r["cmy"], r["cmz"]]), (label, a["iter"])
        assert np.allclose(a["drop"], r["drop"], rtol=0, atol=1e-12), (label, a["iter"], a["drop"], r["drop"])


def test_c3_sweep_batch_matches_oracle_per_instance():
    t0 = time.perf_counter()
    cases = instances(64, 1000)
    for c in cases:
        c["max_iter"] = 1000
    ncells = sum(6 * c["nx"] * c["ny"] * c["nz"] for c in cases)
    print(f"c3 batch: 64 instances, {ncells} cells; building meshes", flush=True)
    with ThreadPoolExecutor(max_workers=8) as pool:
        meshes = list(pool.map(lambda c: E.Mesh.bump(c["nx"], c["ny"], c["nz"], c["h"], 4, (1, 1, 1)), cases))
    assert ncells > 30_000_000
    ctx = E.Context(0)
    ctx.set_params(E.make_params(0.5, 0.0, **OPTS))
    ctx.upload_batch(meshes, [dict(mach=c["mach"], aoa_deg=c["aoa_deg"], cmac=1.0, criterion=-8.0, max_iter=1000)
                              for c in cases])
    ctx.init_freestream()
    done, status = ctx.batch_run(2, use_graph=True)
    assert done == 2 and all(s["iterations"] == 2 and not s["stopped"] for s in status)
    print(f"c3 batch: device done ({time.perf_counter() - t0:.1f}s); oracle per instance", flush=True)
    threads = min(16, os.cpu_count() or 1)
    order = sorted(range(64), key=lambda i: -6 * cases[i]["nx"] * cases[i]["ny"] * cases[i]["nz"])  # largest first
    checked = 0
    with ThreadPoolExecutor(max_workers=threads) as pool:
        futs = {pool.submit(_oracle_run, meshes[i], cases[i], 2, -8.0): i for i in order}
        for f in futs:
            i = futs[f]
            blocks, rows = f.result()
            _check_instance(ctx, i, blocks, rows, ctx.batch_history(i, 10), f"instance {i} (seed {cases[i]['seed']})")
            checked += 1
            if checked % 16 == 0:
                print(f"c3 batch: {checked} instances equal the oracle ({time.perf_counter() - t0:.1f}s)", flush=True)
    ctx.close()
    for m in meshes:
        m.close()


def test_staggered_stops_match_oracle():
    cases = instances(6, 2000, scale=0.15)
    limits = [5, 9, 3, 12, 7, 40]
    for c, n in zip(cases, limits):
        c["max_iter"] = n
    crit = [-15.0] * 5 + [-0.4]  # the last instance stops on its drop criterion
    meshes = [E.Mesh.bump(c["nx"], c["ny"], c["nz"], c["h"], 2, (2, 1, 1) if i % 2 else (1, 1, 1))
              for i, c in enumerate(cases)]
    ctx = E.Context(0)
    ctx.set_params(E.make_params(0.5, 0.0, **OPTS))
    ctx.upload_batch(meshes, [dict(mach=c["mach"], aoa_deg=c["aoa_deg"], cmac=1.0, criterion=k, max_iter=c["max_iter"])
                              for c, k in zip(cases, crit)])
    ctx.set_tile_order(E.TILES_XCD_CHUNKS)
    ctx.init_freestream()
    print(f"staggered stops: {len(cases)} instances, limits {limits}", flush=True)
    done, status = ctx.batch_run(60, use_graph=True)
    for i, c in enumerate(cases):
        blocks, rows = _oracle_run(meshes[i], c, c["max_iter"], crit[i])
        assert status[i]["stopped"] == 1 and status[i]["iterations"] == len(rows), (i, status[i]["iterations"], len(rows))
        assert rows[-1]["stop"] == 1
        _check_instance(ctx, i, blocks, rows, ctx.batch_history(i, 100), f"instance {i}")
    assert status[5]["iterations"] < 40  # stopped by its criterion, not its limit
    ctx.close()
    for m in meshes:
        m.close()
