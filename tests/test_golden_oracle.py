"""The oracle (CPU restatement) against the reference's golden vectors.

Runs the oracle from freestream on the same mesh and parameters as the
reference run of each fixture and checks, bit for bit: the state after
iteration 4, every array right after each stage residual of iteration 5
(gradients, limiters, convective and dissipative residuals, W0, dt, spectral
radius), the state after iteration 5, every monitor row of the whole run and
the per-block monitors of the first iterations.
"""
import numpy as np
import pytest

import oracle as O
from helpers import C1_CASE, SMALL_FIXTURES, Golden, mesh_for, monitor_row, oracle_params, oracle_state


def run_oracle_against(g):
    meta = g.meta
    mesh = mesh_for(g)
    geoms = [mesh.block(b) for b in range(mesh.nblocks)]
    S = O.Solver(geoms, oracle_params(meta))
    S.init_freestream()
    k = meta["stage_iter"]
    checked = []

    def on_stage(it, s, b, B):
        tag = f"it{it}s{s}_b{b}_"
        for what in ("W", "grad", "lim", "conv", "diss") + (("W0", "dt", "specrad") if s == 0 else ()):
            g.check(tag + what, oracle_state(B, what), f"stage {s}")
            checked.append(tag + what)

    def on_iter(it, blocks):
        if it in meta["dumps"]:
            for b, B in enumerate(blocks):
                g.check(f"it{it}_b{b}_W", B.W, f"after iteration {it}")
                checked.append(f"it{it}_b{b}_W")

    rows = S.run(meta["iters"], stage_iter=k, on_stage=on_stage, on_iter=on_iter)
    return rows, checked


@pytest.mark.parametrize("case", SMALL_FIXTURES)
def test_oracle_matches_reference(case):
    g = Golden(case)
    rows, checked = run_oracle_against(g)
    assert len(checked) > 0
    ref = g.monitor()
    assert len(rows) == len(ref)
    for r, m in zip(ref, rows):
        mine = monitor_row(m)
        same = (mine.view(np.int64) == r.view(np.int64)) | (np.isnan(mine) & np.isnan(r))
        assert same.all(), f"iteration {int(r[0])}: monitor {mine[~same]} vs reference {r[~same]}"
    nb = len(rows[0]["block"])
    for it in range(min(3, len(rows))):
        for b in range(nb):
            g.check(f"blockmon_it{it}_b{b}", np.asarray(rows[it]["block"][b], np.float64), "block monitor")


def test_oracle_fused_loop_equals_hooked_loop():
    """e3do_iterate (the C loop) equals the hookable Python loop over the same primitives."""
    g = Golden("bump8_roe2_gg_vk")
    mesh = mesh_for(g)
    geoms = [mesh.block(b) for b in range(mesh.nblocks)]
    a = O.Solver(geoms, oracle_params(g.meta))
    b = O.Solver(geoms, oracle_params(g.meta))
    a.init_freestream()
    b.init_freestream()
    ha = a.iterate(6)
    hb = b.run(6)
    for x, y in zip(ha, hb):
        assert x["rms"] == y["rms"] and x["cl"] == y["cl"] and x["cmz"] == y["cmz"]
    for Ba, Bb in zip(a.blocks, b.blocks):
        assert (Ba.W.view(np.int64) == Bb.W.view(np.int64)).all()


@pytest.mark.slow
def test_oracle_c1_config_200_iterations():
    """The 48x16x16x6-tet channel (73,728 cells), first-order Roe, M=0.5, 200
    iterations from freestream: every array after the last iteration and the
    whole monitor history, bit for bit against the reference run (about a
    minute on one core)."""
    g = Golden(C1_CASE)
    meta = g.meta
    mesh = mesh_for(g)
    S = O.Solver([mesh.block(0)], oracle_params(meta))
    S.init_freestream()
    last = meta["iters"] - 1

    def on_iter(it, blocks):
        if it == last:
            for what in ("W", "W0", "conv", "diss", "dt", "specrad", "grad", "lim"):
                g.check(f"it{it}_b0_{what}", oracle_state(blocks[0], what), f"after iteration {it}")

    rows = S.run(meta["iters"], on_iter=on_iter)
    assert len(rows) == meta["iters"]
    g.check("monitor", np.concatenate([monitor_row(m) for m in rows]), "monitor history")


def test_oracle_threads_do_not_change_results():
    """The oracle spread over threads (contiguous block ranges, as the
    reference's MPI ranks) gives the same bits as the sequential loop."""
    import euler3d_amd as E
    mesh = E.Mesh.bump(12, 4, 4, 0.1, 4, (2, 2, 1))
    geoms = [mesh.block(b) for b in range(mesh.nblocks)]
    runs = []
    for threads in (1, 3):
        S = O.Solver(geoms, O.params(0.675, 1.0, flux=0, order=2, gradient=0, limiter=1, k=5.0, stages=5,
                                     max_iter=100, criterion=-15.0))
        S.init_freestream()
        hist = S.iterate(3, threads=threads)
        runs.append((np.concatenate([b.W for b in S.blocks]), [h["rms"] for h in hist]))
    assert (runs[0][0].view(np.int64) == runs[1][0].view(np.int64)).all()
    assert runs[0][1] == runs[1][1]
