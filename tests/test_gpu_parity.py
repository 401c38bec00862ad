"""GPU parity: the HIP path through the C ABI against the reference and the oracle.

* Granular operators (one C-ABI call per reference operator, driven in the
  reference's Solver::solve order) against the reference's golden digests,
  bit for bit: gradients, limiters, residuals, W0, dt, spectral radius after
  every stage residual of iteration 5, and the state after iterations 4 and 5.
* The fused path (e3d_iterate: two kernels per RK stage) against the oracle:
  every cell's state equal as values (+0 == -0; beta = 0 stages skip the
  dissipation, which only moves the sign of a zero), monitors within 1e-12
  relative (different summation order of the norms and force sums).
"""
import numpy as np
import pytest

import euler3d_amd as E
import oracle as O
from helpers import C1_CASE, SMALL_FIXTURES, Golden, mesh_for, monitor_row, oracle_params, product_params, values_equal

pytestmark = pytest.mark.gpu

ARR = {"W": E.ARR_W, "grad": E.ARR_GRAD, "lim": E.ARR_LIM, "conv": E.ARR_CONV, "diss": E.ARR_DISS,
       "W0": E.ARR_W0, "dt": E.ARR_DT, "specrad": E.ARR_SPECRAD}
MON_RTOL = 1e-12


def gpu_context(mesh, params, ordering=E.ORDER_REFERENCE):
    ctx = E.Context(0)
    ctx.set_params(params)
    ctx.set_ordering(ordering)
    ctx.upload(mesh)
    ctx.init_freestream()
    return ctx


def close_rel(a, b, rtol=MON_RTOL):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300)) or np.allclose(a, b, rtol=rtol, atol=1e-280)


ORDERINGS = {"reference_order": E.ORDER_REFERENCE, "rcm_order": E.ORDER_RCM, "morton_order": E.ORDER_MORTON}


@pytest.mark.parametrize("ordering", list(ORDERINGS))
@pytest.mark.parametrize("case", SMALL_FIXTURES)
def test_granular_operators_match_reference_bitwise(case, ordering):
    """In either memory order (cells and faces renumbered on the device by
    reverse Cuthill-McKee, or not) every array equals the reference's."""
    g = Golden(case)
    meta = g.meta
    mesh = mesh_for(g)
    nb = mesh.nblocks
    ctx = gpu_context(mesh, product_params(meta), ORDERINGS[ordering])
    k = meta["stage_iter"]
    ctx.op("exchange")                      # Solver.cpp:48
    ref_mon = g.monitor()
    for it in range(max(meta["dumps"]) + 1):
        for b in range(nb):                 # Solver.cpp:83-91
            ctx.op("save_w0", b)
            ctx.op("spectral_radius", b)
            ctx.op("timestep", b)
            for s in range(meta["stages"]):
                ctx.op("residual", b, s)
                if it == k:
                    tag = f"it{it}s{s}_b{b}_"
                    for what in ("W", "grad", "lim", "conv", "diss") + (("W0", "dt", "specrad") if s == 0 else ()):
                        g.check(tag + what, ctx.get(b, ARR[what]).reshape(-1), f"GPU stage {s}")
                ctx.op("update", b, s)
                ctx.op("bc", b)
        ctx.op("exchange")                  # Solver.cpp:94
        m = ctx.monitors()
        r = ref_mon[it]
        assert close_rel(m["rms"], r[1:6]), (it, m["rms"], r[1:6])
        assert close_rel([m["cl"], m["cd"], m["cmx"], m["cmy"], m["cmz"]], r[6:11]), it
        if it in meta["dumps"]:
            for b in range(nb):
                g.check(f"it{it}_b{b}_W", ctx.get(b, E.ARR_W).reshape(-1), f"GPU after iteration {it}")
    ctx.close()


def test_c1_config_granular_200_iterations_bitwise():
    """The 48x16x16x6-tet channel (73,728 cells), first-order Roe, M=0.5, 200
    iterations through the granular operators in Solver::solve order: every
    array after the last iteration bit for bit against the reference run, and
    the reference's sampled monitor rows within 1e-12."""
    g = Golden(C1_CASE)
    meta = g.meta
    mesh = mesh_for(g)
    ctx = gpu_context(mesh, product_params(meta))
    print(f"c1 granular: {meta['iters']} iterations", flush=True)
    ctx.op("exchange")
    rows = []
    for it in range(meta["iters"]):
        ctx.op("save_w0", 0)
        ctx.op("spectral_radius", 0)
        ctx.op("timestep", 0)
        ctx.op("runge_kutta", 0)
        ctx.op("exchange")
        rows.append(monitor_row(ctx.monitors()))
    for what in ("W", "W0", "conv", "diss", "dt", "specrad", "grad", "lim"):
        g.check(f"it{meta['iters'] - 1}_b0_{what}", ctx.get(0, ARR[what]).reshape(-1), "GPU granular")
    idx, ref = g.samples("monitor")
    mine = np.concatenate(rows)[idx]
    assert close_rel(mine, ref), np.flatnonzero(~np.isclose(mine, ref, rtol=MON_RTOL, atol=0))[:8]
    ctx.close()


def test_c1_config_fused_200_iterations():
    """The same run through e3d_iterate: sampled states after 200 iterations
    equal to the reference's (as values), sampled monitors within 1e-12, and
    the run stops at iteration 199 on max_iter as the reference did."""
    g = Golden(C1_CASE)
    meta = g.meta
    mesh = mesh_for(g)
    ctx = gpu_context(mesh, product_params(meta))
    print(f"c1 fused: {meta['iters']} iterations", flush=True)
    hist = ctx.iterate(meta["iters"] + 10)
    assert len(hist) == meta["iters"] and hist[-1]["stop"] == 1
    last = meta["iters"] - 1
    for what in ("W", "W0", "conv", "diss", "dt"):
        idx, ref = g.samples(f"it{last}_b0_{what}")
        mine = ctx.get(0, ARR[what]).reshape(-1)[idx]
        assert values_equal(mine, ref).all(), (what, idx[~values_equal(mine, ref)][:5])
    idx, ref = g.samples("monitor")
    mine = np.concatenate([monitor_row(m) for m in hist])[idx]
    assert close_rel(mine, ref)
    ctx.close()


def run_both(mesh, meta, n, ordering=E.ORDER_REFERENCE):
    ctx = gpu_context(mesh, product_params(meta), ordering)
    hist = ctx.iterate(n)
    geoms = [mesh.block(b) for b in range(mesh.nblocks)]
    S = O.Solver(geoms, oracle_params(meta))
    S.init_freestream()
    rh = S.iterate(n)
    return ctx, hist, S, rh


def assert_states_equal(ctx, S):
    for b, B in enumerate(S.blocks):
        W = ctx.get(b, E.ARR_W).reshape(-1)
        same = values_equal(W, B.W)
        assert same.all(), f"block {b}: {np.sum(~same)} cells differ, first {np.flatnonzero(~same)[:5]}"
        d = ctx.get(b, E.ARR_DISS).reshape(-1)
        dref = B.diss.reshape(5, B.nall)[:, :B.nreal].reshape(-1)
        assert values_equal(d, dref).all(), f"block {b}: dissipation differs"


def assert_monitors_close(hist, rh):
    assert len(hist) == len(rh)
    for a, r in zip(hist, rh):
        assert a["iter"] == r["iter"] and a["stop"] == r["stop"]
        assert close_rel(a["rms"], r["rms"]), (a["iter"], a["rms"], r["rms"])
        assert close_rel([a["cl"], a["cd"], a["cmx"], a["cmy"], a["cmz"]],
                         [r["cl"], r["cd"], r["cmx"], r["cmy"], r["cmz"]]), a["iter"]


@pytest.mark.parametrize("ordering", list(ORDERINGS))
@pytest.mark.parametrize("case", SMALL_FIXTURES)
def test_fused_iterations_match_oracle(case, ordering):
    g = Golden(case)
    mesh = mesh_for(g)
    ctx, hist, S, rh = run_both(mesh, g.meta, g.meta["iters"], ORDERINGS[ordering])
    assert_states_equal(ctx, S)
    assert_monitors_close(hist, rh)
    # and against the reference's own monitor history
    ref = g.monitor()
    for a, r in zip(hist, ref):
        assert close_rel(a["rms"], r[1:6]) and close_rel([a["cl"], a["cd"]], r[6:8])
    ctx.close()


@pytest.mark.parametrize("scheme", ["roe1", "roe2_gg_vk", "ausm2_lsq_bj", "roe2_lsq_bj_8b"])
def test_fused_medium_mesh_matches_oracle(scheme):
    """24x8x8 hexes (9,216 tets), several tiles per block, against the oracle."""
    meta = dict(mach=0.675, aoa=1.0, flux=0, order=2, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=8)
    parts = (1, 1, 1)
    if scheme == "roe1":
        meta.update(order=1, mach=0.5, aoa=0.0)
    elif scheme == "ausm2_lsq_bj":
        meta.update(flux=1, grad=1, lim=0, mach=0.7, aoa=3.0)
    elif scheme == "roe2_lsq_bj_8b":
        meta.update(grad=1, lim=0)
        parts = (2, 2, 2)
    mesh = E.Mesh.bump(24, 8, 8, 0.1, 4, parts)
    ctx, hist, S, rh = run_both(mesh, meta, meta["iters"])
    assert_states_equal(ctx, S)
    assert_monitors_close(hist, rh)
    ctx.close()


def test_developed_flow_matches_oracle():
    """300 iterations of the 24x8x8 channel in 8 blocks, Roe + Green-Gauss +
    Venkatakrishnan: by then the limiter's d2 signs and the entropy fix's
    branches mix within wavefronts (the regime where the branch-free limiter
    and the one-pass ghost code differ most from a per-lane reading of the
    reference); every state bit against the oracle."""
    meta = dict(mach=0.675, aoa=1.0, flux=0, order=2, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=300)
    mesh = E.Mesh.bump(24, 8, 8, 0.1, 4, (2, 2, 2))
    ctx, hist, S, rh = run_both(mesh, meta, meta["iters"])
    assert_states_equal(ctx, S)
    assert_monitors_close(hist, rh)
    assert hist[-1]["drop"][0] < -0.5  # the flow has moved well away from the freestream start
    ctx.close()


@pytest.mark.parametrize("scheme", ["ausm2_lsq_bj", "roe2_gg_bj", "roe2_lsq_vk"])
def test_hex_mesh_schemes_match_oracle(scheme):
    """The hexahedral NACA0012 fixture mesh (six face slots per cell, up to 192
    ghost slots per wavefront) with the gradient/limiter pairs its golden run
    does not use, against the oracle."""
    meta = dict(mach=0.8, aoa=1.25, flux=0, order=2, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=4)
    if scheme == "ausm2_lsq_bj":
        meta.update(flux=1, grad=1, lim=0)
    elif scheme == "roe2_gg_bj":
        meta.update(lim=0, omega=0.01)
    elif scheme == "roe2_lsq_vk":
        meta.update(grad=1)
    mesh = mesh_for(Golden("naca_hex_roe2_gg_vk"))
    ctx, hist, S, rh = run_both(mesh, meta, meta["iters"])
    assert_states_equal(ctx, S)
    assert_monitors_close(hist, rh)
    ctx.close()


@pytest.mark.parametrize("tiles", ["linear", "xcd_contiguous", "xcd_chunks"])
def test_fused_tile_orders_match_oracle(tiles):
    """Every workgroup-to-tile order gives the oracle's values: 48x16x16 hexes
    (73,728 tets, 288 tiles) in 8 blocks, so the runs of 16 tiles do not divide
    the tiles evenly over the eight XCDs."""
    meta = dict(mach=0.675, aoa=1.0, flux=0, order=2, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=3)
    mesh = E.Mesh.bump(48, 16, 16, 0.1, 4, (2, 2, 2))
    ctx = gpu_context(mesh, product_params(meta))
    ctx.set_tile_order({"linear": E.TILES_LINEAR, "xcd_contiguous": E.TILES_XCD_CONTIGUOUS,
                        "xcd_chunks": E.TILES_XCD_CHUNKS}[tiles])
    hist = ctx.iterate(meta["iters"])
    S = O.Solver([mesh.block(b) for b in range(mesh.nblocks)], oracle_params(meta))
    S.init_freestream()
    rh = S.iterate(meta["iters"], threads=8)
    assert_states_equal(ctx, S)
    assert_monitors_close(hist, rh)
    ctx.close()


@pytest.mark.parametrize("cells", [64, 128])
@pytest.mark.parametrize("scheme", ["roe_gg_vk", "ausm_lsq_bj"])
def test_fused_tile_sizes_match_oracle_and_each_other(cells, scheme):
    """Tiles of 64 and 128 cells per workgroup (256 is the default of every
    other test) give the oracle's values and, monitors included, the same
    bits as tiles of 256 cells: 48x16x16 hexes in 8 blocks, every tile order."""
    meta = dict(mach=0.675, aoa=1.0, flux=0, order=2, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=3)
    if scheme == "ausm_lsq_bj":
        meta.update(flux=1, grad=1, lim=0, mach=0.7, aoa=3.0)
    mesh = E.Mesh.bump(48, 16, 16, 0.1, 4, (2, 2, 2))
    ref = gpu_context(mesh, product_params(meta))
    href = ref.iterate(meta["iters"])
    S = O.Solver([mesh.block(b) for b in range(mesh.nblocks)], oracle_params(meta))
    S.init_freestream()
    rh = S.iterate(meta["iters"], threads=8)
    assert_states_equal(ref, S)
    assert_monitors_close(href, rh)
    for order in (E.TILES_LINEAR, E.TILES_XCD_CONTIGUOUS, E.TILES_XCD_CHUNKS):
        ctx = E.Context(0)
        ctx.set_params(product_params(meta))
        ctx.upload(mesh)
        ctx.set_tile_cells(cells)
        ctx.set_tile_order(order)
        ctx.init_freestream()
        hist = ctx.iterate(meta["iters"])
        for blk in range(mesh.nblocks):
            for arr in (E.ARR_W, E.ARR_DISS, E.ARR_CONV):
                assert (ctx.get(blk, arr).view(np.int64) == ref.get(blk, arr).view(np.int64)).all(), (order, blk, arr)
        assert hist == href, order
        ctx.close()
    for bad in (32, 192, 512):
        with pytest.raises(E.E3DError):
            ref.set_tile_cells(bad)
    ref.close()


def test_fused_is_deterministic_and_resumable():
    """Two runs give identical bits; 3 + 3 iterations equal 6 in one call."""
    meta = dict(mach=0.675, aoa=1.0, flux=0, order=2, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=50)
    mesh = E.Mesh.bump(16, 6, 6, 0.1, 4, (2, 1, 1))
    a = gpu_context(mesh, product_params(meta))
    ha = a.iterate(6)
    b = gpu_context(mesh, product_params(meta))
    hb = b.iterate(3) + b.iterate(3)
    for blk in range(mesh.nblocks):
        wa, wb = a.get(blk, E.ARR_W), b.get(blk, E.ARR_W)
        assert (wa.view(np.int64) == wb.view(np.int64)).all()
    assert [h["rms"] for h in ha] == [h["rms"] for h in hb]
    a.close()
    b.close()


def test_fused_equals_granular_path():
    """e3d_iterate and the per-operator ABI calls produce the same state."""
    meta = dict(mach=0.7, aoa=3.0, flux=1, order=2, grad=1, lim=0, omega=1e-16, k=5.0, stages=5, iters=50)
    mesh = E.Mesh.bump(12, 4, 4, 0.1, 4, (2, 1, 1))
    a = gpu_context(mesh, product_params(meta))
    a.iterate(3)
    b = gpu_context(mesh, product_params(meta))
    b.op("exchange")
    for _ in range(3):
        for blk in range(mesh.nblocks):
            b.op("save_w0", blk)
            b.op("spectral_radius", blk)
            b.op("timestep", blk)
            b.op("runge_kutta", blk)
        b.op("exchange")
        b.monitors()
    for blk in range(mesh.nblocks):
        assert values_equal(a.get(blk, E.ARR_W), b.get(blk, E.ARR_W)).all()
    a.close()
    b.close()


def test_stop_criterion_and_request():
    meta = dict(mach=0.5, aoa=0.0, flux=0, order=1, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=40)
    mesh = E.Mesh.bump(8, 3, 3, 0.1, 4)
    ctx = gpu_context(mesh, product_params(meta))
    h = ctx.iterate(100)
    assert len(h) == 40 and h[-1]["stop"] == 1 and all(x["stop"] == 0 for x in h[:-1])  # max_iter
    p = product_params(meta)
    p.criterion = -0.2
    ctx2 = gpu_context(mesh, p)
    h2 = ctx2.iterate(100)
    assert h2[-1]["drop"][0] <= -0.2 and all(x["drop"][0] > -0.2 for x in h2[:-1])
    ctx3 = gpu_context(mesh, product_params(meta))
    ctx3.iterate(2)
    E.lib().e3d_request_stop(ctx3._h)
    assert ctx3.iterate(5) == []
    for c in (ctx, ctx2, ctx3):
        c.close()


def test_abi_errors_on_device():
    ctx = E.Context(0)
    with pytest.raises(E.E3DError) as e:
        ctx.op("residual", 0, 0)
    assert e.value.code == -5  # no mesh
    p = E.make_params(0.5)
    p.stages = 7
    with pytest.raises(E.E3DError) as e:
        ctx.set_params(p)
    assert e.value.code == -1
    p = E.make_params(0.5)
    p.smoothing = 2
    with pytest.raises(E.E3DError):
        ctx.set_params(p)
    ctx.set_params(E.make_params(0.5))
    ctx.upload(E.Mesh.bump(4, 2, 2, 0.1, 4))
    with pytest.raises(E.E3DError) as e:
        ctx.op("update", 0, 9)
    assert e.value.code == -1
    with pytest.raises(E.E3DError):
        ctx.op("bc", 3)
    ctx.close()
