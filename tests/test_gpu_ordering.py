"""Memory order of cells and faces on the device (e3d_set_ordering).

Reverse Cuthill-McKee or a Morton curve renumbers each block's real cells and
faces in device memory only; every cell keeps its faces in the reference's order, so every
per-cell value is the same bit for bit as in the reference's numbering. The
residual norms sum over cells in memory order and agree to 1e-12.

The scrambled bump channel (tets in a pseudo-random order) is the case RCM is
for: the same physics as any other numbering of the mesh, but neighbours far
apart in memory. It is checked against the oracle in both orders.
"""
import numpy as np
import pytest

import euler3d_amd as E
import oracle as O
from helpers import product_params, values_equal

pytestmark = pytest.mark.gpu

ARRAYS = (E.ARR_W, E.ARR_W0, E.ARR_CONV, E.ARR_DISS, E.ARR_DT)
SECOND_ORDER = dict(mach=0.675, aoa=1.0, flux=0, order=2, grad=0, lim=1, omega=1e-16, k=5.0, stages=5, iters=50)


def run(mesh, meta, ordering, n):
    ctx = E.Context(0)
    ctx.set_params(product_params(meta))
    ctx.set_ordering(ordering)
    ctx.upload(mesh)
    ctx.init_freestream()
    hist = ctx.iterate(n)
    arrays = {(b, a): ctx.get(b, a) for b in range(mesh.nblocks) for a in ARRAYS}
    ctx.close()
    return arrays, hist


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def rel_close(a, b, rtol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(b)))


@pytest.mark.parametrize("meta", [SECOND_ORDER, dict(SECOND_ORDER, flux=1, grad=1, lim=0, mach=0.7, aoa=3.0)],
                         ids=["roe_gg_vk", "ausm_lsq_bj"])
def test_scrambled_mesh_equal_in_every_order_and_to_oracle(meta):
    mesh = E.Mesh.bump(16, 6, 6, 0.1, -7, (2, 1, 1))
    print("scrambled 16x6x6 mesh, 2 blocks, 4 iterations", flush=True)
    ref_arrays, ref_hist = run(mesh, meta, E.ORDER_REFERENCE, 4)
    rcm_arrays, rcm_hist = run(mesh, meta, E.ORDER_RCM, 4)
    for key, a in ref_arrays.items():
        assert (bits(a) == bits(rcm_arrays[key])).all(), key
    for a, b in zip(ref_hist, rcm_hist):
        assert rel_close(b["rms"], a["rms"]) and rel_close([b["cl"], b["cd"]], [a["cl"], a["cd"]])
    mort_arrays, mort_hist = run(mesh, meta, E.ORDER_MORTON, 4)
    for key, a in ref_arrays.items():
        assert (bits(a) == bits(mort_arrays[key])).all(), key
    for a, b in zip(ref_hist, mort_hist):
        assert rel_close(b["rms"], a["rms"]) and rel_close([b["cl"], b["cd"]], [a["cl"], a["cd"]])
    # the oracle on the same (scrambled) numbering
    geoms = [mesh.block(b) for b in range(mesh.nblocks)]
    S = O.Solver(geoms, O.params(meta["mach"], meta["aoa"], flux=meta["flux"], order=meta["order"],
                                 gradient=meta["grad"], limiter=meta["lim"], omega=meta["omega"], k=meta["k"],
                                 stages=meta["stages"], max_iter=meta["iters"], criterion=-15.0))
    S.init_freestream()
    rh = S.iterate(4)
    for b, B in enumerate(S.blocks):
        assert values_equal(rcm_arrays[(b, E.ARR_W)].reshape(-1), B.W).all(), b
    for a, r in zip(rcm_hist, rh):
        assert rel_close(a["rms"], r["rms"])


def test_batch_equal_in_every_order():
    scheme = dict(flux=E.FLUX_ROE, order=2, gradient=E.GRAD_GREEN_GAUSS, limiter=E.LIM_VENKATAKRISHNAN, k=5.0,
                  stages=5)
    meshes = [E.Mesh.bump(8, 3, 3, 0.1, -3, (1, 1, 1)), E.Mesh.bump(12, 4, 4, 0.06, 4, (2, 1, 1))]
    insts = [dict(mach=0.5, aoa_deg=1.0, criterion=-15.0, max_iter=12),
             dict(mach=0.7, aoa_deg=-1.0, criterion=-15.0, max_iter=12)]
    out = []
    for ordering in (E.ORDER_REFERENCE, E.ORDER_RCM, E.ORDER_MORTON):
        ctx = E.Context(0)
        ctx.set_params(E.make_params(0.5, 0.0, **scheme))
        ctx.set_ordering(ordering)
        ctx.upload_batch(meshes, insts)
        ctx.init_freestream()
        ctx.batch_run(12, use_graph=True)
        out.append([ctx.get(b, E.ARR_W) for b in range(3)])
        ctx.close()
    for a, b, c in zip(*out):
        assert (bits(a) == bits(b)).all() and (bits(a) == bits(c)).all()


def test_set_ordering_rejects_unknown_mode():
    ctx = E.Context(0)
    with pytest.raises(E.E3DError) as e:
        ctx.set_ordering(5)
    assert e.value.code == -1
    ctx.close()
