"""Failure detection: a non-finite state is reported, never run through.

The reference has no such check (log10 of NaN compares false and it runs on to
max_iter). Here a NaN placed in one cell's state from the host makes
e3d_iterate return E3D_ENONFINITE with the iteration in the message, and in a
batch it stops only the instance it belongs to: the other instances go on
and end exactly as they do alone. Nothing on the device faults: the NaN is
ordinary data for every kernel.
"""
import numpy as np
import pytest

import euler3d_amd as E

pytestmark = pytest.mark.gpu

SCHEME = dict(flux=E.FLUX_ROE, order=2, gradient=E.GRAD_GREEN_GAUSS, limiter=E.LIM_VENKATAKRISHNAN, k=5.0,
              stages=5)


def test_nan_state_is_reported_by_iterate():
    mesh = E.Mesh.bump(8, 3, 3, 0.1, 4, (2, 1, 1))
    ctx = E.Context(0)
    ctx.set_params(E.make_params(0.6, 1.0, max_iter=50, criterion=-15.0, **SCHEME))
    ctx.upload(mesh)
    ctx.init_freestream()
    assert len(ctx.iterate(2)) == 2
    W = ctx.get(1, E.ARR_W)
    W[0, 10] = np.nan  # density of real cell 10 of block 1
    ctx.set(1, E.ARR_W, W)
    with pytest.raises(E.E3DError) as e:
        ctx.iterate(5)
    assert e.value.code == -6 and "not finite at iteration 2" in str(e.value)
    ctx.close()


def test_nan_stops_only_its_instance_in_a_batch():
    meshes = [E.Mesh.bump(8, 3, 3, 0.1, 4, (1, 1, 1)), E.Mesh.bump(10, 3, 3, 0.08, 4, (1, 1, 1))]
    insts = [dict(mach=0.5, aoa_deg=1.0, criterion=-15.0, max_iter=20),
             dict(mach=0.6, aoa_deg=-1.0, criterion=-15.0, max_iter=20)]

    def batch(poison):
        ctx = E.Context(0)
        ctx.set_params(E.make_params(0.5, 0.0, **SCHEME))
        ctx.upload_batch(meshes, insts)
        ctx.init_freestream()
        if poison:
            (b,) = ctx.batch_blocks(0)
            W = ctx.get(b, E.ARR_W)
            W[4, 3] = np.nan  # pressure of real cell 3 of instance 0
            ctx.set(b, E.ARR_W, W)
        _, st = ctx.batch_run(20, use_graph=True)
        (b1,) = ctx.batch_blocks(1)
        out = st, ctx.get(b1, E.ARR_W)
        ctx.close()
        return out

    st_ok, w_ok = batch(False)
    st_bad, w_bad = batch(True)
    assert st_ok[0]["iterations"] == 20 and st_ok[1]["iterations"] == 20
    assert st_bad[0]["stopped"] == 1 and st_bad[0]["iterations"] == 1
    assert not np.isfinite(st_bad[0]["last"]["rms"][0])
    assert st_bad[1]["iterations"] == 20
    assert (w_bad.view(np.int64) == w_ok.view(np.int64)).all()
