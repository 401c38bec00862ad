"""GPU: the Green-Gauss stage kernel with slot geometry rebuilt from node
copies (the default on tet meshes) against the same kernel streaming the
stored per-slot copies (E3D_GEOMETRY_STREAMED), bit for bit, single meshes
and batches, and which form each mesh and scheme gets. The default form is
also the one every Green-Gauss tet case of the oracle parity tests runs."""
import numpy as np
import pytest

import euler3d_amd as E
from helpers import Golden, mesh_for

pytestmark = pytest.mark.gpu

GG_VK = dict(flux=E.FLUX_ROE, order=2, gradient=E.GRAD_GREEN_GAUSS, limiter=E.LIM_VENKATAKRISHNAN, k=5.0, stages=5)
GG_BJ = dict(flux=E.FLUX_AUSM, order=2, gradient=E.GRAD_GREEN_GAUSS, limiter=E.LIM_BARTH_JESPERSEN, omega=0.01,
             stages=5)
LSQ_BJ = dict(flux=E.FLUX_ROE, order=2, gradient=E.GRAD_LEAST_SQUARES, limiter=E.LIM_BARTH_JESPERSEN, stages=5)


def run(mesh, scheme, geometry, n=6, tile_cells=256, mach=0.675, aoa=2.0):
    ctx = E.Context(0)
    ctx.set_params(E.make_params(mach, aoa, max_iter=n, criterion=-15.0, **scheme))
    ctx.upload(mesh)
    ctx.set_slot_geometry(geometry)
    ctx.set_tile_cells(tile_cells)
    ctx.init_freestream()
    nodes = ctx.slot_geometry_nodes()
    hist = ctx.iterate(n)
    out = {k: [ctx.get(b, a) for b in range(mesh.nblocks)] for k, a in (("W", E.ARR_W), ("diss", E.ARR_DISS),
                                                                       ("conv", E.ARR_CONV))}
    ctx.close()
    return nodes, out, hist


def assert_same_bits(a, b):
    for k in a:
        for x, y in zip(a[k], b[k]):
            assert np.array_equal(x.view(np.int64), y.view(np.int64)), k


@pytest.mark.parametrize("scheme", ["gg_vk", "gg_bj"])
@pytest.mark.parametrize("mesh_desc", [(24, 8, 8, 0.1, 4, (2, 2, 2)), (18, 6, 6, 0.08, -5, (1, 1, 1))])
@pytest.mark.parametrize("tile_cells", [64, 256])
def test_node_geometry_equals_streamed(scheme, mesh_desc, tile_cells):
    nx, ny, nz, h, brick, parts = mesh_desc
    mesh = E.Mesh.bump(nx, ny, nz, h, brick, parts)
    sc = {"gg_vk": GG_VK, "gg_bj": GG_BJ}[scheme]
    n_nodes, a, ha = run(mesh, sc, E.GEOMETRY_AUTO, tile_cells=tile_cells)
    n_str, b, hb = run(mesh, sc, E.GEOMETRY_STREAMED, tile_cells=tile_cells)
    assert n_nodes and not n_str
    assert_same_bits(a, b)
    assert [r["rms"] for r in ha] == [r["rms"] for r in hb]
    assert [(r["cl"], r["cd"]) for r in ha] == [(r["cl"], r["cd"]) for r in hb]
    mesh.close()


def test_forms_by_mesh_and_scheme():
    """Node copies on tet meshes with Green-Gauss only: least squares and a
    hexahedral mesh stream the stored copies."""
    tets = E.Mesh.bump(8, 3, 3, 0.1, 4, (1, 1, 1))
    assert run(tets, GG_VK, E.GEOMETRY_AUTO, n=1)[0]
    assert not run(tets, LSQ_BJ, E.GEOMETRY_AUTO, n=1)[0]
    assert not run(tets, GG_VK, E.GEOMETRY_STREAMED, n=1)[0]
    hexes = mesh_for(Golden("naca_hex_roe2_gg_vk"))
    assert not run(hexes, GG_VK, E.GEOMETRY_AUTO, n=1, mach=0.8, aoa=1.25)[0]
    ctx = E.Context(0)
    for bad in (1, -2):
        with pytest.raises(E.E3DError):
            ctx.set_slot_geometry(bad)
    ctx.close()
    tets.close()
    hexes.close()


def test_batch_node_geometry_equals_streamed():
    """A batch of three tet meshes with an instance that stops early, through
    the hipGraph step, in both forms."""
    meshes = [E.Mesh.bump(8, 3, 3, 0.1, 4, (1, 1, 1)), E.Mesh.bump(12, 4, 4, 0.06, 4, (2, 1, 1)),
              E.Mesh.bump(10, 3, 3, 0.12, 2, (1, 1, 1))]
    insts = [dict(mach=0.5, aoa_deg=1.0, criterion=-15.0, max_iter=4), dict(mach=0.7, aoa_deg=-1.0, max_iter=12),
             dict(mach=0.35, aoa_deg=5.0, max_iter=12)]
    res = {}
    for geo in (E.GEOMETRY_AUTO, E.GEOMETRY_STREAMED):
        ctx = E.Context(0)
        ctx.set_params(E.make_params(0.5, 0.0, **GG_VK))
        ctx.upload_batch(meshes, insts)
        ctx.set_slot_geometry(geo)
        ctx.init_freestream()
        assert ctx.slot_geometry_nodes() == (geo == E.GEOMETRY_AUTO)
        done, status = ctx.batch_run(12, use_graph=True)
        assert status[0]["stopped"] and status[0]["iterations"] == 4
        res[geo] = ([ctx.get(b, E.ARR_W) for b in range(len(ctx.sizes))],
                    [ctx.batch_history(i, 12) for i in range(3)])
        ctx.close()
    a, b = res[E.GEOMETRY_AUTO], res[E.GEOMETRY_STREAMED]
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x.view(np.int64), y.view(np.int64))
    assert a[1] == b[1]
    for m in meshes:
        m.close()
