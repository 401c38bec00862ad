"""The reference's own multi-block mesh: ONERA M6 in 120 blocks
(Meshes/oneraM6_120B, written by the reference's partitioner), read where the
reference tree is present (skipped elsewhere; nothing is copied from it).

Its block files declare more markers (NMARK) than they list; the reference's
reader runs past the list and its serial build stops with a segmentation
fault while building block 0. This build reads the markers the file holds.
Checked: cell and ghost counts against the files and the topology, every
interface ghost filled from a real cell of its origin block, closed cells
(sum of outward face vectors zero) and positive volumes, and one oracle
iteration with the reference's own input (input_onera120) staying finite.
"""
import os

import numpy as np
import pytest

import euler3d_amd as E
import oracle as O

REF = "/root/reference"
TOP = os.path.join(REF, "Meshes", "oneraM6_120B", "topology_M6.top")

pytestmark = pytest.mark.skipif(not os.path.exists(TOP), reason="reference tree not present")


@pytest.fixture(scope="module")
def onera(tmp_path_factory):
    cwd = os.getcwd()
    os.chdir(REF)  # the topology lists its block files relative to the reference root
    try:
        mesh = E.Mesh.read(TOP)
    finally:
        os.chdir(cwd)
    return mesh


def test_onera_m6_120_blocks_build(onera):
    m = onera
    assert m.nblocks == 120
    n_real = n_ghost = 0
    for b in range(m.nblocks):
        v = m.block(b)
        n_real += v["nreal"]
        n_ghost += v["nall"] - v["nreal"]
        # interface ghosts come after the physical ones and are fed by real origin cells
        assert v["ghost_count"] <= v["nall"]
        off = np.asarray(v["itf_off"])
        assert off[-1] == v["nall"] - v["ghost_count"]
        for i, o in enumerate(v["itf_origin_block"]):
            cells = np.asarray(v["itf_origin_cells"][off[i]:off[i + 1]])
            assert ((cells >= 0) & (cells < m.block(int(o))["nreal"])).all()
        # closed cells: the face vectors of a real cell, oriented outward, sum to zero
        S = np.asarray(v["S"]).reshape(-1, 3)
        f2c = np.asarray(v["face2cell"]).reshape(-1, 2)
        co, c2f = np.asarray(v["cell_off"]), np.asarray(v["cell2face"])
        q = np.arange(v["nreal"])
        tot = np.zeros((v["nreal"], 3))
        for k in range(4):
            f = c2f[co[q] + k]
            sgn = np.where(f2c[f, 0] == q, 1.0, -1.0)[:, None]
            tot += sgn * S[f]
        assert np.abs(tot).max() < 1e-12 * max(1.0, np.abs(S).max())
        assert (np.asarray(v["volume"])[:v["nreal"]] > 0).all()
    assert n_real == 582752


def test_onera_m6_one_oracle_iteration(onera):
    p, _ = E.read_input(os.path.join(REF, "input_onera120"))
    geoms = [onera.block(b) for b in range(onera.nblocks)]
    S = O.Solver(geoms, O.params(p.mach, p.aoa_deg, flux=p.flux, order=p.order, gradient=p.gradient,
                                 limiter=p.limiter, omega=p.omega, k=p.k, max_iter=10, criterion=-15.0))
    S.init_freestream()
    h = S.iterate(1, threads=8)
    assert np.isfinite(h[0]["rms"]).all() and h[0]["rms"][0] > 0
    assert np.isfinite([h[0]["cl"], h[0]["cd"]]).all()
