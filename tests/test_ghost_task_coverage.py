"""The stage kernels form ghost face states and ghost BCs in one pass per
wavefront: each wavefront's ghost slots are dealt to its active lanes, up to
one per lane per round. These checks (host only) pin that the meshes of the
GPU parity fixtures drive every branch of that dealing: a wavefront with more
ghost slots than lanes (several rounds), cells with several ghost slots, and a
partly filled last wavefront of a block (inactive lanes). A wavefront is 64
consecutive cells of a block: blocks start on a 256-cell boundary in device
memory."""
import numpy as np

import euler3d_amd as E
from helpers import Golden, mesh_for


def ghost_slots_per_cell(blk):
    n = blk["nreal"]
    off = blk["cell_off"][: n + 1]
    faces = blk["cell2face"][: off[n]]
    fc = blk["face2cell"].reshape(-1, 2)[faces]
    owner = np.repeat(np.arange(n), np.diff(off))
    other = np.where(fc[:, 0] == owner, fc[:, 1], fc[:, 0])
    return np.bincount(owner, weights=(other >= n).astype(int), minlength=n).astype(int)


def wavefront_stats(mesh):
    most, multi, partial = 0, 0, 0
    for b in range(mesh.nblocks):
        g = ghost_slots_per_cell(mesh.block(b))
        multi = max(multi, int(g.max()))
        for w0 in range(0, len(g), 64):
            most = max(most, int(g[w0:w0 + 64].sum()))
            partial += len(g) - w0 < 64
    return most, multi, partial


def test_hex_fixture_needs_several_rounds():
    most, multi, partial = wavefront_stats(mesh_for(Golden("naca_hex_roe2_gg_vk")))
    assert most > 128 and multi >= 2 and partial >= 1


def test_tet_fixture_has_partial_wavefronts_and_corner_cells():
    most, multi, partial = wavefront_stats(mesh_for(Golden("bump1_roe2_gg_vk")))
    assert 0 < most <= 64 and multi >= 2 and partial >= 1
