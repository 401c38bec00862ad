"""The product's mesh builder against the reference's own builder and metrics.

For every golden case the reference built the mesh from the same SU2 file
(BlockBuilder/src/ConcreteBlockBuilder.cpp:22-742, BlockBuilder/src/BlockBuilder.cpp:147-458,
Metrics/src/MetricsInitializer.cpp:27-800); faces, connectivity, boundary lists
and every geometric array must agree bit for bit.
"""
import numpy as np
import pytest

from helpers import SMALL_FIXTURES, Golden, geometry_arrays, mesh_for


@pytest.mark.parametrize("case", SMALL_FIXTURES)
def test_builder_matches_reference(case):
    g = Golden(case)
    mesh = mesh_for(g)
    nb = 8 if case.startswith("bump8") else 1
    assert mesh.nblocks == nb
    for b in range(nb):
        blk = mesh.block(b)
        counts = g.z[f"g{b}_counts"]
        assert counts[0] == blk["nreal"] and counts[1] == blk["nall"] and counts[2] == blk["nface"]
        assert counts[4] == blk["n_real_boundaries"] and counts[6] == len(blk["wall_faces"])
        for key, arr in geometry_arrays(blk).items():
            g.check(f"g{b}_{key}", arr, "geometry")


def test_builder_bump_structure():
    """Closed-form counts of the 6-tet bump channel: F = (4N + G) / 2, G = 4(nx ny + nx nz + ny nz)."""
    import euler3d_amd as E
    nx, ny, nz = 6, 3, 2
    m = E.Mesh.bump(nx, ny, nz, 0.1, 4)
    b = m.block(0)
    N = 6 * nx * ny * nz
    G = 4 * (nx * ny + nx * nz + ny * nz)
    assert b["nreal"] == N and b["nall"] == N + G and b["nface"] == (4 * N + G) // 2
    # every real cell has 4 faces, every ghost 1; left < right on every face
    off = b["cell_off"]
    assert (np.diff(off)[:N] == 4).all() and (np.diff(off)[N:] == 1).all()
    f2c = b["face2cell"].reshape(-1, 2)
    assert (f2c[:, 0] < f2c[:, 1]).all()
    # cell volumes sum to the channel volume less the faceted bump: the lower wall
    # is piecewise linear in x between the arc's node heights and constant in z
    R = (0.25 + 0.01) / 0.2
    x = 3.0 * np.arange(nx + 1) / nx
    yb = np.where((x > 1) & (x < 2), np.sqrt(np.maximum(R * R - (x - 1.5) ** 2, 0)) - (R - 0.1), 0.0)
    bump_area = np.sum(0.5 * (yb[1:] + yb[:-1]) * np.diff(x))
    assert abs(b["volume"][:N].sum() - (3.0 - bump_area)) < 1e-12
    # closed cells: the area vectors of each cell sum to zero
    S = b["S"].reshape(-1, 3)
    for c in range(0, N, 7):
        fs = b["cell2face"][off[c]:off[c + 1]]
        sgn = np.where(f2c[fs, 0] == c, 1.0, -1.0)
        assert np.abs((S[fs] * sgn[:, None]).sum(axis=0)).max() < 1e-14


def test_multiblock_partition_topology():
    """8-block partition: every interface ghost is fed by a real cell across a cut face."""
    import euler3d_amd as E
    m = E.Mesh.bump(8, 4, 4, 0.1, 4, (2, 2, 2))
    blocks = [m.block(b) for b in range(8)]
    assert sum(b["nreal"] for b in blocks) == 6 * 8 * 4 * 4
    for d, b in enumerate(blocks):
        n_conn = sum(int(b["bnd_off"][j + 1] - b["bnd_off"][j]) for j in range(len(b["bnd_type"]))
                     if b["bnd_type"][j] == 3)
        assert n_conn == len(b["itf_origin_cells"])
        assert b["n_real_boundaries"] == sum(1 for t in b["bnd_type"] if t != 3)
        assert (np.asarray(b["bnd_type"])[b["n_real_boundaries"]:] == 3).all()  # CONNEXION last
        for i, o in enumerate(b["itf_origin_block"]):
            assert o != d
            cells = b["itf_origin_cells"][b["itf_off"][i]:b["itf_off"][i + 1]]
            assert cells.min() >= 0 and cells.max() < blocks[o]["nreal"]


def test_scrambled_bump_is_a_renumbering():
    """brick < 0 lists the same tets in a pseudo-random order: same counts,
    boundary lists and total volume; the same multiset of cell centroids; a
    different cell order; and the same order for the same seed."""
    import euler3d_amd as E
    a = E.Mesh.bump(6, 3, 3, 0.1, 0, (1, 1, 1)).block(0)
    b = E.Mesh.bump(6, 3, 3, 0.1, -11, (1, 1, 1)).block(0)
    b2 = E.Mesh.bump(6, 3, 3, 0.1, -11, (1, 1, 1)).block(0)
    N = a["nreal"]
    assert (b["nreal"], b["nall"], b["nface"]) == (N, a["nall"], a["nface"])
    assert b["n_real_boundaries"] == a["n_real_boundaries"] and len(b["wall_faces"]) == len(a["wall_faces"])
    ca = a["ccenter"].reshape(-1, 3)[:N]
    cb = b["ccenter"].reshape(-1, 3)[:N]
    key = lambda c: np.lexsort(np.round(c, 12).T[::-1])
    assert np.allclose(ca[key(ca)], cb[key(cb)], rtol=0, atol=1e-12)
    assert not np.allclose(ca, cb)
    assert abs(a["volume"][:N].sum() - b["volume"][:N].sum()) < 1e-12
    assert (b["ccenter"] == b2["ccenter"]).all()
