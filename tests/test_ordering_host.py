"""Device memory order computed on the host (e3d_mesh_layout): reverse
Cuthill-McKee or Morton-curve cells, faces by first appearance over them.

Checked against what the definitions fix: both maps are permutations; the
reference order is the identity; faces appear in increasing position when the
cells are walked in their new order; the order is deterministic; and on a
mesh numbered at random, RCM brings face neighbours close together again
(bandwidth and mean neighbour distance against a brute-force evaluation).
"""
import numpy as np
import pytest

import euler3d_amd as E


def neighbour_pairs(blk):
    f2c = blk["face2cell"].reshape(-1, 2)
    inner = (f2c[:, 0] < blk["nreal"]) & (f2c[:, 1] < blk["nreal"])
    return f2c[inner]


def spread(pos, pairs):
    d = np.abs(pos[pairs[:, 0]] - pos[pairs[:, 1]])
    return int(d.max()), float(d.mean())


def test_reference_order_is_identity():
    m = E.Mesh.bump(8, 3, 3, 0.1, 4, (1, 1, 1))
    cp, fp = m.layout(0, E.ORDER_REFERENCE)
    assert (cp == np.arange(len(cp))).all() and (fp == np.arange(len(fp))).all()


@pytest.mark.parametrize("mode", [E.ORDER_RCM, E.ORDER_MORTON], ids=["rcm", "morton"])
def test_order_is_a_permutation_with_faces_in_first_appearance_order(mode):
    m = E.Mesh.bump(10, 4, 3, 0.1, -3, (2, 1, 1))
    for b in range(m.nblocks):
        blk = m.block(b)
        cp, fp = m.layout(b, mode)
        assert sorted(cp) == list(range(blk["nreal"]))
        assert sorted(fp) == list(range(blk["nface"]))
        order = np.argsort(cp)  # reference id of the cell at each position
        off, c2f = blk["cell_off"], blk["cell2face"]
        first = []
        seen = set()
        for q in order:
            for f in c2f[off[q]:off[q + 1]]:
                if f not in seen:
                    seen.add(f)
                    first.append(fp[f])
        assert first == sorted(first) == list(range(len(first)))
        cp2, fp2 = m.layout(b, mode)
        assert (cp == cp2).all() and (fp == fp2).all()


def test_rcm_restores_locality_of_a_scrambled_mesh():
    ordered = E.Mesh.bump(24, 8, 8, 0.1, 0, (1, 1, 1))
    scrambled = E.Mesh.bump(24, 8, 8, 0.1, -5, (1, 1, 1))
    b_ord = ordered.block(0)
    b_scr = scrambled.block(0)
    ident = np.arange(b_scr["nreal"])
    bw_scr, mean_scr = spread(ident, neighbour_pairs(b_scr))
    cp, _ = scrambled.layout(0, E.ORDER_RCM)
    bw_rcm, mean_rcm = spread(cp, neighbour_pairs(b_scr))
    bw_ord, mean_ord = spread(np.arange(b_ord["nreal"]), neighbour_pairs(b_ord))
    # random numbering: neighbours anywhere; RCM: within a level set of the breadth-first sweep
    assert bw_scr > 0.9 * b_scr["nreal"] and mean_scr > 0.25 * b_scr["nreal"]
    assert bw_rcm < bw_scr / 20 and mean_rcm < mean_scr / 20
    assert bw_rcm <= 2 * bw_ord


def test_morton_follows_the_z_curve_of_the_centroids():
    """Brute force on a small scrambled block: interleave the bits of the
    quantised centroids in Python and sort."""
    m = E.Mesh.bump(6, 3, 3, 0.1, -7, (1, 1, 1))
    blk = m.block(0)
    n = blk["nreal"]
    c = blk["ccenter"].reshape(-1, 3)[:n]
    lo = c.min(axis=0)
    ext = (c.max(axis=0) - lo).max()
    q = ((c - lo) * ((2 ** 21 - 1) / ext)).astype(np.uint64)
    keys = []
    for i in range(n):
        k = 0
        for bit in range(21):
            for d in range(3):
                k |= ((int(q[i, d]) >> bit) & 1) << (3 * bit + d)
        keys.append((k, i))
    order = [i for _, i in sorted(keys)]
    cp, _ = m.layout(0, E.ORDER_MORTON)
    assert (np.argsort(cp) == np.array(order)).all()
    # and it restores locality like RCM does
    scr = E.Mesh.bump(24, 8, 8, 0.1, -5, (1, 1, 1))
    pairs = neighbour_pairs(scr.block(0))
    bw_scr, mean_scr = spread(np.arange(scr.block(0)["nreal"]), pairs)
    cpm, _ = scr.layout(0, E.ORDER_MORTON)
    assert spread(cpm, pairs)[1] < mean_scr / 20


def test_layout_rejects_bad_arguments():
    m = E.Mesh.bump(4, 2, 2, 0.1, 4, (1, 1, 1))
    with pytest.raises(E.E3DError):
        m.layout(0, 7)
    with pytest.raises(E.E3DError):
        m.layout(3, E.ORDER_RCM)
