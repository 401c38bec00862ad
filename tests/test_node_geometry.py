"""Node-rebuilt slot geometry (host side, no GPU).

On a mesh of tetrahedra the Green-Gauss stage kernel rebuilds each face
slot's area vector S and r vector from a per-cell copy of the cell's four node
coordinates and a slot code (e3d_mesh_tet_geometry, the same host routine the
upload runs before it enables that form). These tests check, on generated and
on the reference's own meshes:

* the routine accepts every tet block (its own bit-for-bit comparison with the
  stored metrics passes) and rejects a hexahedral mesh;
* the node copy is the cell's element nodes in element order, and each slot
  byte names three distinct nodes of the cell: those of the slot's face;
* rebuilding S and r in numpy from the copy and the codes alone gives the
  block's stored metrics (MetricsInitializer.cpp:64-142, 337-504) bit for
  bit: S oriented as the slot holds it (negated on the face's right cell),
  r = face centre - cell centre on the cell's own side. A wrong node order,
  index or sign bit in the codes fails this check.
* x / 3 by the three-operation FMA form the kernel uses equals the true
  quotient on every face-centre sum of the meshes (numpy's division is the
  correctly rounded quotient; each FMA of the form is evaluated exactly in
  rationals and rounded once).
"""
import tempfile
from fractions import Fraction

import numpy as np
import pytest

import euler3d_amd as E
from helpers import Golden, mesh_for, onera_mesh


def rebuilt_slot_geometry(xyz, code):
    """S [n x 4 x 3] and r [n x 4 x 3] from the node copies and slot codes."""
    n = len(xyz)
    P = xyz.reshape(n, 4, 3)
    S = np.empty((n, 4, 3))
    R = np.empty((n, 4, 3))
    cc = (((0.0 + P[:, 0]) + P[:, 1]) + P[:, 2]) + P[:, 3]
    cc = cc / 4.0
    rows = np.arange(n)
    for k in range(4):
        byte = (code >> np.uint32(8 * k)) & np.uint32(0xFF)
        a, b, c = (byte & 3).astype(int), ((byte >> 2) & 3).astype(int), ((byte >> 4) & 3).astype(int)
        neg = (byte & 64) != 0
        p1, p2, p3 = P[rows, a], P[rows, b], P[rows, c]
        x1, y1, z1 = p1.T
        x2, y2, z2 = p2.T
        x3, y3, z3 = p3.T
        dxy = ((x1 - x2) * (y1 + y2), (x2 - x3) * (y2 + y3), (x3 - x1) * (y3 + y1))
        dyz = ((y1 - y2) * (z1 + z2), (y2 - y3) * (z2 + z3), (y3 - y1) * (z3 + z1))
        dzx = ((z1 - z2) * (x1 + x2), (z2 - z3) * (x2 + x3), (z3 - z1) * (x3 + x1))
        sg = np.where(neg, -1.0, 1.0)
        S[:, k, 0] = 0.5 * (dyz[0] + dyz[1] + dyz[2]) * sg
        S[:, k, 1] = 0.5 * (dzx[0] + dzx[1] + dzx[2]) * sg
        S[:, k, 2] = 0.5 * (dxy[0] + dxy[1] + dxy[2]) * sg
        fc = ((0.0 + p1) + p2) + p3
        R[:, k] = fc / 3.0 - cc
    return S, R


def stored_slot_geometry(blk):
    """The block's stored S (as the slot holds it) and r of each real cell's four slots."""
    n = blk["nreal"]
    off = blk["cell_off"][: n + 1]
    assert np.all(np.diff(off) == 4)
    f = blk["cell2face"][: 4 * n].reshape(n, 4)
    right = blk["face2cell"].reshape(-1, 2)[f, 0] != np.arange(n)[:, None]
    S = blk["S"].reshape(-1, 3)[f] * np.where(right, -1.0, 1.0)[..., None]
    R = np.where(right[..., None], blk["rR"].reshape(-1, 3)[f], blk["rL"].reshape(-1, 3)[f])
    return S, R


def check_mesh(mesh):
    for b in range(mesh.nblocks):
        blk = mesh.block(b)
        xyz, code = mesh.tet_geometry(b)
        n = blk["nreal"]
        # the copy is the element's nodes in element order
        cn = blk["cell_nodes"][blk["cell_node_off"][0]: blk["cell_node_off"][n]].reshape(n, 4)
        assert np.array_equal(xyz, blk["node_xyz"].reshape(-1, 3)[cn].reshape(n, 12))
        # each slot byte names three distinct local nodes: the face's nodes in the face's order
        f = blk["cell2face"][: 4 * n].reshape(n, 4)
        fn = blk["face_nodes"].reshape(-1, 4)[f][..., :3]
        for k in range(4):
            byte = (code >> np.uint32(8 * k)) & np.uint32(0xFF)
            loc = np.stack([(byte >> np.uint32(2 * i)) & np.uint32(3) for i in range(3)], axis=1).astype(int)
            assert np.all(byte < 128)
            assert np.all(loc[:, 0] != loc[:, 1]) and np.all(loc[:, 1] != loc[:, 2]) and np.all(loc[:, 0] != loc[:, 2])
            assert np.array_equal(cn[np.arange(n)[:, None], loc], fn[:, k])
        S, R = rebuilt_slot_geometry(xyz, code)
        Ss, Rs = stored_slot_geometry(blk)
        assert np.array_equal(S.view(np.int64), Ss.view(np.int64)), f"block {b}: S differs"
        assert np.array_equal(R.view(np.int64), Rs.view(np.int64)), f"block {b}: r differs"


@pytest.mark.parametrize("shape", [(12, 4, 4, 0.1, 4, (1, 1, 1)), (24, 8, 8, 0.06, 0, (2, 2, 2)),
                                   (18, 6, 6, 0.12, -7, (3, 1, 2))])
def test_generated_meshes(shape):
    nx, ny, nz, h, brick, parts = shape
    mesh = E.Mesh.bump(nx, ny, nz, h, brick, parts)
    check_mesh(mesh)
    mesh.close()


def test_golden_fixture_meshes():
    for case in ("bump1_roe2_gg_vk", "bump8_roe2_gg_vk"):
        mesh = mesh_for(Golden(case))
        check_mesh(mesh)
        mesh.close()


def test_reference_onera_m6_mesh():
    """The reference's 120-block ONERA M6 mesh as its partitioner wrote it."""
    with tempfile.TemporaryDirectory() as d:
        mesh, _ = onera_mesh(d)
        check_mesh(mesh)
        mesh.close()


def test_hexahedral_mesh_is_refused():
    mesh = mesh_for(Golden("naca_hex_roe2_gg_vk"))
    with pytest.raises(E.E3DError):
        mesh.tet_geometry(0)
    mesh.close()


def fma(a, b, c):
    """a * b + c rounded once (exact rational arithmetic; float() of a Fraction rounds to nearest)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def div3_fma(x):
    z = 1.0 / 3.0
    q = x * z
    r = fma(-q, 3.0, x)
    return fma(r, z, q)


def test_fma_division_by_three():
    """Face-centre sums of a scrambled 2-block mesh and 40,000 random doubles of
    every exponent: the FMA form equals the correctly rounded quotient."""
    rng = np.random.default_rng(7)
    mesh = E.Mesh.bump(12, 4, 4, 0.1, -3, (2, 1, 1))
    sums = []
    for b in range(mesh.nblocks):
        xyz, code = mesh.tet_geometry(b)
        P = xyz.reshape(-1, 4, 3)
        for k in range(4):
            byte = (code >> np.uint32(8 * k)) & np.uint32(0xFF)
            idx = [((byte >> np.uint32(2 * i)) & np.uint32(3)).astype(int) for i in range(3)]
            rows = np.arange(len(P))
            sums.append((((0.0 + P[rows, idx[0]]) + P[rows, idx[1]]) + P[rows, idx[2]]).ravel())
    mesh.close()
    bits = rng.integers(0, 2 ** 63, size=20000, dtype=np.int64) & np.int64(0x7FEFFFFFFFFFFFFF)
    vals = np.concatenate(sums + [bits.view(np.float64), -bits.view(np.float64)])
    vals = vals[np.isfinite(vals)]
    bad = [x for x in vals.tolist() if div3_fma(x) != x / 3.0]
    assert not bad, f"{len(bad)} quotients differ, e.g. {bad[:3]}"
