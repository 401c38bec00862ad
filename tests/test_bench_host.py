"""bench.py's host-side figures, on the CPU: the compulsory-traffic model of
the stage kernels (algorithmic bytes per launch and per cell-stage) against
its closed form on a generated mesh's own counts, and the streaming ceiling
interpolated between measured traffic mixes."""
import os
import sys

import pytest

import euler3d_amd as E

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_algorithmic_bytes_closed_form():
    m = E.Mesh.bump(12, 4, 4, 0.1, 4, (2, 1, 1))
    blocks = [m.block(b) for b in range(m.nblocks)]
    N = sum(b["nreal"] for b in blocks)
    G = sum(b["nall"] - b["nreal"] for b in blocks)
    F = sum(b["nface"] for b in blocks)
    alg, per_cell, n, g, f = bench.algorithmic_bytes(blocks)
    assert (n, g, f) == (N, G, F) and N == 12 * 4 * 4 * 6
    assert alg["interp"] == 208 * (N + G) + 80 * F + 16 * N
    assert alg["flux_update"] == 200 * (N + G) + 80 * F + 120 * N + 176 * N + 112 * G
    assert per_cell == pytest.approx((alg["interp"] + alg["flux_update"] + (232 * N + 24 * F) / 5) / N)
    lsq, _, _, _, _ = bench.algorithmic_bytes(blocks, lsq=True)
    assert lsq["interp"] == 200 * (N + G) + 128 * N + 56 * F + 24 * G
    first, pc1, _, _, _ = bench.algorithmic_bytes(blocks, order=1)
    assert first["interp"] == 0 and first["flux_update"] == 40 * (N + G) + 32 * F + 120 * N + 176 * N + 112 * G
    m.close()


def test_stream_rate_interpolates_between_measured_mixes():
    t = [{"write_fraction": 0.0, "GB/s": 7000.0}, {"write_fraction": 0.125, "GB/s": 6150.0},
         {"write_fraction": 0.5, "GB/s": 6550.0}, {"write_fraction": 1.0, "GB/s": 6650.0}]
    assert bench.stream_rate(t, 0.0) == 7000.0
    assert bench.stream_rate(t, 0.0625) == pytest.approx(6575.0)
    assert bench.stream_rate(t, 0.3125) == pytest.approx(6350.0)
    assert bench.stream_rate(t, 1.0) == 6650.0
    assert bench.stream_rate(t[1:3], 0.05) == 6150.0   # below the measured range: its lowest mix
    assert bench.stream_rate(t[1:3], 0.9) == 6550.0    # above it: its highest
