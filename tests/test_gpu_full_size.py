"""Parity at the full sizes of the measured workloads, against the oracle.

* 192x64x64x6 = 4,718,592 tets in 8 blocks (the throughput workload):
  2nd-order Roe + Green-Gauss + Venkatakrishnan (K = 5), M = 0.675; two fused
  RK iterations.
* 384x128x128x6 = 37,748,736 tets, 2nd-order AUSM + least squares +
  Barth-Jespersen, M = 0.7 at 3 deg; one fused RK iteration. The mesh is split
  in 8 blocks so that the oracle runs one thread per block (the device runs
  every block in the same launches either way).

Every cell's state and dissipation must equal the oracle's as values (+0 ==
-0), the residual norms and force coefficients within 1e-12 relative. At
these sizes the per-slot geometry, face-state records and 32-bit offsets
span hundreds of megabytes to gigabytes, which the small parity cases do not
reach. The oracle runs on the host's cores (at most 16 threads).
"""
import os
import time

import numpy as np
import pytest

import euler3d_amd as E
import oracle as O
from helpers import values_equal

pytestmark = pytest.mark.gpu

CASES = {
    "bump192x64x64_8blocks_roe_gg_vk": dict(n=(192, 64, 64), mach=0.675, aoa=0.0, flux=E.FLUX_ROE,
                                            gradient=E.GRAD_GREEN_GAUSS, limiter=E.LIM_VENKATAKRISHNAN, iters=2),
    "bump384x128x128_8blocks_ausm_lsq_bj": dict(n=(384, 128, 128), mach=0.7, aoa=3.0, flux=E.FLUX_AUSM,
                                                gradient=E.GRAD_LEAST_SQUARES, limiter=E.LIM_BARTH_JESPERSEN,
                                                iters=1),
}


def _rel_close(a, b, rtol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return bool(np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("name", list(CASES))
def test_full_size_fused_iterations_match_oracle(name):
    c = CASES[name]
    t0 = time.perf_counter()
    print(f"{name}: building the mesh", flush=True)
    mesh = E.Mesh.bump(*c["n"], 0.1, 4, (2, 2, 2))
    common = dict(flux=c["flux"], order=2, gradient=c["gradient"], limiter=c["limiter"], omega=1e-16, k=5.0,
                  stages=5, max_iter=c["iters"], criterion=-15.0)
    ctx = E.Context(0)
    ctx.set_params(E.make_params(c["mach"], c["aoa"], **common))
    ctx.upload(mesh)
    ctx.init_freestream()
    hist = ctx.iterate(c["iters"])
    print(f"{name}: device done ({time.perf_counter() - t0:.1f}s); oracle", flush=True)
    S = O.Solver([mesh.block(b) for b in range(mesh.nblocks)], O.params(c["mach"], c["aoa"], **common))
    S.init_freestream()
    rh = S.iterate(c["iters"], threads=min(mesh.nblocks, os.cpu_count() or 1, 16))
    print(f"{name}: oracle done ({time.perf_counter() - t0:.1f}s)", flush=True)
    for b, B in enumerate(S.blocks):
        W = ctx.get(b, E.ARR_W).reshape(-1)
        same = values_equal(W, B.W)
        assert same.all(), f"block {b}: {np.sum(~same)} state values differ, first {np.flatnonzero(~same)[:5]}"
        d = ctx.get(b, E.ARR_DISS).reshape(-1)
        dref = B.diss.reshape(5, B.nall)[:, :B.nreal].reshape(-1)
        same = values_equal(d, dref)
        assert same.all(), f"block {b}: {np.sum(~same)} dissipation values differ"
    assert len(hist) == len(rh)
    for a, r in zip(hist, rh):
        assert a["iter"] == r["iter"]
        assert _rel_close(a["rms"], r["rms"]), (a["iter"], a["rms"], r["rms"])
        assert _rel_close([a["cl"], a["cd"], a["cmx"], a["cmy"], a["cmz"]],
                          [r["cl"], r["cd"], r["cmx"], r["cmy"], r["cmz"]]), a["iter"]
    ctx.close()
    mesh.close()
