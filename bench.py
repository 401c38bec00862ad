#!/usr/bin/env python3
"""Throughput of the residual-and-update path on one MI355X.

Workload (default): the bump channel at 192x64x64 hexes x 6 = 4,718,592 tets,
split into 8 blocks (2x2x2) whose interfaces are exchanged on the device;
second-order Roe with Green-Gauss gradients and the Venkatakrishnan limiter
(K=5), M=0.675, 5-stage RK, local CFL 1.5, FP64. A step is one RK iteration
(5 stages, the interface exchange and the residual/force monitors) over the
whole mesh, launched through e3d_iterate; the mesh and state are resident in
HBM before the timed region starts. By default the timed region is the whole
1,000-iteration run from the freestream start (after 3 untimed iterations):
the flux kernels grow slower as the flow develops (the lanes of a wavefront
stop agreeing on the branches of the Roe entropy fix), so a short run at the
start would overstate the rate.

metric: Mcell-RK-stage/s = real cells x RK stages x steps / wall time.
roofline: the kernel with the largest share of the timed region (interp:
gradient + limiter + face states, with the ghosts' face states in a second
small kernel when the slot geometry is rebuilt from node copies, the two timed
as one; or flux_update: residual + RK update + BCs on a stage with
dissipation), algorithmic bytes per launch / average launch time from HIP
events recorded on the context's stream during the timed region;
"rooflines" carries the same figures for both kernels.
cpu_baseline: the oracle (plain-C restatement of the reference's solver loop)
on the host cores, on a bounded sample of the same workload: all 8 blocks, one
thread per block (the reference's MPI decomposition), 4 RK iterations;
cpu_baseline_1core: block 0 (589,824 cells) alone on one core, 4 iterations;
cpu_baseline_reference: the reference's own serial binary (oracle/_ref) on a
9,216-cell channel of the same scheme, one core, its solve timer.

Other workloads: --scheme ausm-lsq-bj --nx 384 --ny 128 --nz 128 --parts 1 1 1
is the 37,748,736-cell single block with AUSM, least squares and
Barth-Jespersen; --brick -S numbers the tets in a pseudo-random order (seed S)
and --order rcm / morton renumbers cells and faces in device memory by reverse
Cuthill-McKee or along a Morton curve (the same results bit for bit).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE) for the two
# benchmark workloads: (nx, ny, nz, parts, scheme, brick, order) -> summary file.
TRAFFIC_FILES = {
    (192, 64, 64, (2, 2, 2), "roe-gg-vk", 4, "reference"): "fused_kernels_bump192x64x64_traffic.json",
    (384, 128, 128, (1, 1, 1), "ausm-lsq-bj", 4, "reference"): "fused_kernels_bump384x128x128_ausm_lsq_bj_traffic.json",
}

# Schemes of the benchmark workloads: 4.7M-cell channel in 8 blocks (Roe,
# Green-Gauss, Venkatakrishnan K=5, M=0.675) and the 37.7M-cell single block
# (AUSM, least squares, Barth-Jespersen omega=1e-16, M=0.7, inlet angle 3 deg).
SCHEMES = {
    "roe-gg-vk": dict(mach=0.675, aoa=0.0, flux=0, gradient=0, limiter=1, k=5.0, omega=1e-16,
                      label="2nd-order Roe + Green-Gauss + Venkatakrishnan K=5, M=0.675"),
    "ausm-lsq-bj": dict(mach=0.7, aoa=3.0, flux=1, gradient=1, limiter=0, k=5.0, omega=1e-16,
                        label="2nd-order AUSM + least squares + Barth-Jespersen omega=1e-16, M=0.7, 3 deg"),
    "roe-1st": dict(mach=0.5, aoa=0.0, flux=0, gradient=0, limiter=1, k=5.0, omega=1e-16, order=1,
                    label="1st-order Roe, M=0.5"),
    # the reference's own ONERA M6 input (input_onera120)
    "roe-lsq-vk": dict(mach=0.8, aoa=1.25, flux=0, gradient=1, limiter=1, k=5.0, omega=2.16e-16,
                       label="2nd-order Roe + least squares + Venkatakrishnan K=5, M=0.8, 1.25 deg"),
}
ONERA_FIXTURE = os.path.join(ROOT, "tests", "golden", "onera_m6_120b.npz")


def onera_mesh(E):
    """The reference's ONERA M6 mesh in 120 blocks (582,752 tets, its own
    partitioner's output), written from the byte-for-byte copy kept with the
    tests into a scratch directory and read by this build's reader."""
    import tempfile

    import numpy as np
    d = tempfile.mkdtemp(prefix="e3d_onera_")
    z = np.load(ONERA_FIXTURE, allow_pickle=False)
    mdir = os.path.join(d, "Meshes", "oneraM6_120B")
    os.makedirs(mdir)
    for k in z.files:
        if k.startswith("file:"):
            with open(os.path.join(mdir, k[5:]), "wb") as f:
                f.write(z[k].tobytes())
    cwd = os.getcwd()
    os.chdir(d)  # the topology names its block files relative to the reference tree's root
    try:
        return E.Mesh.read(os.path.join(mdir, "topology_M6.top"))
    finally:
        os.chdir(cwd)


def algorithmic_bytes(blocks, lsq=False, order=2):
    """Algorithmic bytes per launch of the two stage kernels, and per cell-stage of the whole step.

    The compulsory-traffic model of this path: every array element counted once
    per kernel that must read or write it (W 40 B/cell, grad 120, lim 40, W0 40,
    conv+diss 80, dt 8, V 8; per face S 24 + rL, rR 48 + left/right 8 = 80 B).
    interp is the gradient + limiter kernel K_A = 208(N+G) + 80F + 16N with
    Green-Gauss, 200(N+G) + 128N + 56F + 24G with least squares (weights 96 B,
    cell-to-cell and cell-to-face lists 16 B each per cell; faces only for r);
    flux_update does the residual kernel K_B = 200(N+G) + 80F + 120N and the
    update + BC kernel K_C = 176N + 112G of that model in one launch. Once per
    iteration: saveW0, spectral radius, time step and monitors, 232N + 24F.
    First order has no interp kernel and its flux kernel reads W of both sides
    instead of face states: 40(N+G) + 32F + 120N, plus K_C."""
    N = sum(b["nreal"] for b in blocks)
    G = sum(b["nall"] - b["nreal"] for b in blocks)
    F = sum(b["nface"] for b in blocks)
    ka = 200 * (N + G) + 128 * N + 56 * F + 24 * G if lsq else 208 * (N + G) + 80 * F + 16 * N
    kb = 200 * (N + G) + 80 * F + 120 * N if order == 2 else 40 * (N + G) + 32 * F + 120 * N
    kc = 176 * N + 112 * G
    if order != 2:
        ka = 0
    per_cell_stage = (ka + kb + kc + (232 * N + 24 * F) / 5) / N
    return {"interp": ka, "flux_update": kb + kc}, per_cell_stage, N, G, F


def cpu_baselines(a, sc, blocks):
    """The oracle (plain-C restatement of the reference's solver loop) on the
    host cores, on a bounded sample of the workload: all blocks of the mesh
    with one thread per block in contiguous ranges (the reference's MPI
    decomposition, one exchange per iteration), and block 0 alone on one core.
    Meshes above 5M cells are sampled by the same scheme on the 96x32x32x6
    channel in 8 blocks (589,824 cells)."""
    import oracle as O
    import euler3d_amd as E
    sample = f"the same mesh ({sum(b['nreal'] for b in blocks)} cells in {len(blocks)} blocks)"
    if sum(b["nreal"] for b in blocks) > 5_000_000:
        m = E.Mesh.bump(96, 32, 32, 0.1, 4, (2, 2, 2))
        blocks = [m.block(b) for b in range(m.nblocks)]
        sample = "the same scheme on the 96x32x32x6 channel (589,824 cells in 8 blocks)"
    prm = O.params(sc["mach"], sc["aoa"], flux=sc["flux"], order=sc.get("order", 2), gradient=sc["gradient"],
                   limiter=sc["limiter"],
                   k=sc["k"], omega=sc["omega"], stages=5, max_iter=10 ** 9, criterion=-1e9)
    threads = max(1, min(len(blocks), os.cpu_count() or 1, 16))
    S = O.Solver(blocks, prm)
    S.init_freestream()
    c0 = time.perf_counter()
    S.iterate(a.cpu_iters, threads=threads)
    c1 = time.perf_counter()
    n_all = sum(b["nreal"] for b in blocks)
    cpu = {"value": n_all * 5 * a.cpu_iters / (c1 - c0) / 1e6, "unit": "Mcell-RK-stage/s", "cores": threads,
           "kind": "port",
           "sample": f"oracle (plain-C restatement) on {sample}, one thread per block in contiguous ranges "
                     f"({threads} threads), {a.cpu_iters} RK iteration(s) of 5 stages, {c1 - c0:.1f}s"}
    print(f"[bench] cpu baseline {threads} threads: {cpu['value']:.3f} Mcell-RK-stage/s", file=sys.stderr,
          flush=True)
    b0 = blocks[0]
    S = O.Solver([b0], prm, interfaces=False)
    S.init_freestream()
    c0 = time.perf_counter()
    S.iterate(a.cpu_iters)
    c1 = time.perf_counter()
    cpu1 = {"value": b0["nreal"] * 5 * a.cpu_iters / (c1 - c0) / 1e6, "unit": "Mcell-RK-stage/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle on block 0 of {sample} ({b0['nreal']} cells), 1 thread, {a.cpu_iters} RK "
                      f"iteration(s), {c1 - c0:.1f}s"}
    print(f"[bench] cpu baseline 1 thread: {cpu1['value']:.3f} Mcell-RK-stage/s", file=sys.stderr, flush=True)
    return cpu, cpu1


FLUX_NAMES = {0: "roe", 1: "ausm"}
GRAD_NAMES = {0: "greengauss", 1: "leastsquares"}
LIM_NAMES = {0: "barthjespersen", 1: "venkatakrishnan"}


def reference_baseline(sc, nx=24, iters=40):
    """The reference's own serial binary (compiled from its sources into
    oracle/_ref by oracle/Makefile; absent: None) on a small channel of the same
    scheme, one core: its mesh setup is quadratic in the face count (about 17 s
    at 9,216 cells), so only its own solve timer (main.cpp:89-95) is used."""
    import re
    import subprocess
    import tempfile
    import euler3d_amd as E
    exe = os.path.join(ROOT, "oracle", "_ref", "euler3D")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        mesh = E.Mesh.bump(nx, nx // 3, nx // 3, 0.1, 4, (1, 1, 1))
        n = mesh.block(0)["nreal"]
        top = mesh.write(d, "bump")
        lines = ["MESH", "topologyfilename", top, "INPUT", "cfl gamma angleattackdeg", f"1.5 1.4 {sc['aoa']!r}",
                 "rkstage mach cmac", f"5 {sc['mach']!r} 1.0", "nbitermax convcriterion", f"{iters} -1e9",
                 "residual smoothing (0-no 1-yes)", "0", "fluxscheme schemeorder",
                 f"{FLUX_NAMES[sc['flux']]} {sc.get('order', 2)}", "gradient limiter",
                 f"{GRAD_NAMES[sc['gradient']]} {LIM_NAMES[sc['limiter']]}", "omega k",
                 f"{sc['omega']!r} {sc['k']!r}", "EXECUTABLE (0-no 1-yes)", "build execute", "1 1",
                 "OUTPUT (0-no 1-yes)", "clalpha cdalpha cmalpha", "0 0 0", "coefficientsconv  residualconv", "1 1",
                 "cpxc machcontour cpcontour machisosurface", "0 0 0 0", "slicecp", "0",
                 "axis xcoord ycoord zcoord", "x 1.00 1.00 1.00"]
        with open(os.path.join(d, "input"), "w") as f:
            f.write("\n".join(lines) + "\n")
        os.makedirs(os.path.join(d, "bin"))
        print(f"[bench] reference binary on the {nx}x{nx // 3}x{nx // 3}x6 channel ({n} cells), {iters} iterations",
              file=sys.stderr, flush=True)
        try:
            r = subprocess.run([exe, "input"], cwd=d, capture_output=True, text=True, timeout=600)
        except (OSError, subprocess.TimeoutExpired) as e:
            print(f"[bench] reference binary did not run: {e}", file=sys.stderr, flush=True)
            return None
        m = re.search(r"Time elapsed solving: ([0-9.eE+-]+)s", r.stdout)
        if r.returncode != 0 or not m:
            print(f"[bench] reference binary failed (status {r.returncode})", file=sys.stderr, flush=True)
            return None
        t = float(m.group(1))
    # the oracle on the identical sample, for the oracle-to-reference speed ratio
    import oracle as O
    S = O.Solver([mesh.block(0)], O.params(sc["mach"], sc["aoa"], flux=sc["flux"], order=sc.get("order", 2),
                                          gradient=sc["gradient"], limiter=sc["limiter"], k=sc["k"],
                                          omega=sc["omega"], stages=5, max_iter=10 ** 9, criterion=-1e9))
    S.init_freestream()
    c0 = time.perf_counter()
    S.iterate(iters)
    t_or = time.perf_counter() - c0
    out = {"value": n * 5 * iters / t / 1e6, "unit": "Mcell-RK-stage/s", "cores": 1, "kind": "reference",
           "oracle_same_sample": n * 5 * iters / t_or / 1e6,
           "sample": f"the reference's serial binary (oracle/_ref, built from its sources) on the same scheme on the "
                     f"{nx}x{nx // 3}x{nx // 3}x6 channel ({n} cells, one block), {iters} RK iterations in {t:.2f}s "
                     f"by its own solve timer; its O(F^2) mesh setup excluded"}
    print(f"[bench] reference binary: {out['value']:.3f} Mcell-RK-stage/s", file=sys.stderr, flush=True)
    return out


STREAM_MIXES = [(1, 0), (7, 1), (3, 1), (2, 1), (1, 1), (1, 2), (0, 1)]  # planes read : planes written


def stream_bandwidth(gib=4, reps=10):
    """Achievable HBM bandwidth on this GPU by traffic mix, with the library's
    own streaming kernel (e3d_stream_bandwidth: 16-byte nontemporal loads and
    stores, consecutive lanes on consecutive addresses) moving gib GiB per
    launch, HIP events around reps launches after 3 warmup launches."""
    import ctypes

    import euler3d_amd as E
    out = []
    for r, w in STREAM_MIXES:
        g = ctypes.c_double()
        E._check(E.lib().e3d_stream_bandwidth(0, gib << 30, r, w, reps, ctypes.byref(g)))
        out.append({"reads": r, "writes": w, "write_fraction": w / (r + w), "GB/s": g.value,
                    "frac_of_peak": g.value / HBM_PEAK_GBS})
        print(f"[bench] stream {r}:{w} (writes {w / (r + w):.3f}): {g.value:.0f} GB/s = "
              f"{g.value / HBM_PEAK_GBS:.3f} of peak", file=sys.stderr, flush=True)
    return out


def stream_rate(table, write_fraction):
    """Streaming rate of a traffic mix: linear in the write fraction between the
    two measured mixes around it."""
    pts = sorted((m["write_fraction"], m["GB/s"]) for m in table)
    if write_fraction <= pts[0][0]:
        return pts[0][1]
    for (x0, y0), (x1, y1) in zip(pts, pts[1:]):
        if x0 <= write_fraction <= x1:
            return y0 + (y1 - y0) * (write_fraction - x0) / (x1 - x0)
    return pts[-1][1]


def scrambled_leg(a):
    """The workload's mesh with its tets in pseudo-random file order (no
    locality in the file's numbering, as from a mesher that does not reorder),
    uploaded in the Morton memory order, against the generator's numbering in
    file order, both with AUSM + least squares + Barth-Jespersen: with
    Venkatakrishnan's limiter the reference method stops on the scrambled
    numbering within three iterations (1 / d2 overflows for a subnormal d2,
    Venkatakrishnan.cpp:227-238)."""
    import euler3d_amd as E
    t0 = time.perf_counter()
    sc = SCHEMES["ausm-lsq-bj"]
    params = E.make_params(sc["mach"], sc["aoa"], flux=sc["flux"], order=2, gradient=sc["gradient"],
                           limiter=sc["limiter"], k=sc["k"], omega=sc["omega"], stages=5, max_iter=10 ** 9,
                           criterion=-1e9)
    runs = {}
    for label, brick, order in (("generator numbering, file order", 4, E.ORDER_REFERENCE),
                                ("pseudo-random numbering (seed 5), Morton order", -5, E.ORDER_MORTON)):
        mesh = E.Mesh.bump(a.nx, a.ny, a.nz, 0.1, brick, tuple(a.parts))
        ctx = E.Context(0)
        ctx.set_params(params)
        ctx.set_ordering(order)
        ctx.upload(mesh)
        ctx.set_tile_cells(a.tile_cells)
        ctx.init_freestream()
        ctx.iterate(3, history=False)
        ctx.op("sync")
        E.lib().e3d_enable_timing(ctx._h, 1)
        steps = 10
        t1 = time.perf_counter()
        ctx.iterate(steps, history=False)
        ctx.op("sync")
        wall = time.perf_counter() - t1
        n = sum(mesh.block(b)["nreal"] for b in range(mesh.nblocks))
        runs[label] = {"ms_per_step": wall / steps * 1e3, "value": n * 5 * steps / wall / 1e6,
                       "kernels_avg_ms": {k: round(ctx.kernel_time(k)[0], 4)
                                          for k in ("interp", "flux_update0", "flux_update", "flux_update_nodiss")}}
        ctx.close()
        mesh.close()
    ms = [r["ms_per_step"] for r in runs.values()]
    out = {"workload": f"bump channel {a.nx}x{a.ny}x{a.nz}x6 tets in {a.parts} blocks, {sc['label']}, 5-stage RK",
           "runs": runs, "scrambled_over_generator": ms[1] / ms[0], "steps": 10, "warmup": 3,
           "wall_seconds": time.perf_counter() - t0}
    print(f"[bench] scrambled mesh, Morton order: {ms[1]:.3f} ms per iteration, {ms[1] / ms[0]:.2f}x the generator "
          f"numbering's {ms[0]:.3f} ms (AUSM/LSQ/BJ)", file=sys.stderr, flush=True)
    return out


# Warm-start leg of the bench line: the learned warm start against the cold
# (freestream) start on a configuration that fits the bench's time budget.
WARM_START = dict(nx=48, n_train=256, train_criterion=-6.0, n_heldout=16, criterion=-8.0, max_iter=30000,
                  steps=20000, width=512, boost=2, deeper=3.0)


def warm_start_leg():
    """Time to an 8-decade residual drop from the surrogate's field vs from
    freestream: training fields from the batched solver, the networks trained
    on the GPU in the run, held-out conditions run cold and warm (drop measured
    from the cold run's iteration-0 residual), both then converged three decades
    further to compare Cl and Cd."""
    from euler3d_amd import warmstart as ws
    c = WARM_START
    t0 = time.perf_counter()
    s = ws.run(n_train=c["n_train"], n_heldout=c["n_heldout"], nx=c["nx"], ny=c["nx"] // 3, nz=c["nx"] // 3, order=1,
               max_iter=c["max_iter"], criterion=c["criterion"], steps=c["steps"], width=c["width"],
               deeper=c["deeper"], train_criterion=c["train_criterion"], boost=c["boost"])
    both = [r for r in s["heldout"] if r["cold_converged"] and r["warm_converged"]]
    out = {
        "config": f"bump channel {c['nx']}x{c['nx'] // 3}x{c['nx'] // 3}x6 tets, 1st-order Roe, 5-stage RK, CFL 1.5; "
                  f"{c['n_train']} training fields (seeds 0-{c['n_train'] - 1}) to a {-c['train_criterion']:g}-decade "
                  f"drop; {1 + c['boost']} networks 25 -> 5x{c['width']} SiLU -> 5 (BF16 MFMA, FP32 accumulation), "
                  f"{c['steps']} Adam steps each; {c['n_heldout']} held-out conditions (seeds 5000-"
                  f"{5000 + c['n_heldout'] - 1})",
        "drop": c["criterion"], "drop_reference": "cold start's iteration-0 residual of the same condition",
        "converged_both": len(both), "heldout": c["n_heldout"],
        "cold_iterations": sum(r["cold_iterations"] for r in both),
        "warm_iterations": sum(r["warm_iterations"] for r in both),
        "iterations_ratio": s["iterations_ratio"],
        "cold_seconds": sum(r["cold_seconds"] for r in both),
        "warm_seconds": sum(r["warm_seconds"] for r in both),
        "time_ratio": s["time_ratio"],
        "time_includes": "device time of each instance in its batch; warm adds its share of the surrogate's inference",
        "iterations_ratio_own_r0": s["iterations_ratio_own_r0"],
        "max_cl_rel_diff_at_drop": s["max_cl_rel_diff"], "max_cd_rel_diff_at_drop": s["max_cd_rel_diff"],
        "max_cl_rel_diff_converged": s["max_cl_deep_rel_diff"], "max_cd_rel_diff_converged": s["max_cd_deep_rel_diff"],
        "converged_drop": c["criterion"] - c["deeper"],
        "train_seconds": s["train_seconds"], "generate_seconds": s["generate_seconds"],
        "train_tflops": s["train_tflops"], "wall_seconds": time.perf_counter() - t0,
    }
    print(f"[bench] warm start: {out['converged_both']}/{c['n_heldout']} held-out reached {c['criterion']:g}; "
          f"cold/warm iterations {out['iterations_ratio']:.2f}x, time {out['time_ratio']:.2f}x; Cl/Cd rel diff "
          f"{out['max_cl_rel_diff_converged']}/{out['max_cd_rel_diff_converged']} at {out['converged_drop']:g}; "
          f"{out['wall_seconds']:.0f}s", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--parts", type=int, nargs=3, default=[2, 2, 2])
    ap.add_argument("--scheme", choices=sorted(SCHEMES), default="roe-gg-vk")
    ap.add_argument("--mesh", choices=["bump", "onera"], default="bump",
                    help="onera: the reference's ONERA M6 mesh in 120 blocks with its own input's scheme "
                         "(implies --scheme roe-lsq-vk)")
    ap.add_argument("--brick", type=int, default=4,
                    help="cell numbering of the generated mesh: bricks of brick^3 hexes (0 lexicographic), "
                         "or < 0 for tets in a pseudo-random order seeded by -brick")
    ap.add_argument("--order", choices=["reference", "rcm", "morton"], default="reference",
                    help="device memory order of cells and faces (results are the same bit for bit)")
    ap.add_argument("--graph", action="store_true",
                    help="advance through the batch API with one instance: stop test on the device and a "
                         "hipGraph of two iterations replayed (no per-kernel timing)")
    ap.add_argument("--tile-cells", type=int, choices=[64, 128, 256], default=256,
                    help="cells per workgroup of the stage kernels (results are the same bit for bit)")
    ap.add_argument("--geometry", choices=["auto", "streamed"], default="auto",
                    help="Green-Gauss slot geometry: rebuilt from per-cell node copies on tet meshes (auto) or "
                         "streamed from the stored per-slot copies (results are the same bit for bit)")
    ap.add_argument("--no-scrambled", action="store_true",
                    help="skip the scrambled-mesh leg (the same mesh with its cells in pseudo-random file order, "
                         "Morton memory order)")
    ap.add_argument("--cpu-iters", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the warm-start leg (time to an 8-decade drop, learned warm start vs cold start)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, one 1 GiB device-to-device copy: the reference kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes")
    ap.add_argument("--stream-bw", action="store_true",
                    help="only measure the achievable HBM bandwidth of streams of 7 read/write mixes and exit")
    a = ap.parse_args()
    if a.stream_bw:
        print(json.dumps({"stream_bandwidth": stream_bandwidth(), "peak_GB/s": HBM_PEAK_GBS}))
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (blocks would shard "
              "across GPUs by contiguous ranges, but that path is not built)", file=sys.stderr)
        sys.exit(2)

    import euler3d_amd as E

    if a.mesh == "onera":
        a.scheme = "roe-lsq-vk"
    sc = SCHEMES[a.scheme]
    t_mesh = time.perf_counter()
    if a.mesh == "onera":
        print("[bench] reading the ONERA M6 mesh (120 blocks)", file=sys.stderr, flush=True)
        mesh = onera_mesh(E)
    else:
        print(f"[bench] building the {a.nx}x{a.ny}x{a.nz}x6 mesh in {a.parts} blocks", file=sys.stderr, flush=True)
        mesh = E.Mesh.bump(a.nx, a.ny, a.nz, 0.1, a.brick, tuple(a.parts))
    blocks = [mesh.block(b) for b in range(mesh.nblocks)]
    t_mesh = time.perf_counter() - t_mesh
    order = sc.get("order", 2)
    alg, model_bytes, N, G, F = algorithmic_bytes(blocks, lsq=sc["gradient"] == 1, order=order)
    params = E.make_params(sc["mach"], sc["aoa"], flux=sc["flux"], order=order, gradient=sc["gradient"],
                           limiter=sc["limiter"], k=sc["k"], omega=sc["omega"], stages=5, max_iter=10 ** 9,
                           criterion=-1e9)
    ctx = E.Context(0)
    ctx.set_params(params)
    ctx.set_ordering({"reference": E.ORDER_REFERENCE, "rcm": E.ORDER_RCM, "morton": E.ORDER_MORTON}[a.order])
    print(f"[bench] mesh built in {t_mesh:.1f}s; uploading", file=sys.stderr, flush=True)
    t_up = time.perf_counter()
    if a.graph:
        ctx.upload_batch([mesh], [dict(mach=sc["mach"], aoa_deg=sc["aoa"], criterion=-1e9, max_iter=10 ** 9)])
    else:
        ctx.upload(mesh)
    ctx.set_tile_cells(a.tile_cells)
    ctx.set_slot_geometry(E.GEOMETRY_AUTO if a.geometry == "auto" else E.GEOMETRY_STREAMED)
    slot_geometry = "rebuilt from per-cell node copies" if ctx.slot_geometry_nodes() else "streamed per-slot copies"
    ctx.init_freestream()
    t_up = time.perf_counter() - t_up
    print(f"[bench] mesh {N} cells, {G} ghosts, {F} faces, {mesh.nblocks} blocks (built in {t_mesh:.1f}s, uploaded "
          f"in {t_up:.1f}s); warmup {a.warmup}", file=sys.stderr, flush=True)
    if a.graph and a.warmup % 2:
        # the graph holds two iterations (an even number of state-buffer swaps): an odd
        # warmup would leave the buffers swapped and re-capture the graph inside the timed region
        a.warmup += 1
    if a.warmup:
        if a.graph:
            ctx.batch_run(a.warmup, use_graph=True)
        else:
            ctx.iterate(a.warmup, history=False)
    ctx.op("sync")
    E.lib().e3d_enable_timing(ctx._h, 1)
    t0 = time.perf_counter()
    if a.graph:
        ctx.batch_run(a.steps, use_graph=True)
        hist = ctx.batch_history(0, a.warmup + a.steps)[-1:]
    else:
        hist = ctx.iterate(a.steps)
    ctx.op("sync")
    t1 = time.perf_counter()
    wall = t1 - t0
    stages = 5
    value = N * stages * a.steps / wall / 1e6
    kt = {}
    for name in ("interp", "flux_update0", "flux_update", "flux_update_nodiss", "exchange", "monitors"):
        ms, n = ctx.kernel_time(name)
        kt[name] = {"avg_ms": ms, "launches": n}
    measured = {}
    tfile = TRAFFIC_FILES.get((a.nx, a.ny, a.nz, tuple(a.parts), a.scheme, a.brick, a.order))
    if tfile and os.path.exists(os.path.join(ROOT, "profiles", tfile)):
        with open(os.path.join(ROOT, "profiles", tfile)) as f:
            measured = json.load(f)["hbm_bytes_per_launch"]
    rooflines = {}
    for name in ("interp", "flux_update"):
        t = kt[name]
        if t["avg_ms"] <= 0 or alg[name] == 0:  # not launched, or launches inside a graph (not timed)
            continue
        ach = alg[name] / (t["avg_ms"] * 1e-3) / 1e9
        rooflines[name] = {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": measured.get(name),
                           "algorithmic_bytes_per_launch": alg[name], "avg_ms": t["avg_ms"],
                           "share_of_step": t["avg_ms"] * t["launches"] / (wall * 1e3)}
    # the kernels' physical HBM rate against what the library's streaming kernel
    # reaches on this GPU for the same read/write mix (measured live, 2 GiB per launch)
    tf = None
    if tfile and os.path.exists(os.path.join(ROOT, "profiles", tfile)):
        with open(os.path.join(ROOT, "profiles", tfile)) as f:
            tf = json.load(f)
    if tf and rooflines:
        import ctypes
        cal = tf["calibration"]
        stream = []
        for r_, w_ in ((7, 1), (3, 1), (2, 1), (1, 1)):
            g = ctypes.c_double()
            E._check(E.lib().e3d_stream_bandwidth(0, 2 << 30, r_, w_, 5, ctypes.byref(g)))
            stream.append({"write_fraction": w_ / (r_ + w_), "GB/s": g.value})
        for name, rl in rooflines.items():
            rd = tf["fetch_size_kib"][name] * 1024 * cal["fetch_factor"]
            wr = tf["write_size_kib"][name] * 1024 * cal["write_factor"]
            rate = (rd + wr) / (rl["avg_ms"] * 1e-3) / 1e9
            ceil = stream_rate(stream, wr / (rd + wr))
            rl.update(hbm_rate_gbs=rate, write_fraction=wr / (rd + wr), stream_ceiling_gbs=ceil,
                      frac_of_stream_ceiling=rate / ceil)
    dominant = max(rooflines, key=lambda k: rooflines[k]["share_of_step"]) if rooflines else None
    rms_ro = hist[-1]["rms"][0] if hist else float("nan")
    print(f"[bench] {a.steps} steps in {wall:.3f}s; rms_ro {rms_ro:.6e}; kernels "
          + ", ".join(f"{k} {v['avg_ms']:.3f}ms x{v['launches']}" for k, v in kt.items()), file=sys.stderr, flush=True)
    ctx.close()
    if a.calib_copy:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library already loaded
        n = ctypes.c_size_t(1 << 30)
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        ok = hip.hipMalloc(ctypes.byref(src), n) == 0 and hip.hipMalloc(ctypes.byref(dst), n) == 0
        ok = ok and hip.hipMemset(src, 1, n) == 0 and hip.hipMemcpy(dst, src, n, 3) == 0
        ok = ok and hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        print(f"[bench] calibration copy of 1 GiB: {'done' if ok else 'FAILED'}", file=sys.stderr, flush=True)

    scrambled = None
    if a.mesh == "bump" and a.brick >= 0 and not a.no_scrambled:
        scrambled = scrambled_leg(a)

    cpu, cpu1, cpu_ref = None, None, None
    if not a.no_cpu_baseline:
        cpu, cpu1 = cpu_baselines(a, sc, blocks)
        cpu_ref = reference_baseline(sc)
    warm = None if a.no_warm_start else warm_start_leg()

    out = {
        "metric": "Mcell-RK-stage/s (FP64, %HBM roof) + time to 8-order residual drop, warm vs cold",
        "value": value, "unit": "Mcell-RK-stage/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"bump channel {a.nx}x{a.ny}x{a.nz}x6 tets ({N} cells)" if a.mesh == "bump" else
                                f"the reference's ONERA M6 mesh ({N} tets, Meshes/oneraM6_120B)")
                               + f", {mesh.nblocks} block(s) on one device, {sc['label']}, 5-stage RK, CFL 1.5",
                   "cells": N, "ghosts": G, "faces": F, "blocks": mesh.nblocks, "stages": stages,
                   "cell_numbering": ("the mesh file's" if a.mesh == "onera" else
                                      "scrambled" if a.brick < 0 else f"brick {a.brick}"),
                   "memory_order": a.order, "launch": "hipGraph, batch API" if a.graph else "e3d_iterate",
                   "tile_cells": a.tile_cells,
                   "slot_geometry": slot_geometry,
                   "tile_order": "automatic by mesh size (XCD-contiguous below 262,144 cells, "
                                 "runs of 4096 cells dealt to the XCDs above)"},
        "model_bytes_per_cell_stage": model_bytes,
        "hbm_roof_frac_overall": value / (HBM_PEAK_GBS * 1e9 / model_bytes / 1e6),
        "roofline": rooflines[dominant] if dominant else None,
        "rooflines": rooflines,
        "kernels": kt,
        "scrambled_mesh": scrambled,
        "cpu_baseline": cpu,
        "cpu_baseline_1core": cpu1,
        "cpu_baseline_reference": cpu_ref,
        "warm_start": warm,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
