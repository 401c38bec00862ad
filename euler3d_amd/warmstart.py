"""Learned warm start of the bump-channel solver.

A coordinate MLP (``euler3d_amd.MLP``: BF16 matrix-core kernels, trained on the
GPU) maps (Mach, inlet angle, bump height, cell centroid, distance to the
wall) to the converged primitive state, as a deviation from the freestream.
Its prediction replaces the uniform freestream start
(DataStructure/src/PrimitiveVariables.cpp:58-69); the solver then converges
the field with its own method and stop test, so the answer stays the
method's. The drop of a warm run is measured from the cold start's
iteration-0 residual of the same case (``rms_ref`` of a batch instance).

Pipeline (``python -m euler3d_amd.warmstart``):
1. training fields: ``--train`` conditions (seeds 0..n-1) solved from
   freestream by the batched solver to the stop criterion or ``--max-iter``;
2. training: ``--steps`` Adam steps on all cells of all training fields
   (``--boost`` further networks fitted to what the ones before leave;
   ``--memory`` keeps what they all leave on each training field, recalled
   for a new condition from its nearest training conditions);
3. evaluation: ``--heldout`` conditions (seeds 5000..) run cold and warm in
   two batches of the same composition; per instance the iterations and the
   device time to the criterion, with the surrogate's inference time added to
   the warm run, and the converged Cl / Cd of both.
Conditions: Mach ~ U[0.3, 0.8], inlet angle ~ U[-2, 6] deg, h ~ U[0.04, 0.12].
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import (ARR_W, FLUX_AUSM, FLUX_ROE, GRAD_GREEN_GAUSS, GRAD_LEAST_SQUARES, LIM_BARTH_JESPERSEN,
               LIM_VENKATAKRISHNAN, MLP, Context, Mesh, make_params)

GAMMA = 1.4
N_FEATURES = 25   # 3 flow parameters, 3 coordinates, wall distance and its log, 17 Fourier terms
TARGET_SCALE = 0.2


def conditions(seeds):
    out = []
    for s in seeds:
        rng = np.random.default_rng(s)
        out.append(dict(seed=int(s), mach=float(rng.uniform(0.3, 0.8)), aoa_deg=float(rng.uniform(-2.0, 6.0)),
                        h=float(rng.uniform(0.04, 0.12))))
    return out


def wall_y(x, h):
    """Height of the lower wall: the circular-arc bump of thickness h over x in [1, 2]."""
    R = (0.25 + h * h) / (2.0 * h)
    y = np.sqrt(np.maximum(R * R - (x - 1.5) ** 2, 0.0)) - (R - h)
    return np.where((x >= 1.0) & (x <= 2.0), np.maximum(y, 0.0), 0.0)


def freestream(mach, aoa_deg):
    a = math.radians(aoa_deg)
    return np.array([1.0, mach * math.sqrt(GAMMA) * math.cos(a), mach * math.sqrt(GAMMA) * math.sin(a), 0.0, 1.0])


def features(block, cond, chunk=1 << 20):
    """Network inputs of every cell of a block (real cells, then ghosts), about [-1, 1]: flow
    parameters, centroid, distance above the wall (linear and logarithmic), and sines and
    cosines of the streamwise coordinate (4 octaves), the wall-normal one (3) and the spanwise
    one (1), so that the network can place the sharp features near the bump. Large blocks
    are done in chunks on a thread pool (numpy releases the GIL inside its loops)."""
    c = block["ccenter"].reshape(-1, 3)
    if len(c) <= chunk:
        return _features(c, cond)
    from concurrent.futures import ThreadPoolExecutor
    parts = [c[i:i + chunk] for i in range(0, len(c), chunk)]
    with ThreadPoolExecutor(16) as pool:
        return np.concatenate(list(pool.map(lambda p: _features(p, cond), parts)))


def _features(c, cond):
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    d = y - wall_y(x, cond["h"])
    f = np.empty((len(c), N_FEATURES), np.float32)
    f[:, 0] = (cond["mach"] - 0.55) / 0.25
    f[:, 1] = (cond["aoa_deg"] - 2.0) / 4.0
    f[:, 2] = (cond["h"] - 0.08) / 0.04
    f[:, 3] = x / 1.5 - 1.0
    f[:, 4] = 2.0 * y - 1.0
    f[:, 5] = 2.0 * z - 1.0
    f[:, 6] = 2.0 * d
    f[:, 7] = np.log10(np.abs(d) + 1e-3) / 3.0 + 0.5
    k = 8
    for coord, octaves in ((x / 3.0, 4), (d, 3), (z, 1)):
        for j in range(octaves):
            w = np.pi * 2.0 ** j
            f[:, k] = np.sin(w * coord)
            f[:, k + 1] = np.cos(w * coord)
            k += 2
    f[:, k] = np.sin(np.pi * 2.0 * x / 3.0) * np.exp(-8.0 * d)  # streamwise shape of the near-wall layer
    assert k + 1 == N_FEATURES
    return f


def targets(W, cond, nreal):
    """Training targets of the real cells: (W - freestream) / scale, [n x 5]."""
    fs = freestream(cond["mach"], cond["aoa_deg"])
    return ((W[:, :nreal].T - fs) / TARGET_SCALE).astype(np.float32)


def state_from_prediction(y, cond):
    """Primitive state [5 x n] from network outputs [n x 5]."""
    return (y.astype(np.float64) * TARGET_SCALE + freestream(cond["mach"], cond["aoa_deg"])).T


def scheme_params(order, flux="roe", gradient="gg", limiter="vk", k=5.0, omega=1e-16):
    return dict(flux=FLUX_AUSM if flux == "ausm" else FLUX_ROE, order=order,
                gradient=GRAD_LEAST_SQUARES if gradient == "lsq" else GRAD_GREEN_GAUSS,
                limiter=LIM_BARTH_JESPERSEN if limiter == "bj" else LIM_VENKATAKRISHNAN, k=k, omega=omega, stages=5)


def solve_batch(meshes, conds, scheme, max_iter, criterion, warm=None, rms_ref=None, use_graph=True, chunk=2000,
                log=sys.stderr, history_every=0, deadline=None, keep_states=True):
    """Runs one batch; warm: per instance [5 x nall] initial states (None: freestream).
    deadline: wall-clock seconds after which no further chunk starts (instances
    still running are reported as not converged). Returns per-instance results
    and the final states (None each when keep_states is false)."""
    ctx = Context(0)
    ctx.set_params(make_params(0.5, 0.0, max_iter=max_iter, criterion=criterion, **scheme))
    insts = []
    for i, c in enumerate(conds):
        d = dict(mach=c["mach"], aoa_deg=c["aoa_deg"], cmac=1.0, criterion=criterion, max_iter=max_iter)
        if rms_ref is not None:
            d["rms_ref"] = list(rms_ref[i])
        insts.append(d)
    ctx.upload_batch(meshes, insts)
    ctx.init_freestream()
    if warm is not None:
        for i, Wi in enumerate(warm):
            (b,) = ctx.batch_blocks(i)
            ctx.set(b, ARR_W, Wi)
    ctx.op("sync")
    t0 = time.perf_counter()
    done = 0
    while done < max_iter:  # in chunks, reporting progress after each
        n, status = ctx.batch_run(min(chunk, max_iter - done), use_graph=use_graph)
        done += n
        active = sum(1 for x in status if not x["stopped"])
        if log is not None:
            print(f"[warmstart]   {len(conds)} instances: {done} iterations, {active} still running "
                  f"({time.perf_counter() - t0:.1f}s)", file=log, flush=True)
        if active == 0 or n == 0:
            break
        if deadline is not None and time.perf_counter() - t0 > deadline:
            print(f"[warmstart]   wall-clock limit of {deadline:.0f}s: {active} instances stop unconverged",
                  file=log, flush=True)
            break
    wall = time.perf_counter() - t0
    out = []
    states = []
    for i, (c, s) in enumerate(zip(conds, status)):
        (b,) = ctx.batch_blocks(i)
        first = ctx.batch_history(i, 1)
        if history_every:
            h = ctx.batch_history(i, s["iterations"])
            hist_rows = [[r["iter"], r["drop"][0], r["cl"], r["cd"]] for r in h[::history_every]]
        else:
            hist_rows = None
        out.append(dict(c, iterations=s["iterations"], stop_seconds=s["stop_seconds"], drop=s["last"]["drop"][0],
                        cl=s["last"]["cl"], cd=s["last"]["cd"], rms0=first[0]["rms"] if first else None,
                        drop0=first[0]["drop"][0] if first else None,
                        converged=bool(s["stopped"] and s["last"]["drop"][0] <= criterion), history=hist_rows))
        states.append(ctx.get(b, ARR_W) if keep_states else None)
    ctx.close()
    return out, states, wall


def predict_chunked(net, x, chunk=1 << 22):
    return np.concatenate([net.predict(x[i:i + chunk]) for i in range(0, len(x), chunk)]) if len(x) else \
        np.zeros((0, 5), np.float32)


def bump_meshes(conds, nx, ny, nz):
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(8) as pool:  # the builder runs in C++ without the GIL
        return list(pool.map(lambda c: Mesh.bump(nx, ny, nz, c["h"], 4, (1, 1, 1)), conds))


def training_set(conds, nx, ny, nz, scheme, max_iter, criterion, log=sys.stderr):
    """Fields from the batched solver for every condition (freestream start, up to
    the criterion or max_iter), as network inputs and targets of their real cells."""
    tm = bump_meshes(conds, nx, ny, nz)
    res, states, t_gen = solve_batch(tm, conds, scheme, max_iter, criterion, log=log)
    X, T, kept = [], [], []
    for m, c, W in zip(tm, conds, states):
        blk = m.block(0)
        if not np.isfinite(W).all():
            continue  # a diverged case is no training data
        X.append(features(blk, c)[:blk["nreal"]])
        T.append(targets(W, c, blk["nreal"]))
        kept.append(c)
    for m in tm:
        m.close()
    X, T = np.concatenate(X), np.concatenate(T)
    print(f"[warmstart] {len(conds)} training fields ({sum(r['converged'] for r in res)} at {criterion:g}) in "
          f"{t_gen:.1f}s; {len(X)} samples", file=log, flush=True)
    return X, T, res, t_gen, kept


class Surrogate:
    """Networks trained in stages on the GPU: the first fits the fields; each
    further one fits what the ones before it leave (rescaled to unit size), so
    the smooth part of the error, which the solver removes slowly, shrinks."""

    def __init__(self, X, T, stages=1, steps=30000, batch=8192, lr=2e-3, width=256, depth=5, log=sys.stderr,
                 var_weights=None):
        """var_weights: relative loss weight of (rho, u, v, w, p), 1 each by default; a variable's
        targets are multiplied by its weight for training and its predictions divided by it,
        so its squared error counts weight^2 times."""
        self.nets, self.scales, self.loss = [], [], None
        t1 = time.perf_counter()
        wv = np.ones(5, np.float32) if var_weights is None else np.asarray(var_weights, np.float32)
        R = T
        for stage in range(stages):
            sc = (np.ones(5, np.float32) if stage == 0 else R.std(axis=0).astype(np.float32) + 1e-12) / wv
            net = MLP(N_FEATURES, 5, width, depth, seed=1 + 10 * stage)
            net.set_data(X, R / sc)
            ls = net.train(steps, batch=batch, lr=lr, seed=2 + 10 * stage)
            self.loss = ls if self.loss is None else self.loss
            self.nets.append(net)
            self.scales.append(sc)
            if stage < stages - 1:
                R = R - predict_chunked(net, X) * sc
                print(f"[warmstart] stage {stage}: loss {ls[0]:.3e} -> {np.mean(ls[-50:]):.3e}; residual targets "
                      f"std {R.std(axis=0)}", file=log, flush=True)
        self.train_seconds = time.perf_counter() - t1
        self.step_ms, flops = self.nets[0].time_step(batch, reps=5)
        self.tflops = flops / self.step_ms / 1e9
        # fit on the training fields themselves (the generalisation gap is this vs the held-out error)
        sub = np.random.default_rng(3).choice(len(X), size=min(len(X), 1 << 20), replace=False)
        err = np.abs(self(X[sub]) - T[sub]) * TARGET_SCALE
        self.train_max_err, self.train_mean_err = float(err.max()), float(err.mean())
        print(f"[warmstart] trained {stages} network(s) x {steps} steps in {self.train_seconds:.1f}s, loss "
              f"{self.loss[0]:.3e} -> {np.mean(self.loss[-50:]):.3e}; step {self.step_ms:.3f} ms = "
              f"{self.tflops:.1f} TFLOP/s; error on training cells: max {self.train_max_err:.4f}, "
              f"mean {self.train_mean_err:.2e}", file=log, flush=True)

    def __call__(self, x, n=None):
        """Prediction of the first n networks (all by default)."""
        k = len(self.nets) if n is None else n
        return sum(predict_chunked(net, x) * sc for net, sc in zip(self.nets[:k], self.scales[:k]))

    def state(self, block, cond, n=None):
        """Initial primitive state [5 x nall] of a block for a condition."""
        return state_from_prediction(self(features(block, cond), n), cond)

    def close(self):
        for net in self.nets:
            net.close()
        self.nets = []


def flow_params(cond):
    """(Mach, inlet angle, bump height) scaled to about [-1, 1], as in the network's first three inputs."""
    return np.array([(cond["mach"] - 0.55) / 0.25, (cond["aoa_deg"] - 2.0) / 4.0, (cond["h"] - 0.08) / 0.04])


class ErrorMemory:
    """What the surrogate leaves on each training field (target minus prediction,
    every real cell), recalled for a new condition as a weighted sum of the
    leftovers of its nearest training conditions in (Mach, angle, height).

    The surrogate's error varies smoothly with the flow parameters, and its
    smooth part is what the solver removes slowly; the leftovers of nearby
    training fields carry most of it. The weights come from one of a few
    interpolants, picked by leave-one-out over the training fields themselves:
    ("idw", k) inverse squared distance over the k nearest, ("linear", k) the
    intercept of a weighted least-squares plane (tricube weights) through the
    k nearest, which is exact when the leftovers vary linearly. Needs the
    training and the new meshes to share their cell numbering (one mesh
    family: the bump height moves the nodes, not the topology)."""

    CANDIDATES = (("idw", 1), ("idw", 2), ("idw", 4), ("idw", 8), ("idw", 16), ("linear", 8), ("linear", 12),
                  ("linear", 16), ("linear", 24))

    def __init__(self, params, R, loo_fields=128, loo_cells=16384, log=sys.stderr):
        self.P = np.asarray(params, np.float64)   # [n x 3]
        self.R = R                                # [n x nreal x 5], float32, target units
        rng = np.random.default_rng(4)
        n, nreal = R.shape[0], R.shape[1]
        fields = rng.choice(n, size=min(n, loo_fields), replace=False)
        cells = np.sort(rng.choice(nreal, size=min(nreal, loo_cells), replace=False))
        Rs = R[:, cells, :]
        base = float(np.mean([np.abs(Rs[j]).mean() for j in fields]))
        self.loo = {"none": base}
        for method in self.CANDIDATES:
            if method[1] >= n:
                continue
            e = []
            for j in fields:
                idx, c = self.weights(self.P[j], method, exclude=j)
                e.append(np.abs(Rs[j] - np.tensordot(c, Rs[idx], axes=1)).mean())
            self.loo[f"{method[0]}{method[1]}"] = float(np.mean(e))
        best = min(self.loo, key=self.loo.get)
        self.method = None if best == "none" else next(m for m in self.CANDIDATES if f"{m[0]}{m[1]}" == best)
        print(f"[warmstart] error memory of {n} fields: leave-one-out mean error "
              + ", ".join(f"{k} {v:.3e}" for k, v in self.loo.items()) + f"; using {best}", file=log, flush=True)

    def weights(self, p, method, exclude=None):
        """Indices of the training fields used for parameters p and their weights (summing to 1)."""
        d = np.sqrt(((self.P - p) ** 2).sum(axis=1))
        if exclude is not None:
            d[exclude] = np.inf
        kind, k = method
        idx = np.argsort(d, kind="stable")[:k]
        dk = d[idx]
        if kind == "idw":
            if dk[0] == 0.0:
                return idx[:1], np.ones(1)
            w = 1.0 / dk ** 2
            return idx, w / w.sum()
        h = dk[-1] * 1.0001 + 1e-12
        w = (1.0 - (dk / h) ** 3) ** 3
        A = np.hstack([np.ones((k, 1)), self.P[idx] - p])
        M = A.T @ (w[:, None] * A) + np.diag([0.0, 1e-8, 1e-8, 1e-8])
        c = np.linalg.solve(M, A.T * w)[0]    # first row of (A'WA)^-1 A'W: the plane's value at p
        return idx, c

    def __call__(self, cond):
        """Correction [nreal x 5] (target units) for a condition; zeros when no interpolant helped."""
        if self.method is None:
            return np.zeros(self.R.shape[1:], np.float32)
        idx, c = self.weights(flow_params(cond), self.method)
        return np.tensordot(c.astype(np.float32), self.R[idx], axes=1)


def run(n_train=128, n_heldout=16, nx=48, ny=16, nz=16, order=1, max_iter=30000, criterion=-8.0, steps=3000,
        batch=8192, lr=2e-3, width=256, depth=5, deeper=3.0, train_criterion=None, boost=0, memory=False,
        var_weights=None, log=sys.stderr):
    scheme = scheme_params(order)
    train_c = conditions(range(n_train))
    held_c = conditions(range(5000, 5000 + n_heldout))
    t0 = time.perf_counter()
    # 1. training fields from the batched solver; they need only be converged
    # well below the network's own error
    tcrit = criterion if train_criterion is None else train_criterion
    X, T, res, t_gen, kept = training_set(train_c, nx, ny, nz, scheme, max_iter, tcrit, log)
    # 2. training on the GPU
    sur = Surrogate(X, T, 1 + boost, steps, batch, lr, width, depth, log, var_weights)
    loss, t_train, ms = sur.loss, sur.train_seconds, sur.step_ms
    flops = sur.tflops * ms * 1e9
    train_max_err, train_mean_err = sur.train_max_err, sur.train_mean_err
    n_samples = len(X)
    mem = None
    if memory:  # what the networks leave on every training field, for recall by nearby conditions
        t3 = time.perf_counter()
        R = (T - sur(X)).reshape(len(kept), len(X) // len(kept), 5)
        mem = ErrorMemory([flow_params(c) for c in kept], R, log=log)
        mem_build_seconds = time.perf_counter() - t3
    del X, T
    # 3. held-out: cold, then warm from the network with the cold iteration-0 norms as reference
    hm = bump_meshes(held_c, nx, ny, nz)
    cold, cold_states, t_cold = solve_batch(hm, held_c, scheme, max_iter, criterion, history_every=50)
    t2 = time.perf_counter()
    net_states = [sur.state(m.block(0), c) for m, c in zip(hm, held_c)]
    warm_states = net_states
    use_mem = mem is not None and mem.method is not None
    if use_mem:
        warm_states = []
        for m, c, Wn in zip(hm, held_c, net_states):
            nr = m.block(0)["nreal"]
            if nr != mem.R.shape[1]:
                raise ValueError("error memory: held-out mesh numbering differs from the training meshes'")
            W = Wn.copy()
            W[:, :nr] += (mem(c).astype(np.float64) * TARGET_SCALE).T
            warm_states.append(W)
    t_infer = time.perf_counter() - t2
    # the surrogates with fewer stages (and all networks without the memory), for comparison (iterations only)
    by_stages = []
    for k in range(1, boost + 1 + (1 if use_mem else 0)):
        ws_k = [sur.state(m.block(0), c, k) for m, c in zip(hm, held_c)]
        wk, _, _ = solve_batch(hm, held_c, scheme, max_iter, criterion, warm=ws_k, rms_ref=[r["rms0"] for r in cold])
        ok = [(a, b) for a, b in zip(cold, wk) if a["converged"] and b["converged"]]
        by_stages.append(dict(networks=k, converged_both=len(ok),
                              iterations_ratio=sum(a["iterations"] for a, _ in ok) / max(1, sum(b["iterations"] for _, b in ok)),
                              mean_initial_drop=float(np.mean([b["drop0"] for b in wk]))))
        print(f"[warmstart] {k} network(s): cold/warm iterations {by_stages[-1]['iterations_ratio']:.3f}x, mean "
              f"initial drop {by_stages[-1]['mean_initial_drop']:.2f}", file=log, flush=True)
    warm, _, t_warm = solve_batch(hm, held_c, scheme, max_iter, criterion, warm=warm_states,
                                  rms_ref=[r["rms0"] for r in cold], history_every=50)
    # converged coefficients: both starts taken `deeper` decades further, where Cl and Cd no longer move
    deep_cold = deep_warm = None
    if deeper > 0:
        deep_cold, _, _ = solve_batch(hm, held_c, scheme, 2 * max_iter, criterion - deeper)
        deep_warm, _, _ = solve_batch(hm, held_c, scheme, 2 * max_iter, criterion - deeper, warm=warm_states,
                                      rms_ref=[r["rms0"] for r in cold], history_every=10)
    rows = []
    for c, w, cs, ws, wn, m in zip(cold, warm, cold_states, warm_states, net_states, hm):
        nr = m.block(0)["nreal"]
        field_err = float(np.max(np.abs(ws[:, :nr] - cs[:, :nr])))  # surrogate vs converged cold state, real cells
        field_mean_err = float(np.mean(np.abs(ws[:, :nr] - cs[:, :nr])))
        rows.append(dict(seed=c["seed"], mach=c["mach"], aoa_deg=c["aoa_deg"], h=c["h"], field_max_err=field_err,
                         field_mean_err=field_mean_err,
                         field_max_err_networks=float(np.max(np.abs(wn[:, :nr] - cs[:, :nr]))),
                         field_mean_err_networks=float(np.mean(np.abs(wn[:, :nr] - cs[:, :nr]))),
                         cold_iterations=c["iterations"], warm_iterations=w["iterations"],
                         cold_seconds=c["stop_seconds"], warm_seconds=w["stop_seconds"] + t_infer / len(held_c),
                         cold_converged=c["converged"], warm_converged=w["converged"],
                         warm_initial_drop=w["drop0"], cl_cold=c["cl"], cl_warm=w["cl"], cd_cold=c["cd"], cd_warm=w["cd"],
                         cl_rel_diff=abs(w["cl"] - c["cl"]) / max(abs(c["cl"]), 1e-300),
                         cd_rel_diff=abs(w["cd"] - c["cd"]) / max(abs(c["cd"]), 1e-300),
                         cold_history=c["history"], warm_history=w["history"]))
    if deeper > 0:
        for r, dc, dw in zip(rows, deep_cold, deep_warm):
            # the reference's own definition measures each run from its own iteration-0 residual
            # (PostProcessing/src/PostProcessing.cpp:240-250): the warm run then has to fall 8 decades below
            # its own, lower start; first recorded iteration (every 10th) where it has
            own = next((it for it, d, *_ in dw["history"] if d - dw["drop0"] <= criterion), None)
            r.update(warm_iterations_own_r0=(own + 1) if own is not None else None)
            r.update(deep_converged=bool(dc["converged"] and dw["converged"]), cl_deep_cold=dc["cl"],
                     cl_deep_warm=dw["cl"], cd_deep_cold=dc["cd"], cd_deep_warm=dw["cd"],
                     cl_deep_rel_diff=abs(dw["cl"] - dc["cl"]) / max(abs(dc["cl"]), 1e-300),
                     cd_deep_rel_diff=abs(dw["cd"] - dc["cd"]) / max(abs(dc["cd"]), 1e-300))
    both = [r for r in rows if r["cold_converged"] and r["warm_converged"]]
    summary = dict(n_train=n_train, n_heldout=n_heldout, mesh=[nx, ny, nz], order=order, criterion=criterion,
                   var_weights=None if var_weights is None else [float(x) for x in var_weights],
                   train_criterion=tcrit, width=width, boost=boost, by_stages=by_stages, train_max_err=train_max_err,
                   train_mean_err=train_mean_err,
                   memory=(dict(method=mem.method, leave_one_out=mem.loo, fields=int(mem.R.shape[0]),
                                build_seconds=mem_build_seconds) if mem is not None else None),
                   max_iter=max_iter, samples=int(n_samples), train_steps=steps, batch=batch,
                   train_seconds=t_train, train_step_ms=ms, train_tflops=flops / ms / 1e9,
                   loss_first=float(loss[0]), loss_last=float(np.mean(loss[-50:])), generate_seconds=t_gen,
                   inference_seconds=t_infer, cold_batch_seconds=t_cold, warm_batch_seconds=t_warm,
                   converged_both=len(both),
                   iterations_ratio=(sum(r["cold_iterations"] for r in both) / max(1, sum(r["warm_iterations"] for r in both))),
                   time_ratio=(sum(r["cold_seconds"] for r in both) / max(1e-12, sum(r["warm_seconds"] for r in both))),
                   iterations_ratio_own_r0=(sum(r["cold_iterations"] for r in both if r.get("warm_iterations_own_r0"))
                                            / max(1, sum(r["warm_iterations_own_r0"] for r in both
                                                         if r.get("warm_iterations_own_r0")))),
                   max_cl_rel_diff=max([r["cl_rel_diff"] for r in both], default=None),
                   max_cd_rel_diff=max([r["cd_rel_diff"] for r in both], default=None),
                   deeper=deeper,
                   max_cl_deep_rel_diff=max([r["cl_deep_rel_diff"] for r in rows if r.get("deep_converged")],
                                            default=None),
                   max_cd_deep_rel_diff=max([r["cd_deep_rel_diff"] for r in rows if r.get("deep_converged")],
                                            default=None),
                   heldout=rows, wall_seconds=time.perf_counter() - t0)
    sur.close()
    return summary


def run_large(n_train=256, nx_train=48, train_iters=8000, steps=30000, stages=3, width=512, nx=384, mach=0.7,
              aoa_deg=3.0, h=0.1, iters=3000, log=sys.stderr):
    """The warm start on one large mesh it has never seen: the surrogate (it
    takes coordinates, not mesh indices) is trained on fields of small channels
    (nx_train x nx_train/3 x nx_train/3 hexes) solved with the same scheme
    (2nd-order AUSM, least squares, Barth-Jespersen) for train_iters
    iterations, then predicts the initial field of the nx x nx/3 x nx/3 channel.
    Cold and warm runs of `iters` iterations each; the drop of both is measured
    from the cold run's iteration-0 residual."""
    scheme = scheme_params(2, "ausm", "lsq", "bj")
    t0 = time.perf_counter()
    X, T, _, t_gen, _ = training_set(conditions(range(n_train)), nx_train, nx_train // 3, nx_train // 3, scheme,
                                  train_iters, -8.0, log)
    sur = Surrogate(X, T, stages, steps, width=width, log=log)
    n_samples = len(X)
    del X, T
    cond = dict(seed=-1, mach=mach, aoa_deg=aoa_deg, h=h)
    print(f"[warmstart] building the {nx}x{nx // 3}x{nx // 3}x6 channel", file=log, flush=True)
    mesh = Mesh.bump(nx, nx // 3, nx // 3, h, 4, (1, 1, 1))
    blk = mesh.block(0)
    t1 = time.perf_counter()
    W = sur.state(blk, cond)
    t_infer = time.perf_counter() - t1
    print(f"[warmstart] surrogate field of {blk['nall']} cells in {t_infer:.1f}s", file=log, flush=True)
    cold, _, t_cold = solve_batch([mesh], [cond], scheme, iters, -20.0, chunk=500, history_every=10, log=log)
    warm, _, t_warm = solve_batch([mesh], [cond], scheme, iters, -20.0, warm=[W], rms_ref=[cold[0]["rms0"]],
                                  chunk=500, history_every=10, log=log)
    ch, wh = cold[0]["history"], warm[0]["history"]
    cold_final = ch[-1][1]
    reach = next((it for it, d, *_ in wh if d <= cold_final), None)
    out = dict(cells=int(blk["nreal"]), mach=mach, aoa_deg=aoa_deg, h=h, iterations=iters, n_train=n_train,
               train_mesh=[nx_train, nx_train // 3, nx_train // 3], train_iterations=train_iters,
               samples=int(n_samples), networks=stages, width=width, train_steps=steps,
               generate_seconds=t_gen, train_seconds=sur.train_seconds, train_step_ms=sur.step_ms,
               train_tflops=sur.tflops, inference_seconds=t_infer, cold_seconds=t_cold, warm_seconds=t_warm,
               cold_final_drop=cold_final, warm_initial_drop=wh[0][1], warm_final_drop=wh[-1][1],
               warm_iterations_to_cold_final=reach,
               iterations_ratio_at_cold_final=(ch[-1][0] + 1) / (reach + 1) if reach is not None else None,
               cl_cold=ch[-1][2], cl_warm=wh[-1][2], cd_cold=ch[-1][3], cd_warm=wh[-1][3],
               cold_history=ch, warm_history=wh, wall_seconds=time.perf_counter() - t0)
    print(f"[warmstart] {blk['nreal']} cells, {iters} iterations: cold drop {cold_final:.2f}, warm starts at "
          f"{wh[0][1]:.2f} and ends at {wh[-1][1]:.2f}; warm reaches the cold run's final drop after {reach} "
          f"iterations", file=log, flush=True)
    sur.close()
    return out


def anatomy(n_train=128, seeds=(5000, 5004, 5008), nx=48, ny=16, nz=16, steps=20000, width=256, max_iter=30000,
            criterion=-8.0, log=sys.stderr):
    """Which part of a surrogate field's error sets the length of the warm run.

    A surrogate (one network, n_train fields) predicts the held-out cases; each
    case is also converged three decades past the criterion (the "exact"
    field U). The warm run is then repeated from U plus parts of the
    prediction's error e: all of it, scaled (1/2, 1/10), its spanwise mean
    (the channel is an extrusion in z) and the rest, one variable at a time,
    near the wall and away from it, and U with its spanwise velocity set to
    zero. All variants of all cases run in one batch, each measured from the
    cold run's iteration-0 residual."""
    scheme = scheme_params(1)
    X, T, _, _, _ = training_set(conditions(range(n_train)), nx, ny, nz, scheme, max_iter, -6.0, log)
    sur = Surrogate(X, T, 1, steps, width=width, log=log)
    del X, T
    held = conditions(seeds)
    hm = bump_meshes(held, nx, ny, nz)
    exact, exact_states, _ = solve_batch(hm, held, scheme, 3 * max_iter, criterion - 3.0, log=log)
    cold, _, _ = solve_batch(hm, held, scheme, max_iter, criterion, log=log)
    names, states, meshes, conds, refs, cases = [], [], [], [], [], []
    for m, c, U, cr in zip(hm, held, exact_states, cold):
        blk = m.block(0)
        nr = blk["nreal"]
        cc = blk["ccenter"].reshape(-1, 3)[:nr]
        Uw = sur.state(blk, c)
        e = np.zeros_like(Uw)
        e[:, :nr] = Uw[:, :nr] - U[:, :nr]
        # spanwise mean over cells with the same (x, y) centroid
        _, grp = np.unique(np.round(cc[:, 0], 7) * 1e3 + np.round(cc[:, 1], 7), return_inverse=True)
        cnt = np.bincount(grp)
        e2 = np.zeros_like(e)
        for v in range(5):
            e2[v, :nr] = (np.bincount(grp, e[v, :nr]) / cnt)[grp]
        d = np.zeros(e.shape[1])
        d[:nr] = cc[:, 1] - wall_y(cc[:, 0], c["h"])
        real = np.arange(e.shape[1]) < nr
        U0w = U.copy()
        U0w[3] = 0.0
        variants = [("network field", Uw), ("U + e", U + e), ("U + e/2", U + 0.5 * e), ("U + e/10", U + 0.1 * e),
                    ("U + spanwise mean of e", U + e2), ("U + (e - spanwise mean)", U + (e - e2))]
        for v, nm in enumerate(("rho", "u", "v", "w", "p")):
            ev = np.zeros_like(e)
            ev[v] = e[v]
            variants.append((f"U + e in {nm} only", U + ev))
        variants += [("U + e where d < 0.15", U + e * (real & (d < 0.15))),
                     ("U + e where d >= 0.15", U + e * (real & (d >= 0.15))), ("U with w := 0", U0w)]
        cases.append(dict(seed=c["seed"], mach=c["mach"], aoa_deg=c["aoa_deg"], h=c["h"],
                          cold_iterations=cr["iterations"], e_mean=float(np.abs(e[:, :nr]).mean()),
                          e_spanwise_mean_part=float(np.abs(e2[:, :nr]).mean()),
                          e_rest=float(np.abs((e - e2)[:, :nr]).mean()),
                          w_exact_mean=float(np.abs(U[3, :nr]).mean()), w_exact_max=float(np.abs(U[3, :nr]).max())))
        for nm, W in variants:
            names.append((c["seed"], nm))
            states.append(W)
            meshes.append(m)
            conds.append(c)
            refs.append(cr["rms0"])
    out, _, _ = solve_batch(meshes, conds, scheme, max_iter, criterion, warm=states, rms_ref=refs, log=log)
    rows = [dict(seed=seed, variant=nm, iterations=o["iterations"], initial_drop=o["drop0"], converged=o["converged"])
            for (seed, nm), o in zip(names, out)]
    for r in rows:
        print(f"[warmstart] {r['seed']} {r['variant']:26s} {r['iterations']:6d} iterations, initial drop "
              f"{r['initial_drop']:.2f}", file=log, flush=True)
    sur.close()
    return dict(n_train=n_train, train_steps=steps, width=width, mesh=[nx, ny, nz], criterion=criterion,
                exact_drop=criterion - 3.0, cases=cases, runs=rows)


C4_MESH = (96, 32, 32)


def run_c4(stage, n_train=128, n_heldout=64, criterion=-2.0, train_criterion=-1.5, train_iters=8000,
           max_iter=25000, steps=20000, width=256, depth=5, boost=1, cold_json=None, deadline=None,
           log=sys.stderr):
    """The learned warm start on the 96x32x32x6 = 589,824-tet bump channel with
    the batched second-order solver (Roe, Green-Gauss, Venkatakrishnan K = 5,
    5-stage RK, CFL 1.5), in two runs that each fit one GPU session:

    stage "cold": the held-out conditions (seeds 5000..) from freestream, one
      batch, to `criterion` decades below their iteration-0 residual or
      max_iter; per instance the iterations, device time, coefficients and
      iteration-0 norms.
    stage "warm": training fields of conditions 0..n_train-1 from the batched
      solver (freestream start, to train_criterion or train_iters), the
      surrogate (1 + boost networks of depth x width SiLU) trained on them,
      then the held-out conditions from the surrogate's field, with the drop
      measured from the cold runs' iteration-0 norms (read from cold_json),
      and the comparison with the cold runs."""
    nx, ny, nz = C4_MESH
    scheme = scheme_params(2)
    held_c = conditions(range(5000, 5000 + n_heldout))
    t0 = time.perf_counter()
    if stage == "cold":
        hm = bump_meshes(held_c, nx, ny, nz)
        cold, _, t_cold = solve_batch(hm, held_c, scheme, max_iter, criterion, history_every=100, deadline=deadline,
                                      keep_states=False, log=log)
        for m in hm:
            m.close()
        return dict(stage="cold", mesh=list(C4_MESH), order=2, criterion=criterion, max_iter=max_iter,
                    n_heldout=n_heldout, batch_seconds=t_cold, heldout=cold, wall_seconds=time.perf_counter() - t0)
    with open(cold_json) as f:
        cold_run = json.load(f)
    cold = cold_run["heldout"]
    if cold_run["criterion"] != criterion or len(cold) != n_heldout:
        raise ValueError("the cold runs were made with another criterion or number of held-out conditions")
    X, T, res, t_gen, kept = training_set(conditions(range(n_train)), nx, ny, nz, scheme, train_iters,
                                          train_criterion, log)
    train_drops = [r["drop"] for r in res]
    sur = Surrogate(X, T, 1 + boost, steps, width=width, depth=depth, log=log)
    n_samples = len(X)
    del X, T
    hm = bump_meshes(held_c, nx, ny, nz)
    t2 = time.perf_counter()
    warm_states = [sur.state(m.block(0), c) for m, c in zip(hm, held_c)]
    t_infer = time.perf_counter() - t2
    warm, _, t_warm = solve_batch(hm, held_c, scheme, max_iter, criterion, warm=warm_states,
                                  rms_ref=[r["rms0"] for r in cold], history_every=100, deadline=deadline,
                                  keep_states=False, log=log)
    del warm_states
    for m in hm:
        m.close()
    rows = []
    for c, w in zip(cold, warm):
        rows.append(dict(seed=c["seed"], mach=c["mach"], aoa_deg=c["aoa_deg"], h=c["h"],
                         cold_iterations=c["iterations"], warm_iterations=w["iterations"],
                         cold_seconds=c["stop_seconds"], warm_seconds=w["stop_seconds"] + t_infer / len(cold),
                         cold_converged=c["converged"], warm_converged=w["converged"], warm_initial_drop=w["drop0"],
                         cl_cold=c["cl"], cl_warm=w["cl"], cd_cold=c["cd"], cd_warm=w["cd"],
                         cl_rel_diff=abs(w["cl"] - c["cl"]) / max(abs(c["cl"]), 1e-300),
                         cd_rel_diff=abs(w["cd"] - c["cd"]) / max(abs(c["cd"]), 1e-300),
                         cold_history=c["history"], warm_history=w["history"]))
    both = [r for r in rows if r["cold_converged"] and r["warm_converged"]]
    summary = dict(stage="warm", mesh=list(C4_MESH), order=2, criterion=criterion, max_iter=max_iter,
                   n_train=n_train, train_criterion=train_criterion, train_iters=train_iters,
                   train_fields_at_criterion=sum(1 for d in train_drops if d <= train_criterion),
                   train_drop_median=float(np.nanmedian(train_drops)), n_heldout=n_heldout, width=width, depth=depth,
                   networks=1 + boost, samples=int(n_samples), train_steps=steps, generate_seconds=t_gen,
                   train_seconds=sur.train_seconds, train_step_ms=sur.step_ms, train_tflops=sur.tflops,
                   train_max_err=sur.train_max_err, train_mean_err=sur.train_mean_err,
                   inference_seconds=t_infer, warm_batch_seconds=t_warm, cold_batch_seconds=cold_run["batch_seconds"],
                   converged_both=len(both), cold_converged=sum(r["cold_converged"] for r in rows),
                   warm_converged=sum(r["warm_converged"] for r in rows),
                   iterations_ratio=(sum(r["cold_iterations"] for r in both) / max(1, sum(r["warm_iterations"] for r in both))),
                   time_ratio=(sum(r["cold_seconds"] for r in both) / max(1e-12, sum(r["warm_seconds"] for r in both))),
                   mean_warm_initial_drop=float(np.mean([r["warm_initial_drop"] for r in rows])),
                   max_cl_rel_diff=max([r["cl_rel_diff"] for r in both], default=None),
                   max_cd_rel_diff=max([r["cd_rel_diff"] for r in both], default=None),
                   heldout=rows, wall_seconds=time.perf_counter() - t0)
    sur.close()
    print(f"[warmstart] C4: {len(both)}/{len(rows)} held-out reached {criterion:g} from both starts; cold/warm "
          f"iterations {summary['iterations_ratio']:.2f}x, time {summary['time_ratio']:.2f}x; warm start's mean "
          f"initial drop {summary['mean_warm_initial_drop']:.2f}", file=log, flush=True)
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=128)
    ap.add_argument("--heldout", type=int, default=16)
    ap.add_argument("--nx", type=int, default=48)
    ap.add_argument("--order", type=int, default=1)
    ap.add_argument("--max-iter", type=int, default=30000)
    ap.add_argument("--criterion", type=float, default=-8.0)
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--deeper", type=float, default=3.0,
                    help="decades beyond the criterion for the converged-coefficient comparison (0: skip)")
    ap.add_argument("--train-criterion", type=float, default=None,
                    help="stop criterion of the training fields (default: --criterion)")
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--boost", type=int, default=0,
                    help="further networks, each fitted to what the ones before leave")
    ap.add_argument("--var-weights", type=lambda t: [float(x) for x in t.split(",")], default=None,
                    help="loss weights of rho,u,v,w,p (default 1,1,1,1,1)")
    ap.add_argument("--memory", action="store_true",
                    help="add the recalled leftovers of the nearest training fields to the networks' field")
    ap.add_argument("--anatomy", action="store_true",
                    help="which part of the field error sets the warm run's length (one network, seeds 5000/5004/5008)")
    ap.add_argument("--large", action="store_true",
                    help="the 37.7M-cell AUSM/LSQ/BJ channel (384x128x128, M=0.7, 3 deg, h=0.1) instead")
    ap.add_argument("--large-iters", type=int, default=3000)
    ap.add_argument("--train-iters", type=int, default=8000)
    ap.add_argument("--c4", choices=["cold", "warm"], default=None,
                    help="the 96x32x32x6 channel at second order: held-out cold runs, or training + warm runs "
                         "(needs --cold-json from the cold stage)")
    ap.add_argument("--cold-json", default=None)
    ap.add_argument("--deadline", type=float, default=None,
                    help="wall-clock seconds after which a held-out batch stops (unfinished instances reported)")
    ap.add_argument("--json")
    a = ap.parse_args(argv)
    if a.c4:
        s = run_c4(a.c4, n_train=a.train, n_heldout=a.heldout, criterion=a.criterion,
                   train_criterion=-1.5 if a.train_criterion is None else a.train_criterion,
                   train_iters=a.train_iters, max_iter=a.max_iter, steps=a.steps, width=a.width,
                   boost=a.boost, cold_json=a.cold_json, deadline=a.deadline)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(s, f, indent=1)
        return s
    if a.anatomy:
        s = anatomy(a.train, nx=a.nx, ny=a.nx // 3, nz=a.nx // 3, steps=a.steps, width=a.width, max_iter=a.max_iter,
                    criterion=a.criterion)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(s, f, indent=1)
        return s
    if a.large:
        s = run_large(a.train, a.nx, a.train_iters, a.steps, 1 + a.boost, a.width, iters=a.large_iters)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(s, f, indent=1)
        return s
    s = run(a.train, a.heldout, a.nx, a.nx // 3, a.nx // 3, a.order, a.max_iter, a.criterion, a.steps, a.batch, a.lr,
            width=a.width, deeper=a.deeper, train_criterion=a.train_criterion, boost=a.boost, memory=a.memory,
            var_weights=a.var_weights)
    print(f"[warmstart] held-out converged in both: {s['converged_both']}/{a.heldout}; cold/warm iterations "
          f"{s['iterations_ratio']:.2f}x (each run from its own iteration-0 residual: "
          f"{s['iterations_ratio_own_r0']:.2f}x), time {s['time_ratio']:.2f}x; Cl/Cd rel diff at the criterion "
          f"{s['max_cl_rel_diff']:.2e}/{s['max_cd_rel_diff']:.2e}, {a.deeper:g} decades deeper "
          f"{s['max_cl_deep_rel_diff']}/{s['max_cd_deep_rel_diff']}", file=sys.stderr, flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(s, f, indent=1)
    return s


if __name__ == "__main__":
    main()
