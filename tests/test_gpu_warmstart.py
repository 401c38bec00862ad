"""The learned warm start end to end, at a size that runs in seconds.

Training fields from the batched solver on a 12x4x4x6 channel, two networks
trained on the GPU (the second fitted to what the first leaves), then held-out
conditions run from freestream and from the surrogate's field with the same
stop test, the drop of both measured from the cold run's iteration-0
residual. The warm runs must stop sooner, and, taken four decades further,
reach the same lift and drag as the cold runs (1e-7): the answer is the
method's, the network only shortens the way there.
"""
import numpy as np
import pytest

from euler3d_amd import warmstart as ws

pytestmark = pytest.mark.gpu


def test_warm_start_reaches_the_criterion_sooner_with_the_same_coefficients():
    print("warm start: 24 training fields on 12x4x4x6, 4 held-out", flush=True)
    s = ws.run(n_train=24, n_heldout=4, nx=12, ny=4, nz=4, order=1, max_iter=20000, criterion=-6.0, steps=3000,
               width=128, deeper=4.0, train_criterion=-6.0, boost=1)
    assert s["converged_both"] == 4
    rows = s["heldout"]
    for r in rows:
        info = {k: r[k] for k in ("seed", "cold_iterations", "warm_iterations", "warm_initial_drop",
                                  "cl_deep_rel_diff", "cd_deep_rel_diff")}
        print(info, flush=True)
        assert r["warm_initial_drop"] < 0.0, info      # the surrogate starts below the cold residual
        assert r["warm_iterations"] < r["cold_iterations"], info
        assert r["deep_converged"], info
        assert r["cl_deep_rel_diff"] <= 1e-7 and r["cd_deep_rel_diff"] <= 1e-7, info
    assert s["iterations_ratio"] > 1.2, s["iterations_ratio"]
    assert np.isfinite(s["train_mean_err"]) and s["train_mean_err"] < 0.01


def test_error_memory_keeps_the_answer_the_methods():
    print("warm start with error memory: 32 training fields on 12x4x4x6, 4 held-out", flush=True)
    s = ws.run(n_train=32, n_heldout=4, nx=12, ny=4, nz=4, order=1, max_iter=20000, criterion=-6.0, steps=3000,
               width=128, deeper=4.0, train_criterion=-6.0, boost=0, memory=True)
    m = s["memory"]
    print({k: m[k] for k in ("method", "leave_one_out", "fields")}, flush=True)
    assert m["fields"] == 32
    # the interpolant chosen is the one with the least leave-one-out error ("none" included)
    best = min(m["leave_one_out"], key=m["leave_one_out"].get)
    assert (m["method"] is None) == (best == "none")
    assert s["converged_both"] == 4
    for r in s["heldout"]:
        info = {k: r[k] for k in ("seed", "cold_iterations", "warm_iterations", "field_mean_err",
                                  "field_mean_err_networks", "cl_deep_rel_diff", "cd_deep_rel_diff")}
        print(info, flush=True)
        assert r["warm_iterations"] < r["cold_iterations"], info
        assert r["deep_converged"], info
        assert r["cl_deep_rel_diff"] <= 1e-7 and r["cd_deep_rel_diff"] <= 1e-7, info


def test_c4_pipeline_on_the_full_mesh(tmp_path):
    """The C4 pipeline (run_c4) at full mesh size, 96x32x32x6 = 589,824 tets,
    second-order Roe / Green-Gauss / Venkatakrishnan, with a few fields: two
    held-out conditions run cold to a shallow criterion (-0.7 decades), four
    training fields from the batched solver, two 5 x 256 networks, then the
    warm runs measured from the cold runs' iteration-0 norms. The surrogate's
    field must start below the cold residual, stay finite and stop sooner."""
    import json
    print("c4 pipeline: 2 held-out cold on 96x32x32x6", flush=True)
    cold = ws.run_c4("cold", n_heldout=2, criterion=-0.7, max_iter=3000)
    assert all(r["converged"] for r in cold["heldout"]), [r["drop"] for r in cold["heldout"]]
    path = str(tmp_path / "cold.json")
    with open(path, "w") as f:
        json.dump(cold, f)
    print("c4 pipeline: 4 training fields, 2 networks, 2 held-out warm", flush=True)
    s = ws.run_c4("warm", n_train=4, n_heldout=2, criterion=-0.7, train_criterion=-0.7, train_iters=600,
                  max_iter=3000, steps=1500, width=256, boost=1, cold_json=path)
    assert s["mesh"] == [96, 32, 32] and s["samples"] == 4 * 589824
    assert s["converged_both"] == 2
    for r in s["heldout"]:
        info = {k: r[k] for k in ("seed", "cold_iterations", "warm_iterations", "warm_initial_drop")}
        print(info, flush=True)
        assert r["warm_initial_drop"] < 0.0, info
        assert r["warm_iterations"] < r["cold_iterations"], info
        assert np.isfinite([r["cl_warm"], r["cd_warm"]]).all(), info
