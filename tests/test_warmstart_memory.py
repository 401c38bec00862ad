"""The warm start's error memory (host logic, no GPU): the recalled leftover is
a weighted sum of the nearest training fields' leftovers, and the interpolant
is chosen by leave-one-out on the training fields.

Pinned by what the mathematics fixes: the least-squares plane reproduces a
field that varies linearly with the flow parameters exactly; inverse-distance
weighting at a training point returns that field; weights sum to one; and on
leftovers that are independent noise no interpolant beats none, so the
memory switches itself off.
"""
import numpy as np

from euler3d_amd import warmstart as ws


def _linear_fields(P, nreal=64, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(nreal, 5))
    b = rng.normal(size=(3, nreal, 5))
    return (a[None] + np.tensordot(P, b, axes=1)).astype(np.float32), a, b


def test_flow_params_match_the_network_inputs():
    c = dict(mach=0.63, aoa_deg=-1.5, h=0.071)
    f = ws._features(np.array([[1.2, 0.3, 0.4]]), c)[0, :3]
    np.testing.assert_allclose(ws.flow_params(c), f, rtol=1e-6)


def test_linear_leftovers_are_recalled_exactly_and_chosen_by_leave_one_out():
    rng = np.random.default_rng(1)
    P = rng.uniform(-1, 1, size=(60, 3))
    R, a, b = _linear_fields(P)
    mem = ws.ErrorMemory(P, R, log=None)
    assert mem.method[0] == "linear", mem.loo
    assert mem.loo[f"linear{mem.method[1]}"] < 1e-4 * mem.loo["none"]
    cond = dict(mach=0.55 + 0.25 * 0.3, aoa_deg=2.0 + 4.0 * -0.2, h=0.08 + 0.04 * 0.45)
    p = ws.flow_params(cond)
    want = a + np.tensordot(p, b, axes=1)
    np.testing.assert_allclose(mem(cond), want, rtol=0, atol=2e-5 * np.abs(want).max())


def test_weights_sum_to_one_and_idw_returns_a_training_field_at_its_point():
    rng = np.random.default_rng(2)
    P = rng.uniform(-1, 1, size=(30, 3))
    R, _, _ = _linear_fields(P, seed=3)
    mem = ws.ErrorMemory(P, R, log=None)
    p = rng.uniform(-1, 1, size=3)
    for method in ws.ErrorMemory.CANDIDATES:
        idx, c = mem.weights(p, method)
        assert len(idx) == method[1]
        np.testing.assert_allclose(c.sum(), 1.0, rtol=1e-9)
    idx, c = mem.weights(P[7], ("idw", 4))
    assert list(idx) == [7] and c[0] == 1.0
    idx, c = mem.weights(P[7], ("idw", 4), exclude=7)
    assert 7 not in idx


def test_independent_noise_switches_the_memory_off():
    rng = np.random.default_rng(5)
    P = rng.uniform(-1, 1, size=(40, 3))
    R = rng.normal(size=(40, 256, 5)).astype(np.float32)
    mem = ws.ErrorMemory(P, R, log=None)
    assert mem.method is None, mem.loo
    assert not mem(dict(mach=0.5, aoa_deg=1.0, h=0.07)).any()
