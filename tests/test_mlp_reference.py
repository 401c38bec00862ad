"""The CPU reference of the warm-start MLP, pinned without a GPU: its
gradients against central finite differences (float64 network), and its
bfloat16 rounding against hand-checked values."""
import numpy as np

from mlp_reference import loss_and_grads, to_bf16


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e38, 1e-40], np.float32)
    got = to_bf16(x)
    want = np.array([1.0, 1.0, 1.0 + 2 * 2.0 ** -7, 1.0, -2.5, got[5], got[6]])
    assert (got[:5] == want[:5]).all(), got
    # halfway cases go to the even mantissa; the top 16 bits are kept
    assert to_bf16(np.float32(1.0 + 2.0 ** -8))[()] == 1.0
    assert (to_bf16(x).astype(np.float32).view(np.uint32) & 0xFFFF == 0).all()


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(7)
    n_in, width, depth, n_out, B = 3, 4, 2, 2, 5
    dims = [n_in] + [width] * depth + [n_out]
    n_par = sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(depth + 1))
    p = rng.normal(0, 0.7, n_par)
    x = rng.uniform(-1, 1, (B, n_in))
    t = rng.uniform(-1, 1, (B, n_out))
    _, g = loss_and_grads(p, x, t, n_in, width, depth, n_out, bf16=False)
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(n_par):
        q = p.copy()
        q[i] += eps
        lp, _ = loss_and_grads(q, x, t, n_in, width, depth, n_out, bf16=False)
        q[i] -= 2 * eps
        lm, _ = loss_and_grads(q, x, t, n_in, width, depth, n_out, bf16=False)
        fd[i] = (lp - lm) / (2 * eps)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-9), np.max(np.abs(g - fd))
