"""Warm-start MLP kernels on the GPU (BF16 matrix cores, FP32 accumulation)
against the CPU reference that rounds to bfloat16 at the same points:
gradients and predictions, training progress and determinism.

Tolerances: the GPU accumulates in FP32 in its own order, which can move a
bfloat16 rounding by one unit here and there; gradients are compared per
layer by relative L2 error (< 2e-2) and direction (cosine > 0.999)."""
import numpy as np
import pytest

import euler3d_amd as E
from mlp_reference import loss_and_grads, predict, unflatten

pytestmark = pytest.mark.gpu

N_IN, N_OUT, WIDTH, DEPTH = 7, 5, 256, 5


def random_net(seed=3):
    net = E.MLP(N_IN, N_OUT, WIDTH, DEPTH, seed=seed)
    rng = np.random.default_rng(seed)
    p = net.get_params()
    dims = [N_IN] + [WIDTH] * DEPTH + [N_OUT]
    o = 0
    for l in range(DEPTH + 1):   # He-scaled weights everywhere (the last layer starts at zero otherwise)
        n = dims[l + 1] * dims[l]
        p[o:o + n] = rng.normal(0, np.sqrt(2.0 / dims[l]), n)
        o += n
        p[o:o + dims[l + 1]] = rng.normal(0, 0.1, dims[l + 1])
        o += dims[l + 1]
    net.set_params(p)
    return net, net.get_params()


def layer_slices():
    dims = [N_IN] + [WIDTH] * DEPTH + [N_OUT]
    out, o = [], 0
    for l in range(DEPTH + 1):
        n = dims[l + 1] * dims[l] + dims[l + 1]
        out.append(slice(o, o + n))
        o += n
    return out


def test_mlp_gradients_match_reference():
    net, p = random_net()
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, (2048, N_IN)).astype(np.float32)
    t = rng.uniform(-1, 1, (2048, N_OUT)).astype(np.float32)
    net.set_data(x, t)
    idx = rng.permutation(2048)[:1024].astype(np.int32)
    print("mlp gradients: batch 1024", flush=True)
    loss, g = net.gradients(idx)
    rloss, rg = loss_and_grads(p.astype(np.float64), x[idx].astype(np.float64), t[idx].astype(np.float64),
                               N_IN, WIDTH, DEPTH, N_OUT)
    assert abs(loss - rloss) <= 1e-3 * rloss, (loss, rloss)
    for l, sl in enumerate(layer_slices()):
        a, b = g[sl].astype(np.float64), rg[sl]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        assert rel < 2e-2 and cos > 0.999, (l, rel, cos)
    net.close()


def test_mlp_predict_matches_reference():
    net, p = random_net(5)
    rng = np.random.default_rng(12)
    x = rng.uniform(-1, 1, (3000, N_IN)).astype(np.float32)   # not a multiple of the chunk: exercises padding
    y = net.predict(x)
    ry = predict(p.astype(np.float64), x.astype(np.float64), N_IN, WIDTH, DEPTH, N_OUT)
    assert y.shape == (3000, N_OUT)
    assert np.max(np.abs(y - ry)) <= 2e-2 * np.max(np.abs(ry)), np.max(np.abs(y - ry))
    net.close()


def test_mlp_untrained_predicts_zero():
    """The output layer starts at zero: an untrained network gives the freestream start."""
    net = E.MLP(N_IN, N_OUT, WIDTH, DEPTH, seed=1)
    y = net.predict(np.random.default_rng(0).uniform(-1, 1, (600, N_IN)).astype(np.float32))
    assert (y == 0).all()
    net.close()


def smooth_data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, N_IN)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] * x[:, 3], np.cos(x[:, 4] + x[:, 5]),
                  x[:, 6] ** 2, 0.5 * x[:, 0] - x[:, 3]], axis=1).astype(np.float32)
    return x, t


def test_mlp_training_fits_and_is_deterministic():
    x, t = smooth_data(32768, 2)
    losses, params = [], []
    for _ in range(2):
        net = E.MLP(N_IN, N_OUT, WIDTH, DEPTH, seed=9)
        net.set_data(x, t)
        print("mlp training: 400 steps", flush=True)
        losses.append(net.train(400, batch=4096, lr=2e-3, seed=4))
        params.append(net.get_params())
        net.close()
    l0 = losses[0]
    assert np.mean(l0[-20:]) < 0.02 * np.mean(l0[:5]), (l0[:5], l0[-5:])
    assert (losses[0] == losses[1]).all() and (params[0] == params[1]).all()
    xv, tv = smooth_data(4096, 3)   # held out
    net = E.MLP(N_IN, N_OUT, WIDTH, DEPTH, seed=9)
    net.set_params(params[0])
    err = np.mean((net.predict(xv) - tv) ** 2)
    assert err < 0.05 * np.mean(tv ** 2), err
    net.close()


def test_mlp_rejects_bad_batches():
    net = E.MLP(N_IN, N_OUT, WIDTH, DEPTH, seed=1)
    x, t = smooth_data(1024, 2)
    net.set_data(x, t)
    with pytest.raises(E.E3DError):
        net.train(1, batch=500)          # not a multiple of 512
    with pytest.raises(E.E3DError):
        net.gradients(np.array([0, 1, 2], np.int32))
    with pytest.raises(E.E3DError):
        E.MLP(40, 5)                     # more than 32 features
    net.close()


# SHA-256 of the float32 parameters after 64 deterministic training steps
# (7 -> 5 x width SiLU -> 5, seed 11, data from numpy's default_rng(2024),
# batch 2048, lr 1e-3, shuffle seed 5), as measured on MI355X. The hidden
# layers run in k_gemm_lds, the input and output layers in k_gemm; a GEMM
# change that alters any accumulation order changes these bits.
PINNED = {512: "91e0cdac6da8c806049ee82cf882a9366789e9f685d7e6cc8d452d77b90636b4",
          256: "c131394b022581d41925bd53bcc3605eaf71f5f29809b97414e81c208becb2fe"}


@pytest.mark.parametrize("width", [512, 256])
def test_mlp_training_bits_are_pinned(width):
    import hashlib
    net = E.MLP(7, 5, width, 5, seed=11)
    rng = np.random.default_rng(2024)
    x = rng.uniform(-1, 1, (8192, 7)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] * x[:, 3], np.cos(x[:, 4]), x[:, 5] ** 2, x[:, 6]], 1)
    net.set_data(x, t.astype(np.float32))
    net.train(64, batch=2048, lr=1e-3, seed=5)
    digest = hashlib.sha256(net.get_params().tobytes()).hexdigest()
    net.close()
    assert digest == PINNED[width], digest
