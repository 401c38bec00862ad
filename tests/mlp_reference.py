"""CPU reference of the warm-start MLP (test infrastructure).

Same network and loss as the GPU kernels of euler3d_amd/csrc/mlp.hip:
features padded to 32 -> `depth` SiLU layers of `width` -> outputs, mean
squared error over the real outputs. With ``bf16=True`` it rounds to bfloat16
(nearest even) exactly where the GPU path does (inputs, weights, activations,
activation gradients), and computes everything else in float64, so the GPU
results differ from it only by FP32 accumulation order. With ``bf16=False`` it
is a plain float64 network, which the finite-difference test pins.
"""
import numpy as np


def to_bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(p, n_in, width, depth, n_out):
    """Flat parameters (per layer W [out x in] then b [out]) -> list of (W, b)."""
    layers, o = [], 0
    dims = [n_in] + [width] * depth + [n_out]
    for l in range(depth + 1):
        fi, fo = dims[l], dims[l + 1]
        W = np.asarray(p[o:o + fo * fi], np.float64).reshape(fo, fi)
        o += fo * fi
        b = np.asarray(p[o:o + fo], np.float64)
        o += fo
        layers.append((W, b))
    assert o == len(p)
    return layers


def flatten(grads):
    return np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def loss_and_grads(p, x, t, n_in, width, depth, n_out, bf16=True):
    """Loss and flat gradients of the mean squared error on a batch."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(p, n_in, width, depth, n_out)
    B = len(x)
    acts, zs = [r(x)], []
    for l in range(depth):
        W, b = layers[l]
        z = acts[-1] @ r(W).T + b
        if bf16:
            z = z.astype(np.float32).astype(np.float64)
        zs.append(z)
        acts.append(r(z * sigmoid(z)))
    W, b = layers[depth]
    out = acts[-1] @ r(W).T + b
    diff = out - t
    loss = np.mean(diff * diff)
    d = r(2.0 * diff / (B * n_out))
    grads = [None] * (depth + 1)
    for l in range(depth, -1, -1):
        W, _ = layers[l]
        grads[l] = (d.T @ acts[l], d.sum(axis=0))
        if l > 0:
            z = zs[l - 1]
            s = sigmoid(z)
            d = r((d @ r(W)) * (s + z * s * (1.0 - s)))
    return loss, flatten(grads)


def predict(p, x, n_in, width, depth, n_out, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(p, n_in, width, depth, n_out)
    a = r(x)
    for l in range(depth):
        W, b = layers[l]
        z = a @ r(W).T + b
        a = r(z * sigmoid(z))
    W, b = layers[depth]
    return a @ r(W).T + b
