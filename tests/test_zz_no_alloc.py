"""e3d_mlp_predict_noalloc, the allocation-free inference entry point, against
e3d_mlp_predict through the C ABI: the same network on the same rows gives
the same bits (both run the same GEMMs in the same accumulation order), on
two calls in a row where the second has fewer rows than the first and than
one scratch chunk, so stale scratch contents from the first call would show.
Device buffers come from torch; the test synchronises torch's stream before
each call and after it."""
import numpy as np
import pytest
import torch

import euler3d_amd as E

pytestmark = pytest.mark.gpu

N_IN, N_OUT, WIDTH, DEPTH = 7, 5, 512, 5


def trained_net():
    net = E.MLP(N_IN, N_OUT, WIDTH, DEPTH, seed=5)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (4096, N_IN)).astype(np.float32)
    t = np.stack([np.sin(3 * x[:, 0]) * x[:, 1], x[:, 2] ** 2, x[:, 3] - x[:, 4], np.cos(x[:, 5]), x[:, 6]], 1)
    net.set_data(x, t.astype(np.float32))
    net.train(20, batch=1024, lr=2e-3)  # moves the weights (and their bf16 copies) off their initial values
    return net


def predict_device(net, x):
    xd = torch.from_numpy(x).to("cuda:0")
    yd = torch.full((len(x), N_OUT), float("nan"), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    net.predict_noalloc(xd.data_ptr(), len(x), yd.data_ptr())
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def test_predict_noalloc_equals_predict_on_two_calls():
    net = trained_net()
    rng = np.random.default_rng(17)
    # first call: 3 chunks of the scratch (3,584 rows each at width 512) and a ragged tail
    for n in (3 * 3584 + 1000, 700):
        x = rng.uniform(-1.2, 1.2, (n, N_IN)).astype(np.float32)
        print(f"noalloc predict: {n} rows, width {WIDTH}", flush=True)
        ref = net.predict(x)
        got = predict_device(net, x)
        assert np.isfinite(ref).all()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (n, np.abs(got - ref).max())
    net.close()


def test_predict_noalloc_rejects_host_memory_and_empty_is_ok():
    net = E.MLP(N_IN, N_OUT, 256, 2, seed=1)
    x = np.zeros((4, N_IN), np.float32)
    y = np.zeros((4, N_OUT), np.float32)
    with pytest.raises(E.E3DError) as e:
        net.predict_noalloc(x.ctypes.data, 4, y.ctypes.data)
    assert e.value.code == -1 and "device memory" in str(e.value)
    xd = torch.zeros((1, N_IN), dtype=torch.float32, device="cuda:0")
    yd = torch.zeros((1, N_OUT), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    net.predict_noalloc(xd.data_ptr(), 0, yd.data_ptr())
    net.close()
