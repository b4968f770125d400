"""GPU MLP kernels (BF16 MFMA, FP32 accumulation) against the FP64 oracle MLP
with the same BF16 rounding points, through the C ABI."""
import ctypes

import numpy as np
import pytest

import oracle
from compound_ray_amd import eyerenderer as er

pytestmark = pytest.mark.gpu

P = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1  # 9473: W1[16][64], b1, W2, b2, W3, b3, W4, b4
IN = er.MLP_IN


def params_of(lib):
    out = np.zeros(P, np.float32)
    assert lib.mlpGetParams(out.ctypes.data_as(ctypes.c_void_p), P) == 0, er.last_error(lib)
    return out


def set_params(lib, p):
    p = np.ascontiguousarray(p, np.float32)
    assert lib.mlpSetParams(p.ctypes.data_as(ctypes.c_void_p), P) == 0, er.last_error(lib)


def predict(lib, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros(x.shape[0], np.float32)
    assert lib.mlpPredict(x.ctypes.data_as(ctypes.c_void_p), x.shape[0], y.ctypes.data_as(ctypes.c_void_p)) == 0
    return y


def loss_grad(lib, x, t):
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    loss = np.zeros(1, np.float32)
    g = np.zeros(P, np.float32)
    assert lib.mlpLossGrad(x.ctypes.data_as(ctypes.c_void_p), t.ctypes.data_as(ctypes.c_void_p), x.shape[0],
                           loss.ctypes.data_as(ctypes.c_void_p), g.ctypes.data_as(ctypes.c_void_p)) == 0
    return float(loss[0]), g


@pytest.fixture(scope="module")
def mlp(lib):
    lib.mlpParamCount.restype = ctypes.c_size_t
    assert lib.mlpParamCount() == P
    assert lib.mlpInit(7) == 0, er.last_error(lib)
    return lib


@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 40000])
def test_forward_matches_oracle(mlp, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, IN)).astype(np.float32)
    p = params_of(mlp)
    y = predict(mlp, x)
    want, _, _ = oracle.mlp_loss_grad(p, x, mode=1)
    scale = max(1.0, float(np.max(np.abs(want))))
    assert np.max(np.abs(y - want)) <= 1e-2 * scale


@pytest.mark.parametrize("n", [16, 50, 513, 8192])
def test_gradient_matches_oracle(mlp, n):
    rng = np.random.default_rng(100 + n)
    x = rng.normal(size=(n, IN)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    p = params_of(mlp)
    loss, g = loss_grad(mlp, x, t)
    _, want, wloss = oracle.mlp_loss_grad(p, x, t, mode=1)
    assert abs(loss - wloss) <= 1e-3 * max(1.0, wloss)
    # per parameter block: relative error of the gradient vector
    w1 = IN * 64
    blocks = [(0, w1), (w1, w1 + 64), (w1 + 64, w1 + 4160), (w1 + 4160, w1 + 4224), (w1 + 4224, w1 + 8320),
              (w1 + 8320, w1 + 8384), (w1 + 8384, w1 + 8448), (w1 + 8448, w1 + 8449)]
    for a, b in blocks:
        num = np.linalg.norm(g[a:b] - want[a:b])
        den = np.linalg.norm(want[a:b]) + 1e-12
        assert num <= 1e-2 * den + 1e-7, (a, b, num / den)


def test_training_fits_a_smooth_function(mlp):
    rng = np.random.default_rng(9)
    n = 20000
    x = rng.uniform(-1, 1, size=(n, IN)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] + 0.5 * np.abs(x[:, 3])).astype(np.float32)
    assert mlp.mlpInit(3) == 0
    losses = np.zeros(30, np.float32)
    rc = mlp.mlpTrain(x.ctypes.data_as(ctypes.c_void_p), t.ctypes.data_as(ctypes.c_void_p), n, 30, 1024,
                      ctypes.c_float(3e-3), 5, losses.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, er.last_error(mlp)
    assert losses[-1] < 0.05 * losses[0], losses
    # the trained network evaluates the same on the oracle
    p = params_of(mlp)
    xt = rng.uniform(-1, 1, size=(2000, IN)).astype(np.float32)
    want, _, _ = oracle.mlp_loss_grad(p, xt, mode=1)
    assert np.max(np.abs(predict(mlp, xt) - want)) < 2e-2
    tt = np.sin(2 * xt[:, 0]) + xt[:, 1] * xt[:, 2] + 0.5 * np.abs(xt[:, 3])
    assert np.mean((predict(mlp, xt) - tt) ** 2) < 0.02
