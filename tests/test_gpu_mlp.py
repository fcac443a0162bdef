"""Surrogate MLP kernels against the CPU MLP oracle (oracle/mlp_oracle.py):
gradients, the Adam step, training progress, FP32 inference and the BF16
matrix-core inference (checked against a numpy emulation of its BF16
rounding, so a wrong operand layout cannot pass)."""

import ctypes

import numpy
import pytest

from imate_amd import _lib
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu

f32 = numpy.float32


def _params(seed=0):
    return M.init(seed).astype(f32)


def _data(B, seed=1):
    rng = numpy.random.default_rng(seed)
    x = rng.uniform(-1, 1, (B, 2)).astype(f32)
    t = (numpy.sin(2 * x[:, 0]) + x[:, 1] ** 2).astype(f32)
    return x, t


def _grad(lib, p, x, t):
    g = numpy.zeros(M.NUM_PARAMS, dtype=f32)
    loss = ctypes.c_float(0)
    _lib.check(lib.slq_mlp_grad(_lib.ptr(p), _lib.ptr(x), _lib.ptr(t), len(t),
                                ctypes.byref(loss), _lib.ptr(g)), 'grad')
    return g, loss.value


def _predict(lib, p, x, bf16):
    y = numpy.zeros(len(x), dtype=f32)
    sec = ctypes.c_double(0)
    _lib.check(lib.slq_mlp_predict(_lib.ptr(p), _lib.ptr(x), len(x), int(bf16),
                                   _lib.ptr(y), ctypes.byref(sec)), 'predict')
    return y, sec.value


def _bf16(a):
    u = numpy.ascontiguousarray(a, dtype=f32).view(numpy.uint32).astype(numpy.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(numpy.uint32).view(f32).astype(numpy.float64)


def test_num_params(gpu_lib):
    assert gpu_lib.slq_mlp_num_params() == M.NUM_PARAMS == 17025


@pytest.mark.parametrize('B', [1, 37, 256])
def test_gradient_matches_oracle(gpu_lib, B):
    p = _params(2)
    x, t = _data(B, B)
    g, loss = _grad(gpu_lib, p, x, t)
    gr = M.grad(p.astype(numpy.float64), x.astype(numpy.float64), t.astype(numpy.float64))
    lr = M.loss(p.astype(numpy.float64), x.astype(numpy.float64), t.astype(numpy.float64))
    assert abs(loss - lr) <= 1e-5 * lr
    # FP32 sums of B terms: error ~ B * 2^-24 of the block's scale
    for lo, hi in [(0, 256), (256, 384), (384, 16768), (16768, 16896),
                   (16896, 17024), (17024, 17025)]:
        scale = numpy.abs(gr[lo:hi]).max()
        assert numpy.abs(g[lo:hi] - gr[lo:hi]).max() <= 2e-5 * scale + 1e-9, (lo, hi)


def test_adam_first_step(gpu_lib):
    p = _params(4)
    x, t = _data(64, 5)
    q = p.copy()
    losses = numpy.zeros(1, dtype=f32)
    lr = 1e-3
    _lib.check(gpu_lib.slq_mlp_train(_lib.ptr(q), _lib.ptr(x), _lib.ptr(t), 64, 1,
                                     f32(lr), _lib.ptr(losses)), 'train')
    g = M.grad(p.astype(numpy.float64), x.astype(numpy.float64), t.astype(numpy.float64))
    # first Adam step: m_hat = g, v_hat = g^2  ->  delta = -lr g / (|g| + eps)
    expect = p - lr * g / (numpy.abs(g) + 1e-8)
    big = numpy.abs(g) > 1e-3 * numpy.abs(g).max()
    assert numpy.abs(q[big] - expect[big]).max() <= 1e-3 * lr + 1e-6
    assert abs(losses[0] - M.loss(p.astype(numpy.float64), x.astype(numpy.float64),
                                  t.astype(numpy.float64))) <= 1e-5 * losses[0]


def test_training_reduces_loss(gpu_lib):
    p = _params(6)
    x, t = _data(512, 7)
    losses = numpy.zeros(2000, dtype=f32)
    _lib.check(gpu_lib.slq_mlp_train(_lib.ptr(p), _lib.ptr(x), _lib.ptr(t), 512, 2000,
                                     f32(2e-3), _lib.ptr(losses)), 'train')
    assert numpy.all(numpy.isfinite(losses))
    assert losses[-1] < 1e-2 * losses[0]
    # the final loss is the oracle's loss at the returned parameters (up to
    # the last step's update)
    final = M.loss(p.astype(numpy.float64), x.astype(numpy.float64), t.astype(numpy.float64))
    assert final < 2 * losses[-1] + 1e-6


@pytest.mark.parametrize('Q', [1, 15, 16, 65, 1007, 4096])
def test_inference_f32(gpu_lib, Q):
    p = _params(8)
    x, _ = _data(Q, 9)
    y, sec = _predict(gpu_lib, p, x, False)
    yr, _ = M.forward(p.astype(numpy.float64), x.astype(numpy.float64))
    assert numpy.abs(y - yr).max() <= 1e-5 * max(1.0, numpy.abs(yr).max())
    assert sec > 0


@pytest.mark.parametrize('Q', [1, 15, 16, 65, 1007, 4096])
def test_inference_bf16_matrix_cores(gpu_lib, Q):
    p = _params(10)
    x, _ = _data(Q, 11)
    y, _ = _predict(gpu_lib, p, x, True)
    w = M.unpack(p.astype(numpy.float64))
    xd = x.astype(numpy.float64)
    # emulate: h1 rounded to BF16 (from FP32 silu), W2 rounded to BF16,
    # products accumulated exactly, the rest in full precision
    h1 = _bf16(M.silu(xd @ w['W1'].T + w['b1']).astype(f32))
    z2 = h1 @ _bf16(p[384:16768].reshape(128, 128)).T + w['b2']
    ye = M.silu(z2) @ w['W3'] + w['b3'][0]
    scale = max(1.0, numpy.abs(ye).max())
    err = numpy.abs(y - ye)
    # agreement to FP32 rounding except where the device's FP32 SiLU and the
    # emulation's FP64 one round h1 to neighbouring BF16 values (rare; one
    # BF16 step of one hidden unit moves y by ~1e-4)
    assert numpy.median(err) <= 1e-6 * scale
    assert numpy.quantile(err, 0.99) <= 2e-5 * scale
    assert err.max() <= 2e-3 * scale
    # and it is a BF16-accurate forward pass of the FP32 network
    yr, _ = M.forward(p.astype(numpy.float64), xd)
    assert numpy.abs(y - yr).max() <= 5e-2 * scale
