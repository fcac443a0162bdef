"""Pins the CPU MLP oracle (CPU only): analytic back-propagation against
central finite differences of the loss, on every parameter block."""

import numpy

from oracle import mlp_oracle as M


def test_backprop_matches_finite_differences():
    rng = numpy.random.default_rng(1)
    p = M.init(3) + 0.01 * rng.standard_normal(M.NUM_PARAMS)
    x = rng.uniform(-1, 1, (40, 2))
    t = rng.standard_normal(40)
    g = M.grad(p, x, t)
    idx = numpy.concatenate([numpy.arange(0, 256, 37), numpy.arange(256, 384, 29),
                             numpy.arange(384, 384 + 16384, 1013),
                             numpy.arange(16768, 16896, 31),
                             numpy.arange(16896, 17024, 23), [17024]])
    h = 1e-6
    for i in idx:
        e = numpy.zeros(M.NUM_PARAMS)
        e[i] = h
        fd = (M.loss(p + e, x, t) - M.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i
    assert M.NUM_PARAMS == 17025
