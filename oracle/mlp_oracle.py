"""ORACLE — test infrastructure only.

CPU restatement (numpy, FP64) of the surrogate network's mathematics, used to
check the GPU training kernels' gradients and the inference kernels: a
2 -> 128 -> 128 -> 1 MLP with SiLU activations and a mean-squared-error loss,
parameters in the flat layout W1[128][2], b1[128], W2[128][128], b2[128],
W3[128], b3. Pinned in tests/test_mlp_oracle.py against central finite
differences of the loss.
"""

import numpy

H = 128
SIZES = [('W1', (H, 2)), ('b1', (H,)), ('W2', (H, H)), ('b2', (H,)),
         ('W3', (H,)), ('b3', (1,))]
NUM_PARAMS = sum(int(numpy.prod(s)) for _, s in SIZES)


def unpack(p):
    p = numpy.asarray(p, dtype=numpy.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(numpy.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return numpy.concatenate([numpy.asarray(d[name], dtype=numpy.float64).ravel()
                              for name, _ in SIZES])


def silu(z):
    return z / (1.0 + numpy.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + numpy.exp(-z))
    return s + z * s * (1.0 - s)


def forward(p, x):
    w = unpack(p)
    z1 = x @ w['W1'].T + w['b1']
    h1 = silu(z1)
    z2 = h1 @ w['W2'].T + w['b2']
    h2 = silu(z2)
    y = h2 @ w['W3'] + w['b3'][0]
    return y, (z1, h1, z2, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(numpy.mean((y - t) ** 2))


def grad(p, x, t):
    """Analytic gradient of the mean-squared-error loss (back-propagation)."""
    w = unpack(p)
    y, (z1, h1, z2, h2) = forward(p, x)
    B = x.shape[0]
    d3 = 2.0 * (y - t) / B
    g = {'W3': h2.T @ d3, 'b3': numpy.array([d3.sum()])}
    d2 = numpy.outer(d3, w['W3']) * dsilu(z2)
    g['W2'] = d2.T @ h1
    g['b2'] = d2.sum(axis=0)
    d1 = (d2 @ w['W2']) * dsilu(z1)
    g['W1'] = d1.T @ x
    g['b1'] = d1.sum(axis=0)
    return pack(g)


def init(seed=0):
    """He-style initialisation used by the surrogate (seeded)."""
    rng = numpy.random.default_rng(seed)
    d = {'W1': rng.standard_normal((H, 2)) * numpy.sqrt(2.0 / 2),
         'b1': numpy.zeros(H),
         'W2': rng.standard_normal((H, H)) * numpy.sqrt(2.0 / H),
         'b2': numpy.zeros(H),
         'W3': rng.standard_normal(H) * numpy.sqrt(1.0 / H),
         'b3': numpy.zeros(1)}
    return pack(d)
