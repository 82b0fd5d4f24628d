"""The oracle's restatement of the closure network (10-64-64-6, GELU), pinned
by central finite differences of its own loss in plain double arithmetic
(the FP32 parameter steps are measured exactly), by the GELU identities
gelu(x) - gelu(-x) = x and gelu'(0) = 1/2 seen through a one-unit path, and by
the bf16 operand rounding staying within bf16's relative precision."""
import numpy as np

from oracle import oracle as O

NI, NH, NO = 10, 64, 6
OW1, OB1 = 0, NH * NI
OW2, OB2 = OB1 + NH, OB1 + NH + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NO * NH


def rand_params(seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(O.MLP_NPAR, dtype=np.float32)
    P[OW1:OB1] = rng.uniform(-0.6, 0.6, NH * NI)
    P[OB1:OW2] = rng.uniform(-0.2, 0.2, NH)
    P[OW2:OB2] = rng.uniform(-0.25, 0.25, NH * NH)
    P[OB2:OW3] = rng.uniform(-0.2, 0.2, NH)
    P[OW3:OB3] = rng.uniform(-0.3, 0.3, NO * NH)
    P[OB3:] = rng.uniform(-0.2, 0.2, NO)
    return P


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    P = rand_params(2)
    X = rng.standard_normal((7, NI)).astype(np.float32)
    T = rng.standard_normal((7, NO)).astype(np.float32)
    _, g = O.mlp_grad(P, X, T, bf16=False)
    h = 2.0 ** -8
    idx = list(rng.choice(O.MLP_NPAR, 60, replace=False)) + [OB1, OW2 + 5, OB2 + 7, OW3 + 3, OB3, OB3 + 5]
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        step = float(Pp[k]) - float(Pm[k])  # the FP32 steps actually taken
        lp, _ = O.mlp_grad(Pp, X, T, bf16=False)
        lm, _ = O.mlp_grad(Pm, X, T, bf16=False)
        fd = (lp - lm) / step
        assert abs(fd - g[k]) <= 2e-4 * max(1.0, abs(g[k])) + 1e-6, (k, fd, g[k])


def test_gelu_identities_through_one_unit():
    """Only unit 0 of each layer is wired: y0 = gelu(gelu(x0)) path; with the
    second layer passing z2 = h1[0], y0 - y0(-x) isolates gelu's odd part."""
    P = np.zeros(O.MLP_NPAR, dtype=np.float32)
    P[OW1 + 0] = 1.0          # z1[0] = x0
    P[OW2 + 0] = 1.0          # z2[0] = h1[0]
    P[OW3 + 0] = 1.0          # y0 = h2[0]
    xs = np.array([0.0, 0.3, -0.3, 1.7, -1.7, 4.0], dtype=np.float32)
    X = np.zeros((len(xs), NI), dtype=np.float32)
    X[:, 0] = xs
    Y = O.mlp_forward(P, X, bf16=False)
    from math import erf, sqrt
    g = lambda v: 0.5 * v * (1 + erf(v / sqrt(2)))
    for v, y in zip(xs, Y[:, 0]):
        assert abs(y - g(g(float(v)))) < 1e-14
    # gelu(x) - gelu(-x) = x for the first layer alone (second and third identity-like)
    assert abs((g(0.3) - g(-0.3)) - 0.3) < 1e-15
    # gelu'(0) = 1/2: d y0 / d W1[0][0] at x0 = 1 with everything else zero is
    # gelu'(z2) gelu'(z1) x0 = gelu'(gelu(1)) gelu'(1)
    X1 = np.zeros((1, NI), dtype=np.float32)
    X1[0, 0] = 1.0
    T1 = np.zeros((1, NO), dtype=np.float32)
    loss, grad = O.mlp_grad(P, X1, T1, bf16=False)
    y0 = g(g(1.0))
    from math import exp, pi
    gd = lambda v: 0.5 * (1 + erf(v / sqrt(2))) + v * exp(-v * v / 2) / sqrt(2 * pi)
    want = 2 * y0 / NO * gd(g(1.0)) * gd(1.0)
    assert abs(grad[OW1 + 0] - want) < 1e-7 * abs(want)  # the output error is formed in FP32
    assert abs(gd(0.0) - 0.5) < 1e-16


def test_bf16_rounding_is_within_bf16_precision():
    rng = np.random.default_rng(3)
    P = rand_params(4)
    X = rng.standard_normal((200, NI)).astype(np.float32)
    y64 = O.mlp_forward(P, X, bf16=False)
    y16 = O.mlp_forward(P, X, bf16=True)
    rel = np.linalg.norm(y16 - y64) / np.linalg.norm(y64)
    assert 1e-4 < rel < 3e-2  # rounded, and only at bf16's 2^-8 level
