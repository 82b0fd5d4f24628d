"""The closure network's MFMA kernels against the oracle's restatement of the
same network with the matrix-core operands rounded to bfloat16: forward
outputs, every parameter gradient of the training kernel, a training run on
a known teacher network, and the field-to-stress inference path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NI, NH, NO = 10, 64, 6
SLICES = {"W1": (0, 640), "b1": (640, 704), "W2": (704, 4800), "b2": (4800, 4864), "W3": (4864, 5248),
          "b3": (5248, 5254)}


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def net(seed, bias=0.1):
    """Glorot weights from the library and non-zero biases"""
    import torch
    from ghost_amd.closure import MLP
    m = MLP(seed)
    P = m.params().cpu().numpy()
    rng = np.random.default_rng(seed + 100)
    for k in ("b1", "b2", "b3"):
        a, b = SLICES[k]
        P[a:b] = rng.uniform(-bias, bias, b - a)
    m.set_params(torch.from_numpy(P))
    return m, P


def test_forward_matches_oracle(gpu):
    import torch
    from oracle import oracle as O
    m, P = net(11)
    rng = np.random.default_rng(12)
    X = rng.standard_normal((1000, NI)).astype(np.float32)  # not a multiple of 16: a partial tile
    Y = m.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    assert rel(Y, O.mlp_forward(P, X, bf16=True)) < 2e-3
    assert rel(Y, O.mlp_forward(P, X, bf16=False)) < 3e-2


def test_gradients_match_oracle(gpu):
    import torch
    from oracle import oracle as O
    m, P = net(21, bias=0.3)
    rng = np.random.default_rng(22)
    n = 777
    X = rng.standard_normal((n, NI)).astype(np.float32)
    T = rng.standard_normal((n, NO)).astype(np.float32)
    loss, g = m.grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    lref, gref = O.mlp_grad(P, X, T, bf16=True)
    assert abs(loss - lref) < 1e-3 * lref
    g = g.cpu().numpy()
    for name, (a, b) in SLICES.items():
        assert rel(g[a:b], gref[a:b]) < 2e-2, name


def test_training_fits_a_teacher(gpu):
    """Targets from a fixed teacher network; 300 Adam steps on 2^16 points
    bring the loss well below its start (and below the target variance)."""
    import torch
    from ghost_amd.closure import MLP
    teacher, _ = net(31, bias=0.2)
    rng = np.random.default_rng(32)
    n = 1 << 16
    X = torch.from_numpy(rng.standard_normal((n, NI)).astype(np.float32)).cuda()
    T = teacher.forward(X)
    var = float(T.var())
    student = MLP(33)
    losses = student.train(X, T, batch=8192, steps=300, lr=3e-3, seed=5)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0] and losses[-1] < 0.1 * var
    l_end, _ = student.grad(X, T)
    assert l_end < 0.1 * var


def test_stress_from_fields_matches_pack_and_forward(gpu):
    import torch
    m, P = net(41)
    n = 5000
    gen = torch.Generator(device="cuda").manual_seed(7)
    grad9 = [torch.randn(n, dtype=torch.float64, device="cuda", generator=gen) * (1 + i) for i in range(9)]
    tau6 = [torch.randn(n, dtype=torch.float64, device="cuda", generator=gen) * 0.1 for _ in range(6)]
    m.fit_norm(grad9, 0.05, tau6)
    norm = m.norm()
    for i in range(9):
        assert abs(norm[i] - float(grad9[i].mean())) < 1e-3 * (1 + i)
        assert abs(norm[NI + i] - float(grad9[i].std())) < 1e-2 * (1 + i)
    assert norm[9] == np.float32(0.05) and norm[NI + 9] == 1.0
    out = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(6)]
    m.stress(grad9, 0.05, out)
    X, _ = m.pack(grad9, 0.05)
    Y = m.forward(X).double().cpu().numpy()
    my, sy = norm[20:26].astype(np.float64), norm[26:32].astype(np.float64)
    for j in range(6):
        assert rel(out[j].cpu().numpy(), Y[:, j] * sy[j] + my[j]) < 1e-6
