"""Pins the CPU restatement of the bracket network's gradient (oracle/mlp.py), which the GPU
training kernels are checked against: central finite differences of the loss on sampled
parameters of every layer, and a planted mistake (the ReLU mask of the second layer dropped from
the backward pass) that the same check must reject."""
import numpy as np

from oracle import mlp


def _fd_check(gradfn, p, X, y, idx, h=1e-6):
    _, g = gradfn(p, X, y)
    worst = 0.0
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(p + e, X, y) - mlp.loss(p - e, X, y)) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / max(1e-6, abs(fd) + abs(g[i])))
    return worst


def _problem(seed=0, B=48):
    rng = np.random.default_rng(seed)
    p = mlp.init(rng)
    p[mlp.N_PARAMS - 2:] = [0.1, -0.3]  # b3: a nonzero scale output
    X = rng.uniform(0, 1, (B, mlp.D_IN))
    y = rng.uniform(-1.5, 0, B)
    # indices spread over every parameter block
    rng2 = np.random.default_rng(seed + 1)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SHAPES])
    idx = np.concatenate([rng2.integers(offs[k], offs[k + 1], 8) for k in range(len(mlp.SHAPES))])
    return p, X, y, idx


def test_gradient_matches_finite_differences():
    p, X, y, idx = _problem()
    assert _fd_check(mlp.grad, p, X, y, idx) < 1e-5


def test_planted_backward_mistake_is_caught():
    p, X, y, idx = _problem()

    def bad_grad(p, X, y):
        P = mlp.unpack(p)
        B = X.shape[0]
        O, (Z1, H1, Z2, H2) = mlp.forward(p, X)
        r = y - O[:, 0]
        ew = np.exp(-O[:, 1])
        dO = np.stack([-np.sign(r) * ew, 1.0 - np.abs(r) * ew], axis=1) / B
        g = {"W3": H2.T @ dO, "b3": dO.sum(0)}
        dZ2 = dO @ P["W3"].T  # mask dropped
        g["W2"] = H1.T @ dZ2
        g["b2"] = dZ2.sum(0)
        dZ1 = (dZ2 @ P["W2"].T) * (Z1 > 0)
        g["W1"] = X.T @ dZ1
        g["b1"] = dZ1.sum(0)
        return 0.0, mlp.pack(g)

    assert _fd_check(bad_grad, p, X, y, idx) > 1e-2
