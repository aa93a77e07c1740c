"""CPU restatement (numpy, float64) of the bracket network's forward pass, loss and gradient, the
checker of the GPU training kernels (excalibur_amd/csrc/learned.hip). Test infrastructure only:
tests and bench.py's checks import it; the library never does.

Network: in 32 -> 256 -> 256 -> 2 with ReLU, parameters in one flat array (W1 32x256, b1 256,
W2 256x256, b2 256, W3 256x2, b3 2; row-major, inputs by rows). Loss: the mean over the batch
of the Laplace negative log-likelihood |y - c| exp(-w) + w of the outputs (c, w).
"""
import numpy as np

D_IN, HID, D_OUT = 32, 256, 2
SHAPES = [("W1", (D_IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, D_OUT)),
          ("b3", (D_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def init(rng):
    """He-scaled random parameters (float64)."""
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.standard_normal(shape) * np.sqrt(2.0 / shape[0])
        else:
            d[name] = np.zeros(shape)
    return pack(d)


def forward(p, X):
    P = unpack(p)
    Z1 = X @ P["W1"] + P["b1"]
    H1 = np.maximum(Z1, 0.0)
    Z2 = H1 @ P["W2"] + P["b2"]
    H2 = np.maximum(Z2, 0.0)
    O = H2 @ P["W3"] + P["b3"]
    return O, (Z1, H1, Z2, H2)


def loss(p, X, y):
    O, _ = forward(p, X)
    r = y - O[:, 0]
    return float(np.mean(np.abs(r) * np.exp(-O[:, 1]) + O[:, 1]))


def grad(p, X, y):
    """(loss, flat gradient) by back-propagation through the three layers."""
    P = unpack(p)
    B = X.shape[0]
    O, (Z1, H1, Z2, H2) = forward(p, X)
    r = y - O[:, 0]
    ew = np.exp(-O[:, 1])
    L = float(np.mean(np.abs(r) * ew + O[:, 1]))
    dO = np.stack([-np.sign(r) * ew, 1.0 - np.abs(r) * ew], axis=1) / B
    g = {"W3": H2.T @ dO, "b3": dO.sum(0)}
    dZ2 = (dO @ P["W3"].T) * (Z2 > 0)
    g["W2"] = H1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ P["W2"].T) * (Z1 > 0)
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    return L, pack(g)
