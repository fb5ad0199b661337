"""FP64 numpy oracle of the lambda warm-start network -- TEST INFRASTRUCTURE.

Only tests/, the smoke entry point and bench.py's cpu_baseline leg use it.

The network has no counterpart in the reference (SPHPC has no learned part);
this file restates the mathematics the HIP kernels implement so their forward
pass, gradients and Adam updates can be checked.  Its own gradient is pinned by
central finite differences in tests/test_mlp_oracle.py.

Network: x (12 features, zero-padded to 16) -> 64 SiLU -> 64 SiLU -> 1.
Parameter vector layout (float32 on the device, float64 here):
  W1[16][64] (row = input, col = hidden), b1[64], W2[64][64], b2[64], w3[64], b3.
Loss: mean squared error 0.5/B * sum (y - t)^2 over a batch.
"""
import numpy as np

N_IN = 12
N_PAD = 16
N_H = 64
SIZES = [("W1", (N_PAD, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
         ("w3", (N_H,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)  # 5313


def unpack(theta):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = theta[o:o + k].reshape(shape)
        o += k
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], np.float64).ravel() for name, _ in SIZES])


def init(seed=0):
    """He-style uniform init; rows of W1 for the padding inputs are zero."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (N_PAD, N_H)) * np.sqrt(3.0 / N_IN)
    W1[N_IN:] = 0.0
    W2 = rng.uniform(-1, 1, (N_H, N_H)) * np.sqrt(3.0 / N_H)
    w3 = rng.uniform(-1, 1, N_H) * np.sqrt(3.0 / N_H)
    return pack(dict(W1=W1, b1=np.zeros(N_H), W2=W2, b2=np.zeros(N_H), w3=w3, b3=np.zeros(1)))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def pad(x):
    x = np.asarray(x, np.float64)
    if x.shape[1] == N_PAD:
        return x
    out = np.zeros((x.shape[0], N_PAD))
    out[:, :x.shape[1]] = x
    return out


def forward(theta, x):
    p = unpack(theta)
    x = pad(x)
    z1 = x @ p["W1"] + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = silu(z2)
    y = h2 @ p["w3"] + p["b3"][0]
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(theta, x, t):
    """0.5/B sum (y-t)^2 and its gradient with respect to theta."""
    y, (x, z1, h1, z2, h2) = forward(theta, x)
    p = unpack(theta)
    B = len(y)
    r = y - np.asarray(t, np.float64)
    loss = 0.5 * np.sum(r * r) / B
    dy = r / B
    g = {}
    g["w3"] = h2.T @ dy
    g["b3"] = np.array([dy.sum()])
    dz2 = np.outer(dy, p["w3"]) * dsilu(z2)
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * dsilu(z1)
    g["W1"] = x.T @ dz1
    g["W1"][N_IN:] = 0.0  # padding inputs are not parameters of the model
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def adam(theta, m, v, g, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba), bias-corrected with the 1-based step."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
