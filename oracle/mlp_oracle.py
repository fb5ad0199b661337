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


def bf16(a):
    """Round to the nearest BF16 value (ties to even) via FP32, as the GPU does."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def normalize(x, mean, std):
    """(x - mean) * (1/std) in FP32 as the device computes it (std 0: input dropped)."""
    x = np.asarray(x, np.float32)[:, :N_IN]
    mean = np.asarray(mean, np.float32)
    std = np.asarray(std, np.float32)
    isd = np.where(std > 0, np.float32(1.0) / np.where(std > 0, std, 1), 0).astype(np.float32)
    return ((x - mean) * isd).astype(np.float64)


def forward_bf16(theta, xn):
    """The inference path's arithmetic: BF16 inputs, hidden-1 activations and
    weights (W1, W2, w3); FP32 biases; exact (here FP64) accumulation."""
    p = unpack(theta)
    xb = bf16(pad(xn))
    z1 = xb @ bf16(p["W1"]) + np.float32(p["b1"]).astype(np.float64)
    h1 = bf16(silu(z1))
    z2 = h1 @ bf16(p["W2"]) + np.float32(p["b2"]).astype(np.float64)
    h2 = silu(z2)
    return h2 @ bf16(p["w3"]) + float(np.float32(p["b3"][0]))


def trilinear(sdf, dims, u, v, w):
    """Trilinear sample at normalised coordinates with texel centres at
    (i + 0.5) / dim and clamp addressing (tex3D linear filtering,
    obstacle_cuda.cu:174-180); sdf is x-fastest."""
    g = np.asarray(sdf, np.float64).reshape(dims[2], dims[1], dims[0])
    out = []
    for a, d in ((u, dims[0]), (v, dims[1]), (w, dims[2])):
        c = np.asarray(a, np.float64) * d - 0.5
        f = np.floor(c)
        i0 = np.clip(f.astype(np.int64), 0, d - 1)
        i1 = np.clip(f.astype(np.int64) + 1, 0, d - 1)
        out.append((i0, i1, c - f))
    (x0, x1, ax), (y0, y1, ay), (z0, z1, az) = out
    c00 = g[z0, y0, x0] + ax * (g[z0, y0, x1] - g[z0, y0, x0])
    c10 = g[z0, y1, x0] + ax * (g[z0, y1, x1] - g[z0, y1, x0])
    c01 = g[z1, y0, x0] + ax * (g[z1, y0, x1] - g[z1, y0, x0])
    c11 = g[z1, y1, x0] + ax * (g[z1, y1, x1] - g[z1, y1, x0])
    c0 = c00 + ay * (c10 - c00)
    c1 = c01 + ay * (c11 - c01)
    return c0 + az * (c1 - c0)


def warm_features(xs, xo, nbr, cnt, lam_prev, h, mass, rho0, eps, sdf=None):
    """FP64 feature rows of the warm-start pass, per particle id.

    xs, xo: predicted and start-of-substep positions (n x 3); nbr, cnt: the
    neighbour lists (n x 60, counts); lam_prev: accumulated multiplier of the
    previous solve.  Density and |grad C|^2 follow ComputeDensities and
    ComputeLambda (particles.c:461-514, 588-666: poly6 without an r floor,
    spiky gradient with r floored at 1e-4 h).  sdf: None or a dict with keys
    data, dims, wmin, wmax, m2w.  Returns (n x 16) rows: features 0-11,
    |grad C|^2 + eps, and the plain lambda."""
    n = len(xs)
    W_norm = 315.0 / (64.0 * np.pi * h ** 9)
    DelW_norm = -45.0 / (np.pi * h ** 6)
    lam_k2 = (mass * DelW_norm / rho0) ** 2
    rmin2 = (1e-4 * h) ** 2
    out = np.zeros((n, 16))
    for i in range(n):
        js = nbr[i, :cnt[i]].astype(np.int64)
        d = xs[i] - xs[js]                       # x_i - x_j
        r2 = np.sum(d * d, 1)
        u = h * h - r2
        w3 = np.where(u > 0, u ** 3, 0.0)
        rho = mass * W_norm * h ** 6 + mass * W_norm * np.sum(w3)
        r = np.sqrt(np.maximum(r2, rmin2))
        g = np.where(h - r > 0, (h - r) ** 2 / r, 0.0)
        sg = np.sum(g[:, None] * d, 0)
        gc2 = lam_k2 * (np.sum(g * g * r2) + sg @ sg)
        denom = gc2 + eps
        cr = rho / rho0 - 1.0
        e = (xs[js] - xo[js]) - (xs[i] - xo[i])
        out[i, 0] = cnt[i] / 32.0
        out[i, 1] = cr
        out[i, 2] = -lam_prev[i] * denom
        out[i, 3] = gc2 / denom
        out[i, 4] = mass / rho0 * DelW_norm * np.sum(g * np.sum(e * d, 1))
        out[i, 5] = 4.0
        mom = -mass * W_norm / (rho * h) * np.sum(w3[:, None] * d, 0)
        out[i, 9:12] = mom
        out[i, 12] = denom
        out[i, 13] = -max(cr, 0.0) / denom
    if sdf is not None:
        lo, hi = np.asarray(sdf["wmin"], np.float64), np.asarray(sdf["wmax"], np.float64)
        inside = np.all((xs > lo) & (xs < hi), 1)
        idx = np.nonzero(inside)[0]
        if len(idx):
            dims = sdf["dims"]
            uvw = (xs[idx] - lo) / (hi - lo)
            dist = trilinear(sdf["data"], dims, *uvw.T)
            out[idx, 5] = np.clip(dist * sdf["m2w"] / h, -4.0, 4.0)
            grad = []
            for a in range(3):
                p, m = uvw.copy(), uvw.copy()
                p[:, a] += 1.0 / dims[a]
                m[:, a] -= 1.0 / dims[a]
                grad.append(trilinear(sdf["data"], dims, *p.T) - trilinear(sdf["data"], dims, *m.T))
            grad = np.stack(grad, 1)
            nrm = np.sqrt(np.sum(grad * grad, 1))
            ok = nrm > 0
            out[idx[ok], 6:9] = grad[ok] / nrm[ok, None]
    return out
