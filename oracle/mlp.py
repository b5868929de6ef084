"""FP64 restatement of the closure network and of its input features.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
kernels in starpolymersmd_amd/csrc/spmd_mlp.hip and of the closure feature
kernel; never by the library.  The closure has no counterpart in the
reference, so this module is pinned by mathematics instead: the analytic
backward pass is checked against central finite differences of the loss,
the forward pass against a direct matrix-product evaluation, and the
features against hand-computed cases.

Parameter layout (flat, FP32 on the device): W0[128][24] b0[128] W1[128][128]
b1[128] W2[128][128] b2[128] W3[6][128] b3[6]; layer y = W x + b; SiLU after
the three hidden layers.  Loss: mean over rows of
1/2 sum_c [(y_c/sdy_c - mu_c)^2 exp(-s_c) + s_c], s clamped to [-12, 12].
"""
from __future__ import annotations

import numpy as np

SIZES = [(128, 24), (128, 128), (128, 128), (6, 128)]
N_PARAMS = sum(o * i + o for o, i in SIZES)


def unpack(p):
    out, k = [], 0
    for o, i in SIZES:
        W = p[k:k + o * i].reshape(o, i)
        k += o * i
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, mu_x=None, sd_x=None):
    """x: (M, 24) raw features -> (M, 6) outputs and the cache for backward."""
    p = np.asarray(p, np.float64)
    a = np.asarray(x, np.float64)
    if mu_x is not None:
        a = (a - mu_x) / sd_x
    cache = [a]
    layers = unpack(p)
    for li, (W, b) in enumerate(layers):
        z = a @ W.T + b
        if li < 3:
            cache.append(z)
            a = silu(z)
            cache.append(a)
        else:
            a = z
    return a, cache


def loss_and_grad(p, x, y, sd_y=(1.0, 1.0, 1.0), mu_x=None, sd_x=None):
    """Mean Gaussian NLL and its gradient with respect to the flat parameters."""
    out, cache = forward(p, x, mu_x, sd_x)
    M = out.shape[0]
    yt = np.asarray(y, np.float64) / np.asarray(sd_y, np.float64)
    mu, s_raw = out[:, :3], out[:, 3:6]
    s = np.clip(s_raw, -12.0, 12.0)
    r = yt - mu
    e = np.exp(-s)
    L = 0.5 * np.sum(r * r * e + s) / M
    dO = np.zeros_like(out)
    dO[:, :3] = -r * e / M
    inside = (s_raw >= -12.0) & (s_raw <= 12.0)
    dO[:, 3:6] = np.where(inside, 0.5 * (1.0 - r * r * e) / M, 0.0)
    layers = unpack(np.asarray(p, np.float64))
    grads = []
    dz = dO
    # cache: [x0, z1, a1, z2, a2, z3, a3]
    acts = [cache[0], cache[2], cache[4], cache[6]]
    zs = [None, cache[1], cache[3], cache[5]]
    for li in range(3, -1, -1):
        W, b = layers[li]
        gW = dz.T @ acts[li]
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if li > 0:
            da = dz @ W
            dz = da * dsilu(zs[li])
    grads.reverse()
    g = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, g


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in SIZES:
        bnd = np.sqrt(6.0 / i)
        parts.append(rng.uniform(-bnd, bnd, size=o * i))
        parts.append(rng.uniform(-0.1, 0.1, size=o))
    return np.concatenate(parts)


def features(pos, vel, mass, btype, bonds, shear):
    """The 24 closure features of one star (see k_closure_features):
    pos, vel (N,3) FP64; mass (N,); btype (N,) 0/1; bonds list of (i, j)
    with i < j (i nearer the core)."""
    pos = np.asarray(pos, np.float64)
    vel = np.asarray(vel, np.float64)
    N = len(pos)
    com, vcom = pos.mean(axis=0), vel.mean(axis=0)
    deg = np.zeros(N, int)
    prv = np.arange(N)
    nxt = np.arange(N)
    for i, j in bonds:
        deg[i] += 1
        deg[j] += 1
        prv[j] = i
        nxt[i] = j
    core = deg > 2
    p32 = pos.astype(np.float32)
    f = np.zeros((N, 24))
    f[:, 0:3] = pos - com
    f[:, 3:6] = vel - vcom
    u = vel.copy()
    u[:, 0] -= shear * pos[:, 1]
    f[:, 6:9] = u
    ip = np.where(core, np.arange(N), prv)
    inn = np.where(core, np.arange(N), nxt)
    f[:, 9:12] = p32[ip] - p32
    f[:, 12:15] = p32[inn] - p32
    d2 = ((p32[:, None, :] - p32[None, :, :]) ** 2).sum(-1)
    f[:, 15] = (d2 < 2.25).sum(axis=1) - 1
    f[:, 16] = btype
    f[:, 17] = core
    f[:, 18] = shear
    v32 = vel.astype(np.float32)
    nb = (d2 < 2.25) & ~np.eye(N, dtype=bool)
    cnt = nb.sum(axis=1)
    vn = (nb.astype(np.float32) @ v32) / np.maximum(cnt, 1)[:, None]
    f[:, 19:22] = np.where(cnt[:, None] > 0, vn - v32, 0.0)
    f[:, 22] = 1.0 / np.asarray(mass, np.float64)
    f[:, 23] = deg
    return f
