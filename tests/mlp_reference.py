"""CPU reference of the surrogate network's loss and gradient (numpy,
float64 arithmetic), with the GPU kernels' bf16 rounding points reproduced
when `emulate_bf16` is set: inputs, hidden activations, weights and output
gradients are rounded to bf16 (round to nearest even) where the matrix-core
products read them; accumulation, biases and losses stay wide."""
import numpy as np


def bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + (((u >> 16) & 1) + 0x7FFF)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def unflatten(p, shapes):
    out, o = [], 0
    for n, k in shapes:
        W = p[o:o + n * k].reshape(n, k)
        o += n * k
        b = p[o:o + n]
        o += n
        out.append((W, b))
    return out


def shapes(n_in, width, depth, n_out):
    dims = [n_in] + [width] * depth + [n_out]
    return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def features(x, in_shift, in_scale, kz_features=False):
    """Network input rows (float32): normalised inputs, then with kz_features
    Re and Im of sqrt(1 - x0^2 - x1^2) of the raw first two inputs."""
    x = np.float32(x)
    f = (x - np.float32(in_shift)) * np.float32(in_scale)
    if kz_features:
        u2 = x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]
        f = np.concatenate([f, np.sqrt(np.maximum(1 - u2, 0))[:, None], np.sqrt(np.maximum(u2 - 1, 0))[:, None]],
                           axis=1)
    return f


def loss_grad(p, shp, x, y, in_shift, in_scale, out_shift, out_scale, emulate_bf16=True, kz_features=False):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(np.asarray(p, np.float64), shp)
    n, nout = y.shape
    X = r(features(x, in_shift, in_scale, kz_features)).astype(np.float64)
    H, Z = [X], []
    for l, (W, b) in enumerate(layers):
        z = H[-1] @ r(W).astype(np.float64).T + b
        Z.append(z)
        if l < len(layers) - 1:
            H.append(r(silu(z)).astype(np.float64))
    t = (np.float64(y) - out_shift) / out_scale
    d = Z[-1] - t
    L = np.sum(d * d) / (n * nout)
    dZ = r(2 * d / (n * nout)).astype(np.float64)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        grads[l] = (dZ.T @ H[l], dZ.sum(axis=0))
        if l > 0:
            dH = dZ @ r(W).astype(np.float64)
            dZ = r(dH * dsilu(Z[l - 1])).astype(np.float64)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return L, flat


def forward(p, shp, x, in_shift, in_scale, out_shift, out_scale, emulate_bf16=True, kz_features=False):
    """De-normalised network outputs y (n, n_out), the forward half of loss_grad."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unflatten(np.asarray(p, np.float64), shp)
    H = r(features(x, in_shift, in_scale, kz_features)).astype(np.float64)
    for l, (W, b) in enumerate(layers):
        z = H @ r(W).astype(np.float64).T + b
        H = r(silu(z)).astype(np.float64) if l < len(layers) - 1 else z
    return H * np.float64(out_scale) + np.float64(out_shift)
