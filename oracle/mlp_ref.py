"""CPU reference of the learned-guess MLP (test infrastructure only).

The same network as the library's polyeri_mlp_* (layer l: Z = A W^T + b,
A = silu(Z), last layer linear; loss sum w (y - t)^2 / sum w), in float64,
with the operands of every matrix product rounded to BF16 the way the device
rounds them before its MFMA products (round to nearest even), so device
gradients can be compared to it at FP32-accumulation accuracy.  With
bf16=False it is the plain float64 network, checked against finite
differences in the tests.  Only tests/ use this module.
"""
import numpy as np


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, dims):
    """Flat parameter vector -> [(W (d_l x d_{l-1}), b (d_l))] in the
    library's order (W_1, b_1, W_2, b_2, ...)."""
    out, k = [], 0
    for l in range(1, len(dims)):
        nw = dims[l] * dims[l - 1]
        W = np.asarray(params[k:k + nw], dtype=np.float64).reshape(dims[l], dims[l - 1])
        k += nw
        b = np.asarray(params[k:k + dims[l]], dtype=np.float64)
        k += dims[l]
        out.append((W, b))
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, dims, X, rnd=True):
    r = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    A = np.asarray(X, dtype=np.float64)
    zs, acts = [], [A]
    layers = unpack(params, dims)
    for l, (W, b) in enumerate(layers):
        Z = r(A) @ r(W).T + b
        A = _silu(Z) if l < len(layers) - 1 else Z
        zs.append(Z)
        acts.append(A)
    return A, zs, acts


def loss_and_grads(params, dims, X, T, Wt, rnd=True):
    """Loss and the flat gradient vector (library order)."""
    r = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    Y, zs, acts = forward(params, dims, X, rnd)
    ws = float(np.sum(Wt))
    d = Y - T
    loss = float(np.sum(Wt * d * d) / ws)
    dZ = 2.0 * Wt * d / ws
    layers = unpack(params, dims)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        gW = r(dZ).T @ r(acts[l])
        gb = dZ.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dA = r(dZ) @ r(W)
            dZ = dA * _silu_grad(zs[l - 1])
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat
