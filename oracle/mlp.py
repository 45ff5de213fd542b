"""CPU oracle of the dense GELU network used by the DOS surrogate -- test
infrastructure only.

``forward_backward(params, widths, x, t, bf16=True)`` follows the GPU
kernels step for step: BF16 rounding (round to nearest even) of every matrix
operand at the same points (inputs, weights, hidden activations, the
back-propagated dZ), FP64 accumulation instead of FP32.  With ``bf16=False``
it is the exact FP64 network, pinned by central finite differences
(``finite_difference_grad``).
"""
import math

import numpy as np

__all__ = ["bf16_round", "unpack", "forward_backward", "finite_difference_grad",
           "gelu", "gelu_d"]


def bf16_round(x):
    """float32 -> bfloat16 (round to nearest even) -> float64 values."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    cdf = 0.5 * (1.0 + _erf(z / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)
    return cdf + z * pdf


def unpack(params, widths):
    out, off = [], 0
    for l in range(len(widths) - 1):
        a, b = widths[l], widths[l + 1]
        W = np.asarray(params[off:off + a * b], dtype=np.float64).reshape(a, b)
        off += a * b
        bias = np.asarray(params[off:off + b], dtype=np.float64)
        off += b
        out.append((W, bias))
    return out


def forward_backward(params, widths, x, t, bf16=True):
    """MSE loss mean((y - t)^2) and its packed gradient."""
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, widths)
    L = len(layers)
    a = [r(x)]
    zs = []
    for l, (W, b) in enumerate(layers):
        z = a[-1] @ r(W) + b
        zs.append(z)
        if l < L - 1:
            a.append(r(gelu(z)))
    y = zs[-1]
    n = y.size
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(d * d) / n)
    dz = 2.0 * d / n
    grads = [None] * L
    for l in range(L - 1, -1, -1):
        W, b = layers[l]
        dzb = r(dz)
        gW = a[l].T @ dzb
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dz = (dzb @ r(W).T) * gelu_d(zs[l - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat, y


def finite_difference_grad(params, widths, x, t, idx, h=1e-6):
    """Central differences of the exact FP64 loss at parameter indices."""
    p = np.asarray(params, dtype=np.float64).copy()
    out = []
    for i in idx:
        p0 = p[i]
        p[i] = p0 + h
        lp = forward_backward(p, widths, x, t, bf16=False)[0]
        p[i] = p0 - h
        lm = forward_backward(p, widths, x, t, bf16=False)[0]
        p[i] = p0
        out.append((lp - lm) / (2 * h))
    return np.array(out)
