"""CPU checker of the warm-start MLP kernels (test infrastructure only).

Restates the forward and backward of the 32-128-128-128-5 SiLU network in numpy with
the quantisation points of the MFMA kernels: BF16 (round-to-nearest-even) inputs,
weights, hidden activations and hidden-layer gradients, FP32 accumulation, FP32
biases.  With quant=False and float64 it is the exact network, for finite-difference
checks of the backward itself.
"""
import numpy as np

SHAPES = [(128, 32), (128, 128), (128, 128), (5, 128)]


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    out, o = [], 0
    for fo, fi in SHAPES:
        W = p[o:o + fo * fi].reshape(fo, fi); o += fo * fi
        b = p[o:o + fo]; o += fo
        out.append((W, b))
    return out


def forward(p, X, in_mean, in_std, quant=True, dtype=np.float32):
    q = bf16 if quant else (lambda a: a)
    layers = unpack(np.asarray(p, dtype))
    istd = np.where(in_std > 0, 1.0 / np.where(in_std > 0, in_std, 1), 0).astype(np.float32)
    if quant:
        x0 = bf16((np.asarray(X, np.float32) - in_mean.astype(np.float32)) * istd)
    else:
        x0 = (np.asarray(X, dtype) - in_mean) * istd
    acts, zs = [x0], []
    h = x0
    for (W, b) in layers[:3]:
        z = (h.astype(dtype) @ q(W).astype(dtype).T + b).astype(dtype)
        zs.append(z)
        s = 1.0 / (1.0 + np.exp(-z))
        h = q((z * s).astype(dtype))
        acts.append(h)
    W4, b4 = layers[3]
    y = (h.astype(dtype) @ q(W4).astype(dtype).T + b4).astype(dtype)
    return y, acts, zs


def loss_grad(p, X, T, M, in_mean, in_std, quant=True, dtype=np.float32):
    q = bf16 if quant else (lambda a: a)
    layers = unpack(np.asarray(p, dtype))
    y, acts, zs = forward(p, X, in_mean, in_std, quant, dtype)
    n = y.size
    d = (y - T) * M
    loss = (d * d).sum() / n
    dy = (2.0 * d / n).astype(dtype)
    grads = [None] * 4
    W4, _ = layers[3]
    grads[3] = (dy.T @ acts[3].astype(dtype), dy.sum(0))
    dh = dy @ q(W4).astype(dtype)
    for l in (2, 1, 0):
        z = zs[l]
        s = 1.0 / (1.0 + np.exp(-z))
        dz = (dh * (s * (1.0 + z * (1.0 - s)))).astype(dtype)
        dzq = q(dz).astype(dtype)
        grads[l] = (dzq.T @ acts[l].astype(dtype), dz.sum(0))
        if l > 0:
            W, _ = layers[l]
            dh = dzq @ q(W).astype(dtype)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
