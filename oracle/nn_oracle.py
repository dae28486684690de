"""Oracle of the surrogate network's arithmetic (test infrastructure only).

float64 numpy restatement of the layers the GPU kernels implement: 3-D
convolution (kernel 3, padding 1, stride 1 or 2, channels-last), dense
layers, ReLU, global average pooling and the Gaussian negative
log-likelihood, with their gradients.  Nothing here comes from the
reference (PMMoTo has no learned component); the arithmetic is pinned by
tests/test_oracle_nn.py against a direct loop convolution on tiny inputs
and against central finite differences of the whole network.
"""

from __future__ import annotations

import numpy as np


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def out_size(n, s):
    return (n - 1) // s + 1


def im2col(x, s):
    """x [N,D,H,W,C] -> patches [N*od*oh*ow, 27*C], k = ((kz*3+ky)*3+kx)*C + c."""
    n, d, h, w, c = x.shape
    od, oh, ow = out_size(d, s), out_size(h, s), out_size(w, s)
    xp = np.zeros((n, d + 2, h + 2, w + 2, c), dtype=np.float64)
    xp[:, 1:-1, 1:-1, 1:-1] = x
    cols = np.empty((n, od, oh, ow, 27, c), dtype=np.float64)
    t = 0
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                cols[:, :, :, :, t] = xp[:, kz:kz + s * (od - 1) + 1:s, ky:ky + s * (oh - 1) + 1:s,
                                         kx:kx + s * (ow - 1) + 1:s]
                t += 1
    return cols.reshape(n * od * oh * ow, 27 * c)


def col2im(dcols, shape, s):
    """Adjoint of im2col: scatter-add the patch gradients onto [N,D,H,W,C]."""
    n, d, h, w, c = shape
    od, oh, ow = out_size(d, s), out_size(h, s), out_size(w, s)
    g = dcols.reshape(n, od, oh, ow, 27, c)
    xp = np.zeros((n, d + 2, h + 2, w + 2, c), dtype=np.float64)
    t = 0
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                xp[:, kz:kz + s * (od - 1) + 1:s, ky:ky + s * (oh - 1) + 1:s, kx:kx + s * (ow - 1) + 1:s] += g[:, :, :, :, t]
                t += 1
    return xp[:, 1:-1, 1:-1, 1:-1]


def conv_fwd(x, w, b, s):
    """x [N,D,H,W,Ci], w [Co,3,3,3,Ci] -> pre-activation [N,od,oh,ow,Co]."""
    n, d, h, wd, c = x.shape
    cols = im2col(x, s)
    y = cols @ w.reshape(w.shape[0], -1).T + b
    return y.reshape(n, out_size(d, s), out_size(h, s), out_size(wd, s), w.shape[0])


def conv_bwd(x, w, dy, s):
    """Gradients of conv_fwd: (dx, dw, db)."""
    cols = im2col(x, s)
    g = dy.reshape(-1, w.shape[0])
    dw = (g.T @ cols).reshape(w.shape)
    db = g.sum(axis=0)
    dx = col2im(g @ w.reshape(w.shape[0], -1), x.shape, s)
    return dx, dw, db


def conv_direct(x, w, b, s):
    """Direct loop convolution (the pin for im2col on tiny inputs)."""
    n, d, h, wd, c = x.shape
    co = w.shape[0]
    od, oh, ow = out_size(d, s), out_size(h, s), out_size(wd, s)
    y = np.zeros((n, od, oh, ow, co))
    for bi in range(n):
        for z in range(od):
            for yy in range(oh):
                for xx in range(ow):
                    for o in range(co):
                        acc = b[o]
                        for kz in range(3):
                            for ky in range(3):
                                for kx in range(3):
                                    iz, iy, ix = z * s - 1 + kz, yy * s - 1 + ky, xx * s - 1 + kx
                                    if 0 <= iz < d and 0 <= iy < h and 0 <= ix < wd:
                                        acc += float(np.dot(w[o, kz, ky, kx], x[bi, iz, iy, ix]))
                        y[bi, z, yy, xx, o] = acc
    return y


def gnll(pred, target):
    """Per-sample loss se/2 + (se e^-s + q s)/2 (se = sum_j (p_j - t_j)^2, s =
    log variance clamped to [-12, 12]), its batch mean and the gradient the
    training uses: the curve by squared error, s by the Gaussian negative
    log-likelihood with se held fixed and the clamp passed straight through."""
    q = target.shape[1]
    s = np.clip(pred[:, q], -12.0, 12.0)
    inv = np.exp(-s)
    e = pred[:, :q] - target
    se = (e * e).sum(axis=1)
    loss = 0.5 * se + 0.5 * (se * inv + q * s)
    n = pred.shape[0]
    d = np.zeros_like(pred)
    d[:, :q] = e / n
    d[:, q] = (-0.5 * se * inv + 0.5 * q) / n
    return loss, loss.mean(), d


def quadratic(pred, target):
    """0.5 mean_b sum_j (pred - target)^2 over all outputs and its gradient
    (a loss whose gradient is exact, for finite-difference checks)."""
    e = pred - target
    return 0.5 * (e * e).sum() / pred.shape[0], e / pred.shape[0]


def network_forward(params, x, strides):
    """The surrogate: conv blocks (conv + bias + ReLU), global average pool,
    dense + ReLU, dense -> [curve, log variance].  Returns (pred, cache)."""
    acts = [x]
    h = x
    for (w, b), s in zip(params["conv"], strides):
        h = np.maximum(conv_fwd(h, w, b, s), 0.0)
        acts.append(h)
    pooled = h.reshape(h.shape[0], -1, h.shape[-1]).mean(axis=1)
    w1, b1 = params["fc1"]
    z1 = np.maximum(pooled @ w1.T + b1, 0.0)
    w2, b2 = params["fc2"]
    pred = z1 @ w2.T + b2
    return pred, (acts, pooled, z1)


def network_backward(params, cache, dpred, strides):
    acts, pooled, z1 = cache
    grads = {}
    w2, _ = params["fc2"]
    grads["fc2"] = (dpred.T @ z1, dpred.sum(axis=0))
    dz1 = (dpred @ w2) * (z1 > 0)
    w1, _ = params["fc1"]
    grads["fc1"] = (dz1.T @ pooled, dz1.sum(axis=0))
    dpool = dz1 @ w1
    h = acts[-1]
    p = int(np.prod(h.shape[1:4]))
    dh = np.repeat(dpool[:, None, :] / p, p, axis=1).reshape(h.shape)
    conv_grads = []
    for li in range(len(params["conv"]) - 1, -1, -1):
        w, _ = params["conv"][li]
        dpre = dh * (acts[li + 1] > 0)
        dx, dw, db = conv_bwd(acts[li], w, dpre, strides[li])
        conv_grads.append((dw, db))
        dh = dx
    grads["conv"] = conv_grads[::-1]
    return grads
