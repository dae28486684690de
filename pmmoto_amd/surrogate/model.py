"""A 3-D CNN that predicts a domain's drainage saturation curve.

Input: a cubic uint8 pore image (1 = pore).  Six convolution blocks
(kernel 3, padding 1, stride 2, bias, ReLU) halve the grid each time; a
global average pool, a dense layer with ReLU and a dense output layer give
q saturation values and a log variance (the network's own uncertainty;
its gradient trains only its own output row, the trunk learns from the
curve error alone).
All arithmetic runs in the library's kernels (BF16 operands on the matrix
cores, FP32 accumulation, FP32 master weights with Adam); torch only holds
the buffers.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .. import _device, _lib

BF16 = torch.bfloat16


def _p(t):
    return _device.ptr(t)


def _L():
    return _lib.lib()


def gemm(m, n, k, a, lda, a_trans, b, ldb, b_trans, c, ldc, accumulate=False, k_splits=1):
    _lib.check(_L().pmx_gemm_bf16(int(m), int(n), int(k), _p(a), int(lda), int(a_trans), _p(b), int(ldb),
                                  int(b_trans), _p(c), int(ldc), int(accumulate or k_splits > 1), int(k_splits),
                                  _device.stream()))


def to_bf16(src: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    out = torch.empty(src.shape, dtype=BF16, device=src.device) if out is None else out
    _lib.check(_L().pmx_to_bf16(_p(src), src.numel(), _p(out), _device.stream()))
    return out


def _splits(k_total: int) -> int:
    return int(max(1, min(1024, k_total // 8192)))


class SurrogateNet:
    """channels: output channels of the six stride-2 blocks; q: curve points."""

    def __init__(self, size=128, channels=(8, 16, 48, 96, 160, 256), hidden=256, q=32, seed=0, device="cuda"):
        _device.require_gpu()
        self.size, self.channels, self.hidden, self.q = size, tuple(channels), hidden, q
        self.strides = (2,) * len(channels)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.params = []  # [name, fp32 master, bf16 copy or None]
        cin = 1
        for i, co in enumerate(self.channels):
            fan_in = 27 * cin
            w = torch.randn((co, 3, 3, 3, cin), generator=g) * math.sqrt(2.0 / fan_in)
            self.params.append([f"conv{i}.w", w.to(device), None])
            self.params.append([f"conv{i}.b", torch.zeros(co).to(device), None])
            cin = co
        w1 = torch.randn((hidden, cin), generator=g) * math.sqrt(2.0 / cin)
        w2 = torch.randn((q + 1, hidden), generator=g) * math.sqrt(1.0 / hidden)
        b2 = torch.zeros(q + 1)
        b2[:q] = 0.5
        b2[q] = -4.0
        self.params += [["fc1.w", w1.to(device), None], ["fc1.b", torch.zeros(hidden).to(device), None],
                        ["fc2.w", w2.to(device), None], ["fc2.b", b2.to(device), None]]
        for p in self.params:
            if p[0].endswith(".w"):
                p[2] = to_bf16(p[1])
        self.m = [torch.zeros_like(p[1]) for p in self.params]
        self.v = [torch.zeros_like(p[1]) for p in self.params]
        self.step_count = 0
        self._infer_bufs = {}
        self._graphs = {}  # (input address, shape) -> (captured inference, its output)

    def n_params(self) -> int:
        return sum(p[1].numel() for p in self.params)

    def param(self, name):
        for p in self.params:
            if p[0] == name:
                return p
        raise KeyError(name)

    def flops_per_sample(self) -> float:
        """Forward multiply-adds x 2 of the convolutions and dense layers."""
        f, n, cin = 0.0, self.size, 1
        for co in self.channels:
            n = (n - 1) // 2 + 1
            f += 2.0 * n ** 3 * co * 27 * cin
            cin = co
        return f + 2.0 * (cin * self.hidden + self.hidden * (self.q + 1))

    # -- forward ---------------------------------------------------------
    def forward(self, x_u8: torch.Tensor, keep=True):
        """x_u8: CUDA uint8 [B, S, S, S].  Returns pred [B, q+1] (FP32) and the
        cache the backward pass needs."""
        b = x_u8.shape[0]
        shape = [b, x_u8.shape[1], x_u8.shape[2], x_u8.shape[3], 1]
        h = torch.empty(tuple(shape), dtype=BF16, device=x_u8.device)
        _lib.check(_L().pmx_u8_to_bf16(_p(x_u8), x_u8.numel(), _p(h), _device.stream()))
        cache = {"acts": [h], "cols": [], "shapes": [tuple(shape)]}
        for i, co in enumerate(self.channels):
            w = self.param(f"conv{i}.w")
            bias = self.param(f"conv{i}.b")[1]
            cin = shape[4]
            kk = 27 * cin
            od, oh, ow = [(v - 1) // 2 + 1 for v in shape[1:4]]
            m = b * od * oh * ow
            cols = torch.empty((m, kk), dtype=BF16, device=x_u8.device)
            _lib.check(_L().pmx_im2col3d(_p(h), (C.c_int32 * 5)(*shape), 2, kk, _p(cols), _device.stream()))
            y = torch.empty((m, co), dtype=torch.float32, device=x_u8.device)
            gemm(m, co, kk, cols, kk, 0, w[2], kk, 1, y, co)
            a = torch.empty((b, od, oh, ow, co), dtype=BF16, device=x_u8.device)
            _lib.check(_L().pmx_bias_act(_p(y), _p(bias), m, co, 1, _p(a), None, _device.stream()))
            del y
            shape = [b, od, oh, ow, co]
            if keep:
                cache["cols"].append(cols)
                cache["acts"].append(a)
                cache["shapes"].append(tuple(shape))
            else:
                del cols
                cache["acts"] = [a]
            h = a
        pred, pooled, z1, p, cf = self._head(h, shape)
        cache.update(pooled=pooled, z1=z1, p=p, cf=cf, b=b)
        return pred, cache

    # -- backward --------------------------------------------------------
    def backward(self, cache, dpred: torch.Tensor):
        """Gradients (FP32, same order as self.params) of a loss whose
        gradient with respect to pred is dpred."""
        dev = dpred.device
        b, p, cf = cache["b"], cache["p"], cache["cf"]
        grads = {}
        # fc2
        dpred_bf = to_bf16(dpred)
        w2 = self.param("fc2.w")
        gw2 = torch.zeros((self.q + 1, self.hidden), dtype=torch.float32, device=dev)
        gemm(self.q + 1, self.hidden, b, dpred_bf, self.q + 1, 1, cache["z1"], self.hidden, 0, gw2, self.hidden,
             accumulate=True)
        gb2 = torch.zeros(self.q + 1, dtype=torch.float32, device=dev)
        _lib.check(_L().pmx_col_sum(_p(dpred), b, self.q + 1, _p(gb2), _device.stream()))
        grads["fc2.w"], grads["fc2.b"] = gw2, gb2
        # the uncertainty output reads the features but does not shape them:
        # only the curve error flows back into the trunk
        trunk_bf = dpred_bf.clone()
        trunk_bf[:, self.q] = 0
        dz1 = torch.empty((b, self.hidden), dtype=torch.float32, device=dev)
        gemm(b, self.hidden, self.q + 1, trunk_bf, self.q + 1, 0, w2[2], self.hidden, 0, dz1, self.hidden)
        # fc1 (ReLU)
        dy1 = torch.empty_like(dz1)
        dy1_bf = torch.empty((b, self.hidden), dtype=BF16, device=dev)
        _lib.check(_L().pmx_relu_backward(_p(dz1), _p(cache["z1"]), dz1.numel(), _p(dy1), _p(dy1_bf),
                                          _device.stream()))
        w1 = self.param("fc1.w")
        gw1 = torch.zeros((self.hidden, cf), dtype=torch.float32, device=dev)
        gemm(self.hidden, cf, b, dy1_bf, self.hidden, 1, cache["pooled"], cf, 0, gw1, cf, accumulate=True)
        gb1 = torch.zeros(self.hidden, dtype=torch.float32, device=dev)
        _lib.check(_L().pmx_col_sum(_p(dy1), b, self.hidden, _p(gb1), _device.stream()))
        grads["fc1.w"], grads["fc1.b"] = gw1, gb1
        dpool = torch.empty((b, cf), dtype=torch.float32, device=dev)
        gemm(b, cf, self.hidden, dy1_bf, self.hidden, 0, w1[2], cf, 0, dpool, cf)
        # pool
        last = cache["shapes"][-1]
        dh = torch.empty(last, dtype=torch.float32, device=dev)
        _lib.check(_L().pmx_avgpool_backward(_p(dpool), b, p, cf, _p(dh), _device.stream()))
        # conv blocks, last to first
        for i in range(len(self.channels) - 1, -1, -1):
            a = cache["acts"][i + 1]
            cols = cache["cols"][i]
            in_shape = cache["shapes"][i]
            co = self.channels[i]
            kk = 27 * in_shape[4]
            m = a.numel() // co
            dz = torch.empty((m, co), dtype=torch.float32, device=dev)
            dz_bf = torch.empty((m, co), dtype=BF16, device=dev)
            _lib.check(_L().pmx_relu_backward(_p(dh), _p(a), dz.numel(), _p(dz), _p(dz_bf), _device.stream()))
            del dh
            gw = torch.zeros((co, kk), dtype=torch.float32, device=dev)
            gemm(co, kk, m, dz_bf, co, 1, cols, kk, 0, gw, kk, accumulate=True, k_splits=_splits(m))
            gb = torch.zeros(co, dtype=torch.float32, device=dev)
            _lib.check(_L().pmx_col_sum(_p(dz), m, co, _p(gb), _device.stream()))
            grads[f"conv{i}.w"] = gw.reshape(self.param(f"conv{i}.w")[1].shape)
            grads[f"conv{i}.b"] = gb
            del dz
            if i > 0:
                w = self.param(f"conv{i}.w")
                dcols = torch.empty((m, kk), dtype=torch.float32, device=dev)
                gemm(m, kk, co, dz_bf, co, 0, w[2], kk, 0, dcols, kk)
                dh = torch.empty(in_shape, dtype=torch.float32, device=dev)
                _lib.check(_L().pmx_col2im3d(_p(dcols), (C.c_int32 * 5)(*in_shape), 2, kk, _p(dh), _device.stream()))
                del dcols
            del dz_bf
        return [grads[p[0]] for p in self.params]

    def loss_and_grad(self, pred: torch.Tensor, target: torch.Tensor):
        b = pred.shape[0]
        loss = torch.empty(b, dtype=torch.float32, device=pred.device)
        dpred = torch.empty_like(pred)
        _lib.check(_L().pmx_gnll_loss(_p(pred), _p(target), b, self.q, _p(loss), _p(dpred), _device.stream()))
        return loss, dpred

    def adam_step(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for (name, w, wb), g, m, v in zip(self.params, grads, self.m, self.v):
            _lib.check(_L().pmx_adam(_p(w), _p(g), _p(m), _p(v), w.numel(), lr, b1, b2, eps, self.step_count,
                                     _p(wb) if wb is not None else None, _device.stream()))

    def train_step(self, x_u8, target, lr=1e-3):
        pred, cache = self.forward(x_u8)
        loss, dpred = self.loss_and_grad(pred, target)
        grads = self.backward(cache, dpred)
        self.adam_step(grads, lr)
        return loss

    # -- inference ---------------------------------------------------------
    def infer(self, x_u8: torch.Tensor):
        """Forward pass without the training cache: the convolutions run as
        fused kernels (the first layer on the uint8 image, implicit GEMM on
        the matrix cores after it; on a width-128 image with 8 and 16
        channels the first two layers in one pass) with bias, ReLU and BF16
        rounding in their epilogue, and the activations between them keep a one-voxel zero
        border so that the next layer's taps need no range checks.  Same
        arithmetic as forward() up to FP32 summation order.  Returns (pred
        [B, q+1] FP32, pooled, z1, p, cf)."""
        b = x_u8.shape[0]
        dev = x_u8.device
        shape = [b, x_u8.shape[1], x_u8.shape[2], x_u8.shape[3], 1]
        kinds = []
        cin = 1
        for i, co in enumerate(self.channels):
            kinds.append("first" if i == 0 and co % 4 == 0 and co <= 32 else "igemm" if cin % 8 == 0 else "im2col")
            cin = co
        # the first two layers in one pass when they are 8 and 16 channels at stride 2 on a width-128 image
        fuse = (len(self.channels) > 1 and kinds[:2] == ["first", "igemm"] and self.channels[:2] == (8, 16)
                and self.strides[:2] == (2, 2) and shape[3] == 128 and shape[2] % 32 == 0 and shape[1] % 4 == 0)
        key = (b,) + tuple(shape[1:4])
        bufs = self._infer_bufs.get(key)
        if bufs is None:
            # padded (zero border, kernels write the interior only) when the next layer reads it directly
            bufs, shp = [], list(shape)
            for i, co in enumerate(self.channels):
                s = self.strides[i]
                od, oh, ow = [(v - 1) // s + 1 for v in shp[1:4]]
                if fuse and i == 0:
                    bufs.append(None)  # stays on chip
                elif i + 1 < len(self.channels) and kinds[i + 1] == "igemm":
                    bufs.append(torch.zeros((b, od + 2, oh + 2, ow + 2, co), dtype=BF16, device=dev))
                else:
                    bufs.append(torch.empty((b, od, oh, ow, co), dtype=BF16, device=dev))
                shp = [b, od, oh, ow, co]
            self._infer_bufs = {key: bufs}  # one batch shape cached at a time
        h, h_pad = None, 0
        for i, co in enumerate(self.channels):
            w = self.param(f"conv{i}.w")
            bias = self.param(f"conv{i}.b")[1]
            s = self.strides[i]
            od, oh, ow = [(v - 1) // s + 1 for v in shape[1:4]]
            if fuse and i == 0:
                image_dims = (C.c_int32 * 5)(*shape)
                shape = [b, od, oh, ow, co]
                continue
            a = bufs[i]
            a_pad = 1 if a.shape[1] == od + 2 else 0
            dims = (C.c_int32 * 5)(*shape)
            if fuse and i == 1:
                w0, b0 = self.param("conv0.w"), self.param("conv0.b")[1]
                _lib.check(_L().pmx_conv3d_first_pair(_p(x_u8), image_dims, _p(w0[2]), _p(b0), _p(w[2]), _p(bias),
                                                      _p(a), a_pad, _device.stream()))
            elif kinds[i] == "first":
                _lib.check(_L().pmx_conv3d_first(_p(x_u8), dims, s, co, _p(w[2]), _p(bias), _p(a), a_pad,
                                                 _device.stream()))
            else:
                if h is None:
                    h = torch.empty(tuple(shape), dtype=BF16, device=dev)
                    _lib.check(_L().pmx_u8_to_bf16(_p(x_u8), x_u8.numel(), _p(h), _device.stream()))
                if kinds[i] == "igemm":
                    _lib.check(_L().pmx_conv3d_bf16(_p(h), dims, h_pad, s, co, _p(w[2]), _p(bias), 1, _p(a), a_pad,
                                                    _device.stream()))
                else:
                    kk = 27 * shape[4]
                    m = b * od * oh * ow
                    cols = torch.empty((m, kk), dtype=BF16, device=dev)
                    _lib.check(_L().pmx_im2col3d(_p(h), dims, s, kk, _p(cols), _device.stream()))
                    y = torch.empty((m, co), dtype=torch.float32, device=dev)
                    gemm(m, co, kk, cols, kk, 0, w[2], kk, 1, y, co)
                    del cols
                    _lib.check(_L().pmx_bias_act(_p(y), _p(bias), m, co, 1, _p(a), None, _device.stream()))
                    del y
            h, h_pad = a, a_pad
            shape = [b, od, oh, ow, co]
        return self._head(h, shape)

    def _head(self, h, shape):
        b, dev = shape[0], h.device
        p = shape[1] * shape[2] * shape[3]
        cf = shape[4]
        pooled = torch.empty((b, cf), dtype=BF16, device=dev)
        _lib.check(_L().pmx_avgpool(_p(h), b, p, cf, None, _p(pooled), _device.stream()))
        w1, b1 = self.param("fc1.w"), self.param("fc1.b")[1]
        y1 = torch.empty((b, self.hidden), dtype=torch.float32, device=dev)
        gemm(b, self.hidden, cf, pooled, cf, 0, w1[2], cf, 1, y1, self.hidden)
        z1 = torch.empty((b, self.hidden), dtype=BF16, device=dev)
        _lib.check(_L().pmx_bias_act(_p(y1), _p(b1), b, self.hidden, 1, _p(z1), None, _device.stream()))
        w2, b2 = self.param("fc2.w"), self.param("fc2.b")[1]
        y2 = torch.empty((b, self.q + 1), dtype=torch.float32, device=dev)
        gemm(b, self.q + 1, self.hidden, z1, self.hidden, 0, w2[2], self.hidden, 1, y2, self.q + 1)
        pred = torch.empty((b, self.q + 1), dtype=torch.float32, device=dev)
        _lib.check(_L().pmx_bias_act(_p(y2), _p(b2), b, self.q + 1, 0, None, _p(pred), _device.stream()))
        return pred, pooled, z1, p, cf

    @torch.no_grad()
    def infer_graphed(self, x_u8: torch.Tensor) -> torch.Tensor:
        """infer()'s prediction [B, q+1] through a captured graph of its
        launches, replayed while the same input tensor (address and shape)
        comes back; a new input is captured once (a few graphs are kept).
        The output tensor is the graph's and is rewritten by the next replay."""
        key = (x_u8.data_ptr(), tuple(x_u8.shape))
        hit = self._graphs.get(key)
        if hit is None:
            self.infer(x_u8)  # eager first: allocates the activation buffers outside the capture
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                out = self.infer(x_u8)[0]
            if len(self._graphs) >= 8:
                self._graphs.clear()
            # the entry holds the activation buffers the graph writes: the eager cache may drop them
            hit = self._graphs[key] = (graph, out, list(self._infer_bufs.values()))
        hit[0].replay()
        return hit[1]

    def predict(self, x_u8: torch.Tensor, batch=64, graph=True):
        """(curve [B, q] clipped to [0, 1], sigma [B]) in FP32.  graph: run
        each batch's launches as a captured graph (same results)."""
        outs = []
        for i in range(0, x_u8.shape[0], batch):
            xb = x_u8[i:i + batch].contiguous()
            outs.append(self.infer_graphed(xb).clone() if graph else self.infer(xb)[0])
        pred = torch.cat(outs)
        curve = pred[:, :self.q].clamp(0.0, 1.0)
        sigma = torch.exp(0.5 * pred[:, self.q].clamp(-12.0, 12.0))
        return curve, sigma

    def state(self):
        return {p[0]: p[1].detach().cpu().numpy().astype(np.float64) for p in self.params}
