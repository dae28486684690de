"""GPU checks of the surrogate's kernels against the oracle
(oracle/nn_oracle.py): the BF16 matrix-core GEMM (all operand layouts,
split K), im2col/col2im, and the whole network's forward values and
gradients with the same BF16 rounding points emulated in float64.
Tolerances are in each assertion (FP32 accumulation vs float64)."""

import numpy as np
import pytest

from oracle import nn_oracle as N

pytestmark = pytest.mark.gpu


def tbf(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(N.bf16(a)).astype(np.float32)).cuda().to(torch.bfloat16)


@pytest.mark.parametrize("m,n,k,splits", [(70, 37, 45, 1), (130, 65, 300, 1), (33, 129, 5000, 7), (1, 1, 1, 1)])
@pytest.mark.parametrize("at,bt", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_bf16_matches_float64(m, n, k, splits, at, bt):
    import torch

    from pmmoto_amd.surrogate.model import gemm
    rng = np.random.default_rng(m * 7 + n + k)
    a = N.bf16(rng.normal(size=(m, k)))
    b = N.bf16(rng.normal(size=(k, n)))
    A = tbf(a.T if at else a)
    B = tbf(b.T if bt else b)
    c = torch.zeros((m, n), dtype=torch.float32, device="cuda")
    print(f"[gemm] m={m} n={n} k={k} at={at} bt={bt} splits={splits}", flush=True)
    gemm(m, n, k, A, m if at else k, at, B, k if bt else n, bt, c, n, accumulate=splits > 1, k_splits=splits)
    want = a @ b
    np.testing.assert_allclose(c.cpu().numpy(), want, rtol=1e-4, atol=1e-4 * np.sqrt(k))


@pytest.mark.parametrize("s", [1, 2])
def test_im2col_col2im_match_oracle(s):
    import ctypes as C

    import torch

    from pmmoto_amd import _device, _lib
    rng = np.random.default_rng(5 + s)
    x = N.bf16(rng.normal(size=(2, 5, 7, 6, 3)))
    cols_want = N.im2col(x, s)
    ldk = cols_want.shape[1]
    cols = torch.empty(cols_want.shape, dtype=torch.bfloat16, device="cuda")
    shape = (C.c_int32 * 5)(*x.shape)
    xd = tbf(x)
    _lib.check(_lib.lib().pmx_im2col3d(_device.ptr(xd), shape, s, ldk, _device.ptr(cols), _device.stream()))
    np.testing.assert_array_equal(cols.float().cpu().numpy(), cols_want)
    g = rng.normal(size=cols_want.shape).astype(np.float32)
    dx = torch.empty(x.shape, dtype=torch.float32, device="cuda")
    gd = torch.from_numpy(g).cuda()
    _lib.check(_lib.lib().pmx_col2im3d(_device.ptr(gd), shape, s, ldk, _device.ptr(dx),
                                       _device.stream()))
    np.testing.assert_allclose(dx.cpu().numpy(), N.col2im(g.astype(np.float64), x.shape, s), rtol=1e-5, atol=1e-5)


def _pad(a):
    """[N,D,H,W,C] -> zero-bordered [N,D+2,H+2,W+2,C]."""
    out = np.zeros((a.shape[0],) + tuple(v + 2 for v in a.shape[1:4]) + (a.shape[4],), dtype=a.dtype)
    out[:, 1:-1, 1:-1, 1:-1] = a
    return out


@pytest.mark.parametrize("co,s,shape", [(8, 2, (2, 9, 7, 10)), (16, 1, (1, 5, 6, 7)), (32, 2, (3, 4, 4, 4)),
                                        (8, 2, (2, 6, 19, 16)), (16, 2, (1, 3, 5, 136)), (8, 1, (1, 3, 4, 130)),
                                        (12, 2, (1, 4, 6, 8)), (24, 1, (2, 3, 3, 5)),
                                        # the LDS-staged form (stride 2, w % 16 == 0): two column blocks,
                                        # a partial row block, odd depth, one and two channel blocks
                                        (8, 2, (2, 7, 35, 256)), (24, 2, (1, 5, 40, 48)), (12, 2, (3, 3, 2, 32))])
@pytest.mark.parametrize("out_pad", [0, 1])
def test_first_layer_conv_matches_oracle(co, s, shape, out_pad):
    """uint8 first-layer convolution on the matrix cores, bias + ReLU + BF16 fused."""
    import ctypes as C

    import torch

    from pmmoto_amd import _device, _lib
    rng = np.random.default_rng(co + s)
    x = (rng.random(shape) < 0.5).astype(np.uint8)
    x[0, 0, 0, :3] = (7, 200, 255)  # bytes other than 0/1 convert exactly too
    w = N.bf16(rng.normal(size=(co, 3, 3, 3, 1)))
    b = rng.normal(size=co).astype(np.float32).astype(np.float64)
    want = N.bf16(np.maximum(N.conv_direct(x[..., None].astype(np.float64), w, b, s), 0.0))
    if out_pad:
        want = _pad(want)
    out = torch.zeros(want.shape, dtype=torch.bfloat16, device="cuda")
    dims = (C.c_int32 * 5)(*shape, 1)
    print(f"[conv_first] co={co} s={s} shape={shape} out_pad={out_pad}", flush=True)
    # operands held in names: a temporary freed before the launch would hand its memory to the next one
    xd, wd, bd = torch.from_numpy(x).cuda(), tbf(w), torch.from_numpy(b.astype(np.float32)).cuda()
    _lib.check(_lib.lib().pmx_conv3d_first(_device.ptr(xd), dims, s, co, _device.ptr(wd), _device.ptr(bd),
                                           _device.ptr(out), out_pad, _device.stream()))
    # FP32 vs float64 sums may round to neighbouring BF16 values; the border stays zero
    np.testing.assert_allclose(out.float().cpu().numpy(), want, rtol=2 ** -7, atol=1e-6)


@pytest.mark.parametrize("c,co,s,shape,relu", [(8, 16, 2, (2, 9, 7, 10), 1), (16, 48, 2, (1, 6, 5, 8), 1),
                                               (8, 24, 1, (1, 5, 4, 6), 0), (48, 96, 2, (2, 4, 4, 4), 1),
                                               (16, 130, 1, (1, 3, 3, 3), 1), (24, 20, 2, (1, 5, 5, 5), 1),
                                               (160, 256, 2, (3, 2, 2, 2), 1)])
@pytest.mark.parametrize("in_pad,out_pad", [(0, 0), (1, 1), (1, 0)])
def test_implicit_gemm_conv_matches_oracle(c, co, s, shape, relu, in_pad, out_pad):
    """Matrix-core convolution without a patch matrix, bias/ReLU/BF16 fused,
    on plain and zero-bordered tensors."""
    import ctypes as C

    import torch

    from pmmoto_amd import _device, _lib
    rng = np.random.default_rng(c * 3 + co)
    x = N.bf16(rng.normal(size=shape + (c,)))
    w = N.bf16(rng.normal(size=(co, 3, 3, 3, c)) / np.sqrt(27 * c))
    b = rng.normal(size=co).astype(np.float32).astype(np.float64)
    y = N.conv_fwd(x, w, b, s)
    want = N.bf16(np.maximum(y, 0.0) if relu else y)
    if out_pad:
        want = _pad(want)
    out = torch.zeros(want.shape, dtype=torch.bfloat16, device="cuda")
    dims = (C.c_int32 * 5)(*shape, c)
    print(f"[conv_igemm] c={c} co={co} s={s} shape={shape} pads={in_pad},{out_pad}", flush=True)
    xd, wd, bd = tbf(_pad(x) if in_pad else x), tbf(w), torch.from_numpy(b.astype(np.float32)).cuda()
    _lib.check(_lib.lib().pmx_conv3d_bf16(_device.ptr(xd), dims, in_pad, s, co, _device.ptr(wd), _device.ptr(bd),
                                          relu, _device.ptr(out), out_pad, _device.stream()))
    np.testing.assert_allclose(out.float().cpu().numpy(), want, rtol=2 ** -7, atol=1e-5)


@pytest.mark.parametrize("c,co,shape", [(8, 16, (8, 64, 64, 64)), (16, 48, (64, 32, 32, 32)),
                                         (48, 96, (128, 16, 16, 16)), (96, 160, (256, 8, 8, 8)),
                                         (160, 256, (512, 4, 4, 4)), (24, 64, (64, 32, 32, 32))])
@pytest.mark.parametrize("in_pad", [1, 0])
def test_implicit_gemm_conv_at_network_sizes_matches_patch_matrix_path(c, co, shape, in_pad):
    """The network's layer shapes at inference batches (64 or 32 positions
    per wave, 1..4 channel blocks): the implicit-GEMM kernel against the
    training forward's patch matrix + GEMM + bias/ReLU on the same BF16
    input.  Both accumulate in FP32 in different orders, so an output may
    round to the neighbouring BF16 value."""
    import ctypes as C

    import torch

    from pmmoto_amd import _device, _lib
    from pmmoto_amd.surrogate.model import gemm
    L = _lib.lib()
    g = torch.Generator(device="cpu").manual_seed(c + co)
    x = torch.relu(torch.randn(shape + (c,), generator=g)).to(torch.bfloat16).cuda()
    w = (torch.randn((co, 27 * c), generator=g) / (27 * c) ** 0.5).to(torch.bfloat16).cuda()
    b = torch.randn(co, generator=g).cuda()
    n, d, h, wd = shape
    od, oh, ow = [(v - 1) // 2 + 1 for v in (d, h, wd)]
    m, kk = n * od * oh * ow, 27 * c
    dims = (C.c_int32 * 5)(n, d, h, wd, c)
    print(f"[conv_igemm] network size c={c} co={co} shape={shape} in_pad={in_pad}", flush=True)
    cols = torch.empty((m, kk), dtype=torch.bfloat16, device="cuda")
    _lib.check(L.pmx_im2col3d(_device.ptr(x), dims, 2, kk, _device.ptr(cols), _device.stream()))
    y = torch.empty((m, co), dtype=torch.float32, device="cuda")
    gemm(m, co, kk, cols, kk, 0, w, kk, 1, y, co)
    want = torch.empty((m, co), dtype=torch.bfloat16, device="cuda")
    _lib.check(L.pmx_bias_act(_device.ptr(y), _device.ptr(b), m, co, 1, _device.ptr(want), None, _device.stream()))
    del cols, y
    xin = torch.nn.functional.pad(x, (0, 0, 1, 1, 1, 1, 1, 1)).contiguous() if in_pad else x
    out = torch.empty((n, od, oh, ow, co), dtype=torch.bfloat16, device="cuda")
    _lib.check(L.pmx_conv3d_bf16(_device.ptr(xin), dims, in_pad, 2, co, _device.ptr(w), _device.ptr(b), 1,
                                 _device.ptr(out), 0, _device.stream()))
    got, ref = out.reshape(m, co).float(), want.float()
    bad = (got - ref).abs() > 2 ** -7 * ref.abs() + 1e-6
    assert int(bad.sum()) == 0, f"{int(bad.sum())} of {m * co} outputs differ beyond one BF16 step"
    assert float((got == ref).float().mean()) > 0.99


@pytest.mark.parametrize("shape,bytes_max", [((2, 16, 32, 128), 1), ((3, 36, 64, 128), 1), ((1, 128, 128, 128), 1),
                                              ((2, 8, 96, 128), 255)])
@pytest.mark.parametrize("out_pad", [1, 0])
def test_fused_first_pair_equals_the_two_layer_calls(shape, bytes_max, out_pad):
    """pmx_conv3d_first_pair (uint8 image -> 8 -> 16 channels, both stride 2)
    against pmx_conv3d_first followed by pmx_conv3d_bf16 on its zero-bordered
    output: the same products in the same order, so bit for bit."""
    import ctypes as C

    import torch

    from pmmoto_amd import _device, _lib
    L = _lib.lib()
    g = torch.Generator(device="cpu").manual_seed(sum(shape) + bytes_max)
    x = torch.randint(0, bytes_max + 1, shape, generator=g, dtype=torch.uint8).cuda()
    w0 = (torch.randn((8, 27), generator=g) / 27 ** 0.5).to(torch.bfloat16).cuda()
    b0 = (0.1 * torch.randn(8, generator=g)).cuda()
    w1 = (torch.randn((16, 216), generator=g) / 216 ** 0.5).to(torch.bfloat16).cuda()
    b1 = (0.1 * torch.randn(16, generator=g)).cuda()
    n, d, h, w = shape
    print(f"[first_pair] shape={shape} bytes<={bytes_max} out_pad={out_pad}", flush=True)
    mid = torch.zeros((n, d // 2 + 2, h // 2 + 2, w // 2 + 2, 8), dtype=torch.bfloat16, device="cuda")
    _lib.check(L.pmx_conv3d_first(_device.ptr(x), (C.c_int32 * 5)(n, d, h, w, 1), 2, 8, _device.ptr(w0),
                                  _device.ptr(b0), _device.ptr(mid), 1, _device.stream()))
    p = out_pad
    want = torch.zeros((n, d // 4 + 2 * p, h // 4 + 2 * p, w // 4 + 2 * p, 16), dtype=torch.bfloat16, device="cuda")
    _lib.check(L.pmx_conv3d_bf16(_device.ptr(mid), (C.c_int32 * 5)(n, d // 2, h // 2, w // 2, 8), 1, 2, 16,
                                 _device.ptr(w1), _device.ptr(b1), 1, _device.ptr(want), p, _device.stream()))
    got = torch.zeros_like(want)
    _lib.check(L.pmx_conv3d_first_pair(_device.ptr(x), (C.c_int32 * 5)(n, d, h, w, 1), _device.ptr(w0),
                                       _device.ptr(b0), _device.ptr(w1), _device.ptr(b1), _device.ptr(got), p,
                                       _device.stream()))
    a, e = got.view(torch.int16).cpu().numpy(), want.view(torch.int16).cpu().numpy()
    assert np.array_equal(a, e), f"{int((a != e).sum())} of {a.size} values differ"
    assert float((want.float() > 0).float().mean()) > 0.2


def test_fused_first_pair_rejects_shapes_it_does_not_tile():
    import ctypes as C

    from pmmoto_amd import _lib
    L = _lib.lib()
    for shape in [(1, 16, 32, 64, 1), (1, 16, 48, 128, 1), (1, 18, 32, 128, 1), (1, 16, 32, 128, 2)]:
        rc = L.pmx_conv3d_first_pair(C.c_void_p(256), (C.c_int32 * 5)(*shape), C.c_void_p(256), C.c_void_p(256),
                                     C.c_void_p(256), C.c_void_p(256), C.c_void_p(256), 1, None)
        assert rc != 0, shape


def test_fused_inference_at_width_128_matches_training_forward():
    """The width-128 network (first two layers fused) against the training
    forward's patch-matrix path."""
    import torch

    from pmmoto_amd.surrogate.model import SurrogateNet
    net = SurrogateNet(size=128, seed=5)
    rng = np.random.default_rng(12)
    x = torch.from_numpy((rng.random((2, 128, 128, 128)) < 0.45).astype(np.uint8)).cuda()
    print("[infer] size=128 default channels", flush=True)
    want, _ = net.forward(x, keep=False)
    got = net.infer(x)[0]
    np.testing.assert_allclose(got.cpu().numpy(), want.cpu().numpy(), rtol=1e-2, atol=1e-2)


def test_fused_inference_matches_training_forward():
    import torch

    from pmmoto_amd.surrogate.model import SurrogateNet
    net = SurrogateNet(size=24, channels=(8, 16, 24), hidden=32, q=6, seed=3)
    rng = np.random.default_rng(11)
    x = torch.from_numpy((rng.random((5, 24, 24, 24)) < 0.45).astype(np.uint8)).cuda()
    print("[infer] size=24 channels=(8,16,24)", flush=True)
    want, _ = net.forward(x, keep=False)
    got = net.infer(x)[0]
    np.testing.assert_allclose(got.cpu().numpy(), want.cpu().numpy(), rtol=1e-2, atol=1e-2)
    # the cached zero-bordered buffers serve a second call, and a new batch size
    again = net.infer(x)[0]
    np.testing.assert_array_equal(again.cpu().numpy(), got.cpu().numpy())
    part = net.infer(x[:2].contiguous())[0]
    np.testing.assert_allclose(part.cpu().numpy(), want[:2].cpu().numpy(), rtol=1e-2, atol=1e-2)


def emulated(net_state, channels, x, target, q):
    """float64 forward/backward with the GPU's BF16 rounding points."""
    W = {k: v for k, v in net_state.items()}
    wb = {k: N.bf16(v) if k.endswith(".w") else v for k, v in W.items()}
    h = x.astype(np.float64)[..., None]
    acts, colss = [h], []
    for i in range(len(channels)):
        y = N.conv_fwd(h, wb[f"conv{i}.w"], W[f"conv{i}.b"], 2)
        h = N.bf16(np.maximum(y, 0.0))
        acts.append(h)
    pooled = N.bf16(h.reshape(h.shape[0], -1, h.shape[-1]).mean(axis=1))
    z1 = N.bf16(np.maximum(pooled @ wb["fc1.w"].T + W["fc1.b"], 0.0))
    pred = z1 @ wb["fc2.w"].T + W["fc2.b"]
    _, _, dpred = N.gnll(pred, target)
    g = {}
    dpb = N.bf16(dpred)
    g["fc2.w"] = dpb.T @ z1
    g["fc2.b"] = dpred.sum(axis=0)
    trunk = dpb.copy()
    trunk[:, q] = 0.0
    dz1 = trunk @ wb["fc2.w"]
    dy1 = dz1 * (z1 > 0)
    dy1b = N.bf16(dy1)
    g["fc1.w"] = dy1b.T @ pooled
    g["fc1.b"] = dy1.sum(axis=0)
    dpool = dy1b @ wb["fc1.w"]
    p = int(np.prod(h.shape[1:4]))
    dh = np.repeat(dpool[:, None, :] / p, p, axis=1).reshape(h.shape)
    for i in range(len(channels) - 1, -1, -1):
        dz = dh * (acts[i + 1] > 0)
        dzb = N.bf16(dz)
        cols = N.im2col(acts[i], 2)
        g[f"conv{i}.w"] = (dzb.reshape(-1, channels[i]).T @ cols).reshape(W[f"conv{i}.w"].shape)
        g[f"conv{i}.b"] = dz.reshape(-1, channels[i]).sum(axis=0)
        if i > 0:
            dh = N.col2im(dzb.reshape(-1, channels[i]) @ wb[f"conv{i}.w"].reshape(channels[i], -1), acts[i].shape, 2)
    return pred, g


def test_network_forward_and_gradients_match_oracle():
    import torch

    from pmmoto_amd.surrogate import SurrogateNet
    rng = np.random.default_rng(11)
    channels, q = (4, 6, 8), 5
    net = SurrogateNet(size=12, channels=channels, hidden=7, q=q, seed=3)
    for name, w, wb in net.params:  # non-zero biases exercise every term
        if name.endswith(".b"):
            w.copy_(torch.from_numpy(rng.normal(size=w.shape).astype(np.float32) * 0.1))
    x = (rng.random(size=(3, 12, 12, 12)) < 0.6).astype(np.uint8)
    target = rng.uniform(size=(3, q)).astype(np.float32)
    pred, cache = net.forward(torch.from_numpy(x).cuda())
    loss, dpred = net.loss_and_grad(pred, torch.from_numpy(target).cuda())
    grads = net.backward(cache, dpred)
    want_pred, want_g = emulated(net.state(), channels, x, target.astype(np.float64), q)
    np.testing.assert_allclose(pred.cpu().numpy(), want_pred, rtol=2e-3, atol=2e-3)
    for (name, _, _), gg in zip(net.params, grads):
        got = gg.cpu().numpy().astype(np.float64)
        want = want_g[name]
        scale = max(np.abs(want).max(), 1e-6)
        np.testing.assert_allclose(got, want, rtol=2e-2, atol=2e-2 * scale, err_msg=name)


def test_training_beats_the_mean_curve_and_fallback_is_exact():
    """A small network trained by the library's own kernels on exactly
    labelled packs predicts held-out curves better than the training-set
    mean curve; domains above the uncertainty threshold get the exact curve."""
    import torch

    from pmmoto_amd.surrogate import SurrogateNet
    from pmmoto_amd.surrogate.data import exact_curves, make_dataset
    from pmmoto_amd.surrogate.train import SurrogatePorosimetry, mae, train
    radii = (12.0, 10.0, 8.0, 6.0, 5.0, 4.0, 3.0, 2.0)
    x, y, _ = make_dataset(512, 5_000_000, size=32, radii=radii)
    xt, yt, bt = make_dataset(64, 6_000_000, size=32, radii=radii)
    net = SurrogateNet(size=32, channels=(8, 16, 32, 32), hidden=32, q=len(radii), seed=1)
    hist = train(net, x, y, epochs=20, batch=16, lr=1e-3)
    print("loss per epoch", [round(h, 3) for h in hist])
    assert hist[-1] < hist[0]
    curve, sigma = net.predict(xt)
    m_net = mae(curve.cpu().numpy(), yt)
    m_mean = mae(np.broadcast_to(y.mean(axis=0), yt.shape), yt)
    print(f"held-out MAE: network {m_net:.4f}, mean curve {m_mean:.4f}")
    assert m_net < m_mean
    tau = float(np.quantile(sigma.cpu().numpy(), 0.5))
    sp = SurrogatePorosimetry(net, tau, radii)
    got, sig, fall = sp.curves(xt, bt)
    assert fall.size > 0
    np.testing.assert_array_equal(got[fall], exact_curves([xt[i] for i in fall], [bt[i] for i in fall], radii))


def test_graphed_inference_equals_eager_on_every_replay():
    """predict() through captured graphs equals the eager launches bit for bit:
    two inputs, a partial last batch (another batch shape), eager work
    between replays, and the first input again after the shape changed."""
    import torch

    from pmmoto_amd.surrogate import SurrogateNet
    net = SurrogateNet(size=32, channels=(8, 16, 32, 32), hidden=32, q=8, seed=3)
    g = torch.Generator(device="cuda").manual_seed(7)
    xa = (torch.rand((40, 32, 32, 32), generator=g, device="cuda") < 0.45).to(torch.uint8)
    xb = (torch.rand((40, 32, 32, 32), generator=g, device="cuda") < 0.55).to(torch.uint8)
    want = {k: [t.clone() for t in net.predict(x, batch=16, graph=False)] for k, x in (("a", xa), ("b", xb))}
    for rep in range(3):
        for k, x in (("a", xa), ("b", xb)):
            print(f"[graphed inference] replay {rep} input {k}", flush=True)
            curve, sigma = net.predict(x, batch=16, graph=True)
            assert torch.equal(curve, want[k][0]) and torch.equal(sigma, want[k][1]), (rep, k)
            torch.zeros(1 << 20, device="cuda").add_(1.0)  # eager work between replays
        torch.cuda.synchronize()


def test_graphed_inference_with_fused_first_pair_equals_eager():
    """The width-128 network, whose first two layers run as one fused kernel,
    through captured graphs: every replay (two batch shapes, eager work in
    between) equals the eager launches bit for bit."""
    import torch

    from pmmoto_amd.surrogate import SurrogateNet
    net = SurrogateNet(size=128, channels=(8, 16, 24), hidden=16, q=4, seed=9)
    g = torch.Generator(device="cuda").manual_seed(8)
    x = (torch.rand((6, 128, 128, 128), generator=g, device="cuda") < 0.5).to(torch.uint8)
    want = [t.clone() for t in net.predict(x, batch=4, graph=False)]
    for rep in range(3):
        print(f"[graphed inference, fused first pair] replay {rep}", flush=True)
        curve, sigma = net.predict(x, batch=4, graph=True)
        assert torch.equal(curve, want[0]) and torch.equal(sigma, want[1]), rep
        torch.zeros(1 << 20, device="cuda").add_(1.0)  # eager work between replays
    torch.cuda.synchronize()
