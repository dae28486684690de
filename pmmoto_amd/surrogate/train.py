"""Training, evaluation and fallback inference of the surrogate.

    python -m pmmoto_amd.surrogate.train --train 20000 --epochs 30 --out weights.npz

trains on samples 5_000_000 + i and reports the held-out error on samples
6_000_000 + i (pmmoto_amd.surrogate.data).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from .calibrate import POROSITY_RANGE, tau_for
from .data import RADII, exact_curves, make_dataset
from .model import SurrogateNet

TRAIN_SEED0, TEST_SEED0 = 5_000_000, 6_000_000


def train(net: SurrogateNet, images, curves, epochs=10, batch=32, lr=2e-3, seed=0, log=None):
    """Adam with a cosine learning-rate decay; returns the mean loss per epoch."""
    n = images.shape[0]
    target = torch.from_numpy(np.ascontiguousarray(curves, dtype=np.float32)).cuda()
    rng = np.random.default_rng(seed)
    steps = epochs * ((n + batch - 1) // batch)
    hist, step = [], 0
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(n)).cuda()
        tot = 0.0
        for i in range(0, n, batch):
            idx = order[i:i + batch]
            cur_lr = lr * 0.5 * (1.0 + np.cos(np.pi * step / max(1, steps)))
            loss = net.train_step(images[idx].contiguous(), target[idx].contiguous(), lr=float(cur_lr))
            tot += float(loss.sum().item())
            step += 1
        hist.append(tot / n)
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4f}")
    return hist


def mae(pred: np.ndarray, want: np.ndarray) -> float:
    return float(np.mean(np.abs(pred - want)))


class SurrogatePorosimetry:
    """Saturation curves from the network; domains whose predicted sigma
    exceeds tau, or whose porosity lies outside the training range, go to
    the exact sweep instead (bcs gives their faces).  ``calibrated`` takes
    tau from the calibration of a target error (pmmoto_amd.surrogate.calibrate)."""

    def __init__(self, net: SurrogateNet, tau: float, radii=RADII, porosity_range=POROSITY_RANGE):
        if len(radii) != net.q:
            raise ValueError("one radius per curve point")
        self.net, self.tau, self.radii = net, tau, tuple(radii)
        self.porosity_range = porosity_range

    @classmethod
    def calibrated(cls, net: SurrogateNet, target: float, radii=RADII):
        return cls(net, tau_for(target), radii)

    def curves(self, images, bcs, batch=512):
        curve, sigma = self.net.predict(images, batch=batch)
        curve = curve.cpu().numpy()
        sigma = sigma.cpu().numpy()
        out = np.zeros(len(sigma), bool)
        if self.porosity_range is not None:
            lo, hi = self.porosity_range
            por = np.array([float(images[i].float().mean()) for i in range(len(sigma))])
            out = (por < lo) | (por > hi)
        fall = np.nonzero((sigma > self.tau) | out)[0]
        if fall.size:
            exact = exact_curves([images[i] for i in fall], [bcs[i] for i in fall], self.radii)
            curve[fall] = exact
        return curve, sigma, fall


def save(net: SurrogateNet, path: str, meta: dict):
    arrays = {p[0]: p[1].detach().cpu().numpy() for p in net.params}
    np.savez(path, meta=json.dumps(meta), **arrays)


def load(path: str) -> SurrogateNet:
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    net = SurrogateNet(size=meta["size"], channels=tuple(meta["channels"]), hidden=meta["hidden"], q=meta["q"])
    from .model import to_bf16
    for p in net.params:
        p[1].copy_(torch.from_numpy(z[p[0]]).cuda())
        if p[2] is not None:
            to_bf16(p[1], p[2])
    return net


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--out", default="surrogate_weights.npz")
    a = ap.parse_args(argv)
    log = lambda s: print(f"[train] {s}", file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    x, y, _ = make_dataset(a.train, TRAIN_SEED0, a.size)
    xt, yt, bt = make_dataset(a.test, TEST_SEED0, a.size)
    log(f"data: {a.train} + {a.test} samples in {time.perf_counter() - t0:.1f} s")
    net = SurrogateNet(size=a.size, q=len(RADII))
    log(f"network: {net.n_params()} parameters, {net.flops_per_sample() / 1e9:.2f} GFLOP per sample forward")
    t0 = time.perf_counter()
    hist = train(net, x, y, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    curve, sigma = net.predict(xt)
    m_net = mae(curve.cpu().numpy(), yt)
    m_mean = mae(np.broadcast_to(y.mean(axis=0), yt.shape), yt)
    meta = {"size": a.size, "channels": list(net.channels), "hidden": net.hidden, "q": net.q, "train": a.train,
            "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "train_s": round(train_s, 1), "loss": hist,
            "test_mae": m_net, "mean_curve_mae": m_mean,
            "sigma_quantiles": [float(np.quantile(sigma.cpu().numpy(), q)) for q in (0.5, 0.9, 0.99)]}
    save(net, a.out, meta)
    print(json.dumps(meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
