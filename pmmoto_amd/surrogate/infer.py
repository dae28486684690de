"""Inference throughput of the trained surrogate.

    python -m pmmoto_amd.surrogate.infer --n 512 --reps 5 [--calibrate-copy]

Times SurrogateNet.infer on n seeded 128^3 uint8 packs (pore fraction 0.45,
independent voxels: the network's cost does not depend on the pattern)
already resident on the GPU, and prints one JSON line.  --calibrate-copy
adds the 1 GiB device copy that pmmoto_amd.profiling scales the HBM
counters by.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import torch

from .train import load

WEIGHTS = os.path.join(os.path.dirname(__file__), "weights_128.npz")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--calibrate-copy", action="store_true")
    a = ap.parse_args(argv)
    net = load(a.weights)
    g = torch.Generator(device="cuda").manual_seed(1234)
    s = net.size
    x = (torch.rand((a.n, s, s, s), generator=g, device="cuda") < 0.45).to(torch.uint8)
    if a.calibrate_copy:
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        del src, dst
    net.infer(x)  # warm-up (allocates the padded activation buffers)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        net.infer(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.reps
    print(json.dumps({"n": a.n, "ms_per_batch": round(dt * 1e3, 3), "curves_per_s": round(a.n / dt, 1),
                      "tflops": round(a.n * net.flops_per_sample() / dt / 1e12, 2)}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
