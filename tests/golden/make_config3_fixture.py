"""Writes tests/golden/config3_psd_<n>.json: the oracle's answer for the
1024^3 drainage-radius sweep, as digests small enough to commit.

Workload: the seeded polydisperse sphere pack (lognormal radii, mean 10
voxels, sigma 0.3, porosity 0.35, seed 3; the same draw sequence as
pmmoto_amd.domain_generation.sphere_pack, rasterised by the oracle), padded
as PMMoTo pads it for x END faces with an inlet at x = 0 and WALL faces in
y and z (shape n x (n+2) x (n+2), zero pad layers); pore_size_distribution
with radii 32..1 as numpy float64, inlet=True, mode "distance".

Recorded: SHA-256 of the unpadded image, SHA-256 of the uint8 radius-code
map over the padded array (code i+1 where radii[i] first invades, 0 where
none does), the code histogram, and the float64 wetting-saturation curve
(exact, as float.hex).  The chain is oracle.psd_codes_exact (pinned against
the oracle's reference-driver restatement in tests/test_oracle_morph_ccl.py).

Run in the development container (CPU only, about 20 minutes and 25 GB of
memory at n = 1024):  python tests/golden/make_config3_fixture.py [--size N]
"""

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.helpers import cpu_sphere_pack  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    n = args.size
    t0 = time.time()
    img, spheres = cpu_sphere_pack((n, n, n), 0.35, "lognormal", 3, mean=10.0, sigma=0.3)
    print(f"pack {n}^3: {spheres.shape[0]} spheres, {time.time() - t0:.1f} s", flush=True)
    pm = np.zeros((n, n + 2, n + 2), np.uint8)
    pm[:, 1:-1, 1:-1] = img
    box = O.Box(bc=((O.BC_END, O.BC_END), (O.BC_WALL, O.BC_WALL), (O.BC_WALL, O.BC_WALL)),
                pad=((0, 0), (1, 1), (1, 1)), inlet=((True, False), (False, False), (False, False)),
                outlet=((False, True), (False, False), (False, False)))
    radii = [np.float64(r) for r in range(32, 0, -1)]
    t0 = time.time()
    codes = O.psd_codes_exact(box, pm, radii, inlet=True, threads=args.threads)
    print(f"sweep: {time.time() - t0:.1f} s", flush=True)
    sw = O.saturation_from_codes(box, pm, codes, len(radii), (n, n, n))
    out = {
        "workload": f"{n}^3 lognormal sphere pack (mean 10, sigma 0.3, porosity 0.35, seed 3), padded "
                    f"{n}x{n + 2}x{n + 2} (x END with inlet x=0, y/z WALL); radii 32..1 float64, "
                    "inlet=True, mode distance",
        "shape": [n, n + 2, n + 2],
        "spheres": int(spheres.shape[0]),
        "image_sha256": hashlib.sha256(img.tobytes()).hexdigest(),
        "code_sha256": hashlib.sha256(codes.tobytes()).hexdigest(),
        "code_counts": np.bincount(codes.ravel(), minlength=33).tolist(),
        "saturation_hex": [float(v).hex() for v in sw],
        "saturation": [float(v) for v in sw],
    }
    path = os.path.join(HERE, f"config3_psd_{n}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}", flush=True)


if __name__ == "__main__":
    main()
