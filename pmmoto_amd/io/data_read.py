"""read_sphere_pack_xyzr_domain (src/pmmoto/io/data_read.py:26-85).

File format: three lines ``lo hi`` (the domain box per axis), then one line
``x y z r`` per sphere.
"""

from __future__ import annotations

import logging
import os
import sys

import numpy as np

__all__ = ["read_sphere_pack_xyzr_domain"]

logger = logging.getLogger(__name__)


def check_file(file_name: str) -> None:
    """io_utils.check_file (io_utils.py:11-18): a missing file logs an error
    and exits with status 1, as the reference does."""
    if not os.path.isfile(file_name):
        logger.error("%s does not exist!", file_name)
        sys.exit(1)


def read_sphere_pack_xyzr_domain(input_file: str):
    """Returns (spheres float64 (N, 4), domain box ((x0, x1), (y0, y1),
    (z0, z1))).  A sphere line without exactly four numbers raises
    ValueError (data_read.py:63-74)."""
    check_file(input_file)
    with open(input_file, "r", encoding="utf-8") as f:
        lines = f.readlines()
    if len(lines) < 3:
        raise ValueError(f"{input_file}: expected three domain lines before the spheres")
    box = np.zeros((3, 2), dtype=np.float64)
    for a in range(3):
        parts = lines[a].split()
        if len(parts) < 2:
            raise ValueError(f"Invalid domain line (expected 2 values): {lines[a].strip()}")
        box[a] = [float(parts[0]), float(parts[1])]
    spheres = np.zeros((len(lines) - 3, 4), dtype=np.float64)
    for i, line in enumerate(lines[3:]):
        parts = line.split()
        if len(parts) != 4:
            raise ValueError(f"Invalid sphere line (expected 4 values): {line.strip()}")
        try:
            spheres[i] = [float(p) for p in parts]
        except ValueError as exc:
            raise ValueError(f"Non-numeric value in sphere line: {line.strip()}") from exc
    return spheres, tuple(map(tuple, box))
