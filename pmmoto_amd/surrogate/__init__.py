"""Learned surrogate of the porosimetry sweep (a 3-D CNN on the library's
BF16 matrix-core kernels), its training data and its exact-path fallback."""

from .model import SurrogateNet  # noqa: F401
