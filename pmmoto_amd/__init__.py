"""pmmoto_amd: PMMoTo's voxel-morphology filters (distance, morphology,
connected components, porosimetry) as hand-written HIP kernels for the
AMD Instinct MI355X, behind the pmmoto.filters call surface."""

from . import core, filters, io
from .core import BoundaryType, initialize

__all__ = ["BoundaryType", "core", "filters", "initialize", "io"]
