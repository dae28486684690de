"""Interchange with the reference's files: the xyzr sphere-pack reader and
the VTK image writer (host code; images may be device tensors)."""

from . import data_read, output

__all__ = ["data_read", "output"]
