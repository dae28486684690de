"""Boundary types of a domain face (src/pmmoto/core/boundary_types.py:11-17)."""

from enum import Enum

from .. import _lib


class BoundaryType(str, Enum):
    """Allowable boundary types."""

    END = "end"
    WALL = "wall"
    PERIODIC = "periodic"
    INTERNAL = "internal"


def to_code(bt: BoundaryType) -> int:
    """libpmx face code (PMX_BC_*)."""
    bt = BoundaryType(bt)
    if bt == BoundaryType.WALL:
        return _lib.BC_WALL
    if bt == BoundaryType.PERIODIC:
        return _lib.BC_PERIODIC
    if bt == BoundaryType.END:
        return _lib.BC_END
    raise ValueError("INTERNAL faces only occur between ranks; one GPU holds the whole domain")
