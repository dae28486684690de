from .euclidean_distance_transform import edt, edt2d, edt3d, edt_squared_int

__all__ = ["edt", "edt2d", "edt3d", "edt_squared_int"]
