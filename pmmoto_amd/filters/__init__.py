from . import distance

__all__ = ["distance"]
