from . import connected_components, distance, morphological_operators, porosimetry

__all__ = ["connected_components", "distance", "morphological_operators", "porosimetry"]
