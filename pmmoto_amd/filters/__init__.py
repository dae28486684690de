from . import connected_components, distance, equilibrium_distribution, morphological_operators, porosimetry

__all__ = ["connected_components", "distance", "equilibrium_distribution", "morphological_operators", "porosimetry"]
