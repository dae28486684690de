from .boundary_types import BoundaryType
from .subdomain import Domain, PaddedSubdomain, initialize

__all__ = ["BoundaryType", "Domain", "PaddedSubdomain", "initialize"]
