from .domain_generation import (gen_img_random_binary, gen_mp_constant, gen_pm_inkbottle, gen_pm_spheres_domain,
                                pack_radii, sphere_pack)
from .multiphase import Multiphase
from .porousmedia import PorousMedia, gen_pm

__all__ = ["Multiphase", "PorousMedia", "gen_img_random_binary", "gen_mp_constant", "gen_pm",
           "gen_pm_inkbottle", "gen_pm_spheres_domain", "pack_radii", "sphere_pack"]
