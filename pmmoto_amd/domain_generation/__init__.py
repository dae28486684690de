from .domain_generation import gen_img_random_binary, pack_radii, sphere_pack
from .porousmedia import PorousMedia, gen_pm

__all__ = ["PorousMedia", "gen_img_random_binary", "gen_pm", "pack_radii", "sphere_pack"]
