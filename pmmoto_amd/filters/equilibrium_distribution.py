"""Morphological drainage on the GPU.

Same call surface as src/pmmoto/filters/equilibrium_distribution.py:25-289.
For each capillary pressure (ascending) the probe radius r = 2 gamma cos(theta)
/ pc gives a porosimetry image restricted to inlet-connected non-wetting
phase; wetting phase (2) connected to the outlet is displaced by it, the
reservoir is reset to phase 1 and the wetting saturation recorded.  The
multiphase image stays on the GPU for the whole sweep.
"""

from __future__ import annotations

import logging

import numpy as np
import torch

from .. import _device, _lib
from ..domain_generation.multiphase import update_reservoir_device
from . import connected_components as cc
from . import porosimetry as po

__all__ = ["drainage"]


def _radii(method, mp, pc, gamma, contact_angle):
    if method == "extended_contact_angle":
        # equilibrium_distribution.py:232-289
        er = mp.get_probe_radius(pc, gamma, contact_angle)
        dr = mp.get_probe_radius(pc, gamma, contact_angle) * np.abs(np.cos(np.deg2rad(contact_angle)))
        return [er, dr]
    # _standard_method :133-179 / _contact_angle_method :182-229
    return mp.get_probe_radius(pc, gamma, contact_angle if method == "contact_angle" else 0)


def _morph(method, mp, pc, gamma, contact_angle, mode, pm_t, mp_t):
    sd = mp.subdomain
    radius = _radii(method, mp, pc, gamma, contact_angle)
    rlist = radius if isinstance(radius, list) else [radius]
    if min(rlist) < min(sd.domain.resolution) or max(rlist) > mp.porous_media.max_distance:
        return torch.zeros_like(pm_t)
    er, dr = po._split_radius(radius)
    po.morph.check_radii(sd, [er, dr])
    mp_use = mp_t if int(_count(mp_t, 1)) > 0 else None  # utils.phase_exists(multiphase.img, 1)
    return po.porosimetry_device(sd, pm_t, er, dr, inlet=True, mp_t=mp_use, mode=mode)


def _count(t, value):
    """Voxels equal to value over the whole array (utils.phase_exists counts all)."""
    import ctypes as C
    cnt = torch.empty(1, dtype=torch.int64, device=t.device)
    shape = tuple(t.shape)
    _lib.check(_lib.lib().pmx_box_op(_device.ptr(t), (C.c_int64 * 3)(*shape), (C.c_int64 * 3)(0, 0, 0),
                                     (C.c_int64 * 3)(*shape), 1, int(value), _device.ptr(cnt), _device.stream()))
    return int(cnt.item())


def drainage(multiphase, capillary_pressures, gamma=1, contact_angle=0, method="standard", mode="hybrid",
             save=False, out_folder="drainage_out"):
    """equilibrium_distribution.py:25-130: wetting saturation per pressure;
    multiphase.img holds the final distribution."""
    if isinstance(capillary_pressures, (int, float)):
        capillary_pressures = [capillary_pressures]
    elif isinstance(capillary_pressures, list):
        s = sorted(capillary_pressures)
        if capillary_pressures != s:
            logging.warning("The capillary pressure must be monotonically increasing. Sorting.")
            capillary_pressures = s
    if method == "standard":
        if contact_angle != 0:
            raise ValueError("The standard approach requires a zero contact angle!")
    elif method in ("contact_angle", "extended_contact_angle"):
        if contact_angle == 0:
            logging.warning("The contact angle is zero. This will yield same results as the standard approach.")
    else:
        raise ValueError(f"{method} is not implemented. ")

    sd = multiphase.subdomain
    pm_t, _ = _device.as_device_u8(multiphase.porous_media.img)
    mp_t, as_np = _device.as_device_u8(multiphase.img)
    mp_t = mp_t.clone()
    w_sat = np.zeros(len(capillary_pressures))
    for n, pc in enumerate(capillary_pressures):
        morph = _morph(method, multiphase, pc, gamma, contact_angle, mode, pm_t, mp_t)
        # outlet_connected_img(subdomain, multiphase.img, 2)
        labels, count = cc.label_device(mp_t, sd)
        nl = int(count.item())
        fout = cc._face_flags(sd, labels, nl, "outlet")
        wet = cc.keep_device(mp_t, labels, fout, None, 0, 2)
        _lib.check(_lib.lib().pmx_drain_update(_device.ptr(morph), _device.ptr(wet), mp_t.numel(),
                                               _device.ptr(mp_t), _device.stream()))
        update_reservoir_device(mp_t, sd, 1)
        multiphase.img = mp_t
        if save:
            # equilibrium_distribution.py:109-118: one image per pressure, the
            # porosimetry image alongside
            from ..io.output import save_img
            file_out = str(out_folder) + f"/capillary_pressure_{pc:.3f}".replace(".", "_")
            save_img(file_out, sd, mp_t, additional_img={"morph": morph})
        w_sat[n] = multiphase.get_saturation(2, mp_t)
        logging.info("Wetting phase saturation at capillary pressure of %f: %f", pc, w_sat[n])
    multiphase.img = mp_t.cpu().numpy() if as_np else mp_t
    return w_sat
