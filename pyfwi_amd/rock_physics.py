"""Rock-physics relations used by the PCS gradient parameterisations.

Subset of PyFWI's ``PyFWI.rock_physics`` (src/PyFWI/rock_physics.py) that
``grad_switcher``'s porosity / clay / saturation (PCS) chain rules depend on.
NumPy, elementwise.
"""
import numpy as np

__all__ = ["weighted_average", "vrh", "drained_moduli", "delta_biot_gassmann",
           "biot_gassmann", "Han", "pcs2dv_gassmann", "pcs2dv_vrh", "pcs2dv_han"]


def weighted_average(prop1, prop2, volume1):
    """prop1 * v1 + prop2 * (1 - v1) (rock_physics.py:365-387)."""
    return prop1 * volume1 + prop2 * (1 - volume1)


def vrh(prop1, prop2, volume1, method):
    """Voigt, Reuss or Voigt-Reuss-Hill average of two constituents
    (rock_physics.py:390-424); method in 'Voigt', 'Reuss', 'VRH'/'vrh'."""
    volume2 = 1 - volume1
    voigt = volume1 * prop1 + volume2 * prop2
    reuss = 1 / (volume1 / prop1 + volume2 / prop2)
    if method == "Voigt":
        return voigt
    if method == "Reuss":
        return reuss
    if method in ("VRH", "vrh"):
        return 0.5 * (voigt + reuss)
    return None


def _fraction(phi):
    """Porosity given in percent (any value >= 1) is converted to a fraction,
    as the reference does (rock_physics.py:335-336, 360-361)."""
    phi = np.asarray(phi)
    if (phi >= 1).any():
        return phi / 100
    return phi


def drained_moduli(phi, k_s, g_s, cs):
    """Drained bulk and shear moduli with consolidation parameter cs
    (rock_physics.py:303-339)."""
    phi = _fraction(phi)
    k_d = k_s * ((1 - phi) / (1 + cs * phi))
    g_d = g_s * ((1 - phi) / (1 + 1.5 * cs * phi))
    return k_d, g_d


def delta_biot_gassmann(phi, k_f, k_s, k_d):
    """Gassmann's Delta (rock_physics.py:359-362)."""
    phi = _fraction(phi)
    return ((1 - phi) / phi) * (k_f / k_s) * (1 - k_d / (k_s - k_s * phi))


def biot_gassmann(phi, k_f, k_s, k_d):
    """Undrained bulk modulus and Biot-Gassmann C, M (rock_physics.py:348-356)."""
    delta = delta_biot_gassmann(phi, k_f, k_s, k_d)
    denom = phi * (1 + delta)
    k_u = (phi * k_d + (1 - (1 + phi) * (k_d / k_s)) * k_f) / denom
    return k_u, k_f * (1 - k_d / k_s) / denom, k_f / denom


def Han(phi, cc, a1=5.5, a2=6.9, a3=2.2, b1=3.4, b2=4.7, b3=1.8):
    """Han's linear regressions vp = a1 - a2 phi - a3 C, vs = b1 - b2 phi - b3 C
    (km/s), float32 (rock_physics.py:235-264)."""
    vp = np.asarray(a1 - a2 * phi - a3 * cc).astype(np.float32)
    vs = np.asarray(b1 - b2 * phi - b3 * cc).astype(np.float32)
    return vp, vs


def _unpack(rock):
    keys = ("k_q", "k_c", "k_w", "k_h", "mu_q", "mu_c", "rho_q", "rho_c", "rho_w", "rho_h")
    return [rock[k] for k in keys]


def pcs2dv_gassmann(pcs_model, rock_properties, method):
    """(phi, cc, sw) -> (vp, vs, rho) through Gassmann's fluid substitution with
    an averaged skeleton (rock_physics.py:555-598).  As in the reference,
    vp = sqrt((k_u + 0.75 mu_d) / rho)."""
    phi, cc, sw = pcs_model["phi"], pcs_model["cc"], pcs_model["sw"]
    k_q, k_c, k_w, k_h, mu_q, mu_c, rho_q, rho_c, rho_w, rho_h = _unpack(rock_properties)
    cs = rock_properties["cs"]
    rho_f = weighted_average(rho_w, rho_h, sw)
    k_f = weighted_average(k_w, k_h, sw)
    k_s = vrh(k_c, k_q, cc, method)
    mu_s = vrh(mu_c, mu_q, cc, method)
    rho_s = weighted_average(rho_c, rho_q, cc)
    k_d, mu_d = drained_moduli(phi, k_s, mu_s, cs)
    k_u, _, _ = biot_gassmann(phi, k_f, k_s, k_d)
    rho = weighted_average(rho_f, rho_s, phi)
    vp = np.sqrt((k_u + 0.75 * mu_d) / rho)
    vs = np.sqrt(mu_d / rho)
    return {"vp": vp.astype(np.float32), "vs": vs.astype(np.float32),
            "rho": np.asarray(rho).astype(np.float32)}


def pcs2dv_vrh(pcs_model, rock_properties):
    """(phi, cc, sw) -> (vp, vs, rho) with VRH-averaged moduli of all four
    constituents; the Reuss shear bound is zero (rock_physics.py:506-552)."""
    phi, cc, sw = pcs_model["phi"], pcs_model["cc"], pcs_model["sw"]
    k_q, k_c, k_w, k_h, mu_q, mu_c, rho_q, rho_c, rho_w, rho_h = _unpack(rock_properties)
    kv = (1 - phi) * (k_c * cc + k_q * (1 - cc)) + phi * (k_w * sw + k_h * (1 - sw))
    kr = 1 / ((1 - phi) * (cc / k_c + (1 - cc) / k_q) + phi * (sw / k_w + (1 - sw) / k_h))
    k = 0.5 * (kv + kr)
    mu = 0.5 * ((1 - phi) * (mu_c * cc + mu_q * (1 - cc)))
    rho = weighted_average(weighted_average(rho_w, rho_h, sw),
                           weighted_average(rho_c, rho_q, cc), phi)
    vp = np.sqrt((k + 0.75 * mu) / rho)
    vs = np.sqrt(mu / rho)
    return {"vp": vp.astype(np.float32), "vs": vs.astype(np.float32),
            "rho": np.asarray(rho).astype(np.float32)}


def pcs2dv_han(pcs_model, rock_properties):
    """(phi, cc, sw) -> (vp, vs, rho) with Han's regressions
    (rock_physics.py:601-622)."""
    phi, cc, sw = pcs_model["phi"], pcs_model["cc"], pcs_model["sw"]
    rho_f = weighted_average(rock_properties["rho_w"], rock_properties["rho_h"], sw)
    rho_s = weighted_average(rock_properties["rho_c"], rock_properties["rho_q"], cc)
    rho = weighted_average(rho_f, rho_s, phi)
    vp, vs = Han(phi, cc)
    return {"vp": vp.astype(np.float32), "vs": vs.astype(np.float32),
            "rho": np.asarray(rho).astype(np.float32)}
