"""Chain rules between gradient parameterisations.

Mirror of PyFWI's ``PyFWI.grad_switcher`` (src/PyFWI/grad_switcher.py).  The
functions are elementwise and accept NumPy arrays or torch tensors (CPU or
GPU); they use only arithmetic operators so a device tensor stays on its
device.
"""
import numpy as np

__all__ = ["grad_lmd_to_vd", "grad_vd_to_lmd", "grad_dv2kmd"]


def _sqrt(x):
    return x ** 0.5


def _as_f32(x):
    if isinstance(x, np.ndarray):
        return x.astype(np.float32)
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.to(torch.float32)
    except ImportError:  # pragma: no cover - torch is a hard dependency elsewhere
        pass
    return np.float32(x)


def grad_lmd_to_vd(glam, gmu, grho, lam, mu, rho):
    """Gradient w.r.t. (lambda, mu, rho) -> (vp, vs, rho), with
    ``vp = sqrt((lam+2mu)/rho)``, ``vs = sqrt(mu/rho)``
    (src/PyFWI/grad_switcher.py:331-383):

        gvp  = 2 rho vp glam
        gvs  = -4 rho vs glam + 2 rho vs gmu
        grho = (vp^2 - 2 vs^2) glam + vs^2 gmu + grho
    """
    vp = _sqrt((lam + 2 * mu) / rho)
    vs = _sqrt(mu / rho)
    vs2 = vs ** 2
    vpvs = vp ** 2 - 2 * vs ** 2
    gvp = glam * 2 * vp * rho + gmu * 0 + grho * 0
    gvs = glam * (-4 * rho * vs) + gmu * 2 * vs * rho + grho * 0
    grho_out = glam * vpvs + gmu * vs2 + grho
    return _as_f32(gvp), _as_f32(gvs), _as_f32(grho_out)


def grad_vd_to_lmd(gvp, gvs, grho, vp, vs, rho):
    """Gradient w.r.t. (vp, vs, rho) -> (lambda, mu, rho)
    (src/PyFWI/grad_switcher.py:403-448), the formulas as published there:

        glam = gvp / (2 rho vp)
        gmu  = gvp / (rho vp) + gvs / (2 rho vs)
        grho = -gvp vp / (2 rho) - gvs vs / (2 rho) + grho
    """
    glam = gvp / (2 * rho * vp) + gvs * 0 + grho * 0
    gmu = gvp / (rho * vp) + gvs / (2 * rho * vs) + grho * 0
    grho_out = gvp * (-vp / 2 / rho) + gvs * (-vs / 2 / rho) + grho
    return _as_f32(glam), _as_f32(gmu), _as_f32(grho_out)


def grad_dv2kmd(gdv, vp, vs, rho):
    """Gradient w.r.t. (vp, vs, rho) -> (K, mu, rho)
    (src/PyFWI/grad_switcher.py:246-266), as published there:

        gk   = gvp / (2 rho vp)
        gmu  = 2 gvp / (3 rho vp) + gvs / (2 rho vs)
        grho = -gvp vp / (2 rho) - gvs vs / (2 rho) + grho
    """
    gvp, gvs, grho = gdv["vp"], gdv["vs"], gdv["rho"]
    gk = gvp * (1 / (2 * rho * vp)) + gvs * 0 + grho * 0
    gmu = gvp * (2 / (3 * rho * vp)) + gvs * (1 / (2 * rho * vs)) + grho * 0
    grho_kmd = gvp * (-vp / (2 * rho)) + gvs * (-vs / (2 * rho)) + grho * 1
    return gk, gmu, grho_kmd


# ----------------------------------------------------- PCS parameterisations
# Gradients w.r.t. porosity (phi), clay content (cc) and water saturation (sw)
# from a (vp, vs, rho) gradient, through PyFWI's rock-physics models
# (src/PyFWI/grad_switcher.py:4-243, 269-328, 386-400).  The published
# formulas are reproduced as they are, including where they are not the exact
# derivative of the matching forward model in rock_physics (e.g. vp there uses
# k_u + 0.75 mu while grad_dv2kmd assumes k + 4/3 mu).

from . import rock_physics as rp  # noqa: E402

__all__ += ["PcsParameterization", "grad_dv2pcs_gassmann", "grad_dv2pcs_vrh",
            "grad_dv2pcs_han", "grad_vrh", "grad_kmd2phi_gassmann", "grad_kmd2cc_gassmann",
            "grad_kmd2sw_gassmann", "grad_kmd2phi_vrh", "grad_kmd2cc_vrh", "grad_kmd2sw_vrh",
            "grad_dv2phi_han", "grad_dv2cc_han", "grad_dv2sw_han"]


class PcsParameterization:
    """Selects the rock-physics model of the PCS chain rule
    (src/PyFWI/grad_switcher.py:4-17)."""

    def __init__(self, rp_model="gassmann", grad_args={}, model_args={}):
        self.grad_args = grad_args
        self.model_args = model_args
        self.rp_model = rp_model

    def grad_dv2pcs(self, *args, **kargs):
        return grad_dv2pcs_gassmann(*args, **kargs, **self.grad_args)

    def pcs2dv(self, *args, **kargs):
        return rp.pcs2dv_gassmann(*args, **kargs, **self.model_args)


def _combine(g3, d3):
    """sum_i g_i * d_i for the three parameter directions."""
    return g3[0] * d3[0] + g3[1] * d3[1] + g3[2] * d3[2]


def grad_vrh(prop1, prop2, volume1, method):
    """d/d(volume1) of the Voigt, Reuss or VRH average
    (src/PyFWI/grad_switcher.py:386-400)."""
    volume2 = 1 - volume1
    gv = prop1 - prop2
    gr = (1 / prop2 - 1 / prop1) / (volume1 / prop1 + volume2 / prop2) ** 2
    if method == "Voigt":
        return gv
    if method == "Reuss":
        return gr
    if method == "VRH":
        return 0.5 * (gv + gr)
    return None


def grad_kmd2phi_gassmann(phi, ks, mus, rhos, kf, rhof, cs):
    gkd = -ks * ((1 + cs) / (1 + cs * phi) ** 2)
    gmud = -mus * ((1 + cs * 3 / 2) / (1 + cs * phi * 3 / 2) ** 2)
    gdelta = (-(kf / (ks * phi ** 2)) * (1 - (1 / (1 + cs * phi)))
              + ((kf * (1 - phi)) / (ks * phi)) * (cs / (1 + cs * phi) ** 2))
    kd, _ = rp.drained_moduli(phi, ks, mus, cs)
    delta = rp.delta_biot_gassmann(phi, kf, ks, kd)
    kfks = kf / ks
    pp = 1 + phi
    pdl = phi * (1 + delta)
    gk = ((kd + phi * gkd - kfks * (kd + pp * gkd)) / pdl
          - (phi * kd + kf - kfks * pp * kd) * (1 + delta + phi * gdelta) / (pdl ** 2))
    return gk, gmud, rhof - rhos


def grad_kmd2cc_gassmann(phi, cc, kc, muc, rhoc, kq, muq, rhoq, kf, cs, method):
    ks = rp.vrh(kc, kq, cc, method)
    mus = rp.vrh(muc, muq, cc, method)
    gks = grad_vrh(kc, kq, cc, method)
    gmus = grad_vrh(muc, muq, cc, method)
    gkd = gks * (1 - phi) / (1 + cs * phi)
    gmud = gmus * (1 - phi) / (1 + cs * phi * 3 / 2)
    kfks2 = kf / ks ** 2
    gdelta = ((phi - 1) * kfks2 / phi) * (1 - 1 / (1 + cs * phi)) * gks
    kd, _ = rp.drained_moduli(phi, ks, mus, cs)
    delta = rp.delta_biot_gassmann(phi, kf, ks, kd)
    kfks = kf / ks
    pp = 1 + phi
    pdl = phi * (1 + delta)
    gk = ((phi * gkd + kfks2 * pp * (gks * kd - gkd * ks)) / pdl
          - (phi * kd + kf - kfks * pp * kd) * (phi * gdelta) / (pdl ** 2))
    return gk, gmud, (1 - phi) * (rhoc - rhoq)


def grad_kmd2sw_gassmann(phi, cc, kc, kw, kh, muc, rhow, kq, muq, rhoh, kf, cs):
    # the skeleton here is the Voigt average, as published
    ks = rp.weighted_average(kc, kq, cc)
    mus = rp.weighted_average(muc, muq, cc)
    gkf = kw - kh
    kd, _ = rp.drained_moduli(phi, ks, mus, cs)
    delta = rp.delta_biot_gassmann(phi, kf, ks, kd)
    gdelta = gkf * delta / kf
    pp = 1 + phi
    pdl = phi * (1 + delta)
    num = 1 - (pp * kd / ks)
    gk = num * gkf / pdl - (phi * kd + kf * num) * phi * gdelta / (pdl ** 2)
    return gk, 0, phi * (rhow - rhoh)


def grad_dv2pcs_gassmann(gdv, rock_properties, pcs_model, method):
    """(vp, vs, rho) gradient -> (phi, cc, sw) through Gassmann
    (src/PyFWI/grad_switcher.py:113-161)."""
    r = rock_properties
    dv = rp.pcs2dv_gassmann(pcs_model, r, method)
    phi, cc, sw = pcs_model["phi"], pcs_model["cc"], pcs_model["sw"]
    k_s = rp.vrh(r["k_c"], r["k_q"], cc, method)
    mu_s = rp.vrh(r["mu_c"], r["mu_q"], cc, method)
    rho_s = rp.weighted_average(r["rho_c"], r["rho_q"], cc)
    rho_f = rp.weighted_average(r["rho_w"], r["rho_h"], sw)
    k_f = rp.weighted_average(r["k_w"], r["k_h"], sw)
    g = grad_dv2kmd(gdv, dv["vp"], dv["vs"], dv["rho"])
    d_phi = grad_kmd2phi_gassmann(phi, k_s, mu_s, rho_s, k_f, rho_f, r["cs"])
    d_cc = grad_kmd2cc_gassmann(phi, cc, r["k_c"], r["mu_c"], r["rho_c"], r["k_q"], r["mu_q"],
                                r["rho_q"], k_f, r["cs"], method)
    d_sw = grad_kmd2sw_gassmann(phi, cc, r["k_c"], r["k_w"], r["k_h"], r["mu_c"], r["rho_w"],
                                r["k_q"], r["mu_q"], r["rho_h"], k_f, r["cs"])
    return {"phi": _combine(g, d_phi), "cc": _combine(g, d_cc), "sw": _combine(g, d_sw)}


def grad_kmd2phi_vrh(kr, kc, kq, kw, kh, muc, muq, cc, sw, rhos, rhof):
    gkv = -kc * cc - kq * (1 - cc) + kw * sw + kh * (1 - sw)
    gkr = kr ** 2 * (cc / kc + (1 - cc) / kq - sw / kw - (1 - sw) / kh)
    gmuv = -muc * cc - muq * (1 - cc)
    return 0.5 * (gkv + gkr), 0.5 * (gmuv + 0), rhof - rhos


def grad_kmd2cc_vrh(kr, phi, kc, muc, rhoc, kq, muq, rhoq):
    gkv = (1 - phi) * (kc - kq)
    gkr = (1 - phi) * kr ** 2 * (1 / kq - 1 / kc)
    gmuv = (1 - phi) * (muc - muq)
    return 0.5 * (gkv + gkr), 0.5 * (gmuv + 0), (1 - phi) * (rhoc - rhoq)


def grad_kmd2sw_vrh(kr, phi, kw, kh, rhow, rhoh):
    gk = 0.5 * (phi * (kw - kh) + phi * kr ** 2 * (1 / kh - 1 / kw))
    return gk, np.zeros(np.shape(gk)), phi * (rhow - rhoh)


def grad_dv2pcs_vrh(gdv, rock_properties, pcs_model, method=None):
    """(vp, vs, rho) gradient -> (phi, cc, sw) with VRH-averaged moduli
    (src/PyFWI/grad_switcher.py:21-68).  The published function evaluates the
    model with ``pcs2dv_gassmann(pcs_model, rock_properties)`` and so raises
    TypeError for the missing ``method``; that stays the behaviour when
    ``method`` is not given."""
    if method is None:
        raise TypeError("pcs2dv_gassmann() missing 1 required positional argument: 'method'")
    r = rock_properties
    dv = rp.pcs2dv_gassmann(pcs_model, r, method)
    phi, cc, sw = pcs_model["phi"], pcs_model["cc"], pcs_model["sw"]
    rho_s = rp.weighted_average(r["rho_c"], r["rho_q"], cc)
    rho_f = rp.weighted_average(r["rho_w"], r["rho_h"], sw)
    kr = 1 / ((1 - phi) * (cc / r["k_c"] + (1 - cc) / r["k_q"])
              + phi * (sw / r["k_w"] + (1 - sw) / r["k_h"]))
    g = grad_dv2kmd(gdv, dv["vp"], dv["vs"], dv["rho"])
    d_phi = grad_kmd2phi_vrh(kr, r["k_c"], r["k_q"], r["k_w"], r["k_h"], r["mu_c"], r["mu_q"],
                             cc, sw, rho_s, rho_f)
    d_cc = grad_kmd2cc_vrh(kr, phi, r["k_c"], r["mu_c"], r["rho_c"], r["k_q"], r["mu_q"], r["rho_q"])
    d_sw = grad_kmd2sw_vrh(kr, phi, r["k_w"], r["k_h"], r["rho_w"], r["rho_h"])
    return {"phi": _combine(g, d_phi), "cc": _combine(g, d_cc), "sw": _combine(g, d_sw)}


def grad_dv2phi_han(rhof, rhos, a2=6.9, b2=4.7):
    return -a2, -b2, rhof - rhos


def grad_dv2cc_han(phi, rho_c, rho_q, a3=2.2, b3=1.8):
    return -a3, -b3, (1 - phi) * (rho_c - rho_q)


def grad_dv2sw_han(phi, rho_w, rho_h):
    return 0, 0, phi * (rho_w - rho_h)


def grad_dv2pcs_han(gdv, rock_properties, pcs_model):
    """(vp, vs, rho) gradient -> (phi, cc, sw) through Han's regressions
    (src/PyFWI/grad_switcher.py:269-328)."""
    r = rock_properties
    g = (gdv["vp"], gdv["vs"], gdv["rho"])
    phi, cc, sw = pcs_model["phi"], pcs_model["cc"], pcs_model["sw"]
    rho_s = rp.weighted_average(r["rho_c"], r["rho_q"], cc)
    rho_f = rp.weighted_average(r["rho_w"], r["rho_h"], sw)
    return {"phi": _combine(g, grad_dv2phi_han(rho_f, rho_s)),
            "cc": _combine(g, grad_dv2cc_han(phi, r["rho_c"], r["rho_q"])),
            "sw": _combine(g, grad_dv2sw_han(phi, r["rho_w"], r["rho_h"]))}
