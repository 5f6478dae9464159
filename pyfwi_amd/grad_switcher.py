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
