"""Write tests/golden/pcs_vectors.npz: input/output vectors of PyFWI's rock-physics
(PCS) gradient parameterisations.

These chain rules follow PyFWI's specific rock-physics formulas (not exact
derivatives of the matching forward models), so no closed form pins them; the
vectors come from the reference itself.  Run in the container that holds the
reference (read from /root/reference/src); only the resulting data file is
committed.  Functions sampled: rock_physics.pcs2dv_gassmann / pcs2dv_vrh /
pcs2dv_han / drained_moduli / biot_gassmann / vrh, and grad_switcher.
grad_dv2pcs_gassmann (VRH, Voigt, Reuss) / grad_dv2pcs_han / grad_vrh /
grad_dv2pcs_vrh (which raises TypeError in the reference).
"""
import os
import sys

import numpy as np

REF = "/root/reference/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcs_vectors.npz")


def inputs():
    rng = np.random.default_rng(2024)
    shape = (4, 6)
    pcs = {"phi": rng.uniform(0.05, 0.35, shape), "cc": rng.uniform(0.0, 0.5, shape),
           "sw": rng.uniform(0.1, 1.0, shape)}
    rock = {"k_q": 37.0, "k_c": 21.0, "k_w": 2.25, "k_h": 0.8, "mu_q": 44.0, "mu_c": 7.0,
            "rho_q": 2.65, "rho_c": 2.55, "rho_w": 1.0, "rho_h": 0.7, "cs": 20.0}
    gdv = {k: rng.standard_normal(shape) for k in ("vp", "vs", "rho")}
    return pcs, rock, gdv


def main():
    sys.path.insert(0, REF)
    import PyFWI.rock_physics as rp
    import PyFWI.grad_switcher as gs
    pcs, rock, gdv = inputs()
    out = {}
    for k, v in pcs.items():
        out["in_" + k] = v
    for k, v in gdv.items():
        out["in_g" + k] = v
    for method in ("VRH", "Voigt", "Reuss"):
        dv = rp.pcs2dv_gassmann(pcs, rock, method)
        for k, v in dv.items():
            out[f"pcs2dv_gassmann_{method}_{k}"] = v
        g = gs.grad_dv2pcs_gassmann(gdv, rock, pcs, method)
        for k, v in g.items():
            out[f"grad_dv2pcs_gassmann_{method}_{k}"] = v
        out[f"grad_vrh_{method}"] = gs.grad_vrh(rock["k_c"], rock["k_q"], pcs["cc"], method)
        out[f"vrh_{method}"] = rp.vrh(rock["k_c"], rock["k_q"], pcs["cc"], method)
    for k, v in rp.pcs2dv_vrh(pcs, rock).items():
        out["pcs2dv_vrh_" + k] = v
    for k, v in rp.pcs2dv_han(pcs, rock).items():
        out["pcs2dv_han_" + k] = v
    for k, v in gs.grad_dv2pcs_han(gdv, rock, pcs).items():
        out["grad_dv2pcs_han_" + k] = v
    kd, gd = rp.drained_moduli(pcs["phi"], 30.0, 20.0, rock["cs"])
    out["drained_kd"], out["drained_gd"] = kd, gd
    ku, C, M = rp.biot_gassmann(pcs["phi"], 2.0, 30.0, kd)
    out["biot_ku"], out["biot_C"], out["biot_M"] = ku, C, M
    try:
        gs.grad_dv2pcs_vrh(gdv, rock, pcs)
        out["grad_dv2pcs_vrh_raises"] = np.array(0)
    except TypeError:
        out["grad_dv2pcs_vrh_raises"] = np.array(1)
    np.savez(OUT, **out)
    print(f"wrote {OUT} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
