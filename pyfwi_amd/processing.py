"""Residual preparation (src/PyFWI/processing.py:142-168)."""
import numpy as np

__all__ = ["prepare_residual", "COMPONENTS", "COMPONENT_MASK"]

COMPONENTS = ("vx", "vz", "taux", "tauz", "tauxz")


def prepare_residual(res, s=1.0):
    """Return all five components of a record: present ones scaled by ``s``,
    missing ones as float32 zeros of the same shape."""
    shape = res[next(iter(res))].shape
    data = {}
    for par in COMPONENTS:
        if par in res:
            data[par] = s * res[par]
        else:
            data[par] = np.zeros(shape, np.float32)
    return data


# components (acquisition.py:313-333) -> bits of pyfwi_misfit_gradient's
# comp_mask over (vx, vz, txx, tzz, txz); 0 compares the pressure
COMPONENT_MASK = {0: 0, 1: 0b00100, 2: 0b00011, 3: 0b11100, 4: 0b11111}
