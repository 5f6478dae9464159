"""Residual preparation (src/PyFWI/processing.py:142-168)."""
import numpy as np

__all__ = ["prepare_residual", "COMPONENTS"]

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
