"""Shared helpers of the test suite (oracle set-up from a workload)."""
import numpy as np

from oracle import oracle as O


def oracle_setup(w, chpr=None, components=None):
    src_loc = np.c_[w.src.i, w.src.j].astype(np.float64) * w.inpa["dh"]
    return O.Setup(w.inpa, src_loc, w.rec_loc, w.model_shape, w.src.w, w.src.component,
                   chpr=w.chpr if chpr is None else chpr,
                   components=w.components if components is None else components)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
