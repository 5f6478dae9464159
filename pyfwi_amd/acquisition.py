"""Acquisition geometry, source wavelets and seismogram sections.

Host-side mirror of PyFWI's ``PyFWI.acquisition`` for the surface-seismic hot
path (src/PyFWI/acquisition.py).  Same names, argument meaning and outputs.
"""
import numpy as np

__all__ = ["acq_parameters", "surface_seismic", "crosswell", "Source",
           "seismic_section"]


def acq_parameters(ns, rec_dis, offsetx, depth, dh, sdo, acq_type):
    """Source and receiver positions for an acquisition type
    (src/PyFWI/acquisition.py:5-59).

    Returns ``(src_loc, rec_loc, n_surface_rec, n_well_rec)`` in metres.
    Only ``acq_type == 1`` (surface) is propagated by this package; the
    crosswell geometry is still produced for callers that only need positions.
    """
    if rec_dis < dh:
        raise Exception("Receiver distance should be larger than spatial sampling")
    if acq_type == 0:
        src_loc, rec_loc = crosswell(ns, rec_dis, offsetx, depth, dh, sdo)
        return src_loc, rec_loc, 0, rec_loc.shape[0]
    if acq_type == 1:
        src_loc, rec_loc = surface_seismic(ns, rec_dis, offsetx, dh, sdo)
        return src_loc, rec_loc, rec_loc.shape[0], 0
    if acq_type == 2:
        src_loc, rec_top = surface_seismic(ns, rec_dis, offsetx, dh, sdo)
        _, rec_right = crosswell(ns, rec_dis, offsetx, depth, dh, sdo)
        # left well: the surface receivers' depth column, the right well's depths
        left_x = np.full((rec_right.shape[0], 1), np.float32(rec_top[0, 1]), np.float32)
        rec_left = np.hstack((left_x, rec_right[:, 1].reshape(-1, 1)))
        rec_loc = np.vstack((rec_left, rec_top, rec_right))
        return src_loc, rec_loc, rec_top.shape[0], rec_right.shape[0]
    raise ValueError("acq_type must be 0, 1 or 2")


def crosswell(ns, rec_dis, offsetx, depth, dh, sdo):
    """Vertical receiver line at the right edge, sources at the left edge
    (src/PyFWI/acquisition.py:62-124)."""
    rec_z = np.arange(2 * rec_dis, depth - 2 * rec_dis, rec_dis).reshape(-1, 1)
    nr = rec_z.size
    src_loc = np.empty((ns, 2), dtype=np.float32)
    src_loc[:, 0] = (sdo / 2 + 2) * dh
    if ns == 1:
        src_loc[:, 1] = depth / 2
    else:
        src_loc[:, 1] = np.linspace(4 * dh, depth - 9 * dh, ns)
    rec_loc = np.hstack((np.ones((nr, 1)) * offsetx - (sdo / 2 + 2) * dh, rec_z))
    return src_loc.astype(np.float32), rec_loc.astype(np.float32)


def surface_seismic(ns, rec_dis, offsetx, dh, sdo):
    """Surface line of receivers and sources (src/PyFWI/acquisition.py:127-185).

    Receivers at ``x = arange(2*rec_dis, offsetx - 2*rec_dis, rec_dis)`` and
    depth ``(sdo/2 + 2)*dh``; one source in the middle or ``ns`` sources on
    ``linspace(4*dh, offsetx - 9*dh, ns)`` at the same depth.
    """
    rec_x = np.arange(2 * rec_dis, offsetx - 2 * rec_dis, rec_dis).reshape(-1, 1)
    nr = rec_x.size
    src_loc = np.empty((ns, 2), dtype=np.float32)
    src_loc[:, 1] = (sdo / 2 + 2) * dh
    if ns == 1:
        src_loc[:, 0] = offsetx / 2
    else:
        src_loc[:, 0] = np.linspace(4 * dh, offsetx - 9 * dh, ns)
    rec_loc = np.hstack((rec_x, np.ones((nr, 1)) * (sdo / 2 + 2) * dh))
    return src_loc, rec_loc


class Source:
    """Point sources sharing one wavelet (src/PyFWI/acquisition.py:188-257).

    ``src_type``: 0 explosive (adds to both normal stresses), 1 force along x
    (vx), 2 force along z (vz).  ``i``/``j`` are the grid column/row of each
    source, truncated from ``src_loc/dh`` in the array's own precision.
    """

    def __init__(self, src_loc, dh, dt, src_type=0):
        self.dh = dh
        self.i = np.int32(src_loc[:, 0] / self.dh)
        self.j = np.int32(src_loc[:, 1] / self.dh)
        self.dt = dt
        self.component = np.zeros(5, dtype=np.float32)
        if src_type == 0:
            self.component[2:4] = np.float32(1)
        elif src_type == 1:
            self.component[0] = np.float32(1)
        elif src_type == 2:
            self.component[1] = np.float32(1)
        else:
            raise ValueError("src_type must be 0 (explosive), 1 (x force) or 2 (z force)")

    def __call__(self, ind=None):
        """Source sample ``ind`` times the component vector; zero past the
        end of the wavelet."""
        if ind < self.w.size:
            return self.w[ind] * self.component
        return np.float32(0.0) * self.component

    def Ricker(self, fdom):
        """Ricker wavelet on ``t = arange(-1/fdom, 1/fdom + dt/3, dt)``,
        stored in float32 (src/PyFWI/acquisition.py:231-244)."""
        self.t = np.arange(-1.0 / fdom, 1.0 / fdom + self.dt / 3, self.dt)
        arg = (np.pi * fdom * self.t) ** 2
        self.w = np.float32((1.0 - 2.0 * arg) * np.exp(-arg))

    def delta(self):
        """Single unit spike (src/PyFWI/acquisition.py:246-257)."""
        self.w = np.float32(np.array([1]))

    def samples(self, n):
        """``[n, 5]`` float32 table of ``self(k)`` for ``k < n`` (what the
        propagator injects at each recorded sample)."""
        out = np.zeros((n, 5), np.float32)
        m = min(n, self.w.size)
        out[:m] = self.w[:m, None] * self.component[None, :]
        return out


_SEIS_PLAN = {
    1: ["taux"],
    2: ["vx", "vz"],
    3: ["taux", "tauz", "tauxz"],
    4: ["vx", "vz", "taux", "tauz", "tauxz"],
}


def seismic_section(seismo, components=0, shape="3d"):
    """Select output components of a ``{field: [nt, nr, ns]}`` record
    (src/PyFWI/acquisition.py:313-333).

    ``components``: 0 pressure ``(taux + tauz)/2`` returned under both 'taux'
    and 'tauz'; 1 taux; 2 vx, vz; 3 stresses; 4 all.  ``shape='2d'`` reshapes
    each to ``[nt, nr*ns]`` in Fortran order (column = shot*nr + receiver).
    """
    data = {}
    if components != 0:
        for par in _SEIS_PLAN[int(components)]:
            data[par] = seismo[par]
    else:
        p = (seismo["taux"] + seismo["tauz"]) / 2
        data["taux"] = p
        data["tauz"] = (seismo["taux"] + seismo["tauz"]) / 2
    if shape == "2d":
        nt, nr, ns = data[next(iter(data))].shape
        for par in data:
            data[par] = np.reshape(data[par], (nt, nr * ns), "F")
    return data
