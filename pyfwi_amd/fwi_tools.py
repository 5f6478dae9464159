"""Finite-difference constants, absorbing-layer profiles, model padding and the
L2 misfit used on the propagator path.

Host-side mirror of the hot-path parts of PyFWI's ``PyFWI.fwi_tools``
(src/PyFWI/fwi_tools.py:452-647, 776-811, 943-1063, 1348-1422).
"""
import copy

import numpy as np

__all__ = ["Fdm", "expand_model", "pml_counstruction", "pml_construction",
           "pml_delta_calculation", "pml_profiles_1d", "CostFunction",
           "fd_coefficients", "kernel_constant"]


def fd_coefficients(order):
    """Staggered-grid coefficients (c1, c2, c3, c4) for order 4 or 8
    (src/PyFWI/fwi_tools.py:464-477, wave_propagation.py:423-432)."""
    if order == 4:
        return 9 / 8, -1 / 24, 0.0, 0.0
    if order == 8:
        return 1715 / 1434, -114 / 1434, 14 / 1434, -1 / 1434
    raise AssertionError("Order of the derivative has be either 4 or 8!")


def kernel_constant(x):
    """The value a kernel sees for a host constant: the reference pastes
    constants into its kernel source with ``%f`` (6 decimals) and the device
    compiler parses that text (src/PyFWI/wave_propagation.py:434-490)."""
    return float("%f" % float(x))


class Fdm:
    """Staggered finite-difference operators of order 4 or 8
    (src/PyFWI/fwi_tools.py:452-647)."""

    def __init__(self, order):
        self._order = order
        self._c1, self._c2, self._c3, self._c4 = fd_coefficients(order)
        self.dh_n = {4: 4, 8: 3}[order]

    @property
    def order(self):
        return self._order

    @property
    def c1(self):
        return self._c1

    @property
    def c2(self):
        return self._c2

    @property
    def c3(self):
        return self._c3

    @property
    def c4(self):
        return self._c4

    def _coef(self):
        return (self._c1, self._c2, self._c3, self._c4)[: self._order // 2]

    def _apply(self, x, dx, axis, plus):
        h = self._order // 2
        y = np.zeros(x.shape)
        n = x.shape[axis]
        acc = 0.0
        for k, c in enumerate(self._coef()):
            a = k + 1 if plus else k
            b = -k if plus else -(k + 1)
            sl_a = [slice(h, -h), slice(h, -h)]
            sl_b = [slice(h, -h), slice(h, -h)]
            sl_a[axis] = slice(h + a, n - h + a)
            sl_b[axis] = slice(h + b, n - h + b)
            acc = acc + c * (x[tuple(sl_a)] - x[tuple(sl_b)])
        y[h:-h, h:-h] = acc / dx
        return y

    def dxp(self, x, dx):
        return self._apply(x, dx, 1, True)

    def dxm(self, x, dx):
        return self._apply(x, dx, 1, False)

    def dzp(self, x, dx):
        return self._apply(x, dx, 0, True)

    def dzm(self, x, dx):
        return self._apply(x, dx, 0, False)

    def dot_product_test_derivatives(self, seed=None):
        """``<x, D+ y> + <D- x, y>`` for random zero-bordered fields; both
        errors vanish to rounding (src/PyFWI/fwi_tools.py:621-632)."""
        rng = np.random.default_rng(seed)
        x = rng.random((100, 100))
        y = rng.random((100, 100))
        for a in (x, y):
            a[:4, :] = a[-4:, :] = a[:, :4] = a[:, -4:] = 0
        ex = np.sum(x * self.dxp(y, 1)) - np.sum(-self.dxm(x, 1) * y)
        ez = np.sum(x * self.dzp(y, 1)) - np.sum(-self.dzm(x, 1) * y)
        return ex, ez

    def dt_computation(self, vp_max, dx, dz=None):
        """Largest stable time step (src/PyFWI/fwi_tools.py:634-647)."""
        if dz is None:
            dz = dx
        c_sum = abs(self._c1) + abs(self._c2) + abs(self._c3) + abs(self._c4)
        a = 1 / dx / dx * c_sum + 1 / dz / dz * c_sum
        return 2 / vp_max / np.sqrt(a * (1 + 4.0))


def expand_model(parameter, TNz, TNx, n_pml=10):
    """Pad an ``[nz, nx]`` property to ``[TNz, TNx]`` by replicating the edge
    rows, then the edge columns, into the absorbing layer; float32
    (src/PyFWI/fwi_tools.py:776-811)."""
    out = np.zeros((TNz, TNx), np.float32)
    out[n_pml:TNz - n_pml, n_pml:TNx - n_pml] = np.asarray(parameter, np.float32)
    out[:n_pml, :] = out[n_pml, :]
    out[TNz - n_pml:, :] = out[TNz - n_pml - 1, :]
    out[:, TNx - n_pml:] = out[:, TNx - n_pml - 1].reshape(-1, 1)
    out[:, :n_pml] = out[:, n_pml].reshape(-1, 1)
    return out


def pml_delta_calculation(dh, n_pml=10, pml_r=1e-5):
    """Damping ramp ``d(r) = -(3/(2L)) (r/L)^2 log10(R)`` on
    ``r = 0, dh, ..., (n_pml-1) dh`` with ``L = n_pml*dh``, float32
    (src/PyFWI/fwi_tools.py:1016-1063).  The products are formed in float64,
    as NumPy 1.x does for these scalar operands."""
    L = float(n_pml) * float(dh)
    r = np.arange(n_pml) * np.float64(dh)
    if L != 0:
        return np.float32(-(3 / (2 * L)) * ((r / L) ** 2) * np.log10(pml_r))
    return np.array([], np.float32)


def pml_profiles_1d(TNz, TNx, dh, n_pml=10, pml_r=1e-5, pml_dir=3):
    """Column profile ``dx[TNx]`` and row profile ``dz[TNz]`` of the absorbing
    layer.  ``pml_counstruction`` broadcasts these to full grids; every damping
    array of the reference is constant along one axis.
    ``pml_dir``: 0 no x damping, 1 no z damping, 2 all four sides, 3 no damping
    on the top side (src/PyFWI/fwi_tools.py:991-1013)."""
    a = pml_delta_calculation(dh, n_pml, pml_r)
    dx = np.zeros(TNx, np.float32)
    dz = np.zeros(TNz, np.float32)
    if n_pml > 0:
        dx[TNx - n_pml:] = a
        dx[:n_pml] = a[::-1]
        dz[TNz - n_pml:] = a
        dz[:n_pml] = a[::-1]
    if pml_dir == 0:
        dx[:] = 0
    elif pml_dir == 1:
        dz[:] = 0
    elif pml_dir == 3:
        dz[:n_pml] = 0
    return dx, dz


def pml_counstruction(TNz, TNx, dh, n_pml=10, pml_r=1e-5, pml_dir=3):
    """Full-grid damping arrays ``(dx_pml, dz_pml)``, float32
    (src/PyFWI/fwi_tools.py:943-1013; the spelling is the reference's)."""
    dx, dz = pml_profiles_1d(TNz, TNx, dh, n_pml, pml_r, pml_dir)
    dx_pml = np.ascontiguousarray(np.broadcast_to(dx[None, :], (TNz, TNx)), np.float32)
    dz_pml = np.ascontiguousarray(np.broadcast_to(dz[:, None], (TNz, TNx)), np.float32)
    return dx_pml, dz_pml


pml_construction = pml_counstruction


class CostFunction:
    """Misfit and adjoint source (src/PyFWI/fwi_tools.py:1348-1422).  Only the
    L2 misfit is on the propagator path."""

    def __init__(self, cost_function_type="l2"):
        if cost_function_type != "l2":
            raise NotImplementedError("only the 'l2' cost is provided")
        self.cost_function_type = cost_function_type

    def __call__(self, dest, dobs):
        return self.l2(dest, dobs)

    @staticmethod
    def list2dict(x):
        return {"vx": x[0], "vz": x[1], "taux": x[2], "tauz": x[3], "tauxz": x[4]}

    def l2(self, dest0, dobs0):
        """``0.5*sum(res^2)`` with ``res = float32(dest - dobs)`` per component;
        returns ``(rms, res)``."""
        is_arr = isinstance(dest0, np.ndarray)
        dest = self.list2dict(dest0) if is_arr else copy.deepcopy(dest0)
        dobs = self.list2dict(dobs0) if is_arr else copy.deepcopy(dobs0)
        rms = 0
        res = {}
        for par in dest:
            res[par] = np.float32(dest[par] - dobs[par])
            r = res[par].reshape(-1)
            rms += 0.5 * (r.T @ r)
        if is_arr:
            return rms, np.array(list(res.values()))
        return rms, res
