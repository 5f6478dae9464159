"""Finite-difference constants, absorbing-layer profiles, model padding, the
L2 misfit with its multiscale (lowpass) preparation, regularisation and the
model-vector helpers of the inversion driver.

Host-side mirror of PyFWI's ``PyFWI.fwi_tools`` (src/PyFWI/fwi_tools.py:
17-447 derivative operators and Regularization, 452-647 Fdm, 776-811
expand_model, 943-1063 PML, 1066-1097 model vectors, 1153-1345 cost
preparation, lowpass and cost_seismic, 1348-1422 CostFunction).
"""
import copy

import numpy as np
import scipy.sparse as sp
from scipy.ndimage import gaussian_filter
from scipy.signal import butter, freqz

__all__ = ["Fdm", "expand_model", "pml_counstruction", "pml_construction",
           "pml_delta_calculation", "pml_profiles_1d", "CostFunction",
           "fd_coefficients", "kernel_constant", "derivative", "Regularization",
           "vel_dict2vec", "vec2vel_dict", "cost_preparation", "lowpass", "adj_lowpass",
           "adj_cost_preparation", "source_weighting", "cost_seismic", "dict_diff",
           "dict_summation"]


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


# ------------------------------------------------------------------ model vectors

def vel_dict2vec(m0):
    """{'vp','vs','rho': [nz, nx]} -> [vp, vs, rho] flattened, float64
    (src/PyFWI/fwi_tools.py:1066-1073)."""
    nz, nx = m0[next(iter(m0))].shape
    return np.concatenate([np.asarray(m0[k], np.float64).reshape(-1)
                           for k in ("vp", "vs", "rho")]).reshape(3 * nz * nx)


def vec2vel_dict(m0, nz, nx):
    """Inverse of vel_dict2vec, as views of `m0` (fwi_tools.py:1076-1097)."""
    n = nz * nx
    return {"vp": m0[:n].reshape(nz, nx), "vs": m0[n:2 * n].reshape(nz, nx),
            "rho": m0[2 * n:].reshape(nz, nx)}


# ------------------------------------------------------------------ derivative operators

def derivative(nx, nz, dx, dz, order):
    """Sparse first- (order=1) or second-order (order=2) derivative matrices
    (Dx, Dz) over a row-major nz x nx grid (src/PyFWI/fwi_tools.py:17-134).

    Order 1: one-sided differences on the first and last row/column, centred
    ones inside.  Order 2: the three-point stencil, shifted inwards on the
    edges.  As in the reference, Dz is scaled by 1/dx (1/dx^2) and Dx by 1/dz
    (1/dz^2); PyFWI always has dx = dz."""
    n = nz * nx
    idx = np.arange(n).reshape(nz, nx)
    if order == 1:
        # z: forward on row 0, backward on the last row, centred inside
        lo_z = np.vstack([idx[:1], idx[:-2], idx[-2:-1]])
        hi_z = np.vstack([idx[1:2], idx[2:], idx[-1:]])
        wz = np.concatenate([np.ones(nx), np.full((nz - 2) * nx, 0.5), np.ones(nx)]) / dx
        lo_x = np.hstack([idx[:, :1], idx[:, :-2], idx[:, -2:-1]])
        hi_x = np.hstack([idx[:, 1:2], idx[:, 2:], idx[:, -1:]])
        wx = np.tile(np.concatenate([[1.0], np.full(nx - 2, 0.5), [1.0]]), nz) / dz
        rows = np.repeat(np.arange(n), 2)

        def build(lo, hi, w):
            cols = np.stack([lo.reshape(-1), hi.reshape(-1)], 1).reshape(-1)
            vals = np.stack([-w, w], 1).reshape(-1)
            return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
        return build(lo_x, hi_x, wx), build(lo_z, hi_z, wz)
    # second order: centre row/column clamped to 1 .. n-2
    cz = np.clip(np.arange(nz), 1, nz - 2)
    cx = np.clip(np.arange(nx), 1, nx - 2)
    rows = np.repeat(np.arange(n), 3)
    stencil = np.array([1.0, -2.0, 1.0])

    def build(centre, step, scale):
        cols = (centre.reshape(-1, 1) + step * np.array([-1, 0, 1])).reshape(-1)
        vals = np.tile(stencil * scale, n)
        return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    Dz = build(idx[cz, :], nx, 1 / (dx * dx))
    Dx = build(idx[:, cx], 1, 1 / (dz * dz))
    return Dx, Dz


class Regularization:
    """Regularisation terms of the inversion (src/PyFWI/fwi_tools.py:137-447):
    total variation, Tikhonov of order 1 and 0, relations between parameters
    and a prior model.  Model vectors hold n_params blocks of nz*nx."""

    def __init__(self, nx, nz, dx, dz):
        self.idx, self.idz = 1 / dx, 1 / dz
        self.idx2, self.idz2 = 1 / (dx * dx), 1 / (dz * dz)
        self.dx, self.dz = dx, dz
        self.nx, self.nz = nx, nz
        self.n_elements = nz * nx
        self.Bx2, self.Bz2 = derivative(nx, nz, dx, dz, 2)
        self.Bx1, self.Bz1 = derivative(nx, nz, dx, dz, 1)
        self.D2 = self.Bx2.T @ self.Bx2 + self.Bz2.T @ self.Bz2

    def cost_regularization(self, x0, tv_properties=None, tikhonov_properties=None,
                            tikhonov0_properties=None):
        """Weighted sum of the requested terms; each properties dict holds
        'lambda_weight' (and 'az', 'ax' for tv / tikhonov).  As in the
        reference the order-0 Tikhonov term adds its weight, not a gradient,
        to the gradient (fwi_tools.py:193-197)."""
        x = np.copy(x0)
        rms = 0
        grad = np.zeros(x.shape)
        if tv_properties:
            f, g = self.tv(x, 1e-7, alpha_z=tv_properties["az"], alpha_x=tv_properties["ax"])
            rms += tv_properties["lambda_weight"] * f
            grad += tv_properties["lambda_weight"] * g
        if tikhonov_properties:
            f, g = self.tikhonov(x, alpha_z=tikhonov_properties["az"],
                                 alpha_x=tikhonov_properties["ax"])
            rms += tikhonov_properties["lambda_weight"] * f
            grad += tikhonov_properties["lambda_weight"] * g
        if tikhonov0_properties:
            f, _ = self.tikhonov_0(x)
            rms += tikhonov0_properties["lambda_weight"] * f
            grad += tikhonov0_properties["lambda_weight"]
        return rms, grad

    def _blocks(self, x, op):
        n = self.n_elements
        out = np.zeros(len(x))
        for i in range(len(x) // n):
            out[i * n:(i + 1) * n] = op(x[i * n:(i + 1) * n])
        return out

    def _gradient_image(self, m, alpha_z, alpha_x):
        mz1 = self.Bz1 @ m
        mz1[:17 * self.nx] = 0.0  # the reference skips the first 17 rows
        return alpha_x * (self.Bx1 @ m) + alpha_z * mz1

    def tv(self, x0, eps, alpha_z, alpha_x):
        """l1 of alpha_x Dx m + alpha_z Dz m per block (fwi_tools.py:202-237).
        The returned gradient is that of l1 with respect to the derivative
        image, as in the reference."""
        x1 = self._blocks(np.copy(x0), lambda m: self._gradient_image(m, alpha_z, alpha_x))
        return self.l1(x1, eps)

    @staticmethod
    def l1(x0, eps=1e-7):
        """sum x^2 / max(|x|, eps) and its gradient 2 x / max(|x|, eps)
        (fwi_tools.py:239-253)."""
        x = np.copy(x0)
        w = 1 / np.sqrt(np.maximum(np.abs(x), eps))
        wx = w * x
        return wx @ wx, 2 * wx * w

    @staticmethod
    def l2(x0):
        x = np.copy(x0)
        return x @ x

    def tikhonov(self, x0, alpha_z, alpha_x):
        """||alpha_x Dx m + alpha_z Dz m||^2 per block, with the reference's
        gradient alpha_x Dx'Dx m + alpha_z Dz'Dz m (fwi_tools.py:262-305)."""
        x = np.copy(x0)
        x1 = self._blocks(x, lambda m: self._gradient_image(m, alpha_z, alpha_x))
        grad = self._blocks(x, lambda m: alpha_x * (self.Bx1.T @ (self.Bx1 @ m))
                            + alpha_z * (self.Bz1.T @ (self.Bz1 @ m)))
        return self.l2(x1), grad

    def tikhonov_0(self, x0):
        """l1 of the model itself (fwi_tools.py:307-338)."""
        return self.l1(np.copy(x0))

    def parameter_relation(self, m0, models, k0, kend, freq):
        """0.5 lam ||m_b - model(m_a)||^2 for each relation 'ab' (1-based
        parameter numbers) active at `freq`; its gradient on block b is
        smoothed by a unit Gaussian (fwi_tools.py:340-385).  Like the
        reference, a relation not active at `freq` ends the loop."""
        rms = 0
        grad = np.zeros(m0.shape)
        n = self.n_elements
        for param in models:
            a, b = (int(c) for c in str(param))
            spec = models[param]
            if freq not in np.array(spec["freqs"]).reshape(-1):
                return 0.0, grad
            if b in np.arange(k0 + 1, kend + 1):
                pred = spec["model"](m0[(a - 1) * n:a * n])
                dm = m0[(b - 1) * n:b * n] - pred
                rms += 0.5 * spec["lam"] * np.dot(dm, dm)
                grad[(b - 1) * n:b * n] = gaussian_filter(spec["lam"] * dm * 1, 1)
        return rms, grad

    def priori_regularization(self, m0, regularization_dict, k0, kend, freq):
        """0.5 lam (m - m_p)' W (m - m_p) with W = I / var(m_p) over blocks
        k0..kend-1 (Asnaashari et al. 2013, eq. 5; fwi_tools.py:387-447)."""
        if regularization_dict is None:
            return 0.0, np.zeros(m0.shape, np.float64)
        m = np.copy(m0[:kend * self.n_elements])
        if freq not in np.array(regularization_dict["freqs"]).reshape(-1):
            return 0.0, np.zeros(m.shape, np.float64)
        lam = regularization_dict["lam"]
        mp_dict = regularization_dict["mp"]
        keys = list(mp_dict)
        n = self.n_elements
        if len(m) != (kend - k0) * n:
            raise ValueError("the model vector must hold exactly the blocks k0..kend-1")
        mp = np.zeros(m.shape)
        for i in range(kend - k0):
            mp[i * n:(i + 1) * n] = mp_dict[keys[k0 + i]].reshape(-1)
        diff = m - mp
        wdiff = diff / np.var(mp)
        return lam * 0.5 * float(wdiff @ diff), lam * wdiff


# ------------------------------------------------------------------ multiscale cost

_PAD = 512  # zero samples appended before filtering (fwi_tools.py:1193-1195)


def _butter_response(highcut, fn, order, n):
    """Frequency response of the order-`order` Butterworth lowpass with the
    reference's cut-off 2*highcut/fn, on n points of the whole circle."""
    b, a = butter(order, 2 * highcut / fn, btype="lowpass", output="ba")
    return freqz(b, a, worN=n, whole=True)[1]


def _zero_phase(x1, highcut, fn, order, first_conj):
    x = np.asarray(x1)
    x = np.concatenate((x, np.zeros((x.shape[0], _PAD, x.shape[2]))), axis=1)
    n = x.shape[1]
    h = _butter_response(highcut, fn, order, n)[None, :, None]
    spec = np.fft.fft(x, n=n, axis=1)
    first, second = (np.conjugate(h), h) if first_conj else (h, np.conjugate(h))
    spec = (first * spec) * second
    return np.fft.ifft(spec, n=n, axis=1).real[:, :-_PAD, :]


def lowpass_taps(highcut, fn, order, n, pad=_PAD, tol=1e-12):
    """Taps k[0..L] of the symmetric kernel of lowpass / adj_lowpass on traces
    of n samples: both multiply the (n + pad)-point spectrum of the padded
    trace by |H|^2 (H and conj(H) in either order), a circular convolution
    with k = ifft(|H|^2), real and even.  L is the last lag up to half the
    circle where |k| exceeds tol * k[0].  For the device filter of
    pyfwi_misfit_gradient (pyfwi_set_record_filter), which sums lags -L..L:
    on an even circle lags +-N/2 are one residue, so that tap is halved."""
    N = int(n) + int(pad)
    h = _butter_response(highcut, fn, order, N)
    k = np.fft.ifft(np.abs(h) ** 2).real
    half = N // 2
    big = np.nonzero(np.abs(k[:half + 1]) > tol * abs(k[0]))[0]
    L = int(big[-1]) if big.size else 0
    taps = np.array(k[:L + 1], np.float64)
    if N % 2 == 0 and L == half:
        taps[L] *= 0.5
    return taps


def lowpass(x1, highcut, fn, order=1, axis=1, show=False):
    """Zero-phase Butterworth lowpass along time of records [ncomp, nt, ntr]:
    pad 512 zero samples, multiply the spectrum by H and then by conj(H), and
    drop the padding (src/PyFWI/fwi_tools.py:1189-1222).  Only axis=1 is
    meaningful, as in the reference."""
    if axis != 1:
        raise ValueError("records are filtered along axis 1 (time)")
    return _zero_phase(x1, highcut, fn, order, first_conj=False)


def adj_lowpass(x, highcut, fn, order, axis=1):
    """Adjoint of lowpass: the spectrum is multiplied by conj(H) and then by
    H (src/PyFWI/fwi_tools.py:1225-1255)."""
    if axis != 1:
        raise ValueError("records are filtered along axis 1 (time)")
    return _zero_phase(x, highcut, fn, order, first_conj=True)


def cost_preparation(dpre, dobs, fn, freq=None, order=None, axis=None, params_oh=None):
    """Lowpass both data sets at `freq` (multiscale inversion, Bunks et al.
    1995) and weight them by params_oh (fwi_tools.py:1153-1186)."""
    x_pre = copy.deepcopy(dpre)
    x_obs = copy.deepcopy(dobs)
    if freq:
        x_obs = lowpass(x_obs, freq, fn, order=order, axis=axis)
        x_pre = lowpass(x_pre, freq, fn, order=order, axis=axis)
    if params_oh is not None:
        x_obs = params_oh * x_obs
        x_pre = params_oh * x_pre
    return x_pre, x_obs


def adj_cost_preparation(res, fn, freq=False, order=None, axis=None, params_oh=None):
    """Adjoint of cost_preparation applied to the residual
    (fwi_tools.py:1258-1272).  As in the reference, params_oh scales a copy
    that the lowpass branch then does not use."""
    x_res = np.copy(res)
    if params_oh is not None:
        x_res = params_oh * x_res
    if freq:
        x_res = adj_lowpass(res, freq, fn, order=order, axis=axis)
    return x_res


def source_weighting(d_pre, d_obs, ns, nr):
    """Per-shot least-squares amplitude ratio <d_pre, d_obs> / <d_pre, d_pre>
    over the shot's receiver columns (fwi_tools.py:1275-1287)."""
    x_pre = np.copy(d_pre)
    alpha = np.zeros(x_pre.shape)
    for k in range(ns):
        cols = slice(k * nr, (k + 1) * nr)
        p = x_pre[:, cols].reshape(-1)
        alpha[:, cols] = (p @ d_obs[:, cols].reshape(-1)) / (p @ p)
    return alpha


def cost_seismic(d_pre0, d_obs0, fun, fn=None, freq=None, order=None, axis=None,
                 sourc_weight=False, ns=None, nr=None, params_oh=None):
    """Misfit of predicted against observed records (dicts of the five
    components, [nt, nr*ns] each), after the multiscale preparation, and the
    adjoint source to inject (src/PyFWI/fwi_tools.py:1290-1345).  Returns
    (rms, {component: float32 [nt, nr*ns]})."""
    dpre = np.array(list(copy.deepcopy(d_pre0).values()))
    dobs = np.array(list(copy.deepcopy(d_obs0).values()))
    x_pre, x_obs = cost_preparation(dpre, dobs, fn, freq=freq, order=order, axis=axis,
                                    params_oh=params_oh)
    alpha = 1.0
    if sourc_weight:
        alpha = source_weighting(dpre, dobs, ns, nr)
    rms, res = fun(alpha * x_pre, x_obs)
    adj = adj_cost_preparation(res, fn, freq=freq, order=order, axis=axis, params_oh=params_oh)
    names = ("vx", "vz", "taux", "tauz", "tauxz")
    return rms, {k: adj[i].astype(np.float32) for i, k in enumerate(names)}


# ------------------------------------------------------------------ model dictionaries

def dict_diff(dict1, dict2, positivity=False):
    """dict1 - dict2 per key of dict1; negative differences set to zero when
    `positivity` (src/PyFWI/fwi_tools.py:1537-1558)."""
    out = {}
    for k in dict1:
        out[k] = dict1[k] - dict2[k]
        if positivity:
            out[k][out[k] < 0] = 0
    return out


def dict_summation(dict1, dict2, division=1.0):
    """(dict1 + dict2) / division per key of dict1 (fwi_tools.py:1561-1580)."""
    return {k: (dict1[k] + dict2[k]) / division for k in dict1}
