"""Full-waveform inversion driver with the interface of PyFWI's ``FWI``
(src/PyFWI/fwi.py:12-282): same constructor, ``__call__`` signature, ``inpa``
options and optimiser settings.

Structure of this module:
- ``Window``: the slice of the stacked model vector [p_1, p_2, p_3] (each
  nz*nx, e.g. vp, vs, rho) that one optimisation updates, the other
  parameters held fixed;
- ``Objective``: the misfit of one window -- data term plus the three
  regularisation terms of ``fwi_tools.Regularization`` -- with its gradient,
  evaluated once per point for the optimiser's separate f and f' calls;
- ``OPTIMIZERS``: the two scipy solvers and their settings;
- ``FWI.data_term``: forward modelling, cost and adjoint gradient of the
  whole model on the GPU.  When the cost is L2 with unit data weight and the
  observed data are as the plan expects them, it is one
  ``pyfwi_misfit_gradient`` call (records, residual, misfit and gradient
  stay on the device, a lowpass band included); otherwise
  forward_modeling -> cost_seismic -> gradient.  Both give the same misfit
  and gradient.

The driver runs the frequency bands in order and, in each band, the windows
k_0 .. k_end in steps of n_params (fwi.py:109-158, 174-227).
"""
from dataclasses import dataclass

import numpy as np
from scipy.optimize import fmin_cg, fmin_l_bfgs_b

from . import fwi_tools as tools
from .fwi_tools import Regularization
from .processing import prepare_residual
from .wave_propagation import WavePropagator

__all__ = ["FWI", "Window", "Objective", "OPTIMIZERS"]


# ------------------------------------------------------------------ optimisers
# Settings of the reference's solvers (fwi.py:189-191, 215-218).  Each entry
# maps (f, f', x0, iterations) to (x_opt, final misfit).

def _run_cg(f, fprime, x0, iterations):
    x, fopt, _, _, _ = fmin_cg(f, x0, fprime, gtol=1e-8, maxiter=iterations, full_output=True,
                               disp=False)
    return x, fopt


def _run_lbfgs(f, fprime, x0, iterations):
    x, fopt, _ = fmin_l_bfgs_b(f, x0, fprime, m=10, factr=1e7, pgtol=1e-8, bounds=None,
                               maxfun=15000, maxiter=iterations, maxls=10)
    return x, fopt


OPTIMIZERS = {"cg": _run_cg, "lbfgs": _run_lbfgs}
_METHOD_NAMES = {1: "cg", "cg": "cg", "CG": "cg", 2: "lbfgs", "lbfgs": "lbfgs"}


def method_name(method):
    """The reference's method codes (fwi.py:160-172): 1/'cg'/'CG', 2/'lbfgs';
    steepest descent (0/'sd'/'SD') is not provided."""
    if method in (0, "sd", "SD"):
        raise TypeError("Steepest descent is not provided yet.")
    try:
        return _METHOD_NAMES[method]
    except (KeyError, TypeError):
        raise ValueError(f"unknown optimisation method {method!r}") from None


# ------------------------------------------------------------------ windows

@dataclass(frozen=True)
class Window:
    """Parameters first .. first+count-1 (0-based) of a stacked model vector
    of `cells` values per parameter."""
    first: int
    count: int
    cells: int

    @property
    def lo(self):
        return self.first * self.cells

    @property
    def hi(self):
        return (self.first + self.count) * self.cells

    def take(self, x):
        return x[self.lo:self.hi]

    def put(self, x, part):
        """x with the window replaced by `part` (a new float32 vector, as the
        reference stacks its model, fwi.py:255)."""
        out = np.array(x, np.float32)
        out[self.lo:self.hi] = part
        return out

    @classmethod
    def sweep(cls, k_0, k_end, n_params, cells):
        """Windows of one band: k_0, k_0 + n_params, ... below k_end (0-based
        bounds, fwi.py:181, 207)."""
        return [cls(int(k), int(n_params), cells) for k in range(k_0, k_end, n_params)]


class Objective:
    """Misfit and gradient of one window of the model vector: data term of
    the whole model plus total-variation / Tikhonov terms of the window,
    parameter relations and the prior (fwi.py:254-282).  The value and
    gradient of the last point are kept, so the optimiser's f and f' calls
    cost one evaluation."""

    def __init__(self, fwi, full, window, freq):
        self.fwi, self.full, self.window, self.freq = fwi, full, window, freq
        self._x = None
        self._fg = None
        self.evaluations = 0

    def _eval(self, x):
        if self._x is None or not np.array_equal(x, self._x):
            self._x = np.array(x, copy=True)
            self._fg = self.evaluate(self._x)
            self.evaluations += 1
        return self._fg

    def value(self, x):
        return self._eval(x)[0]

    def gradient(self, x):
        return self._eval(x)[1]

    def evaluate(self, part):
        fwi, w = self.fwi, self.window
        model = w.put(self.full, part)
        k0, kend = w.first, w.first + w.count
        reg = fwi.regularization
        data, g_data = fwi.fprime(model, self.freq)
        smooth, g_smooth = reg.cost_regularization(part, tv_properties=fwi.tv_properties,
                                                   tikhonov_properties=fwi.tikhonov_properties)
        relation, g_relation = reg.parameter_relation(model, fwi.param_relation, k0, kend,
                                                      self.freq)
        prior, g_prior = reg.priori_regularization(part, fwi.prior_model, k0, kend, self.freq)
        total = data + smooth + relation + prior
        grad = w.take(g_data) + g_smooth + w.take(g_relation) + g_prior
        fwi.last_terms = {"freq": self.freq, "total": total, "data": data,
                          "regularization": smooth, "relation": relation, "prior": prior}
        if fwi.verbose:
            print(f"FWI band {self.freq}: misfit {total:.6e} (data {data:.6e}, "
                  f"regularization {smooth:.3e}, relation {relation:.3e}, prior {prior:.3e}); "
                  f"window [{part.min():.4g}, {part.max():.4g}]", flush=True)
        return total, grad


# ------------------------------------------------------------------ driver

# inpa options of the driver: key, attribute, default (fwi.py:64-103)
_OPTIONS = (("sd", "sd", 1.0),                     # data scaling (Virieux and Operto 2009)
            ("param_relation", "param_relation", None),
            ("prior_model", "prior_model", None),
            ("tv", "tv_properties", None),
            ("tikhonov", "tikhonov_properties", None),
            ("grad_coeff", "grad_coeff", None))


class FWI(WavePropagator):
    """Inversion of observed data ``d_obs`` (a seismic-section dict).  Options
    in ``inpa``: 'sd', 'param_relation', 'prior_model', 'tv', 'tikhonov',
    'cost_function_type' (default 'l2'), 'grad_coeff' (default [1, 1, 1]).
    ``param_functions`` ('dict2vec', 'vec2dict', 'model_to_dv',
    'grad_from_dv', optional 'args') invert another parameterisation; the
    default is (vp, vs, rho)."""

    def __init__(self, d_obs, inpa, src, rec_loc, model_shape, components, chpr,
                 n_well_rec=0, param_functions=None):
        super().__init__(inpa, src, rec_loc, model_shape, n_well_rec=n_well_rec, chpr=chpr,
                         components=components)
        self.regularization = Regularization(self.nx, self.nz, self.dh, self.dh)
        pf = param_functions or {}
        self.dict2vec = pf.get("dict2vec", tools.vel_dict2vec)
        self.vec2dict = pf.get("vec2dict", tools.vec2vel_dict)
        self.model_to_dv = pf.get("model_to_dv", lambda model, args: model)
        self.grad_from_dv = pf.get("grad_from_dv", lambda grad, args, model: grad)
        self.param_functions_args = pf.get("args", [])
        for key, attr, default in _OPTIONS:
            setattr(self, attr, inpa.get(key, default))
        if self.param_relation is None:
            self.param_relation = {}
        if self.grad_coeff is None:
            self.grad_coeff = [1.0, 1.0, 1.0]
        self.CF = tools.CostFunction(inpa.get("cost_function_type", "l2"))
        self.d_obs = prepare_residual(d_obs, 1)
        self.n_elements = self.nz * self.nx
        self.device_misfit = True  # the one-call device path where it applies
        self.verbose = True
        self.last_terms = None

    # ------------------------------------------------------------ driver
    def __call__(self, m0, method, iter, freqs, n_params, k_0, k_end):
        """Invert from ``m0`` with `method`, ``iter[b]`` iterations in band
        ``freqs[b]``, over the parameters k_0 .. k_end - 1 (1-based, as the
        reference) in windows of n_params.  Returns (model dict, misfits
        [band][window])."""
        k_0, k_end = k_0 - 1, k_end - 1
        if k_0 + n_params > k_end:
            raise Exception("k_0 + n_params can't be larger than k_end!!")
        solver = OPTIMIZERS[method_name(method)]
        x = np.asarray(self.dict2vec(m0))
        history = []
        for band, freq in enumerate(freqs):
            x, misfits = self._band(solver, x, iter[band], freq, n_params, k_0, k_end)
            history.append(misfits)
        return self.vec2dict(x, self.nz, self.nx), np.array(history)

    def _band(self, solver, x, iterations, freq, n_params, k_0, k_end):
        misfits = []
        for w in Window.sweep(k_0, k_end, n_params, self.n_elements):
            if self.verbose:
                print(f"FWI band {freq}: parameters {w.first + 1} .. {w.first + w.count}",
                      flush=True)
            obj = Objective(self, x, w, freq)
            part, fmin = solver(obj.value, obj.gradient, np.array(w.take(x)), iterations)
            misfits.append(fmin)
            x = np.concatenate((x[:w.lo], part, x[w.hi:]))
        return x, misfits

    def cg(self, m0, ITER, freq, n_params=1, k0=1, k_end=2):
        """One band of nonlinear conjugate gradients (fwi.py:174-199); k0 and
        k_end 0-based as the reference passes them."""
        return self._band(_run_cg, np.asarray(m0), ITER, freq, n_params, k0, k_end)

    def lbfgs(self, m0, ITER, freq, n_params=1, k0=1, k_end=2):
        """One band of L-BFGS-B (fwi.py:201-227)."""
        return self._band(_run_lbfgs, np.asarray(m0), ITER, freq, n_params, k0, k_end)

    # ------------------------------------------------------------ objective
    def fprime(self, m0, freq):
        """Data misfit of the model vector ``m0`` and its gradient vector in
        the inverted parameterisation, each parameter's gradient scaled by
        grad_coeff (fwi.py:229-252)."""
        model = self.vec2dict(np.array(m0), self.nz, self.nx)
        rms, grad_dv = self.data_term(self.model_to_dv(model, self.param_functions_args),
                                      self.d_obs, freq)
        grad = self.grad_from_dv(grad_dv, self.param_functions_args, model)
        for coeff, key in zip(self.grad_coeff, list(grad)):
            grad[key] = grad[key] * coeff
        return rms, self.dict2vec(grad)

    def fprime_single(self, m0, m_1, m1, freq):
        """Misfit and gradient of the window m0 between the fixed parts m_1
        and m1 (the reference's per-window objective, fwi.py:254-282)."""
        n = self.n_elements
        w = Window(len(m_1) // n, len(m0) // n, n)
        return Objective(self, np.hstack((m_1, m0, m1)), w, freq).evaluate(np.asarray(m0))

    def data_term(self, model_dv, d_obs, freq):
        """(misfit, gradient dict in vp/vs/rho) of a model dict against
        ``d_obs``: one pyfwi_misfit_gradient call where it applies, else
        forward_modeling, cost_seismic (multiscale lowpass, cost function,
        data weight) and gradient."""
        if self.device_misfit and self.device_misfit_applies(d_obs, freq, self.sd, self.CF):
            return self.misfit_gradient_device(model_dv, d_obs, freq=freq)
        d_est = prepare_residual(self.forward_modeling(model_dv, show=False), self.sd)
        rms, adj_src = tools.cost_seismic(d_est, d_obs, fun=self.CF, fn=self.fn, freq=freq,
                                          order=3, axis=1)
        return rms, self.gradient(adj_src, parameterization="dv")

    def _device_path_applies(self, freq):
        return self.device_misfit and self.device_misfit_applies(self.d_obs, freq, self.sd,
                                                                 self.CF)
