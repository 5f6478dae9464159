"""Full-waveform inversion driver (src/PyFWI/fwi.py:12-282).

``FWI`` is a ``WavePropagator`` that minimises the misfit of its observed data
over the model with scipy's L-BFGS-B or nonlinear conjugate gradients, one
frequency band after another (multiscale: ``cost_seismic`` lowpasses both data
sets at each band's frequency), one group of parameters at a time, with the
regularisation terms of ``fwi_tools.Regularization``.

Each misfit evaluation (``fprime``) runs the forward modelling, the cost and
the adjoint gradient on the GPU.  When the band needs no lowpass, the cost is
L2 with unit data weight and the observed data are as the plan expects them,
the three run as one ``pyfwi_misfit_gradient`` call; otherwise as the
reference's sequence of forward_modeling, cost_seismic and gradient.  Both
give the same misfit and gradient.
"""
import numpy as np
from scipy.optimize import fmin_cg, fmin_l_bfgs_b
from scipy.optimize._optimize import MemoizeJac

from . import fwi_tools as tools
from .fwi_tools import Regularization
from .processing import prepare_residual
from .wave_propagation import WavePropagator

__all__ = ["FWI"]


class FWI(WavePropagator):
    """Parameters as the reference (fwi.py:12-107): observed data ``d_obs``
    (a seismic section dict), ``inpa`` (with optional 'sd', 'param_relation',
    'prior_model', 'tv', 'tikhonov', 'cost_function_type', 'grad_coeff'),
    source, receivers, model shape, components, chpr, n_well_rec and optional
    ``param_functions`` ('dict2vec', 'vec2dict', 'model_to_dv',
    'grad_from_dv', 'args') for other parameterisations."""

    def __init__(self, d_obs, inpa, src, rec_loc, model_shape, components, chpr,
                 n_well_rec=0, param_functions=None):
        super().__init__(inpa, src, rec_loc, model_shape, n_well_rec=n_well_rec, chpr=chpr,
                         components=components)
        self.regularization = Regularization(self.nx, self.nz, self.dh, self.dh)
        if param_functions is None:
            self.dict2vec = tools.vel_dict2vec
            self.vec2dict = tools.vec2vel_dict
            self.model_to_dv = lambda a, args: a
            self.grad_from_dv = lambda a, args, b: a
            self.param_functions_args = []
        else:
            self.dict2vec = param_functions["dict2vec"]
            self.vec2dict = param_functions["vec2dict"]
            self.model_to_dv = param_functions["model_to_dv"]
            self.grad_from_dv = param_functions["grad_from_dv"]
            self.param_functions_args = param_functions.get("args", [])
        self.sd = inpa.get("sd", 1.0)  # Virieux and Operto 2009
        self.param_relation = inpa.get("param_relation", {})
        self.prior_model = inpa.get("prior_model", None)
        self.tv_properties = inpa.get("tv", None)
        self.tikhonov_properties = inpa.get("tikhonov", None)
        self.d_obs = prepare_residual(d_obs, 1)
        self.CF = tools.CostFunction(inpa.get("cost_function_type", "l2"))
        self.grad_coeff = inpa.get("grad_coeff", [1.0, 1.0, 1.0])
        self.n_elements = self.nz * self.nx
        self.device_misfit = True  # use pyfwi_misfit_gradient where it applies

    # ------------------------------------------------------------ driver
    def __call__(self, m0, method, iter, freqs, n_params, k_0, k_end):
        """Invert from ``m0`` ({'vp','vs','rho'}) with `method` ('lbfgs'/2 or
        'cg'/1), iter[i] iterations at frequency freqs[i], over parameters
        k_0..k_end (1-based) in groups of n_params (fwi.py:109-158).  Returns
        (model dict, misfit history)."""
        m = self.dict2vec(m0)
        run = self._method(method)
        k_0 -= 1
        k_end -= 1
        if k_0 + n_params > k_end:
            raise Exception("k_0 + n_params can't be larger than k_end!!")
        rms = []
        for c, freq in enumerate(freqs):
            m, rms0 = run(m, iter[c], freq, n_params, k_0, k_end)
            rms.append(rms0)
        return self.vec2dict(m, self.nz, self.nx), np.array(rms)

    def _method(self, user_method):
        """fwi.py:160-172."""
        if user_method in [0, "SD", "sd"]:
            raise TypeError("Steepest descent is not provided yet.")
        if user_method in [1, "CG", "cg"]:
            return self.cg
        if user_method in [2, "lbfgs"]:
            return self.lbfgs
        raise ValueError(f"unknown optimisation method {user_method!r}")

    def _groups(self, m0, k0, k_end, n_params):
        n = self.nz * self.nx
        for k in np.arange(k0, k_end, n_params):
            print("Parameter number {} to {}".format(k + 1, k + n_params))
            yield k, n

    def cg(self, m0, ITER, freq, n_params=1, k0=1, k_end=2):
        """Nonlinear CG per parameter group (fwi.py:174-199)."""
        mtotal = np.copy(m0)
        hist_all = []
        fun = MemoizeJac(self.fprime_single)
        for k, n in self._groups(m0, k0, k_end, n_params):
            m_1, m_opt, m1 = (mtotal[:k * n], mtotal[k * n:(k + n_params) * n],
                              mtotal[(k + n_params) * n:])
            m_opt, hist, _, _, _ = fmin_cg(fun, m_opt, fun.derivative, args=(m_1, m1, freq),
                                           gtol=1e-8, maxiter=ITER, full_output=True, disp=False)
            hist_all.append(hist)
            mtotal = np.hstack((m_1, m_opt, m1))
        return mtotal, hist_all

    def lbfgs(self, m0, ITER, freq, n_params=1, k0=1, k_end=2):
        """L-BFGS-B per parameter group, memory 10, at most 10 line-search
        steps (fwi.py:201-227)."""
        mtotal = np.copy(m0)
        hist_all = []
        fun = MemoizeJac(self.fprime_single)
        for k, n in self._groups(m0, k0, k_end, n_params):
            m_1, m_opt, m1 = (mtotal[:k * n], mtotal[k * n:(k + n_params) * n],
                              mtotal[(k + n_params) * n:])
            m_opt, hist, _ = fmin_l_bfgs_b(fun, m_opt, fun.derivative, args=(m_1, m1, freq),
                                           m=10, factr=1e7, pgtol=1e-8, bounds=None,
                                           maxfun=15000, maxiter=ITER, callback=None, maxls=10)
            hist_all.append(hist)
            mtotal = np.hstack((m_1, m_opt, m1))
        return mtotal, hist_all

    # ------------------------------------------------------------ misfit and gradient
    def fprime(self, m0, freq):
        """Data misfit and its gradient over the whole model vector
        (fwi.py:229-252)."""
        mtotal = np.copy(m0)
        m_old = self.vec2dict(mtotal, self.nz, self.nx)
        m_new = self.model_to_dv(m_old, self.param_functions_args)
        if self._device_path_applies(freq):
            rms_data, grad_dv = self._misfit_gradient_device(m_new)
        else:
            d_est = self.forward_modeling(m_new, show=False)
            d_est = prepare_residual(d_est, self.sd)
            rms_data, adj_src = tools.cost_seismic(d_est, self.d_obs, fun=self.CF, fn=self.fn,
                                                   freq=freq, order=3, axis=1)
            grad_dv = self.gradient(adj_src, parameterization="dv")
        grad = self.grad_from_dv(grad_dv, self.param_functions_args, m_old)
        params = list(grad)
        for i in range(3):
            grad[params[i]] *= self.grad_coeff[i]
        return rms_data, self.dict2vec(grad)

    def fprime_single(self, m0, m_1, m1, freq):
        """Misfit (data + regularisation) and gradient of one parameter group
        m0 between the fixed parts m_1 and m1 (fwi.py:254-282)."""
        mtotal = np.float32(np.hstack((m_1, m0, m1)))
        shape_1, shape0 = np.shape(m_1)[0], np.shape(m0)[0]
        k0 = np.int32(shape_1 / self.n_elements)
        kend = np.int32(k0 + shape0 / self.n_elements)
        rms_data, grad_data = self.fprime(mtotal, freq)
        rms_reg, grad_reg = self.regularization.cost_regularization(
            m0, tv_properties=self.tv_properties, tikhonov_properties=self.tikhonov_properties)
        rms_rel, grad_rel = self.regularization.parameter_relation(
            mtotal, self.param_relation, k0, kend, freq)
        rms_mp, grad_mp = self.regularization.priori_regularization(
            m0, self.prior_model, k0, kend, freq)
        rms = rms_data + rms_reg + rms_rel + rms_mp
        grad = (grad_data[shape_1:shape_1 + shape0] + grad_reg
                + grad_rel[shape_1:shape_1 + shape0] + grad_mp)
        print(m0.min(), m0.max())
        print(" for f= {}: rms is: {} with rms_reg: {}, and rms_data: {}, rms_mp: {}, "
              "rms_model_relation: {}".format(freq, rms, rms_reg, rms_data, rms_mp, rms_rel))
        return rms, grad

    # ------------------------------------------------------------ one-call path
    def _device_path_applies(self, freq):
        return self.device_misfit and self.device_misfit_applies(self.d_obs, freq, self.sd,
                                                                 self.CF)

    def _misfit_gradient_device(self, m_new):
        return self.misfit_gradient_device(m_new, self.d_obs)
