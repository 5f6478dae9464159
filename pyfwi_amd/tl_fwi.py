"""Time-lapse full-waveform inversion (src/PyFWI/tl_fwi.py:13-584).

``TimeLapse`` inverts a baseline and a monitor data set for the model change
between them, with the strategies of the reference:
- 'pa' parallel (independent baseline and monitor inversions),
- 'cc' cascaded (the monitor from the inverted baseline),
- 'dd' double difference (Asnaashari et al. 2011),
- 'sim' simultaneous (both models in one L-BFGS problem, coupled by a small
  model-difference penalty),
- 'wa' weighted-averaged cascaded, 'cu' cross updating (Maharramov and Biondi
  2014), 'cd' central difference and 'cj' cascaded joint.
Every misfit evaluation runs on the GPU through the ``FWI`` driver (or, for
the simultaneous strategy, ``tl_fprime``), with the one-call
pyfwi_misfit_gradient path where it applies.
"""
import copy
import timeit

import numpy as np
from scipy.optimize import fmin_l_bfgs_b
from scipy.optimize._optimize import MemoizeJac

from . import fwi_tools as tools
from .fwi import FWI
from .processing import prepare_residual
from .wave_propagation import WavePropagator

__all__ = ["TimeLapse"]


class TimeLapse(WavePropagator):
    """Parameters as the reference (tl_fwi.py:13-56): baseline and monitor
    observed data, ``inpa``, source, receivers, model shape, components,
    chpr, n_well_rec and optional ``param_functions`` (see ``FWI``)."""

    def __init__(self, b_dobs, m_dobs, inpa, src, rec_loc, model_shape, components, chpr,
                 n_well_rec=0, param_functions=None):
        super().__init__(inpa, src, rec_loc, model_shape, n_well_rec=n_well_rec, chpr=chpr,
                         components=components)
        self.inv_obj = FWI(b_dobs, inpa, src, rec_loc, model_shape=model_shape,
                           n_well_rec=n_well_rec, chpr=chpr, components=components,
                           param_functions=param_functions)
        self.b_dobs = copy.deepcopy(b_dobs)
        self.m_dobs = copy.deepcopy(m_dobs)
        self.n_elements = self.nz * self.nx
        self.CF = self.inv_obj.CF
        self.fn = self.inv_obj.fn
        self.device_misfit = True

    def __call__(self, b_m0, iter, freqs, tl_method, n_params, k_0, k_end):
        """Run the strategy `tl_method` from the initial model b_m0; returns
        (models and differences, misfit histories) (tl_fwi.py:58-112)."""
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        runs = {
            "pa": lambda: self.parallel_inversion(b_m0, freqs, iter, **kw),
            "cc": lambda: self.cascaded_inversion(b_m0, freqs, iter, **kw),
            "dd": lambda: self.double_difference_inversion(b_m0, freqs, iter, **kw),
            "sim": lambda: self.sim_inversion(b_m0, freqs, iter, data_to_invert=self.b_dobs, **kw),
            "wa": lambda: self.weighted_averaged_cascaded(b_m0, freqs, iter, **kw),
            "cj": lambda: self.cascaded_joint(b_m0, freqs, iter, **kw),
            "cd": lambda: self.central_difference(b_m0, freqs, iter, **kw),
            "cu": lambda: self.cross_updating(b_m0, freqs, iter, **kw),
        }
        if tl_method not in runs:
            raise ValueError(f"unknown time-lapse method {tl_method!r}")
        return runs[tl_method]()

    # ------------------------------------------------------------ helpers
    def _invert(self, m0, d_obs, freqs, iter, n_params, k_0, k_end):
        """One FWI (L-BFGS) of `d_obs` from m0."""
        self.inv_obj.d_obs = d_obs
        return self.inv_obj(m0, method=2, freqs=freqs, iter=iter, n_params=n_params, k_0=k_0,
                            k_end=k_end)

    def _vintage(self, model_pcs, d_obs, freq):
        """Data misfit and gradient vector of one vintage (tl_fwi.py:170-205)."""
        inv = self.inv_obj
        model = inv.model_to_dv(model_pcs, inv.param_functions_args)
        if self.device_misfit and self.device_misfit_applies(d_obs, freq, inv.sd, self.CF):
            rms, grad_dv = self.misfit_gradient_device(model, d_obs)
        else:
            dpre = prepare_residual(self.forward_modeling(model, show=False), inv.sd)
            rms, adj = tools.cost_seismic(d_pre0=dpre, d_obs0=d_obs, fun=self.CF, fn=self.fn,
                                          freq=freq, order=3, axis=1, sourc_weight=None,
                                          ns=self.ns, nr=self.nr)
            grad_dv = self.gradient(adj)
        grad = inv.grad_from_dv(grad_dv, inv.param_functions_args, model_pcs)
        return rms, inv.dict2vec(grad)

    # ------------------------------------------------------------ simultaneous
    def tl_lbfgs(self, x0, iter, freq, n_params, k0, k_end):
        """L-BFGS-B over [baseline, monitor] parameter groups, memory 30
        (tl_fwi.py:114-147).  k0, k_end are 1-based here, as in the
        reference."""
        n = self.nz * self.nx
        self.xrest = np.copy(x0)
        hist_all = []
        fun = MemoizeJac(self.tl_fprime)
        for k in np.arange(k0 - 1, k_end - 1, n_params):
            print("Parameter number {} to {}".format(k + 1, k + n_params))
            m_opt = np.hstack((x0[k * n:(k + n_params) * n],
                               x0[(3 + k) * n:3 * n + (k + n_params) * n]))
            m_opt, hist, _ = fmin_l_bfgs_b(fun, m_opt, fun.derivative, args=(freq, k, n_params),
                                           m=30, factr=1e-10, pgtol=1e-12, bounds=None,
                                           maxfun=15000, maxiter=iter, callback=None, maxls=20)
            print(m_opt.max(), m_opt.min())
            hist_all.append(hist)
            x0[k * n:(k + n_params) * n] = np.float32(m_opt[:n_params * n])
            x0[(3 + k) * n:(3 + k + n_params) * n] = np.float32(m_opt[n_params * n:])
        return x0, hist_all

    def tl_fprime(self, x, freq, k, n_params):
        """Misfit of both vintages plus gamma/2 ||m - b||^2 (gamma = 1e-10)
        and its gradient for parameter group k (tl_fwi.py:149-222).  The
        working vector is kept in ``xrest``."""
        n = self.n_elements
        x1 = self.xrest
        x1[k * n:(k + n_params) * n] = x[:n_params * n]
        x1[3 * n + k * n:3 * n + (k + n_params) * n] = x[n_params * n:2 * n_params * n]
        b_pcs = self.inv_obj.vec2dict(x1[:3 * n], self.nz, self.nx)
        m_pcs = self.inv_obj.vec2dict(x1[3 * n:], self.nz, self.nx)
        b_rms, b_grad = self._vintage(b_pcs, self.b_dobs, freq)
        b = x1[k * n:(k + n_params) * n]
        m = x1[(3 + k) * n:(3 + k + n_params) * n]
        diff = m - b
        rms_diff = 0.5 * np.dot(diff, diff)
        gamma = 1e-10
        m_rms, m_grad = self._vintage(m_pcs, self.m_dobs, freq)
        rms_data = b_rms + m_rms
        rms = rms_data + gamma * rms_diff
        print(k, np.abs(diff).min(), np.abs(diff).max())
        print("TL fprime_single_parameter for f= {}: rms is: {} for data: {} and model: {}"
              .format(freq, rms, rms_data, rms_diff))
        grad = np.hstack((b_grad[k * n:(k + n_params) * n] + gamma * (b - m),
                          m_grad[k * n:(k + n_params) * n] + gamma * (m - b)))
        return rms, np.float64(grad)

    def sim_inversion(self, b_m0, freqs, iter, n_params, k_0, k_end, data_to_invert=None):
        """Simultaneous inversion, optionally from a first inversion of the
        driver's current data (tl_fwi.py:505-552).  As in the reference,
        `data_to_invert` is stored but the first inversion uses the data the
        FWI driver holds."""
        tic = timeit.default_timer()
        if data_to_invert is not None:
            self.dobs = data_to_invert
            ref, _ = self.inv_obj(b_m0, method=2, freqs=freqs, iter=iter, n_params=n_params,
                                  k_0=k_0, k_end=k_end)
            b0, m0 = copy.deepcopy(ref), copy.deepcopy(ref)
        else:
            b0, m0 = copy.deepcopy(b_m0), copy.deepcopy(b_m0)
        x0 = np.hstack((self.inv_obj.dict2vec(b0), self.inv_obj.dict2vec(m0)))
        rms2 = []
        for c, freq in enumerate(freqs):
            self.restForTest = x0
            x0, hist1 = self.tl_lbfgs(x0, iter[c], freq, n_params, k_0, k_end)
            rms2.append(hist1[0])
        models = self.ins_TL_output(x0, False)
        inverted = {"dm": self.ins_TL_output(x0, True), "b1": models["model_b"],
                    "m1": models["model_m"], "pt": timeit.default_timer() - tic}
        return inverted, {"sim": rms2}

    def ins_TL_output(self, xk, diff=True):
        """[baseline, monitor] vector -> their difference (diff) or both
        models as dicts (tl_fwi.py:554-584)."""
        half = int(xk.shape[0] / 2)
        if diff:
            return self.inv_obj.vec2dict(xk[half:] - xk[:half], self.nz, self.nx)
        return {"model_b": self.inv_obj.vec2dict(xk[:half], self.nz, self.nx),
                "model_m": self.inv_obj.vec2dict(xk[half:2 * half], self.nz, self.nx)}

    # ------------------------------------------------------------ sequential strategies
    def parallel_inversion(self, b_m0, freqs, iter, n_params, k_0, k_end):
        """Baseline and monitor each from b_m0 (tl_fwi.py:402-428)."""
        tic = timeit.default_timer()
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        b1, rms1 = self._invert(b_m0, self.b_dobs, freqs, iter, **kw)
        print("===== Inversion of baseline is done ======")
        m1, rms2 = self._invert(b_m0, self.m_dobs, freqs, iter, **kw)
        inverted = {"b1": b1, "m1": m1, "dm": tools.dict_diff(m1, b1, positivity=False),
                    "pt": timeit.default_timer() - tic}
        print("Parallel")
        return inverted, {"b1": rms1, "m1": rms2}

    def cascaded_inversion(self, b_m0, freqs, iter, n_params, k_0, k_end):
        """The monitor from the inverted baseline (tl_fwi.py:471-503)."""
        tic = timeit.default_timer()
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        b1, rms1 = self._invert(b_m0, self.b_dobs, freqs, iter, **kw)
        initial = b1.copy()
        print("===== Inversion of baseline is done ======")
        m1, rms2 = self._invert(b1, self.m_dobs, freqs, iter, **kw)
        inverted = {"b1": b1, "m1": m1, "dm": tools.dict_diff(m1, initial, positivity=False),
                    "pt": timeit.default_timer() - tic}
        print("Cascaded")
        return inverted, {"b1": rms1, "m1": rms2}

    def double_difference_inversion(self, b_m0, freqs, iter, n_params, k_0, k_end):
        """Invert the composite data d_ref + (d_monitor - d_baseline), with
        d_ref modelled in the inverted baseline (Asnaashari et al. 2011;
        tl_fwi.py:430-469)."""
        tic = timeit.default_timer()
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        b1, rms1 = self._invert(b_m0, self.b_dobs, freqs, iter, **kw)
        print("===== Inversion of baseline is done ======")
        d_ref = prepare_residual(self.forward_modeling(b1, show=False), s=self.inv_obj.sd)
        d_comp = tools.dict_summation(tools.dict_diff(self.m_dobs, self.b_dobs), d_ref)
        composite, rms2 = self._invert(b_m0.copy(), d_comp, freqs, iter, **kw)
        inverted = {"b1": b1, "m1": composite,
                    "dm": tools.dict_diff(composite, b1, positivity=False),
                    "pt": timeit.default_timer() - tic}
        print("double_difference_inversion")
        return inverted, {"b1": rms1, "m1": rms2}

    def weighted_averaged_cascaded(self, b_m0, freqs, iter, n_params, k_0, k_end):
        """Baseline, monitor from it, baseline again from the monitor
        (tl_fwi.py:224-255)."""
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        print("*** First inversion ***")
        b1, rms_b1 = self._invert(copy.deepcopy(b_m0), self.b_dobs, freqs, iter, **kw)
        print("*** Second inversion ***")
        m1, rms_m1 = self._invert(b1, self.m_dobs, freqs, iter, **kw)
        print("*** Third inversion ***")
        b2, rms_b2 = self._invert(m1, self.b_dobs, freqs, iter, **kw)
        inverted = {"dm11": tools.dict_diff(m1, b1), "dm12": tools.dict_diff(m1, b2),
                    "m1": m1, "b1": b1, "b2": b2}
        print("Weighted average")
        return inverted, {"b1": rms_b1, "m1": rms_m1, "b2": rms_b2}

    def cross_updating(self, b_m0, freqs, iter, n_params, k_0, k_end):
        """Four alternating inversions (Maharramov and Biondi 2014;
        tl_fwi.py:257-305)."""
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        tic = timeit.default_timer()
        print("*** First inversion ***")
        b1, rms_b1 = self._invert(copy.deepcopy(b_m0), self.b_dobs, freqs, iter, **kw)
        print("*** Second inversion ***")
        m1, rms_m1 = self._invert(b1, self.m_dobs, freqs, iter, **kw)
        t_cc = timeit.default_timer() - tic
        print("*** Third inversion ***")
        b2, rms_b2 = self._invert(m1, self.b_dobs, freqs, iter, **kw)
        t_wa = timeit.default_timer() - tic
        print("*** Fourth inversion ***")
        m2, rms_m2 = self._invert(b2, self.m_dobs, freqs, iter, **kw)
        t_cu = timeit.default_timer() - tic
        inverted = {"dm11": tools.dict_diff(m1, b1), "dm12": tools.dict_diff(m1, b2),
                    "dm21": tools.dict_diff(m2, b1), "dm22": tools.dict_diff(m2, b2),
                    "b1": b1, "m1": m1, "b2": b2, "m2": m2,
                    "pt_cc": t_cc, "pt_wa": t_wa, "pt_cu": t_cu}
        print("cross_updating")
        return inverted, {"b1": rms_b1, "m1": rms_m1, "b2": rms_b2, "m2": rms_m2}

    def central_difference(self, b_m0, freqs, iter, n_params, k_0, k_end):
        """Baseline then monitor, and monitor then baseline
        (tl_fwi.py:307-353)."""
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        m0 = copy.deepcopy(b_m0)
        tic = timeit.default_timer()
        print("*** Central time lapse *** ")
        b1, rms_b1 = self._invert(m0, self.b_dobs, freqs, iter, **kw)
        print("*** Second INVERSION ***")
        b1m, rms_b1m = self._invert(b1, self.m_dobs, freqs, iter, **kw)
        print("*** Third INVERSION ***")
        m1, rms_m1 = self._invert(m0, self.m_dobs, freqs, iter, **kw)
        print("*** Fourth INVERSION ***")
        m1b, rms_m1b = self._invert(m1, self.b_dobs, freqs, iter, **kw)
        inverted = {"dm0": tools.dict_diff(m1, b1), "dm1": tools.dict_diff(b1m, m1b),
                    "dm3": tools.dict_diff(b1m, b1), "dm4": tools.dict_diff(m1, m1b),
                    "dm34": tools.dict_summation(tools.dict_diff(b1m, b1),
                                                 tools.dict_diff(m1, m1b), division=2),
                    "b1": b1, "m1": m1, "m2": b1m, "b2": m1b,
                    "pt": timeit.default_timer() - tic}
        print("central difference")
        return inverted, {"b1": rms_b1, "m1": rms_m1, "m2": rms_b1m, "b2": rms_m1b}

    def cascaded_joint(self, m0, freqs, iter, n_params, k_0, k_end):
        """Two simultaneous inversions, the second from the average of the
        first's models (tl_fwi.py:355-400)."""
        kw = dict(n_params=n_params, k_0=k_0, k_end=k_end)
        tic = timeit.default_timer()
        print("*** FIRST INVERSION ***")
        r1, rms1 = self.sim_inversion(m0, freqs, iter=iter, data_to_invert=None, **kw)
        b1, m1 = r1["b1"], r1["m1"]
        ave = tools.dict_summation(b1, m1, division=2)
        print("*** SECOND INVERSION ***")
        r2, rms2 = self.sim_inversion(ave, freqs, iter=iter, data_to_invert=None, **kw)
        b1b, b1m = r2["b1"], r2["m1"]
        inverted = {"dm0": tools.dict_diff(m1, b1), "dm7": tools.dict_diff(b1m, b1b),
                    "dm3": tools.dict_diff(b1m, b1), "dm6": tools.dict_diff(m1, b1b),
                    "b1": b1, "m1": m1, "b2": b1b, "m2": b1m,
                    "pt": timeit.default_timer() - tic}
        print("joint_cascaded")
        return inverted, {"sim1": rms1, "sim2": rms2}
