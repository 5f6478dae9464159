"""Time-lapse full-waveform inversion with the interface of PyFWI's
``TimeLapse`` (src/PyFWI/tl_fwi.py:13-584): same constructor, ``__call__``
signature, strategy codes, result and history keys, and optimiser settings.

Every strategy is a row of ``STRATEGIES``: a list of inversion legs and the
outputs built from their models.  A leg inverts one data set (the baseline,
the monitor, or the double-difference composite) from a start model (the
initial model, an earlier leg's result, or the mean of two), or runs the
joint problem over both vintages.  ``TimeLapse.run_plan`` executes a row:
legs in order, the clock read where a row asks for it, then the outputs.

  'pa' parallel          b1 <- m0 | base;  m1 <- m0 | monitor
  'cc' cascaded          b1 <- m0 | base;  m1 <- b1 | monitor
  'dd' double difference b1 <- m0 | base;  m1 <- m0 | monitor - base + F(b1)
                         (Asnaashari et al. 2011)
  'wa' weighted average  b1, m1 <- b1 | monitor, b2 <- m1 | base
  'cu' cross updating    as 'wa', then m2 <- b2 | monitor (Maharramov and
                         Biondi 2014)
  'cd' central diff.     b1 <- m0 | base, m2 <- b1 | monitor,
                         m1 <- m0 | monitor, b2 <- m1 | base
  'sim' simultaneous     the joint problem from m0 (after an inversion of the
                         driver's current data, as the reference does)
  'cj' cascaded joint    joint from m0, then joint from the mean of its two
                         models

The joint problem (tl_fwi.py:114-222) minimises both vintages' data misfits
plus gamma/2 ||m - b||^2 over one window of parameters of both models at once
(L-BFGS-B, memory 30).  Every misfit evaluation runs on the GPU through
``FWI.data_term`` (the one-call pyfwi_misfit_gradient path where it applies).
"""
import copy
import timeit
from dataclasses import dataclass, field

import numpy as np
from scipy.optimize import fmin_l_bfgs_b

from . import fwi_tools as tools
from .fwi import FWI, Window
from .processing import prepare_residual
from .wave_propagation import WavePropagator

__all__ = ["TimeLapse", "STRATEGIES", "Leg", "Plan"]

GAMMA = 1e-10  # weight of the model-difference coupling of the joint problem


@dataclass(frozen=True)
class Leg:
    """One inversion: `name` <- inversion of `data` ('base', 'monitor',
    'composite' or 'current': whatever the driver holds) from `start` ('m0',
    a leg name, or a pair of names for their mean).  kind 'joint' runs the
    joint problem and yields `name`_b and `name`_m."""
    name: str
    start: object
    data: str = "base"
    kind: str = "fwi"


@dataclass
class Plan:
    legs: list
    # output key -> ('model', leg) | ('diff', a, b) | ('mean_diff', (a, b), (c, d))
    outputs: dict
    # history key -> leg name, or ('sim', leg) for {'sim': history of leg}
    hist: dict
    clocks: dict = field(default_factory=dict)  # output key -> leg after which to read
    banner: str = ""


STRATEGIES = {
    "pa": Plan([Leg("b1", "m0", "base"), Leg("m1", "m0", "monitor")],
               {"b1": ("model", "b1"), "m1": ("model", "m1"), "dm": ("diff", "m1", "b1")},
               {"b1": "b1", "m1": "m1"}, {"pt": "m1"}, "parallel"),
    "cc": Plan([Leg("b1", "m0", "base"), Leg("m1", "b1", "monitor")],
               {"b1": ("model", "b1"), "m1": ("model", "m1"), "dm": ("diff", "m1", "b1")},
               {"b1": "b1", "m1": "m1"}, {"pt": "m1"}, "cascaded"),
    "dd": Plan([Leg("b1", "m0", "base"), Leg("m1", "m0", "composite")],
               {"b1": ("model", "b1"), "m1": ("model", "m1"), "dm": ("diff", "m1", "b1")},
               {"b1": "b1", "m1": "m1"}, {"pt": "m1"}, "double difference"),
    "wa": Plan([Leg("b1", "m0", "base"), Leg("m1", "b1", "monitor"), Leg("b2", "m1", "base")],
               {"dm11": ("diff", "m1", "b1"), "dm12": ("diff", "m1", "b2"),
                "m1": ("model", "m1"), "b1": ("model", "b1"), "b2": ("model", "b2")},
               {"b1": "b1", "m1": "m1", "b2": "b2"}, {}, "weighted average"),
    "cu": Plan([Leg("b1", "m0", "base"), Leg("m1", "b1", "monitor"), Leg("b2", "m1", "base"),
                Leg("m2", "b2", "monitor")],
               {"dm11": ("diff", "m1", "b1"), "dm12": ("diff", "m1", "b2"),
                "dm21": ("diff", "m2", "b1"), "dm22": ("diff", "m2", "b2"),
                "b1": ("model", "b1"), "m1": ("model", "m1"), "b2": ("model", "b2"),
                "m2": ("model", "m2")},
               {"b1": "b1", "m1": "m1", "b2": "b2", "m2": "m2"},
               {"pt_cc": "m1", "pt_wa": "b2", "pt_cu": "m2"}, "cross updating"),
    "cd": Plan([Leg("b1", "m0", "base"), Leg("b1m", "b1", "monitor"),
                Leg("m1", "m0", "monitor"), Leg("m1b", "m1", "base")],
               {"dm0": ("diff", "m1", "b1"), "dm1": ("diff", "b1m", "m1b"),
                "dm3": ("diff", "b1m", "b1"), "dm4": ("diff", "m1", "m1b"),
                "dm34": ("mean_diff", ("b1m", "b1"), ("m1", "m1b")),
                "b1": ("model", "b1"), "m1": ("model", "m1"), "m2": ("model", "b1m"),
                "b2": ("model", "m1b")},
               {"b1": "b1", "m1": "m1", "m2": "b1m", "b2": "m1b"}, {"pt": "m1b"},
               "central difference"),
    "sim": Plan([Leg("ref", "m0", "current"), Leg("j", "ref", kind="joint")],
                {"dm": ("diff", "j_m", "j_b"), "b1": ("model", "j_b"), "m1": ("model", "j_m")},
                {"sim": "j"}, {"pt": "j"}, "simultaneous"),
    "cj": Plan([Leg("j1", "m0", kind="joint"), Leg("j2", ("j1_b", "j1_m"), kind="joint")],
               {"dm0": ("diff", "j1_m", "j1_b"), "dm7": ("diff", "j2_m", "j2_b"),
                "dm3": ("diff", "j2_m", "j1_b"), "dm6": ("diff", "j1_m", "j2_b"),
                "b1": ("model", "j1_b"), "m1": ("model", "j1_m"), "b2": ("model", "j2_b"),
                "m2": ("model", "j2_m")},
               {"sim1": ("sim", "j1"), "sim2": ("sim", "j2")}, {"pt": "j2"},
               "cascaded joint"),
}


class JointObjective:
    """Both vintages' data misfits plus GAMMA/2 ||m_w - b_w||^2 over one
    window w of the two stacked model vectors x = [b; m] (tl_fwi.py:150-222):
    the unknowns are [b_w; m_w].  The last point's value and gradient are
    kept for the optimiser's separate f and f' calls."""

    def __init__(self, tl, x, window, freq):
        self.tl, self.x, self.w, self.freq = tl, np.array(x), window, freq
        self._p = None
        self._fg = None

    def _eval(self, p):
        if self._p is None or not np.array_equal(p, self._p):
            self._p = np.array(p, copy=True)
            self._fg = self.evaluate(self._p)
        return self._fg

    def value(self, p):
        return self._eval(p)[0]

    def gradient(self, p):
        return self._eval(p)[1]

    def split(self, p):
        """Full model vectors of both vintages with the window set from p."""
        n3 = 3 * self.w.cells
        k = self.w.count * self.w.cells
        b = self.w.put(self.x[:n3], p[:k])
        m = self.w.put(self.x[n3:], p[k:2 * k])
        return b, m

    def evaluate(self, p):
        tl, w = self.tl, self.w
        inv = tl.inv_obj
        b, m = self.split(p)
        grads, rms_data = [], 0.0
        for vec, d_obs in ((b, tl.b_dobs), (m, tl.m_dobs)):
            pcs = inv.vec2dict(vec, tl.nz, tl.nx)
            rms, g_dv = tl.data_term(inv.model_to_dv(pcs, inv.param_functions_args), d_obs,
                                     self.freq)
            grads.append(w.take(inv.dict2vec(inv.grad_from_dv(g_dv, inv.param_functions_args,
                                                              pcs))))
            rms_data += rms
        diff = w.take(m).astype(np.float64) - w.take(b)
        total = rms_data + GAMMA * 0.5 * float(diff @ diff)
        tl.xrest = np.concatenate((b, m))
        if tl.verbose:
            print(f"TL joint band {self.freq}, window {w.first + 1}: misfit {total:.6e} "
                  f"(data {rms_data:.6e}); |m - b| in [{np.abs(diff).min():.4g}, "
                  f"{np.abs(diff).max():.4g}]", flush=True)
        grad = np.concatenate((grads[0] - GAMMA * diff, grads[1] + GAMMA * diff))
        return total, np.float64(grad)


class TimeLapse(WavePropagator):
    """Baseline and monitor observed data, ``inpa``, source, receivers, model
    shape, components, chpr, n_well_rec and optional ``param_functions``
    (see ``FWI``)."""

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
        self.verbose = True
        self.xrest = None  # both vintages at the joint problem's last evaluation

    def __call__(self, b_m0, iter, freqs, tl_method, n_params, k_0, k_end):
        """Run strategy `tl_method` (see STRATEGIES) from the initial model
        b_m0; returns (models, differences and timings; misfit histories)."""
        if tl_method not in STRATEGIES:
            raise ValueError(f"unknown time-lapse method {tl_method!r}")
        return self.run_plan(STRATEGIES[tl_method], b_m0, freqs, iter, n_params, k_0, k_end)

    # ------------------------------------------------------------ plans
    def run_plan(self, plan, m0, freqs, iter, n_params, k_0, k_end):
        models = {"m0": copy.deepcopy(m0)}
        hist, clock = {}, {}
        tic = timeit.default_timer()
        for leg in plan.legs:
            start = self._start_model(models, leg.start)
            if self.verbose:
                print(f"TL {plan.banner}: leg {leg.name} ({leg.kind}, data {leg.data})",
                      flush=True)
            if leg.kind == "joint":
                b, m, h = self.joint_inversion(start, freqs, iter, n_params, k_0, k_end)
                models[leg.name + "_b"], models[leg.name + "_m"] = b, m
                hist[leg.name] = h
            else:
                models[leg.name], hist[leg.name] = self._invert(
                    start, self._data(leg.data, models), freqs, iter, n_params, k_0, k_end)
            clock[leg.name] = timeit.default_timer() - tic
        out = {}
        for key, spec in plan.outputs.items():
            if spec[0] == "model":
                out[key] = models[spec[1]]
            elif spec[0] == "diff":
                out[key] = tools.dict_diff(models[spec[1]], models[spec[2]])
            else:  # mean of two differences
                (a, b), (c, d) = spec[1], spec[2]
                out[key] = tools.dict_summation(tools.dict_diff(models[a], models[b]),
                                                tools.dict_diff(models[c], models[d]),
                                                division=2)
        for key, leg in plan.clocks.items():
            out[key] = clock[leg]
        hists = {}
        for key, leg in plan.hist.items():
            hists[key] = {leg[0]: hist[leg[1]]} if isinstance(leg, tuple) else hist[leg]
        return out, hists

    def _start_model(self, models, start):
        if isinstance(start, tuple):
            return tools.dict_summation(models[start[0]], models[start[1]], division=2)
        return copy.deepcopy(models[start])

    def _data(self, which, models):
        if which == "base":
            return self.b_dobs
        if which == "monitor":
            return self.m_dobs
        if which == "current":
            return self.inv_obj.d_obs
        # double difference: the monitor - baseline difference added to the
        # data of the inverted baseline (tl_fwi.py:429-433)
        d_ref = prepare_residual(self.forward_modeling(models["b1"], show=False),
                                 s=self.inv_obj.sd)
        return tools.dict_summation(tools.dict_diff(self.m_dobs, self.b_dobs), d_ref)

    def _invert(self, m0, d_obs, freqs, iter, n_params, k_0, k_end):
        """One L-BFGS inversion of `d_obs` from m0 with the FWI driver."""
        self.inv_obj.d_obs = d_obs
        return self.inv_obj(m0, method=2, freqs=freqs, iter=iter, n_params=n_params, k_0=k_0,
                            k_end=k_end)

    # ------------------------------------------------------------ joint problem
    def data_term(self, model_dv, d_obs, freq):
        """One vintage's data misfit and gradient dict, on this object's own
        propagator (as the reference's tl_fprime)."""
        inv = self.inv_obj
        if self.device_misfit and self.device_misfit_applies(d_obs, freq, inv.sd, self.CF):
            return self.misfit_gradient_device(model_dv, d_obs, freq=freq)
        dpre = prepare_residual(self.forward_modeling(model_dv, show=False), inv.sd)
        rms, adj = tools.cost_seismic(d_pre0=dpre, d_obs0=d_obs, fun=self.CF, fn=self.fn,
                                      freq=freq, order=3, axis=1, sourc_weight=None, ns=self.ns,
                                      nr=self.nr)
        return rms, self.gradient(adj)

    def joint_inversion(self, m0, freqs, iter, n_params, k_0, k_end):
        """Both vintages from m0 by the joint problem, band after band;
        returns (baseline, monitor, misfit of the first window per band)."""
        vec = np.asarray(self.inv_obj.dict2vec(m0))
        x = np.concatenate((vec, vec))
        per_band = []
        for band, freq in enumerate(freqs):
            x, misfits = self.tl_lbfgs(x, iter[band], freq, n_params, k_0, k_end)
            per_band.append(misfits[0])
        n3 = 3 * self.n_elements
        v2d = self.inv_obj.vec2dict
        return v2d(x[:n3], self.nz, self.nx), v2d(x[n3:], self.nz, self.nx), per_band

    def tl_lbfgs(self, x0, iter, freq, n_params, k0, k_end):
        """One band of the joint problem: L-BFGS-B with memory 30 over the
        windows k0 .. k_end - 1 (1-based) of both vintages (tl_fwi.py:114-148).
        Returns (x, misfit per window)."""
        x = np.array(x0, np.float64)
        misfits = []
        for w in Window.sweep(k0 - 1, k_end - 1, n_params, self.n_elements):
            obj = JointObjective(self, x, w, freq)
            n3 = 3 * self.n_elements
            p0 = np.concatenate((w.take(x[:n3]), w.take(x[n3:])))
            p, fmin, _ = fmin_l_bfgs_b(obj.value, p0, obj.gradient, m=30, factr=1e-10,
                                       pgtol=1e-12, bounds=None, maxfun=15000, maxiter=iter,
                                       maxls=20)
            b, m = obj.split(np.float32(p))
            x = np.concatenate((b, m)).astype(np.float64)
            misfits.append(fmin)
        return x, misfits

    def tl_fprime(self, x, freq, k, n_params):
        """Joint misfit and gradient at the window values x = [b_w; m_w] of
        window k (0-based) with the rest of both models from ``xrest``
        (tl_fwi.py:150-222)."""
        obj = JointObjective(self, self.xrest, Window(int(k), int(n_params), self.n_elements),
                             freq)
        return obj.evaluate(np.asarray(x))

    def sim_inversion(self, b_m0, freqs, iter, n_params, k_0, k_end, data_to_invert=None):
        """The simultaneous strategy; with ``data_to_invert`` the joint problem
        starts from an inversion of the driver's current data, which the
        reference runs without installing ``data_to_invert`` (tl_fwi.py:
        476-491)."""
        sim = STRATEGIES["sim"]
        if data_to_invert is None:
            plan = Plan([Leg("j", "m0", kind="joint")], sim.outputs, sim.hist, sim.clocks,
                        sim.banner)
        else:
            self.dobs = data_to_invert
            plan = sim
        return self.run_plan(plan, b_m0, freqs, iter, n_params, k_0, k_end)

    def ins_TL_output(self, xk, diff=True):
        """Model dicts of a stacked [b; m] vector: the difference m - b, or
        {'model_b', 'model_m'} (tl_fwi.py:555-584)."""
        half = int(xk.shape[0] / 2)
        v2d = self.inv_obj.vec2dict
        if diff:
            return v2d(xk[half:] - xk[:half], self.nz, self.nx)
        return {"model_b": v2d(xk[:half], self.nz, self.nx),
                "model_m": v2d(xk[half:2 * half], self.nz, self.nx)}
