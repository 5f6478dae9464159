"""Uneven shot batches: shots that differ in duration, receiver aperture and
wavelet, run as one batch of the C ABI.

PyFWI's ``WavePropagator`` (src/PyFWI/wave_propagation.py:617-623) holds one
recording time, one receiver line and one wavelet for all of its shots, so a
survey whose shots differ takes one propagator per geometry there, each run
one shot at a time.  ``UnevenBatch`` keeps every shot's own rules -- timing
(wave_propagation.py:28-35), geometry (:86-131) and wavelet
(acquisition.py:188-244) come from one ``WavePropagator`` per shot, which
never touches the GPU -- and hands all shots to one plan: one shot table, one
concatenated source table and one record buffer with per-shot offsets.  The
shots must share the grid (model shape, dh, npml, order) and the time step.

Records of shot k are ``[5][nt_ext_k][nr_k]`` (vx, vz, txx, tzz, txz) at
``record_offset(k)`` of the flat record buffer.
"""
from dataclasses import dataclass

import numpy as np

from . import _native
from .acquisition import Source
from .fwi_tools import expand_model
from .processing import COMPONENT_MASK
from .wave_propagation import WavePropagator, checkpoint_steps

__all__ = ["ShotSpec", "UnevenBatch"]


@dataclass
class ShotSpec:
    """One shot: explosive point source at ``src_loc`` (x, z in metres), a
    surface receiver line ``rec_loc`` [nr, 2] (metres, uniform spacing),
    recording time ``t`` (s) and a Ricker wavelet of dominant frequency
    ``fdom`` (Hz)."""
    src_loc: tuple
    rec_loc: np.ndarray
    t: float
    fdom: float


class UnevenBatch:
    def __init__(self, inpa, specs, model_shape, components=2):
        if not specs:
            raise ValueError("no shots")
        if inpa.get("acq_type", 1) != 1:
            raise ValueError("UnevenBatch takes surface shots (acq_type 1)")
        self.inpa = dict(inpa)
        self.specs = list(specs)
        self.components = components
        self.props = []
        dh, dt = float(inpa["dh"]), float(inpa["dt"])
        for sp in self.specs:
            rec = np.asarray(sp.rec_loc, np.float64)
            p = dict(inpa, t=float(sp.t))
            if rec.shape[0] > 1:
                p["rec_dis"] = float(rec[1, 0] - rec[0, 0])
            src = Source(np.array([sp.src_loc], np.float64), dh, dt)
            src.Ricker(sp.fdom)
            self.props.append(WavePropagator(p, src, rec, model_shape, components=components,
                                             chpr=100))
        P0 = self.props[0]
        for P in self.props[1:]:
            if (P.tnz, P.tnx, P.dt, P.dt_scale, P.npml, P.sdo) != \
                    (P0.tnz, P0.tnx, P0.dt, P0.dt_scale, P0.npml, P0.sdo):
                raise ValueError("shots of one batch must share the grid and the time step")
        self.base = P0
        self.tnz, self.tnx, self.npml = int(P0.tnz), int(P0.tnx), int(P0.npml)
        self.nt = [int(P.nt) for P in self.props]
        self.nt_ext = [int(P.nt_ext) for P in self.props]
        self.nr = [int(P.nr) for P in self.props]
        self.nt_max = max(self.nt)
        # shot table with offsets into the concatenated source table and records
        self._shot_tab = []
        tabs = []
        src_off = rec_off = 0
        self._rec_off = []
        for P in self.props:
            sh = P._shots([0])[0]
            tab = P._source_table()
            sh.src_offset = src_off
            sh.rec_offset = rec_off
            self._shot_tab.append(sh)
            self._rec_off.append(rec_off)
            tabs.append(tab.reshape(-1))
            src_off += tab.size
            rec_off += 5 * int(P.nt_ext) * int(P.nr)
        self.src_table = np.ascontiguousarray(np.concatenate(tabs), np.float32)
        self.n_records = rec_off

    # ------------------------------------------------------------ layout
    @property
    def plan(self):
        return self.base._plan

    @property
    def ns(self):
        return len(self.specs)

    def shots(self, ids=None):
        """Shot table entries (with their batch offsets) of `ids` (all)."""
        ids = range(self.ns) if ids is None else ids
        return [self._shot_tab[k] for k in ids]

    def record_offset(self, k):
        return self._rec_off[k]

    def records(self, flat, k):
        """View [5, nt_ext_k, nr_k] of shot k in a flat record buffer."""
        n = 5 * self.nt_ext[k] * self.nr[k]
        o = self._rec_off[k]
        return flat[o:o + n].reshape(5, self.nt_ext[k], self.nr[k])

    def cell_updates(self):
        """Sum over shots of tnz*tnx*nt_s."""
        return self.tnz * self.tnx * sum(self.nt)

    # ------------------------------------------------------------ model and runs
    def set_model(self, model0):
        """Upload the model ``{'vp','vs','rho': [nz, nx]}`` (the reference's
        forward_modeling preamble, wave_propagation.py:866-875)."""
        model = {k: expand_model(v, self.tnz, self.tnx, self.npml) for k, v in model0.items()}
        self.base.pml_preparation(model["vp"].max())
        self.base.elastic_buffers(model)

    def forward(self, seis=None, mem=_native.MEM_HOST, src_table=None):
        """Records of all shots into `seis` (flat, n_records floats).  With
        mem = MEM_DEVICE, `seis` and `src_table` (a device copy of
        ``self.src_table``) are device buffers."""
        if seis is None:
            seis = np.zeros(self.n_records, np.float32)
        src = self.src_table if src_table is None else src_table
        self.plan.forward(self.shots(), src, seis, np.zeros(0, np.int32), 0, mem)
        return seis

    def misfit_gradient(self, obs, grad=None, seis=None, flags=0, mem=_native.MEM_HOST,
                        src_table=None):
        """L2 misfit of the components selected by ``components`` against
        `obs` (flat records) and its reference-semantics gradient with every
        forward state kept (chpr = 100, each shot its own nt).  Returns
        (misfit, grad[6, tnz, tnx]) as pyfwi_misfit_gradient.  With mem =
        MEM_DEVICE all buffers, `src_table` included, are device buffers."""
        if grad is None:
            grad = np.zeros((6, self.tnz, self.tnx), np.float32)
        chp = checkpoint_steps(self.nt_max, 100)
        src = self.src_table if src_table is None else src_table
        m = self.plan.misfit_gradient(self.shots(), src, obs, seis, grad, chp,
                                      COMPONENT_MASK[self.components], flags, mem)
        return m, grad
