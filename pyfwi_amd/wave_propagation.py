"""Drop-in ``WavePropagator`` backed by the gfx950 HIP propagator.

Mirror of PyFWI's ``PyFWI.wave_propagation`` (src/PyFWI/wave_propagation.py):
same constructor, ``forward_modeling`` and ``gradient`` signatures, same
attributes (``nt``, ``dt``, ``W``, ``seismogram``, ``lam``, ``mu``, ...).  The
OpenCL program build, buffers and per-step kernel enqueues of the reference
are replaced by one native plan (libpyfwi_hip.so) that runs all shots of a
call as one batch on the GPU.

Supported: surface, crosswell and wells + surface acquisition (``acq_type``
1, 0, 2), orders 4 and 8, every ``pml_dir``, ``components`` 0-4,
``'2d'``/``'3d'`` records, checkpoint ratios ``chpr`` 0-100,
``energy_balancing``, ``g_smooth``.
"""
import math

import numpy as np

from . import _native
from .acquisition import seismic_section
from .fwi_tools import (Fdm, expand_model, fd_coefficients, kernel_constant,
                        pml_counstruction, pml_profiles_1d)
from .grad_switcher import grad_lmd_to_vd
from .processing import COMPONENT_MASK, COMPONENTS, prepare_residual

__all__ = ["WavePreparation", "WavePropagator"]


def checkpoint_steps(nt, chpr):
    """Steps after which the forward state is kept
    (src/PyFWI/wave_propagation.py:95-100): ``int(chpr*nt/100)`` points on
    ``linspace(0, nt-1)``; fewer than two points become ``[1, nt-1]``."""
    n = int(chpr * nt / 100)
    chp = np.linspace(0, nt - 1, n, dtype=np.int32)
    if len(chp) < 2:
        chp = np.array([1, nt - 1], dtype=np.int32)
    return chp


class WavePreparation:
    """Geometry, timing and device plan (src/PyFWI/wave_propagation.py:16-591)."""

    def __init__(self, inpa, src, rec_loc, model_shape, components=0,
                 n_well_rec=0, chpr=10):
        keys = list(inpa)
        # timing (wave_propagation.py:28-35)
        self.dt_scale = np.ceil(inpa["dt"] / 0.0006)
        self.t = inpa["t"]
        self.dt_ext = inpa["dt"]
        self.dt = inpa["dt"] / self.dt_scale
        self.fn = 1.0 / 2.0 / 0.0006
        self.nt = int(1 + self.t // self.dt)
        self.nt_ext = int(1 + self.t // self.dt_ext)

        self.nx = np.int32(model_shape[1])
        self.nz = np.int32(model_shape[0])
        self.seismo_shape = inpa.get("seimogram_shape", "2d")
        self.g_smooth = inpa.get("g_smooth", 0)
        if "npml" in keys:
            self.npml = inpa["npml"]
            pmlR = inpa["pmlR"]
            pml_dir = inpa["pml_dir"]
        else:
            self.npml, pmlR, pml_dir = 0, 0, 2
        self.sdo = np.int32(inpa["sdo"] / 2) if "sdo" in keys else 2
        if "rec_dis" in keys:
            rec_dis = inpa["rec_dis"]
        else:
            rec_dis = (rec_loc[1, :] - rec_loc[0, :]).max()
        self.acq_type = inpa.get("acq_type", 1)
        self.energy_balancing = inpa.get("energy_balancing", False)
        if self.acq_type not in (0, 1, 2):
            raise ValueError("acq_type must be 0 (crosswell), 1 (surface) or 2 (both)")

        self.tnx = np.int32(self.nx + 2 * self.npml)
        self.tnz = np.int32(self.nz + 2 * self.npml)
        self.dh = np.float32(inpa["dh"])

        self.srcx = np.int32(src.i + self.npml)
        self.srcz = np.int32(src.j + self.npml)
        self.src = src
        self.ns = np.int32(src.i.size)
        # NumPy 1.x forms this quotient in float64
        self.dxr = np.int32(float(rec_dis) / float(self.dh))
        if self.dxr < 1:
            raise ValueError("receiver spacing must be at least one grid cell")

        self.chp = checkpoint_steps(self.nt, chpr)
        self.nchp = len(self.chp)
        self.rec_loc = rec_loc
        self.nr = rec_loc.shape[0]
        self.n_well_rec = n_well_rec
        self.n_surface_rec = self.nr - 2 * n_well_rec
        if n_well_rec == 0 and self.acq_type == 2:
            raise Exception(" Number of geophons in the wells is not defined")

        def cell(k):  # (column, row) of receiver k, absorbing layer included
            return (np.int32(float(rec_loc[k, 0]) / float(self.dh) + self.npml),
                    np.int32(float(rec_loc[k, 1]) / float(self.dh) + self.npml))

        # receiver macros (wave_propagation.py:116-143)
        zero = np.int32(0)
        self.rec_top_left_const = self.rec_top_left_var = zero
        self.rec_top_right_const = self.rec_top_right_var = zero
        self.rec_surface_const = self.rec_surface_var = zero
        if self.acq_type == 0:
            self.rec_top_right_const, self.rec_top_right_var = cell(0)
            self.src_cts = src.i[0]
        elif self.acq_type == 1:
            self.rec_surface_var, self.rec_surface_const = cell(0)
            self.src_cts = src.j[0]
        else:
            self.rec_top_right_const, self.rec_top_right_var = cell(self.nr - n_well_rec)
            self.rec_surface_var, self.rec_surface_const = cell(n_well_rec)
            self.rec_top_left_const, self.rec_top_left_var = cell(0)
            self.src_cts = src.j[0]

        self._pml_dir = pml_dir
        self._pmlR = pmlR
        self.dx_pml_1d, self.dz_pml_1d = pml_profiles_1d(
            int(self.tnz), int(self.tnx), self.dh, self.npml, pmlR, pml_dir)
        self.D = Fdm(order=int(self.sdo) * 2)
        self.components = components

        self._rec_table = None
        if self.acq_type == 1:
            # receivers the reference's gather reaches: j % dxr == 0 for j < tnx
            # (src/PyFWI/elastic_surface.cl:30-32)
            n_live = min(int(self.n_surface_rec), -(-int(self.tnx) // int(self.dxr)))
            last = int(self.rec_surface_var) + (n_live - 1) * int(self.dxr)
            if n_live > 0 and (last >= self.tnx or self.rec_surface_const >= self.tnz
                               or self.rec_surface_var < 0 or self.rec_surface_const < 0):
                raise ValueError("receiver line leaves the grid")
            self._n_live = n_live
        else:
            self._n_live = int(self.nr)
            self._rec_table = self.receiver_table()

        grid = _native.Grid()
        grid.tnz, grid.tnx = int(self.tnz), int(self.tnx)
        grid.npml, grid.sdo = int(self.npml), int(self.sdo)
        c = fd_coefficients(int(self.sdo) * 2)
        for k in range(4):
            grid.c[k] = kernel_constant(c[k])
        grid.dt = kernel_constant(self.dt)
        grid.dx = grid.dz = kernel_constant(self.dh)
        self._grid = grid
        self._plan_obj = None
        self._seis = None
        self._forward_done = False
        self.per_shot = bool(inpa.get("per_shot", False))

        self.seismogram = {f: np.zeros((self.nt_ext, self.nr, self.ns), np.float32)
                           for f in COMPONENTS}

    def receiver_table(self):
        """Receiver cells ``[2, nr]`` (row*tnx + column, -1 where no gather
        reaches) of a well acquisition: row 0 the cell each record column
        samples, row 1 the cell each residual column is injected at.

        The gathers (elastic_crosswell.cl:24-31; elastic.cl:62-98, 316-363)
        test ``i % dxr == 0`` over the rows i < tnz (wells) or ``j % dxr == 0``
        over the columns j < tnx (surface); receiver ir = i/dxr of a well
        sits at row rec_top_*_var + i of column rec_top_*_const, column
        order left well, surface, right well.  Two reference rules shape the
        table:
          - forward_propagator reverses the left well's columns of the records
            (wave_propagation.py:647-649), while gradient injects residual
            column k at left-well receiver k (:911-913 reverse only the stored
            records), so row 0 holds the left well reversed and row 1 not;
          - the crosswell and right-well gathers do not bound ir: columns past
            nr land in the next record row, which the next step overwrites,
            and the adjoint reads past the residual row.  Only the nr columns
            are kept here.
        """
        tnz, tnx, dxr, nr = int(self.tnz), int(self.tnx), int(self.dxr), int(self.nr)
        cells = np.full(nr, -1, np.int64)

        def well(first, count, col, row0):
            ir = np.arange(min(count, -(-tnz // dxr), max(nr - first, 0)))
            cells[first + ir] = (ir * dxr + int(row0)) * tnx + int(col)

        def surface(first, count):
            ir = np.arange(min(count, -(-tnx // dxr), max(nr - first, 0)))
            cells[first + ir] = int(self.rec_surface_const) * tnx + ir * dxr \
                + int(self.rec_surface_var)

        nw, nsurf = int(self.n_well_rec), int(self.n_surface_rec)
        if self.acq_type == 0:
            well(0, nr, self.rec_top_right_const, self.rec_top_right_var)
        else:
            well(0, nw, self.rec_top_left_const, self.rec_top_left_var)
            surface(nw, nsurf)
            well(nw + nsurf, nw, self.rec_top_right_const, self.rec_top_right_var)
        if np.any((cells != -1) & ((cells < 0) | (cells >= tnz * tnx))):
            raise ValueError("a receiver lies outside the grid")
        record = cells.copy()
        if self.acq_type == 2:
            record[:nw] = cells[:nw][::-1]
        return np.stack([record, cells]).astype(np.int32)

    # ------------------------------------------------------------ helpers
    @property
    def _plan(self):
        # created on first use so that the geometry can be inspected without a GPU
        if self._plan_obj is None:
            plan = _native.Plan(self._grid)
            if self._rec_table is not None:
                plan.set_receivers(self._rec_table.ravel())
            self._plan_obj = plan
        return self._plan_obj

    @property
    def n_injections(self):
        return int(math.ceil(self.nt / float(self.dt_scale)))

    def _shots(self, shot_ids):
        out = []
        for k, s in enumerate(shot_ids):
            sh = _native.Shot()
            sh.src_x, sh.src_z = int(self.srcx[s]), int(self.srcz[s])
            sh.rec_z, sh.rec_x0 = int(self.rec_surface_const), int(self.rec_surface_var)
            sh.rec_dx, sh.n_rec, sh.rec_stride = int(self.dxr), self._n_live, int(self.nr)
            if self._rec_table is not None:
                sh.rec_layout, sh.rec_table = _native.REC_TABLE, 0
                sh.rec_flags = _native.REC_VZ_FROM_VX if self.acq_type == 0 else 0
            sh.nt, sh.dt_scale, sh.n_samples = int(self.nt), int(self.dt_scale), int(self.nt_ext)
            sh.src_offset = 0
            sh.rec_offset = k * 5 * int(self.nt_ext) * int(self.nr)
            out.append(sh)
        return out

    def _source_table(self):
        return np.ascontiguousarray(self.src.samples(self.n_injections), np.float32)

    def _batches(self):
        ids = list(range(int(self.ns)))
        if self.per_shot:
            return [[s] for s in ids]
        return [ids]

    def pml_preparation(self, v_max):
        """Damping scaled by the model's maximum P velocity
        (src/PyFWI/wave_propagation.py:400-407)."""
        self.vdx_pml = (self.dx_pml_1d * v_max).astype(np.float32)
        self.vdz_pml = (self.dz_pml_1d * v_max).astype(np.float32)

    def elastic_buffers(self, model):
        """mu = rho vs^2, lam = rho vp^2 - 2 mu, b = 1/rho in float32 and
        upload (src/PyFWI/wave_propagation.py:570-591)."""
        self.mu = model["rho"] * (model["vs"] ** 2)
        self.lam = model["rho"] * (model["vp"] ** 2) - 2 * self.mu
        self.rho = model["rho"]
        self.vp = model["vp"]
        self.vs = model["vs"]
        self.rho_b = np.ascontiguousarray(1 / model["rho"], np.float32)
        self._plan.set_model(np.ascontiguousarray(self.lam, np.float32),
                             np.ascontiguousarray(self.mu, np.float32), self.rho_b,
                             self.vdx_pml, self.vdz_pml, _native.MEM_HOST)

    @property
    def W(self):
        """Kept forward states ``W[f][tnz, tnx, ns, nchp]``
        (src/PyFWI/wave_propagation.py:150-156), read back from the GPU."""
        out = {f: np.zeros((self.tnz, self.tnx, self.ns, self.nchp), np.float32)
               for f in COMPONENTS}
        if not self._forward_done:
            return out
        buf = np.zeros((5, int(self.tnz), int(self.tnx)), np.float32)
        for b, ids in enumerate(self._batches()):
            if self.per_shot and b != len(self._batches()) - 1:
                continue  # per-shot runs keep only the last shot's states
            for k, s in enumerate(ids):
                for c in range(self.nchp):
                    self._plan.snapshot(k, c, buf, _native.MEM_HOST)
                    for f, name in enumerate(COMPONENTS):
                        out[name][:, :, s, c] = buf[f]
        return out

    def keeps_every_step(self):
        """True when the checkpoint list is every time step (chpr = 100)."""
        chp = np.asarray(self.chp)
        return chp.size >= int(self.nt) and np.array_equal(chp[:int(self.nt)],
                                                          np.arange(int(self.nt)))

    def state_bytes(self):
        """Device bytes of one shot's state: five fields in the plan's padded
        layout (row pitch roundup(tnx + 8, 64), 8 zero rows above and below)."""
        pitch = -(-(int(self.tnx) + 8) // 64) * 64
        field = -(-((int(self.tnz) + 16) * pitch + 64) // 64) * 64
        return 5 * field * 4

    def gradient_window(self):
        """Rows [z0, z1) and columns [x0, x1) of the cropped model where
        gradient_reading keeps the gradient (src/PyFWI/wave_propagation.py:
        303-311): below the source row for surface acquisitions, right of the
        source well for crosswell (acq_type 0), sdo cells off the other
        edges."""
        sdo, nz, nx = int(self.sdo), int(self.nz), int(self.nx)
        r0 = sdo + int(self.src_cts) + 2
        if self.acq_type == 0:
            return (sdo, nz - sdo), (r0, nx - sdo)
        return (r0, nz - sdo), (sdo, nx - sdo)

    def gradient_reading(self, G):
        """Crop the absorbing layer, mask the source rows and edges and, with
        energy balancing, divide by sqrt of the preconditioners normalised to a
        maximum of 1 (src/PyFWI/wave_propagation.py:279-331).
        G: [6, tnz, tnx] = gmu, gmu_pre, glam, glam_pre, grho, grho_pre."""
        npml, sdo = self.npml, int(self.sdo)
        tnz, tnx = int(self.tnz), int(self.tnx)
        if self.energy_balancing:
            pre = {"mu": G[1], "lam": G[3], "rho": G[5]}
        else:
            one = np.ones((tnz, tnx), np.float32)
            pre = {"mu": one, "lam": one, "rho": one}

        (z0, z1), (x0, x1) = self.gradient_window()

        def factor_of(precond):
            factor = np.zeros((int(self.nz), int(self.nx)), np.float32)
            denom = np.sqrt(precond)[npml + z0:npml + z1, npml + x0:npml + x1]
            with np.errstate(divide="ignore", invalid="ignore"):
                factor[z0:z1, x0:x1] = 1 / denom
                factor = factor / np.abs(factor).max()
            return factor

        crop = (slice(npml, tnz - npml), slice(npml, tnx - npml))
        grho = G[4][crop] * factor_of(pre["rho"])
        gmu = G[0][crop] * factor_of(pre["mu"])
        glam = G[2][crop] * factor_of(pre["lam"])
        return glam, gmu, grho


class WavePropagator(WavePreparation):
    """Forward modelling and adjoint-state gradient
    (src/PyFWI/wave_propagation.py:594-938)."""

    def __init__(self, inpa, src, rec_loc, model_shape, components=0,
                 n_well_rec=0, chpr=0):
        WavePreparation.__init__(self, inpa=inpa, src=src, rec_loc=rec_loc,
                                 model_shape=model_shape, components=components,
                                 n_well_rec=n_well_rec, chpr=chpr)

    def forward_propagator(self, model):
        """Run every shot and fill ``self.seismogram``
        (src/PyFWI/wave_propagation.py:625-650, 699-733, 334-366)."""
        src_tab = self._source_table()
        nrec = int(self.nt_ext) * int(self.nr)
        for ids in self._batches():
            seis = np.zeros((len(ids), 5, int(self.nt_ext), int(self.nr)), np.float32)
            self._plan.forward(self._shots(ids), src_tab, seis, self.chp, 0, _native.MEM_HOST)
            assert seis.size == len(ids) * 5 * nrec
            for k, s in enumerate(ids):
                for f, name in enumerate(COMPONENTS):
                    self.seismogram[name][:, :, s] = seis[k, f]
        self._forward_done = True
        return self.seismogram

    def forward_modeling(self, model0, show=False):
        """Seismic section of the model ``{'vp','vs','rho': [nz, nx]}``
        (src/PyFWI/wave_propagation.py:859-887)."""
        model = {k: expand_model(v, int(self.tnz), int(self.tnx), self.npml)
                 for k, v in model0.items()}
        self.pml_preparation(model["vp"].max())
        self.elastic_buffers(model)
        seismo = self.forward_propagator(model)
        return seismic_section(seismo, self.components, shape=self.seismo_shape)

    def _residual_records(self, res, ids):
        """[len(ids), 5, nt_ext, nr] residual records from the reference's
        ``{f: [nt_ext, nr*ns]}`` layout (column = shot*nr + receiver)."""
        out = np.zeros((len(ids), 5, int(self.nt_ext), int(self.nr)), np.float32)
        nr = int(self.nr)
        for f, name in enumerate(COMPONENTS):
            r = np.asarray(res[name], np.float32)
            if r.ndim == 3:  # [nt, nr, ns]
                r = np.reshape(r, (r.shape[0], -1), "F")
            for k, s in enumerate(ids):
                out[k, f] = r[: int(self.nt_ext), s * nr:(s + 1) * nr]
        return out

    def raw_gradient(self, res, mode="pyfwi", ckpt_every=0):
        """Gradient buffers ``[6, tnz, tnx]`` summed over shots (see
        pyfwi_gradient in include/pyfwi_hip.h)."""
        res = prepare_residual(res, 1)
        mode_id = _native.MODE_PYFWI if mode == "pyfwi" else _native.MODE_EXACT
        flags = _native.F_PRECOND if self.energy_balancing else 0
        G = np.zeros((6, int(self.tnz), int(self.tnx)), np.float32)
        src_tab = self._source_table()
        batches = self._batches()
        for b, ids in enumerate(batches):
            if self.per_shot and mode == "pyfwi" and len(batches) > 1:
                # the kept states belong to the last forward batch only
                seis = np.zeros((1, 5, int(self.nt_ext), int(self.nr)), np.float32)
                self._plan.forward(self._shots(ids), src_tab, seis, self.chp, 0,
                                   _native.MEM_HOST)
            rr = np.ascontiguousarray(self._residual_records(res, ids))
            g = np.zeros_like(G)
            self._plan.gradient(self._shots(ids), src_tab, rr, g, mode_id, ckpt_every, flags,
                                _native.MEM_HOST)
            G += g
        return G

    def gradient(self, res, show=False, parameterization="dv", mode="pyfwi"):
        """Adjoint-state gradient for residuals ``{f: [nt_ext, nr*ns]}``
        (src/PyFWI/wave_propagation.py:889-938).  ``mode='pyfwi'`` is the
        reference's adjoint; ``mode='exact'`` the exact discrete adjoint
        (no mask), returned as {'lam','mu','rho'} or {'vp','vs','rho'}."""
        if not self._forward_done:
            raise RuntimeError("forward_modeling must run before gradient")
        self._unflip_left_well()
        if mode == "exact":
            res = _section_vjp(res, self.components)
        G = self.raw_gradient(res, mode=mode)
        return self._finish_gradient(G, parameterization, mode)

    def _unflip_left_well(self):
        """The reference's gradient puts the left well's records of
        ``seismogram`` back in receiver order (wave_propagation.py:911-913);
        the residual it is given keeps the reversed order."""
        if self.acq_type == 2:
            nw = int(self.n_well_rec)
            for name in self.seismogram:
                self.seismogram[name][:, :nw] = np.flip(self.seismogram[name][:, :nw], axis=1)

    def _finish_gradient(self, G, parameterization="dv", mode="pyfwi"):
        """Read-out of the accumulators G [6, tnz, tnx]: mask and balancing
        (pyfwi mode) or padding fold (exact mode), chain rule, smoothing
        (wave_propagation.py:900-938)."""
        if mode == "pyfwi":
            glam, gmu, grho0 = self.gradient_reading(G)
        else:
            glam, gmu, grho0 = (_fold_padding(G[k], int(self.npml)) for k in (2, 0, 4))
        crop = (slice(self.npml, int(self.tnz) - self.npml),
                slice(self.npml, int(self.tnx) - self.npml))
        if parameterization == "dv":
            gvp, gvs, grho = grad_lmd_to_vd(glam, gmu, grho0, self.lam[crop],
                                            self.mu[crop], self.rho[crop])
            out = {"vp": gvp, "vs": gvs, "rho": grho}
        else:
            out = {"lam": glam, "mu": gmu, "rho": grho0}
        if self.g_smooth:
            from scipy.ndimage import gaussian_filter
            out = {k: gaussian_filter(v, self.g_smooth) for k, v in out.items()}
        return out

    # ------------------------------------------------ one-call misfit + gradient
    def device_misfit_applies(self, d_obs, freq=None, sd=1.0, cost=None):
        """True when ``cost_seismic`` followed by ``gradient`` reduces to
        pyfwi_misfit_gradient: L2 cost, unit data weight, all shots in one
        batch, and for the pressure (components=0) equal 'taux' and 'tauz'
        observations (the plan compares the pressure with one record).  A
        lowpass band runs on the device too (pyfwi_set_record_filter)."""
        if self.per_shot:
            return False
        if cost is not None and cost.cost_function_type != "l2":
            return False
        if np.ndim(sd) or float(sd) != 1.0:
            return False
        if self.components == 0:
            return bool(np.array_equal(d_obs["taux"], d_obs["tauz"]))
        return True

    def _obs_records(self, d_obs):
        """[ns, 5, nt_ext, nr] records of an observed-data dict, converted
        once per dict."""
        cache = self.__dict__.setdefault("_obs_cache", {})
        hit = cache.get(id(d_obs))
        if hit is None or hit[0] is not d_obs:
            rec = self._residual_records(prepare_residual(d_obs, 1), list(range(int(self.ns))))
            hit = cache[id(d_obs)] = (d_obs, np.ascontiguousarray(rec))
        return hit[1]

    def misfit_gradient_device(self, model0, d_obs, parameterization="dv", freq=None):
        """L2 misfit of ``model0`` against ``d_obs`` and its gradient in one
        pyfwi_misfit_gradient call; the same values as forward_modeling,
        cost_seismic (L2, lowpass band ``freq`` with the order-3 filter of
        fwi.py:239-240) and gradient (fwi.py:229-252).  Updates
        ``seismogram``; no stand-alone ``gradient`` may follow."""
        from .fwi_tools import lowpass_taps
        self._plan.set_record_filter(
            lowpass_taps(freq, self.fn, 3, int(self.nt_ext)) if freq else None)
        model = {k: expand_model(np.asarray(v, np.float32), int(self.tnz), int(self.tnx),
                                 self.npml) for k, v in model0.items()}
        self.pml_preparation(model["vp"].max())
        self.elastic_buffers(model)
        obs = self._obs_records(d_obs)
        ids = list(range(int(self.ns)))
        G = np.zeros((6, int(self.tnz), int(self.tnx)), np.float32)
        seis = np.zeros_like(obs)
        flags = _native.F_PRECOND if self.energy_balancing else 0
        rms = self._plan.misfit_gradient(self._shots(ids), self._source_table(), obs, seis, G,
                                         self.chp, COMPONENT_MASK[self.components], flags,
                                         _native.MEM_HOST)
        for f, name in enumerate(COMPONENTS):
            self.seismogram[name][:] = seis[:, f].transpose(1, 2, 0)
        self._unflip_left_well()  # as forward_modeling + gradient leave it
        # the plan keeps only the last pass's states
        self._forward_done = False
        return rms, self._finish_gradient(G, parameterization)

    # ------------------------------------------------ device-resident path
    # Used by pyfwi_amd.torchfwi: model, records and gradients stay in torch
    # tensors on the GPU; the plan reads and writes them in place.

    def _tensor_setup(self, vp, vs, rho):
        import torch
        from torch.nn.functional import pad
        dev = vp.device
        npml = int(self.npml)

        def expand(t):
            t = t.detach().to(device=dev, dtype=torch.float32).contiguous()
            if npml == 0:
                return t
            return pad(t[None, None], (npml, npml, npml, npml), mode="replicate")[0, 0]

        m = {"vp": expand(vp), "vs": expand(vs), "rho": expand(rho)}
        vmax = np.float32(m["vp"].max().item())
        self.pml_vmax = vmax
        self.pml_preparation(vmax)
        self.mu = m["rho"] * (m["vs"] ** 2)
        self.lam = m["rho"] * (m["vp"] ** 2) - 2 * self.mu
        self.rho, self.vp, self.vs = m["rho"], m["vp"], m["vs"]
        self.rho_b = (1 / m["rho"]).contiguous()
        plan = self._plan
        plan.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        self._dev = dev
        self._vdx_t = torch.from_numpy(self.vdx_pml).to(dev)
        self._vdz_t = torch.from_numpy(self.vdz_pml).to(dev)
        plan.set_model(self.lam.contiguous(), self.mu.contiguous(), self.rho_b,
                       self._vdx_t, self._vdz_t, _native.MEM_DEVICE)
        self._src_t = torch.from_numpy(self._source_table()).to(dev)
        return dev

    def forward_modeling_tensor(self, vp, vs, rho, ckpt_every=0, keep_all=False):
        """Device version of ``forward_modeling`` for torch tensors on the GPU;
        returns the seismic section as device tensors.  ``ckpt_every`` > 0 also
        keeps checkpoints for an exact-mode gradient; ``keep_all`` keeps every
        step's state (whatever chpr), which an exact-mode gradient then reads
        instead of recomputing them."""
        import torch
        dev = self._tensor_setup(vp, vs, rho)
        ns, nt, nr = int(self.ns), int(self.nt_ext), int(self.nr)
        seis = torch.empty(ns * 5 * nt * nr, dtype=torch.float32, device=dev)
        chp = checkpoint_steps(int(self.nt), 100) if keep_all else self.chp
        self._plan.forward(self._shots(list(range(ns))), self._src_t, seis, chp,
                           ckpt_every, _native.MEM_DEVICE)
        self._forward_done = True
        self._seis_t = seis.view(ns, 5, nt, nr)
        seismo = {f: self._seis_t[:, k].permute(1, 2, 0) for k, f in enumerate(COMPONENTS)}
        return _section_tensor(seismo, self.components, self.seismo_shape)

    def _records_tensor(self, res):
        """{f: [nt_ext, nr*ns] or [nt_ext, nr, ns]} tensors -> [ns, 5, nt, nr]."""
        import torch
        ns, nt, nr = int(self.ns), int(self.nt_ext), int(self.nr)
        out = torch.zeros(ns, 5, nt, nr, dtype=torch.float32, device=self._dev)
        for k, f in enumerate(COMPONENTS):
            if f not in res:
                continue
            r = res[f].to(device=self._dev, dtype=torch.float32)
            if r.dim() == 2:
                r = r.reshape(nt, ns, nr).permute(1, 0, 2)
            else:
                r = r.permute(2, 0, 1)
            out[:, k] = r
        return out.contiguous()

    def gradient_tensor(self, res, parameterization="dv", mode="pyfwi", ckpt_every=0):
        """Device version of ``gradient``: residual tensors in, gradient
        tensors out.  ``mode='exact'`` returns the exact derivative w.r.t. the
        unpadded model (absorbing-layer cells folded onto the edge cells; the
        damping's dependence on max(vp) is held fixed)."""
        import torch
        if not self._forward_done:
            raise RuntimeError("forward_modeling must run before gradient")
        tnz, tnx, npml = int(self.tnz), int(self.tnx), int(self.npml)
        if mode == "exact":
            res = _section_vjp(res, self.components)
        rec = self._records_tensor(res)
        G = torch.zeros(6, tnz, tnx, dtype=torch.float32, device=self._dev)
        mode_id = _native.MODE_PYFWI if mode == "pyfwi" else _native.MODE_EXACT
        flags = _native.F_PRECOND if self.energy_balancing else 0
        self._plan.gradient(self._shots(list(range(int(self.ns)))), self._src_t, rec, G,
                            mode_id, ckpt_every, flags, _native.MEM_DEVICE)
        crop = (slice(npml, tnz - npml), slice(npml, tnx - npml))
        if mode == "pyfwi":
            glam, gmu, grho0 = _gradient_reading_tensor(self, G)
        else:
            glam, gmu, grho0 = (_fold_padding(G[k], npml) for k in (2, 0, 4))
        if parameterization == "dv":
            gvp, gvs, grho = grad_lmd_to_vd(glam, gmu, grho0, self.lam[crop], self.mu[crop],
                                            self.rho[crop])
            out = {"vp": gvp, "vs": gvs, "rho": grho}
        else:
            out = {"lam": glam, "mu": gmu, "rho": grho0}
        if self.g_smooth:
            from scipy.ndimage import gaussian_filter
            out = {k: torch.from_numpy(gaussian_filter(v.cpu().numpy(), self.g_smooth)).to(self._dev)
                   for k, v in out.items()}
        return out


def _section_tensor(seismo, components, shape):
    """seismic_section for tensors {f: [nt, nr, ns]} (acquisition.py:313-333)."""
    plan = {1: ["taux"], 2: ["vx", "vz"], 3: ["taux", "tauz", "tauxz"],
            4: ["vx", "vz", "taux", "tauz", "tauxz"]}
    if components == 0:
        p = (seismo["taux"] + seismo["tauz"]) / 2
        data = {"taux": p, "tauz": p.clone()}
    else:
        data = {k: seismo[k] for k in plan[int(components)]}
    out = {}
    for k, v in data.items():
        if shape == "2d":
            nt, nr, ns = v.shape
            v = v.permute(0, 2, 1).reshape(nt, ns * nr)
        out[k] = v.contiguous()
    return out


def _section_vjp(res, components):
    """Exact adjoint of the component selection: for the pressure output
    (components 0) both returned records are (txx + tzz)/2, so each stress
    receives half of the sum of both adjoints."""
    if components != 0:
        return res
    s = 0
    for k in ("taux", "tauz"):
        if k in res:
            s = s + res[k]
    return {"taux": s / 2, "tauz": s / 2}


def _fold_padding(g, npml):
    """Transpose of expand_model: sum the absorbing-layer cells onto the edge
    cells they replicate (NumPy array or torch tensor)."""
    if npml == 0:
        return g
    g = g.copy() if isinstance(g, np.ndarray) else g.clone()
    g[:, npml] += g[:, :npml].sum(1)
    g[:, -npml - 1] += g[:, -npml:].sum(1)
    g[npml, :] += g[:npml, :].sum(0)
    g[-npml - 1, :] += g[-npml:, :].sum(0)
    out = g[npml:-npml, npml:-npml]
    return np.ascontiguousarray(out) if isinstance(out, np.ndarray) else out.contiguous()


def _gradient_reading_tensor(W, G):
    """gradient_reading for device tensors (wave_propagation.py:279-331)."""
    import torch
    npml = int(W.npml)
    tnz, tnx, nz, nx = int(W.tnz), int(W.tnx), int(W.nz), int(W.nx)
    (z0, z1), (x0, x1) = W.gradient_window()

    def factor_of(pre):
        f = torch.zeros(nz, nx, dtype=torch.float32, device=G.device)
        if pre is None:
            f[z0:z1, x0:x1] = 1.0
            return f
        d = torch.sqrt(pre)[npml + z0:npml + z1, npml + x0:npml + x1]
        f[z0:z1, x0:x1] = 1 / d
        return f / f.abs().max()

    eb = W.energy_balancing
    crop = (slice(npml, tnz - npml), slice(npml, tnx - npml))
    grho = G[4][crop] * factor_of(G[5] if eb else None)
    gmu = G[0][crop] * factor_of(G[1] if eb else None)
    glam = G[2][crop] * factor_of(G[3] if eb else None)
    return glam, gmu, grho

