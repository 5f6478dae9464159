"""CPU oracle of the PyFWI elastic propagator path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker (or the timed
CPU baseline).  The product package ``pyfwi_amd`` never imports it.

It restates, independently of ``pyfwi_amd``, the host logic of PyFWI v0.1.10
around the kernels and drives the C restatement of the kernels in
``liborc.so`` (elastic_oracle.c):
  geometry and timing       src/PyFWI/wave_propagation.py:28-144
  left-well reordering      src/PyFWI/wave_propagation.py:647-649
  damping profiles          src/PyFWI/fwi_tools.py:943-1063
  model padding             src/PyFWI/fwi_tools.py:776-811
  Lame parameters           src/PyFWI/wave_propagation.py:570-591
  source samples            src/PyFWI/acquisition.py:188-244
  record layout             src/PyFWI/acquisition.py:313-333
  gradient read-out + mask  src/PyFWI/wave_propagation.py:279-331
  chain rule to vp, vs, rho src/PyFWI/grad_switcher.py:331-383

Parity status of each part is stated in elastic_oracle.c's header.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
FIELDS = ("vx", "vz", "taux", "tauz", "tauxz")


class OGrid(ctypes.Structure):
    _fields_ = [("nz", ctypes.c_int), ("nx", ctypes.c_int), ("sdo", ctypes.c_int),
                ("npml", ctypes.c_int), ("c1", ctypes.c_double), ("c2", ctypes.c_double),
                ("c3", ctypes.c_double), ("c4", ctypes.c_double), ("dt", ctypes.c_double),
                ("dx", ctypes.c_double), ("dz", ctypes.c_double)]


class OAcq(ctypes.Structure):
    _fields_ = [("ns", ctypes.c_int), ("nr", ctypes.c_int), ("n_main_rec", ctypes.c_int),
                ("dxr", ctypes.c_int), ("rec_row", ctypes.c_int), ("rec_col0", ctypes.c_int),
                ("srcx", ctypes.POINTER(ctypes.c_int)), ("srcz", ctypes.POINTER(ctypes.c_int)),
                ("nt", ctypes.c_int), ("dt_scale", ctypes.c_int), ("nt_ext", ctypes.c_int),
                ("acq_type", ctypes.c_int), ("n_well_rec", ctypes.c_int),
                ("rec_left_col", ctypes.c_int), ("rec_left_row0", ctypes.c_int),
                ("rec_right_col", ctypes.c_int), ("rec_right_row0", ctypes.c_int)]


_lib = None


def build():
    """Compile liborc.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def mutant(k):
    """The oracle built with planted mistake k (ORC_MUTANT, elastic_oracle.c),
    for showing that a pinning check catches it."""
    path = os.path.join(HERE, "_mut", f"liborc_mut{k}.so")
    subprocess.run(["make", "-s", "-C", HERE, os.path.join("_mut", f"liborc_mut{k}.so")],
                   check=True)
    return ctypes.CDLL(path)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _pct(x):
    """A constant as the reference's kernels see it: printed with %f."""
    return float("%f" % float(x))


# ------------------------------------------------------------ host restatement

def coefficients(order):
    if order == 4:
        return (9 / 8, -1 / 24, 0.0, 0.0)
    if order == 8:
        return (1715 / 1434, -114 / 1434, 14 / 1434, -1 / 1434)
    raise ValueError(order)


def damping_ramp(dh, npml, R):
    L = float(npml) * float(dh)
    if L == 0:
        return np.zeros(0, np.float32)
    r = np.arange(npml, dtype=np.float64) * float(dh)
    return np.float32(-(3.0 / (2.0 * L)) * (r / L) ** 2 * np.log10(R))


def damping_grids(tnz, tnx, dh, npml, R, pml_dir):
    a = damping_ramp(dh, npml, R)
    dx = np.zeros((tnz, tnx), np.float32)
    dz = np.zeros((tnz, tnx), np.float32)
    if npml:
        dx[:, tnx - npml:] = a
        dx[:, :npml] = a[::-1]
        dz[tnz - npml:, :] = a[:, None]
        dz[:npml, :] = a[::-1][:, None]
    if pml_dir == 0:
        dx[:] = 0
    elif pml_dir == 1:
        dz[:] = 0
    elif pml_dir == 3:
        dz[:npml, :] = 0
    return dx, dz


def pad_edge(a, npml):
    """expand_model (fwi_tools.py:776-811): edge replication; float32 unless
    given float64 (FP64 finite-difference checks)."""
    a = np.asarray(a)
    dt = np.float64 if a.dtype == np.float64 else np.float32
    return np.pad(a.astype(dt), npml, mode="edge")


class Setup:
    """The reference's WavePreparation state (wave_propagation.py:18-144) for
    surface (acq_type 1), crosswell (0) and wells + surface (2) acquisition."""

    def __init__(self, inpa, src_loc, rec_loc, model_shape, wavelet, component,
                 chpr=0, components=0, src_dh=None, n_well_rec=0):
        self.inpa = dict(inpa)
        self.dt_scale = int(math.ceil(inpa["dt"] / 0.0006))
        self.dt = inpa["dt"] / self.dt_scale
        self.nt = int(1 + inpa["t"] // self.dt)
        self.nt_ext = int(1 + inpa["t"] // inpa["dt"])
        self.nz, self.nx = int(model_shape[0]), int(model_shape[1])
        self.npml = int(inpa.get("npml", 0))
        self.pmlR = inpa.get("pmlR", 0)
        self.pml_dir = inpa.get("pml_dir", 2)
        self.sdo = int(inpa["sdo"] / 2) if "sdo" in inpa else 2
        self.order = 2 * self.sdo
        self.tnz, self.tnx = self.nz + 2 * self.npml, self.nx + 2 * self.npml
        dh32 = np.float32(inpa["dh"])
        self.dh = dh32
        sdh = inpa["dh"] if src_dh is None else src_dh
        src_loc = np.asarray(src_loc)
        self.src_i = np.int32(src_loc[:, 0] / sdh)
        self.src_j = np.int32(src_loc[:, 1] / sdh)
        self.srcx = (self.src_i + self.npml).astype(np.int32)
        self.srcz = (self.src_j + self.npml).astype(np.int32)
        self.ns = len(self.srcx)
        rec_dis = inpa.get("rec_dis", None)
        if rec_dis is None:
            rec_dis = (rec_loc[1, :] - rec_loc[0, :]).max()
        self.dxr = int(np.int32(float(rec_dis) / float(dh32)))
        self.nr = rec_loc.shape[0]
        self.acq_type = int(inpa.get("acq_type", 1))
        self.n_well_rec = int(n_well_rec)
        self.n_main_rec = self.nr - 2 * self.n_well_rec
        if self.acq_type == 2 and self.n_well_rec == 0:
            raise ValueError("acq_type 2 needs n_well_rec")

        def cell(k):  # (column, row) of receiver k incl. the absorbing layer
            return tuple(int(np.int32(float(rec_loc[k, a]) / float(dh32) + self.npml))
                         for a in (0, 1))

        # receiver macros (wave_propagation.py:116-143)
        self.rec_row = self.rec_col0 = 0
        self.rec_left_col = self.rec_left_row0 = 0
        self.rec_right_col = self.rec_right_row0 = 0
        if self.acq_type == 0:
            self.rec_right_col, self.rec_right_row0 = cell(0)
            self.src_cts = int(self.src_i[0])
        elif self.acq_type == 1:
            self.rec_col0, self.rec_row = cell(0)
            self.src_cts = int(self.src_j[0])
        else:
            self.rec_right_col, self.rec_right_row0 = cell(self.nr - self.n_well_rec)
            self.rec_col0, self.rec_row = cell(self.n_well_rec)
            self.rec_left_col, self.rec_left_row0 = cell(0)
            self.src_cts = int(self.src_j[0])
        n = int(chpr * self.nt / 100)
        chp = np.linspace(0, self.nt - 1, n, dtype=np.int32)
        if len(chp) < 2:
            chp = np.array([1, self.nt - 1], np.int32)
        self.chp = chp
        self.components = components
        self.energy_balancing = inpa.get("energy_balancing", False)
        self.seismo_shape = inpa.get("seimogram_shape", "2d")
        c = coefficients(self.order)
        self.grid = OGrid(self.tnz, self.tnx, self.sdo, self.npml, _pct(c[0]), _pct(c[1]),
                          _pct(c[2]), _pct(c[3]), _pct(self.dt), _pct(dh32), _pct(dh32))
        ninj = -(-self.nt // self.dt_scale)
        w = np.asarray(wavelet, np.float32)
        comp = np.asarray(component, np.float32)
        tab = np.zeros((ninj, 5), np.float32)
        m = min(ninj, w.size)
        tab[:m] = w[:m, None] * comp[None, :]
        self.src_table = tab
        self.dpx, self.dpz = damping_grids(self.tnz, self.tnx, dh32, self.npml, self.pmlR,
                                           self.pml_dir)

    def acq(self, shots=None):
        ids = list(range(self.ns)) if shots is None else list(shots)
        self._sx = np.ascontiguousarray(self.srcx[ids], np.int32)
        self._sz = np.ascontiguousarray(self.srcz[ids], np.int32)
        n_main = self.nr if self.acq_type == 1 else self.n_main_rec
        return OAcq(len(ids), self.nr, n_main, self.dxr, self.rec_row, self.rec_col0,
                    self._sx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                    self._sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                    self.nt, self.dt_scale, self.nt_ext, self.acq_type, self.n_well_rec,
                    self.rec_left_col, self.rec_left_row0, self.rec_right_col,
                    self.rec_right_row0)

    def receiver_cells(self):
        """Flat cell each record column reads (None where no gather reaches),
        from the C restatement of the gathers."""
        out = np.zeros(self.nr, np.int64)
        rc = lib().orc_receiver_cells(ctypes.byref(self.grid), ctypes.byref(self.acq()),
                                      _p(out))
        assert rc == 0
        no_cell = -(2 ** 31)
        return [None if c == no_cell else int(c) for c in out]

    def flip_wells(self, seis):
        """forward_propagator's reordering of the left well's receivers
        (wave_propagation.py:647-649) on records [ns, 5, nt, nr]."""
        if self.acq_type != 2:
            return seis
        out = seis.copy()
        out[..., :self.n_well_rec] = seis[..., :self.n_well_rec][..., ::-1]
        return out

    def medium(self, model, dtype=np.float32):
        """Padded lam, mu, b and damping grids in `dtype` (float32 host
        arithmetic as the reference when dtype is float32)."""
        m = {k: pad_edge(v, self.npml) for k, v in model.items()}
        vmax = m["vp"].max()
        if dtype == np.float32:
            mu = m["rho"] * (m["vs"] ** 2)
            lam = m["rho"] * (m["vp"] ** 2) - 2 * mu
            b = (1 / m["rho"]).astype(np.float32)
            dpx, dpz = self.dpx * vmax, self.dpz * vmax
        else:
            vp, vs, rho = (m[k].astype(np.float64) for k in ("vp", "vs", "rho"))
            mu = rho * vs ** 2
            lam = rho * vp ** 2 - 2 * mu
            b = 1 / rho
            dpx = self.dpx.astype(np.float64) * float(vmax)
            dpz = self.dpz.astype(np.float64) * float(vmax)
        cast = lambda a: np.ascontiguousarray(a, dtype)
        return dict(lam=cast(lam), mu=cast(mu), b=cast(b), dpx=cast(dpx), dpz=cast(dpz),
                    rho=m["rho"], vp=m["vp"], vs=m["vs"], mu32=m["rho"] * m["vs"] ** 2,
                    lam32=m["rho"] * m["vp"] ** 2 - 2 * (m["rho"] * m["vs"] ** 2))


def _suffix(dtype):
    return "f" if dtype == np.float32 else "d"


def forward(setup, model, dtype=np.float32, shots=None, keep_states=True, medium=None):
    """Records [ns, 5, nt_ext, nr] and kept states [ns, nchp, 5, tnz, tnx].
    ``medium`` overrides the lam/mu/b/damping arrays derived from ``model``."""
    L = lib()
    md = setup.medium(model, dtype) if medium is None else medium
    acq = setup.acq(shots)
    ns = acq.ns
    seis = np.zeros((5, ns, setup.nt_ext, setup.nr), dtype)
    nchp = len(setup.chp)
    W = np.zeros((ns, nchp, 5, setup.tnz, setup.tnx), dtype) if keep_states else None
    src = np.ascontiguousarray(np.broadcast_to(setup.src_table.astype(dtype),
                                               (ns,) + setup.src_table.shape))
    chp = np.ascontiguousarray(setup.chp, np.int32)
    fn = getattr(L, "orc_forward_" + _suffix(dtype))
    rc = fn(ctypes.byref(setup.grid), ctypes.byref(acq), _p(md["lam"]), _p(md["mu"]),
            _p(md["b"]), _p(md["dpx"]), _p(md["dpz"]), _p(src), _p(chp), ctypes.c_int(nchp),
            _p(seis), _p(W) if W is not None else None)
    assert rc == 0
    return np.ascontiguousarray(seis.transpose(1, 0, 2, 3)), W


def seismogram_dict(seis):
    """[ns, 5, nt, nr] -> {f: [nt, nr, ns]} (the reference's self.seismogram)."""
    return {f: np.ascontiguousarray(seis[:, k].transpose(1, 2, 0)) for k, f in enumerate(FIELDS)}


def section(seis, components, shape="2d"):
    """Component selection and layout (src/PyFWI/acquisition.py:313-333)."""
    sd = seismogram_dict(seis)
    plan = {1: ["taux"], 2: ["vx", "vz"], 3: ["taux", "tauz", "tauxz"], 4: list(FIELDS)}
    if components == 0:
        p = (sd["taux"] + sd["tauz"]) / 2
        data = {"taux": p, "tauz": p.copy()}
    else:
        data = {k: sd[k] for k in plan[components]}
    if shape == "2d":
        for k in data:
            nt, nr, ns = data[k].shape
            data[k] = np.reshape(data[k], (nt, nr * ns), "F")
    return data


def residual_records(setup, res2d, dtype=np.float32):
    """{f: [nt, nr*ns]} (missing components zero) -> [5, ns, nt, nr]."""
    out = np.zeros((5, setup.ns, setup.nt_ext, setup.nr), dtype)
    for k, f in enumerate(FIELDS):
        if f in res2d:
            r = np.asarray(res2d[f])
            out[k] = r.reshape(setup.nt_ext, setup.ns, setup.nr).transpose(1, 0, 2)
    return out


def raw_gradient(setup, model, res2d, W, dtype=np.float32):
    """The six accumulators [6, tnz, tnx] of the reference's adjoint loop."""
    L = lib()
    md = setup.medium(model, dtype)
    acq = setup.acq()
    res = np.ascontiguousarray(residual_records(setup, res2d, dtype))
    G = np.zeros((6, setup.tnz, setup.tnx), dtype)
    chp = np.ascontiguousarray(setup.chp, np.int32)
    fn = getattr(L, "orc_gradient_" + _suffix(dtype))
    rc = fn(ctypes.byref(setup.grid), ctypes.byref(acq), _p(md["lam"]), _p(md["mu"]),
            _p(md["b"]), _p(md["dpx"]), _p(md["dpz"]), _p(chp), ctypes.c_int(len(chp)),
            _p(np.ascontiguousarray(W, dtype)), _p(res), _p(G))
    assert rc == 0
    return G


def read_gradient(setup, G):
    """Crop, mask and optional energy balancing (wave_propagation.py:279-331)."""
    npml, sdo, tnz, tnx = setup.npml, setup.sdo, setup.tnz, setup.tnx
    if setup.energy_balancing:
        pre = (G[1], G[3], G[5])
    else:
        one = np.ones((tnz, tnx), np.float32)
        pre = (one, one, one)

    def factor(p):
        f = np.zeros((setup.nz, setup.nx), np.float32)
        r0 = sdo + setup.src_cts + 2
        with np.errstate(divide="ignore", invalid="ignore"):
            if setup.acq_type == 0:  # columns right of the source well
                d = np.sqrt(p)[npml + sdo:tnz - npml - sdo, npml + r0:tnx - npml - sdo]
                f[sdo:setup.nz - sdo, r0:setup.nx - sdo] = 1 / d
            else:
                d = np.sqrt(p)[npml + r0:tnz - npml - sdo, npml + sdo:tnx - npml - sdo]
                f[r0:setup.nz - sdo, sdo:setup.nx - sdo] = 1 / d
            return f / np.abs(f).max()

    crop = (slice(npml, tnz - npml), slice(npml, tnx - npml))
    gmu = G[0][crop] * factor(pre[0])
    glam = G[2][crop] * factor(pre[1])
    grho = G[4][crop] * factor(pre[2])
    return glam, gmu, grho


def to_vd(glam, gmu, grho, lam, mu, rho):
    """Chain rule (lam, mu, rho) -> (vp, vs, rho) (grad_switcher.py:331-383)."""
    vp = np.sqrt((lam + 2 * mu) / rho)
    vs = np.sqrt(mu / rho)
    gvp = glam * 2 * vp * rho
    gvs = glam * (-4 * rho * vs) + gmu * 2 * vs * rho
    gr = glam * (vp ** 2 - 2 * vs ** 2) + gmu * vs ** 2 + grho
    return gvp.astype(np.float32), gvs.astype(np.float32), gr.astype(np.float32)


def gradient(setup, model, res2d, W, parameterization="dv", dtype=np.float32):
    G = raw_gradient(setup, model, res2d, W, dtype)
    glam, gmu, grho = read_gradient(setup, G.astype(np.float32))
    if parameterization == "dv":
        md = setup.medium(model, np.float32)
        c = (slice(setup.npml, setup.tnz - setup.npml), slice(setup.npml, setup.tnx - setup.npml))
        gvp, gvs, gr = to_vd(glam, gmu, grho, md["lam32"][c], md["mu32"][c], md["rho"][c])
        return {"vp": gvp, "vs": gvs, "rho": gr}
    return {"lam": glam, "mu": gmu, "rho": grho}


def exact_gradient(setup, model, res_records, dtype=np.float64, shots=None, medium=None):
    """dJ/d(lam, mu, b) [3, tnz, tnx] for J = sum(res * records), all shots."""
    L = lib()
    md = setup.medium(model, dtype) if medium is None else medium
    acq = setup.acq(shots)
    ns = acq.ns
    src = np.ascontiguousarray(np.broadcast_to(setup.src_table.astype(dtype),
                                               (ns,) + setup.src_table.shape))
    res = np.ascontiguousarray(res_records, dtype)  # [5, ns, nt_ext, nr]
    g3 = np.zeros((3, setup.tnz, setup.tnx), dtype)
    fn = getattr(L, "orc_exact_gradient_" + _suffix(dtype))
    rc = fn(ctypes.byref(setup.grid), ctypes.byref(acq), _p(md["lam"]), _p(md["mu"]),
            _p(md["b"]), _p(md["dpx"]), _p(md["dpz"]), _p(src), _p(res), _p(g3))
    assert rc == 0
    return g3


def step(name, grid, F, md, dtype=np.float64, A=None, G=None, L=None):
    """Single-step primitives for unit tests: name in fwd, fwd_sv, rev, adj, grad.
    ``L``: another build of the library (a planted-mistake build), else liborc.so."""
    L = lib() if L is None else L
    sfx = _suffix(dtype)
    args = (_p(md["lam"]), _p(md["mu"]), _p(md["b"]), _p(md["dpx"]), _p(md["dpz"]))
    if name in ("fwd", "fwd_sv", "rev"):
        getattr(L, {"fwd": "orc_fwd_step_", "fwd_sv": "orc_fwd_step_sv_",
                    "rev": "orc_rev_step_"}[name] + sfx)(ctypes.byref(grid), _p(F), *args)
    elif name == "adj":
        getattr(L, "orc_adj_step_" + sfx)(ctypes.byref(grid), _p(A), *args)
    elif name == "grad":
        getattr(L, "orc_grad_acc_" + sfx)(ctypes.byref(grid), _p(F), _p(A), _p(md["b"]), _p(G))
    else:
        raise ValueError(name)
