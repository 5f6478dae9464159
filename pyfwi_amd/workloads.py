"""Seeded synthetic workloads for tests and benchmarks.

Each builder returns a ``Workload`` holding the ``inpa`` dictionary, a
``Source``, receiver positions and models in the layout PyFWI's
``WavePropagator`` takes (src/PyFWI/wave_propagation.py:617-623).  Models are
random-layered with Gardner density ``rho = 310 vp^0.25``; all randomness comes
from ``numpy.random.default_rng(seed)``.  The absorbing layer is 20 cells with
``pmlR = 1e-5`` on all four sides (``pml_dir = 2``) and the stencil is of
order 4, as in PyFWI's examples (example/fwi_example_surface.py:24-32).
"""
from dataclasses import dataclass, field

import numpy as np

from .acquisition import Source

__all__ = ["Workload", "layered3", "random_layered", "gaussian_anomaly",
           "plain_small", "plain_medium", "crosswell_medium", "wells_surface_medium",
           "differentiation", "large_step",
           "t_for_nt", "gardner", "UnevenWorkload", "uneven_batch", "uneven_small", "subset"]


def gardner(vp):
    return (310.0 * np.asarray(vp, np.float64) ** 0.25).astype(np.float32)


def t_for_nt(nt, dt):
    """Record length giving exactly ``nt`` samples under
    ``nt = int(1 + t // dt)`` (src/PyFWI/wave_propagation.py:34-35)."""
    return (nt - 1) * dt + dt / 2


@dataclass
class Workload:
    name: str
    inpa: dict
    src: Source
    rec_loc: np.ndarray
    model: dict
    model_obs: dict = None
    components: int = 0
    chpr: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def model_shape(self):
        return self.model["vp"].shape

    @property
    def cells(self):
        npml = self.inpa.get("npml", 0)
        nz, nx = self.model_shape
        return (nz + 2 * npml) * (nx + 2 * npml)


def _base_inpa(dh, dt, nt, npml=20):
    return {"dh": dh, "dt": dt, "t": t_for_nt(nt, dt), "npml": npml, "pmlR": 1e-5,
            "pml_dir": 2, "sdo": 4, "acq_type": 1}


def layered3(nz=100, nx=100):
    """Three equal horizontal layers, vp 1500/2000/2500, vs = vp/1.73."""
    vp = np.empty((nz, nx), np.float32)
    b1, b2 = nz // 3 + 1, 2 * nz // 3 + 1
    vp[:b1], vp[b1:b2], vp[b2:] = 1500, 2000, 2500
    return {"vp": vp, "vs": (vp / 1.73).astype(np.float32), "rho": gardner(vp)}


def random_layered(nz, nx, nlayers, seed, vp_range=(1800.0, 4000.0), ratio=(1.6, 2.0)):
    """Horizontal layers with sorted uniform interface depths, vp uniform in
    ``vp_range``, vs = vp / U(ratio), Gardner density."""
    rng = np.random.default_rng(seed)
    tops = np.sort(rng.integers(1, nz, nlayers - 1))
    vp_l = rng.uniform(*vp_range, nlayers)
    r_l = rng.uniform(*ratio, nlayers)
    layer = np.searchsorted(tops, np.arange(nz), side="right")
    vp = np.repeat(vp_l[layer][:, None], nx, axis=1).astype(np.float32)
    vs = np.repeat((vp_l / r_l)[layer][:, None], nx, axis=1).astype(np.float32)
    return {"vp": vp, "vs": vs, "rho": gardner(vp)}


def gaussian_anomaly(model, z0, x0, sigma, amp, keep_vs=True):
    """vp * (1 + amp * exp(-r^2 / (2 sigma^2))) around cell (z0, x0)."""
    nz, nx = model["vp"].shape
    z, x = np.mgrid[0:nz, 0:nx]
    g = np.exp(-((z - z0) ** 2 + (x - x0) ** 2) / (2.0 * sigma ** 2))
    out = {k: v.copy() for k, v in model.items()}
    out["vp"] = (model["vp"] * (1 + amp * g)).astype(np.float32)
    return out


def _surface_geometry(ns, nx, dh, nrec, rec_dis_cells, src_z_cells, rec_z_cells, x0_cells):
    src_loc = np.zeros((ns, 2), np.float32)
    src_loc[:, 0] = np.linspace(4 * dh, nx * dh - 9 * dh, ns)
    src_loc[:, 1] = src_z_cells * dh
    rec_loc = np.zeros((nrec, 2), np.float64)
    rec_loc[:, 0] = (x0_cells + rec_dis_cells * np.arange(nrec)) * dh
    rec_loc[:, 1] = rec_z_cells * dh
    return src_loc, rec_loc


def plain_small():
    """nz=nx=100, dh=1 m, 3 layers, 30 Hz Ricker, dt=0.1 ms, nt=1000, 5 surface
    explosive shots (z=4 cells, x=linspace(4, 91)), 50 receivers (z=2 cells,
    x=1+2k), vx/vz records; observed data from a +5% Gaussian vp anomaly
    (sigma 8 cells) at the centre; gradient w.r.t. lambda/mu/rho."""
    dh, dt, nt = 1.0, 1e-4, 1000
    inpa = _base_inpa(dh, dt, nt)
    inpa["rec_dis"] = 2 * dh
    model = layered3(100, 100)
    src_loc, rec_loc = _surface_geometry(5, 100, dh, 50, 2, 4, 2, 1)
    src = Source(src_loc, dh, dt)
    src.Ricker(30)
    obs = gaussian_anomaly(model, 50, 50, 8.0, 0.05)
    return Workload("plain_small_100x100", inpa, src, rec_loc, model, obs, components=2,
                    chpr=100, meta={"nt": nt, "shots": 5, "receivers": 50})


def _cfl_dt(model, dh, cfl=0.45):
    return cfl * dh / float(model["vp"].max())


def plain_medium(seed=2, ns=20, nt=2000):
    """nz=500, nx=1000, dh=5 m, 20-layer random model, 20 Hz Ricker, dt from
    CFL 0.45, nt=2000, 20 evenly spaced shots, 400 receivers (10 m spacing)."""
    dh = 5.0
    model = random_layered(500, 1000, 20, seed)
    dt = _cfl_dt(model, dh)
    inpa = _base_inpa(dh, dt, nt)
    inpa["rec_dis"] = 2 * dh
    src_loc, rec_loc = _surface_geometry(ns, 1000, dh, 400, 2, 4, 2, 100)
    src = Source(src_loc, dh, dt)
    src.Ricker(20)
    return Workload("plain_medium_500x1000", inpa, src, rec_loc, model, None, components=0,
                    chpr=0, meta={"nt": nt, "shots": ns, "receivers": 400, "seed": seed})


def _wells_medium(acq_type, name, seed, ns, nt, nz, nx):
    from .acquisition import acq_parameters
    dh = 5.0
    model = random_layered(nz, nx, 20, seed)
    dt = _cfl_dt(model, dh)
    inpa = _base_inpa(dh, dt, nt)
    inpa.update(rec_dis=2 * dh, acq_type=acq_type)
    src_loc, rec_loc, n_surface, n_well = acq_parameters(ns, 2 * dh, nx * dh, nz * dh, dh,
                                                         inpa["sdo"], acq_type)
    src = Source(src_loc, dh, dt)
    src.Ricker(20)
    return Workload(name, inpa, src, rec_loc.astype(np.float64), model, None, components=0,
                    chpr=0, meta={"nt": nt, "shots": ns, "receivers": int(rec_loc.shape[0]),
                                  "n_well_rec": int(n_well) if acq_type != 1 else 0,
                                  "seed": seed})


def crosswell_medium(seed=2, ns=20, nt=2000):
    """nz=1000, nx=500 (the plain_medium cell count, upright), dh=5 m,
    20-layer random model, 20 Hz, nt=2000, with the reference's crosswell
    geometry (acq_type 0, acquisition.py:62-124): 20 sources down the left
    well, 496 receivers 10 m apart down the right one, 2.46 km away."""
    return _wells_medium(0, "crosswell_medium_1000x500", seed, ns, nt, 1000, 500)


def wells_surface_medium(seed=2, ns=20, nt=2000):
    """The plain_medium model with both wells and the surface line
    (acq_type 2, acquisition.py:46-57): 246 receivers in each well and 496
    on the surface line, surface sources."""
    return _wells_medium(2, "wells_surface_medium_500x1000", seed, ns, nt, 500, 1000)


def differentiation(seed=3, ns=10, nt=2500):
    """nz=300, nx=600, dh=5 m, random layers plus 3 Gaussian anomalies of
    +-10% vp, 10 shots, nt=2500."""
    dh = 5.0
    rng = np.random.default_rng(seed + 1000)
    model = random_layered(300, 600, 20, seed)
    obs = model
    for _ in range(3):
        obs = gaussian_anomaly(obs, rng.uniform(60, 240), rng.uniform(100, 500),
                               rng.uniform(10, 25), rng.choice([-0.1, 0.1]))
    dt = _cfl_dt(obs, dh)
    inpa = _base_inpa(dh, dt, nt)
    inpa["rec_dis"] = 2 * dh
    src_loc, rec_loc = _surface_geometry(ns, 600, dh, 280, 2, 4, 2, 20)
    src = Source(src_loc, dh, dt)
    src.Ricker(20)
    return Workload("differentiation_300x600", inpa, src, rec_loc, model, obs, components=0,
                    chpr=0, meta={"nt": nt, "shots": ns, "seed": seed})


# the time-lapse reservoir patch of large_step: 200 x 600 cells
LARGE_STEP_PATCH = (slice(1100, 1300), slice(2700, 3300))


def large_step(seed=5, ns=32, nt=6000):
    """nz=2000, nx=6000, dh=2.5 m, random layers with 5 anomalies, 32 shots,
    nt=6000, 15 Hz, 3000 receivers.  ``model`` holds the layers with the
    anomalies (the baseline); ``model_obs`` the same (observed data of the
    baseline); ``meta['initial']`` the layers alone (the inversions' start)
    and ``meta['monitor']`` the baseline with vp -8 % in the 200 x 600-cell
    patch LARGE_STEP_PATCH."""
    dh = 2.5
    rng = np.random.default_rng(seed + 1000)
    layers = random_layered(2000, 6000, 20, seed)
    model = layers
    for _ in range(5):
        model = gaussian_anomaly(model, rng.uniform(300, 1700), rng.uniform(500, 5500),
                                 rng.uniform(30, 80), rng.choice([-0.1, 0.1]))
    monitor = {k: v.copy() for k, v in model.items()}
    monitor["vp"][LARGE_STEP_PATCH] *= np.float32(0.92)
    dt = _cfl_dt(model, dh)
    inpa = _base_inpa(dh, dt, nt)
    inpa["rec_dis"] = 2 * dh
    src_loc, rec_loc = _surface_geometry(ns, 6000, dh, 3000, 2, 4, 2, 0)
    src = Source(src_loc, dh, dt)
    src.Ricker(15)
    return Workload("large_step_2000x6000", inpa, src, rec_loc, model, model, components=0,
                    chpr=0, meta={"nt": nt, "shots": ns, "seed": seed, "initial": layers,
                                  "monitor": monitor})


def subset(w, ids, nt=None):
    """The workload restricted to shots `ids` (same model, receivers and
    wavelet), optionally with ``nt`` time steps."""
    from copy import deepcopy
    dh, dt = w.inpa["dh"], w.inpa["dt"]
    loc = np.c_[w.src.i, w.src.j][list(ids)].astype(np.float64) * dh
    src = Source(loc, dh, dt)
    src.component = w.src.component.copy()
    src.w = w.src.w.copy()
    out = deepcopy(w)
    out.src = src
    if nt is not None:
        out.inpa["t"] = t_for_nt(nt, dt)
    out.meta = dict(w.meta, shots=len(ids), shot_ids=list(ids))
    return out


@dataclass
class UnevenWorkload:
    """Shots that differ in duration, aperture and wavelet (for
    ``pyfwi_amd.batch.UnevenBatch``)."""
    name: str
    inpa: dict
    model: dict
    model_obs: dict
    specs: list
    components: int = 2
    meta: dict = field(default_factory=dict)

    @property
    def model_shape(self):
        return self.model["vp"].shape


def _uneven_specs(rng, ns, nx, dh, dt, nt_range, aperture, fdom, rec_dis_cells,
                  src_z_cells=4, rec_z_cells=2):
    from .batch import ShotSpec
    specs = []
    for _ in range(ns):
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        nrec = int(rng.integers(aperture[0], aperture[1] + 1))
        sx = int(rng.integers(4, nx - 9))
        span = (nrec - 1) * rec_dis_cells
        x0 = int(np.clip(sx - span // 2, 1, nx - span - 2))
        rec = np.zeros((nrec, 2), np.float64)
        rec[:, 0] = (x0 + rec_dis_cells * np.arange(nrec)) * dh
        rec[:, 1] = rec_z_cells * dh
        specs.append(ShotSpec((sx * dh, src_z_cells * dh), rec, t_for_nt(nt, dt),
                              float(rng.uniform(*fdom))))
    return specs


def uneven_batch(seed=4, ns=64, nz=1000, nx=2000, dh=4.0, nt_range=(1500, 4000),
                 aperture=(200, 1600), fdom=(10.0, 25.0)):
    """nz=1000, nx=2000, dh=4 m, 20-layer random model, dt from CFL 0.45; 64
    shots with source x ~ U, nt ~ U[1500, 4000], receivers U{200..1600}
    (one-cell spacing, centred on the source, clipped to the grid) and a
    Ricker of fdom ~ U(10, 25) Hz; vx/vz records; observed data from a +5 %
    Gaussian vp anomaly (sigma 60 cells) at the centre."""
    model = random_layered(nz, nx, 20, seed)
    dt = _cfl_dt(model, dh)
    inpa = _base_inpa(dh, dt, nt_range[1])
    rng = np.random.default_rng(seed + 1000)
    specs = _uneven_specs(rng, ns, nx, dh, dt, nt_range, aperture, fdom, 1)
    obs = gaussian_anomaly(model, nz / 2, nx / 2, 60.0, 0.05)
    return UnevenWorkload(f"uneven_{nz}x{nx}_{ns}shots", inpa, model, obs, specs,
                          meta={"seed": seed, "shots": ns, "nt_range": nt_range,
                                "aperture": aperture, "fdom": fdom})


def uneven_small(seed=5, ns=5, nt_range=(150, 320), aperture=(15, 40), fdom=(20.0, 40.0)):
    """The uneven batch at the size of plain_small (100x100 cells, dh=1 m,
    dt=0.1 ms, 3 layers): per-shot nt, aperture (two-cell spacing) and
    Ricker frequency."""
    dh, dt = 1.0, 1e-4
    model = layered3(100, 100)
    inpa = _base_inpa(dh, dt, nt_range[1])
    rng = np.random.default_rng(seed)
    specs = _uneven_specs(rng, ns, 100, dh, dt, nt_range, aperture, fdom, 2)
    obs = gaussian_anomaly(model, 50, 50, 8.0, 0.05)
    return UnevenWorkload("uneven_small_100x100", inpa, model, obs, specs,
                          meta={"seed": seed, "shots": ns})
