"""GPU parity of well acquisitions -- crosswell (acq_type 0) and wells +
surface (acq_type 2) -- against the CPU oracle: records, the reference
gradient (kept states and time reversal), the exact gradient and the
one-call misfit + gradient.

Tolerances as the surface parity tests: records <= 1e-5 relative L2,
gradients <= 1e-4 relative L2 (FP32 kernels against the oracle's float
storage with double intermediates)."""
import numpy as np
import pytest

from oracle import oracle as O
from pyfwi_amd.acquisition import Source, acq_parameters
from pyfwi_amd.fwi_tools import CostFunction
from pyfwi_amd.wave_propagation import WavePropagator

from _helpers import rel

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4
DH, DT, NZ, NX, NPML = 5.0, 5e-4, 60, 80, 10


def wells(acq_type, order=4, nt=400, ns=3, chpr=100, components=4, dt=DT):
    inpa = {"dh": DH, "dt": dt, "t": (nt - 1) * dt + dt / 2, "npml": NPML, "pmlR": 1e-5,
            "pml_dir": 2, "sdo": order, "rec_dis": 2 * DH, "acq_type": acq_type}
    src_loc, rec_loc, _, nw = acq_parameters(ns, 2 * DH, NX * DH, NZ * DH, DH, order, acq_type)
    src = Source(src_loc, DH, dt)
    src.Ricker(25)
    rng = np.random.default_rng(11)
    z, x = np.mgrid[0:NZ, 0:NX]
    vp = (2000 + 8 * z + 150 * np.sin(0.11 * x) + 40 * rng.random((NZ, NX))).astype(np.float32)
    model = {"vp": vp, "vs": (vp / 1.8).astype(np.float32),
             "rho": (1900 + 0.1 * vp).astype(np.float32)}
    bump = np.exp(-((z - 30) ** 2 + (x - 40) ** 2) / 50.0)
    model_obs = dict(model, vp=(vp * (1 + 0.05 * bump)).astype(np.float32))
    W = WavePropagator(inpa, src, rec_loc, (NZ, NX), components=components, n_well_rec=nw,
                       chpr=chpr)
    st = O.Setup(inpa, src_loc, rec_loc, (NZ, NX), src.w, src.component, chpr=chpr,
                 components=components, n_well_rec=nw)
    return W, st, model, model_obs


def oracle_section(st, model, components=4):
    seis, states = O.forward(st, model, np.float32)
    return O.section(st.flip_wells(seis), components), states


@pytest.mark.parametrize("acq_type", [0, 2])
@pytest.mark.parametrize("order", [4, 8])
def test_forward_wells(gpu, acq_type, order):
    print(f"wells forward acq_type={acq_type} order={order}", flush=True)
    W, st, model, _ = wells(acq_type, order=order)
    d = W.forward_modeling(model)
    ref, _ = oracle_section(st, model)
    for k in ref:
        e = rel(d[k], ref[k])
        print(f"  {k}: {e:.2e}")
        assert e < SEIS_TOL, (k, e)
    if acq_type == 0:  # the crosswell vz record holds vx
        np.testing.assert_array_equal(d["vz"], d["vx"])


@pytest.mark.parametrize("acq_type", [0, 2])
def test_kept_records_wells_dt_scale(gpu, acq_type):
    """With every state kept, table receivers' records are gathered from the
    kept slots after the time loop; without, before each step.  Both give the
    same records, bit for bit, here with two internal steps a sample, and
    both match the oracle."""
    print(f"wells kept records acq_type={acq_type} dt_scale=2", flush=True)
    # 0.24 s: the direct wave crosses the 400 m between the wells
    W, st, model, _ = wells(acq_type, nt=300, dt=8e-4)
    assert W.dt_scale == 2
    d_kept = W.forward_modeling(model)
    W0, _, _, _ = wells(acq_type, nt=300, dt=8e-4, chpr=0)
    d_step = W0.forward_modeling(model)
    ref, _ = oracle_section(st, model)
    for k in ref:
        np.testing.assert_array_equal(d_kept[k], d_step[k])
        e = rel(d_kept[k], ref[k])
        print(f"  {k}: {e:.2e}")
        assert e < SEIS_TOL, (k, e)


@pytest.mark.parametrize("acq_type", [0, 2])
@pytest.mark.parametrize("chpr,order", [(100, 4), (0, 4), (100, 8)])
def test_gradient_wells(gpu, acq_type, chpr, order):
    """Reference gradient for the residual of CostFunction.l2; for acq_type 2
    the residual's left-well columns are injected in the records' reversed
    order, as the reference does."""
    print(f"wells gradient acq_type={acq_type} chpr={chpr} order={order}", flush=True)
    W, st, model, model_obs = wells(acq_type, order=order, chpr=chpr)
    d_est = W.forward_modeling(model)
    d_obs, _ = oracle_section(st, model_obs)
    _, res = CostFunction("l2")(d_est, d_obs)
    g = W.gradient(res, parameterization="lmd")
    _, states = O.forward(st, model, np.float32)
    gr = O.gradient(st, model, res, states, parameterization="lmd")
    for k in gr:
        e = rel(g[k], gr[k])
        print(f"  {k}: {e:.2e}")
        assert e < GRAD_TOL, (k, e)
    if acq_type == 2:  # gradient puts the stored left-well records back in order
        seis, _ = O.forward(st, model, np.float32)
        nw = st.n_well_rec
        a = W.seismogram["vx"][:, :nw, :]
        b = O.seismogram_dict(seis)["vx"][:, :nw, :]
        assert rel(a, b) < SEIS_TOL


@pytest.mark.parametrize("chpr", [0, 100])
@pytest.mark.parametrize("acq_type", [0, 2])
def test_exact_gradient_wells(gpu, acq_type, chpr):
    """The exact adjoint injects each record column where it was sampled
    (the crosswell vz column into vx); with chpr = 100 it reads the kept
    states of the forward instead of recomputing them."""
    print(f"wells exact gradient acq_type={acq_type} chpr={chpr}", flush=True)
    W, st, model, _ = wells(acq_type, nt=300, chpr=chpr)
    d = W.forward_modeling(model)
    rng = np.random.default_rng(5)
    res = {k: (rng.standard_normal(v.shape) * np.abs(v).max()).astype(np.float32)
           for k, v in d.items()}
    G = W.raw_gradient(res, mode="exact", ckpt_every=0 if chpr else 37)
    rr = O.residual_records(st, res, np.float64)          # record (reordered) columns
    g3 = O.exact_gradient(st, model, st.flip_wells(rr), np.float64)
    md = st.medium(model, np.float64)
    grho = -(md["b"] ** 2) * g3[2]
    for name, a, b in (("lam", G[2], g3[0]), ("mu", G[0], g3[1]), ("rho", G[4], grho)):
        e = rel(a, b)
        print(f"  {name}: {e:.2e}")
        assert e < GRAD_TOL, (name, e)


@pytest.mark.parametrize("acq_type", [0, 2])
def test_misfit_gradient_wells(gpu, acq_type):
    """pyfwi_misfit_gradient equals forward_modeling -> CostFunction.l2 ->
    gradient for well receivers (records in the reordered layout, residual
    injected through the injection cells)."""
    print(f"wells misfit+gradient acq_type={acq_type}", flush=True)
    W, st, model, model_obs = wells(acq_type)
    d_obs, _ = oracle_section(st, model_obs)
    rms_dev, g_dev = W.misfit_gradient_device(model, d_obs, parameterization="lmd")
    W2, _, _, _ = wells(acq_type)
    d_est = W2.forward_modeling(model)
    rms, res = CostFunction("l2")(d_est, d_obs)
    g = W2.gradient(res, parameterization="lmd")
    assert abs(rms_dev - rms) < 1e-5 * abs(rms), (rms_dev, rms)
    for k in g:
        e = rel(g_dev[k], g[k])
        print(f"  {k}: {e:.2e}")
        assert e < 1e-5, (k, e)
    for name in W.seismogram:
        assert rel(W.seismogram[name], W2.seismogram[name]) < SEIS_TOL


def test_receiver_table_checks(gpu):
    """The plan rejects table entries outside the grid and shots whose
    entries run past the table, before any kernel runs."""
    from pyfwi_amd import _native
    W, st, model, _ = wells(0, nt=50, ns=1)
    plan = W._plan
    ncell = int(W.tnz) * int(W.tnx)
    for bad in ([0, ncell], [-2, 5]):
        with pytest.raises(_native.NativeError):
            plan.set_receivers(np.array(bad, np.int32))
    W.forward_modeling(model)                      # the valid table is still in place
    plan.set_receivers(np.zeros(0, np.int32))      # cleared: table shots now fail
    with pytest.raises(_native.NativeError):
        W.forward_modeling(model)
    plan.set_receivers(W.receiver_table().ravel())
    sh = W._shots([0])
    sh[0].rec_table = 1                            # one entry past the end
    with pytest.raises(_native.NativeError):
        plan.forward(sh, W._source_table(), np.zeros(5 * int(W.nt_ext) * int(W.nr), np.float32),
                     np.zeros(0, np.int32), 0, _native.MEM_HOST)


@pytest.mark.parametrize("acq_type", [0, 2])
def test_fwi_wells_gpu_tensors(gpu, acq_type):
    """Fwi.apply with well acquisitions: the GPU-tensor path (device
    gradient read-out, src/PyFWI/wave_propagation.py:279-331, whose mask for
    crosswell keeps the columns right of the source well) equals the
    CPU-tensor path (the host read-out) and the oracle's 'dv' gradient."""
    import torch
    from pyfwi_amd.torchfwi import Fwi
    print(f"Fwi.apply wells acq_type={acq_type}", flush=True)
    W0, st, model, model_obs = wells(acq_type, nt=300, components=0)
    d_obs, _ = oracle_section(st, model_obs, components=0)
    grads = {}
    outs = {}
    for dev in (gpu, torch.device("cpu")):
        W, _, _, _ = wells(acq_type, nt=300, components=0)
        m = {k: torch.tensor(v, device=dev, requires_grad=True) for k, v in model.items()}
        tx, tz = Fwi.apply(W, m["vp"], m["vs"], m["rho"])
        ox = torch.tensor(d_obs["taux"], device=dev)
        oz = torch.tensor(d_obs["tauz"], device=dev)
        (0.5 * ((tx - ox) ** 2).sum() + 0.5 * ((tz - oz) ** 2).sum()).backward()
        grads[dev.type] = {k: m[k].grad.cpu().numpy() for k in m}
        outs[dev.type] = (tx.detach().cpu().numpy(), tz.detach().cpu().numpy())
    res = {"taux": outs["cuda"][0] - d_obs["taux"], "tauz": outs["cuda"][1] - d_obs["tauz"]}
    _, states = O.forward(st, model, np.float32)
    gr = O.gradient(st, model, res, states, parameterization="dv")
    for k in ("vp", "vs", "rho"):
        e_cpu = rel(grads["cuda"][k], grads["cpu"][k])
        e_orc = rel(grads["cuda"][k], gr[k])
        print(f"  {k}: vs CPU tensors {e_cpu:.2e}, vs oracle {e_orc:.2e}")
        assert e_cpu < 1e-5, (k, e_cpu)
        assert e_orc < GRAD_TOL, (k, e_orc)
    if acq_type == 0:  # the crosswell mask: nothing left of the source well
        sdo, src_col = int(W0.sdo), int(W0.src_cts)
        for k in ("vp", "vs", "rho"):
            assert not grads["cuda"][k][:, :sdo + src_col + 2].any(), k
