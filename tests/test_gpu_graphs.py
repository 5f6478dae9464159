"""hipGraph replay of the time loops (pyfwi_plan_set_graphs) against eager
launches of the same calls: records, gradients and misfits equal bit for bit
on every replay, with eager work and synchronisations between the replays
(the graph must not depend on state the eager calls change), for the kept-
state misfit + gradient, the forward + reference gradient with time reversal
between checkpoints, and receiver tables (wells + surface)."""
import numpy as np
import pytest

from pyfwi_amd import _native, workloads
from pyfwi_amd.fwi_tools import expand_model
from pyfwi_amd.wave_propagation import WavePropagator, checkpoint_steps

pytestmark = pytest.mark.gpu


def _device(w, torch, chpr=100, components=2, n_well_rec=0):
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=components,
                       n_well_rec=n_well_rec, chpr=chpr)
    m = {k: expand_model(v, int(W.tnz), int(W.tnx), W.npml) for k, v in w.model.items()}
    W.pml_preparation(m["vp"].max())
    W.elastic_buffers(m)
    W._plan.set_stream(torch.cuda.current_stream().cuda_stream)
    dev = torch.device("cuda:0")
    ns = int(W.ns)
    n = ns * 5 * int(W.nt_ext) * int(W.nr)
    b = dict(shots=W._shots(list(range(ns))), src=torch.from_numpy(W._source_table()).to(dev),
             obs=torch.zeros(n, device=dev), seis=torch.zeros(n, device=dev),
             res=torch.zeros(n, device=dev),
             grad=torch.zeros(6 * int(W.tnz) * int(W.tnx), device=dev))
    # observed data: the anomaly model's records
    mo = {k: expand_model(v, int(W.tnz), int(W.tnx), W.npml) for k, v in w.model_obs.items()}
    W.elastic_buffers(mo)
    W._plan.forward(b["shots"], b["src"], b["obs"], None, 0, _native.MEM_DEVICE)
    W.elastic_buffers(m)
    torch.cuda.synchronize()
    return W, b


def _misfit_step(W, b):
    chp = checkpoint_steps(int(W.nt), 100)
    return W._plan.misfit_gradient(b["shots"], b["src"], b["obs"], b["seis"], b["grad"], chp,
                                   3, 0, _native.MEM_DEVICE)


def _reverse_step(W, b):
    W._plan.forward(b["shots"], b["src"], b["seis"], W.chp, 0, _native.MEM_DEVICE)
    b["res"].copy_(b["seis"]).sub_(b["obs"])
    W._plan.gradient(b["shots"], b["src"], b["res"], b["grad"], _native.MODE_PYFWI, 0,
                     _native.F_PRECOND, _native.MEM_DEVICE)
    return float(b["res"].square().sum().item())


def _compare(W, b, step, torch, replays=3):
    eager_m = step(W, b)
    torch.cuda.synchronize()
    want = (b["seis"].clone(), b["grad"].clone())
    W._plan.set_graphs(True)
    c0, r0 = W._plan.graph_stats()
    for k in range(replays + 1):  # the first call captures, the others replay
        b["seis"].fill_(float("nan"))
        b["grad"].fill_(float("nan"))
        m = step(W, b)
        torch.cuda.synchronize()
        assert torch.equal(b["seis"], want[0]), k
        assert torch.equal(b["grad"], want[1]), k
        # the misfit's block sums meet in double atomics: order-dependent bits
        assert m == pytest.approx(eager_m, rel=1e-12), k
        # eager work between the replays, then a synchronisation
        b["res"].mul_(0.5).add_(1.0)
        torch.cuda.synchronize()
    c1, r1 = W._plan.graph_stats()
    print(f"graphs: {c1 - c0} loops captured, {r1 - r0} replayed", flush=True)
    assert c1 > c0 and r1 >= replays * (c1 - c0)
    W._plan.set_graphs(False)


def test_graph_misfit_gradient_kept_states(gpu):
    import torch
    w = workloads.plain_small()
    w.inpa["t"] = workloads.t_for_nt(300, w.inpa["dt"])
    W, b = _device(w, torch)
    _compare(W, b, _misfit_step, torch)


def test_graph_forward_gradient_reversal(gpu):
    import torch
    w = workloads.plain_small()
    w.inpa["t"] = workloads.t_for_nt(300, w.inpa["dt"])
    W, b = _device(w, torch, chpr=0)
    _compare(W, b, _reverse_step, torch)


def test_graph_receiver_tables(gpu):
    """Wells + surface receivers (receiver table, gathers from the kept
    slots, residual injection launches inside the captured loop)."""
    import torch
    from pyfwi_amd.acquisition import Source, acq_parameters
    dh, dt, nz, nx, nt = 5.0, 5e-4, 60, 80, 250
    inpa = {"dh": dh, "dt": dt, "t": workloads.t_for_nt(nt, dt), "npml": 10, "pmlR": 1e-5,
            "pml_dir": 2, "sdo": 4, "rec_dis": 2 * dh, "acq_type": 2}
    src_loc, rec_loc, _, nw = acq_parameters(3, 2 * dh, nx * dh, nz * dh, dh, 4, 2)
    src = Source(src_loc, dh, dt)
    src.Ricker(25)
    vp = np.full((nz, nx), 2200.0, np.float32)
    vp[30:] = 2600.0
    model = {"vp": vp, "vs": (vp / 1.8).astype(np.float32), "rho": (1900 + 0.1 * vp).astype(np.float32)}
    obs = dict(model, vp=(vp * 1.03).astype(np.float32))
    w = workloads.Workload("wells", inpa, src, rec_loc, model, obs, components=4)
    W, b = _device(w, torch, chpr=100, components=4, n_well_rec=nw)
    _compare(W, b, _misfit_step, torch)
