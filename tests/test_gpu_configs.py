"""GPU parity at the benchmark configurations' real grid sizes (through the C
ABI), against the CPU oracle -- float storage with double intermediates (the
reference's arithmetic) and FP64 -- on shot subsets or truncated time axes
the oracle finishes in seconds, plus size-independent properties at full
size:

  C1 plain_small     100x100, 5 shots, nt=1000: records and the reference
                     gradient (every state kept), as the configuration says
  C2 plain_medium    500x1000 (540x1040 with the absorbing layer), 2 of the
                     20 shots, nt=2000, time reversal between the reference's
                     two checkpoints (chpr=0)
  C3 differentiation 300x600, 2 of the 10 shots, nt=2500: reference-mode
                     gradient (chpr=0)
  C4 uneven_batch    1000x2000 (1040x2040): the two shortest uneven shots, each
                     with its own nt (truncated), receivers and wavelet,
                     batched with every state kept -- the benchmark's path;
                     at full nt all 64 shots batched against one shot a pass
  C5 large_step      2000x6000 (2040x6040): one shot, nt=200 (chpr=0); four
                     shots split into passes of 1, 2 and 4

Bounds: records <= 1e-5 relative L2, gradients <= 1e-4, against the oracle and
against FP64 (the reference-mode gradient's FP64 error is printed and
asserted for every case).
"""
import copy

import numpy as np
import pytest

from oracle import oracle as O
from pyfwi_amd import _native, workloads
from pyfwi_amd.wave_propagation import WavePropagator, checkpoint_steps

from _helpers import oracle_setup, rel

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4


def _release(*plans):
    """Give the plans' device buffers back at once (kept states of the large
    grids take most of the HBM), with torch's cached blocks."""
    import gc
    import torch
    for p in plans:
        p.close()
    gc.collect()
    torch.cuda.empty_cache()


def _parity(w, chpr, components, res_scale=0.5, label=""):
    """Forward + reference gradient of workload w on the GPU against the
    float and FP64 oracles; the residual is res_scale times the records."""
    print(f"{label}: {w.name}, {len(w.src.i)} shots, grid {w.model_shape}, chpr={chpr}",
          flush=True)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=components,
                       chpr=chpr)
    d = W.forward_modeling(w.model)
    res = {k: np.float32(res_scale * v) for k, v in d.items()}
    g = W.gradient(res, parameterization="lmd")
    _release(W._plan)
    st = oracle_setup(w, chpr=chpr, components=components)
    for dtype in (np.float32, np.float64):
        seis, states = O.forward(st, w.model, dtype)
        ref = O.section(seis, components)
        for k in ref:
            e = rel(d[k], ref[k])
            print(f"  records {k} vs {np.dtype(dtype).name} oracle: {e:.2e}", flush=True)
            assert e < SEIS_TOL, (label, k, dtype, e)
        gr = O.gradient(st, w.model, res, states, parameterization="lmd", dtype=dtype)
        del states
        for k in gr:
            e = rel(g[k], gr[k])
            print(f"  gradient {k} vs {np.dtype(dtype).name} oracle: {e:.2e}", flush=True)
            assert e < GRAD_TOL, (label, k, dtype, e)


def test_c1_gradient_full_length(gpu):
    w = workloads.plain_small()
    assert w.meta["nt"] == 1000
    _parity(w, chpr=100, components=2, label="C1")


def test_c2_two_shots_full_grid(gpu):
    w = workloads.subset(workloads.plain_medium(), [0, 13])
    _parity(w, chpr=0, components=0, label="C2")


def test_c3_reference_gradient(gpu):
    w = workloads.subset(workloads.differentiation(), [2, 7])
    _parity(w, chpr=0, components=0, label="C3")


def test_c5_one_shot_truncated(gpu):
    w = workloads.subset(workloads.large_step(), [16], nt=200)
    _parity(w, chpr=0, components=0, res_scale=1.0, label="C5")


def test_c5_pass_split_invariance(gpu):
    """Four C5 shots (nt=150) in passes of 4, 2 and 1: the records are the
    same bit for bit; the gradient differs by the order of the float sums of
    the shot groups only."""
    import torch
    w = workloads.subset(workloads.large_step(), [0, 9, 20, 31], nt=150)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=0, chpr=100)
    from pyfwi_amd.fwi_tools import expand_model
    m = {k: expand_model(v, int(W.tnz), int(W.tnx), W.npml) for k, v in w.model.items()}
    W.pml_preparation(m["vp"].max())
    W.elastic_buffers(m)
    plan = W._plan
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    shots = W._shots(list(range(4)))
    dev = torch.device("cuda:0")
    src = torch.from_numpy(W._source_table()).to(dev)
    n = 4 * 5 * int(W.nt_ext) * int(W.nr)
    obs = torch.zeros(n, dtype=torch.float32, device=dev)
    chp = checkpoint_steps(int(W.nt), 100)
    outs = []
    for g in (4, 2, 1):
        plan.set_tuning(g, False, 0)
        seis = torch.zeros(n, dtype=torch.float32, device=dev)
        grad = torch.zeros(6 * int(W.tnz) * int(W.tnx), dtype=torch.float32, device=dev)
        misfit = plan.misfit_gradient(shots, src, obs, seis, grad, chp, 0, 0,
                                      _native.MEM_DEVICE)
        torch.cuda.synchronize()
        assert plan.last_passes() == [g] * (4 // g)
        outs.append((seis, grad, misfit))
        print(f"C5 passes of {g}: misfit {misfit:.6e}", flush=True)
    _release(plan)
    for seis, grad, misfit in outs[1:]:
        assert torch.equal(seis, outs[0][0])
        e = rel(grad.cpu().numpy(), outs[0][1].cpu().numpy())
        print(f"  gradient vs passes of 4: {e:.2e}")
        assert e < 1e-5 and torch.isfinite(grad).all()
        assert misfit == pytest.approx(outs[0][2], rel=1e-12)


def _uneven_pair(nts):
    """The two shortest shots of uneven_batch with their own receivers and
    wavelets, truncated to nts steps."""
    from pyfwi_amd.batch import UnevenBatch
    wl = workloads.uneven_batch()
    dt = wl.inpa["dt"]
    order = np.argsort([s.t for s in wl.specs])[:2]
    specs = [copy.copy(wl.specs[int(k)]) for k in order]
    for sp, nt in zip(specs, nts):
        sp.t = workloads.t_for_nt(nt, dt)
    return wl, specs, UnevenBatch(wl.inpa, specs, wl.model_shape, components=2)


def test_c4_uneven_pair_kept_states(gpu):
    """The benchmark's path (pyfwi_misfit_gradient, every state kept, shots
    with their own nt, receivers and wavelet in one batch) on the full C4 grid
    against the oracle run shot by shot; observed data zero, so the residual
    is the records."""
    import torch
    wl, specs, B = _uneven_pair((300, 240))
    print(f"C4 pair: nt {B.nt}, receivers {B.nr}, fdom "
          f"{[round(s.fdom, 2) for s in specs]}", flush=True)
    B.set_model(wl.model)
    dev = torch.device("cuda:0")
    obs = torch.zeros(B.n_records, dtype=torch.float32, device=dev)
    seis = torch.zeros_like(obs)
    grad = torch.zeros(6 * B.tnz * B.tnx, dtype=torch.float32, device=dev)
    B.plan.set_stream(torch.cuda.current_stream().cuda_stream)
    src = torch.from_numpy(B.src_table).to(dev)
    misfit, _ = B.misfit_gradient(obs, grad=grad, seis=seis, mem=_native.MEM_DEVICE,
                                  src_table=src)
    torch.cuda.synchronize()
    seis, G = seis.cpu().numpy(), grad.cpu().numpy().reshape(6, B.tnz, B.tnx)
    assert B.plan.last_passes() == [2]
    _release(B.plan)
    for dtype in (np.float32, np.float64):
        G_ref = np.zeros((6, B.tnz, B.tnx), np.float64)
        misfit_ref = 0.0
        for k, sp in enumerate(specs):
            inpa = dict(wl.inpa, t=sp.t, rec_dis=float(sp.rec_loc[1, 0] - sp.rec_loc[0, 0]))
            P = B.props[k]
            st = O.Setup(inpa, np.array([sp.src_loc]), sp.rec_loc, wl.model_shape, P.src.w,
                         P.src.component, chpr=100, components=2)
            rec, states = O.forward(st, wl.model, dtype)
            e = rel(B.records(seis, k), rec[0])
            print(f"  shot {k}: records vs {np.dtype(dtype).name} oracle {e:.2e}", flush=True)
            assert e < SEIS_TOL, (k, dtype, e)
            res2d = O.section(rec.astype(np.float32), 2)  # residual = records (vx, vz)
            misfit_ref += 0.5 * sum(float(np.sum(np.float64(v) ** 2)) for v in res2d.values())
            G_ref += O.raw_gradient(st, wl.model, res2d, states, dtype)
            del states
        for slot, name in ((0, "mu"), (2, "lam"), (4, "rho")):
            win = (slice(20 + 2, B.tnz - 20 - 2), slice(20 + 2, B.tnx - 20 - 2))
            e = rel(G[slot][win], G_ref[slot][win])
            print(f"  gradient {name} vs {np.dtype(dtype).name} oracle: {e:.2e}", flush=True)
            assert e < GRAD_TOL, (name, dtype, e)
        assert misfit == pytest.approx(misfit_ref, rel=1e-4)


def test_c4_all_shots_batched_equals_per_shot(gpu):
    """All 64 uneven shots at full nt: batched (shots packed first-fit into
    passes whose time-major kept states fill the free HBM) against one shot a
    pass -- records bit for bit, gradient to the float
    rounding of the shot groups' sums, misfit to 1e-9."""
    import torch
    from pyfwi_amd.batch import UnevenBatch
    wl = workloads.uneven_batch()
    B = UnevenBatch(wl.inpa, wl.specs, wl.model_shape, components=wl.components)
    dev = torch.device("cuda:0")
    B.plan.set_stream(torch.cuda.current_stream().cuda_stream)
    src = torch.from_numpy(B.src_table).to(dev)
    B.set_model(wl.model_obs)
    obs = B.forward(seis=torch.zeros(B.n_records, dtype=torch.float32, device=dev),
                    mem=_native.MEM_DEVICE, src_table=src)
    B.set_model(wl.model)
    outs = []
    for per_pass in (0, 1):
        B.plan.set_tuning(per_pass, False, 0)
        seis = torch.zeros_like(obs)
        grad = torch.zeros(6 * B.tnz * B.tnx, dtype=torch.float32, device=dev)
        misfit, _ = B.misfit_gradient(obs, grad=grad, seis=seis, mem=_native.MEM_DEVICE,
                                      src_table=src)
        torch.cuda.synchronize()
        passes = B.plan.last_passes()
        order = B.plan.last_order()
        print(f"C4 64 shots, passes {passes}: misfit {misfit:.8e}", flush=True)
        # the passes: consecutive runs of a permutation of the shots, each by
        # decreasing nt; packed first-fit, the longest shot shares its pass
        assert sorted(order) == list(range(B.ns)) and sum(passes) == B.ns
        k = 0
        for n in passes:
            nts = [B.nt[j] for j in order[k:k + n]]
            assert nts == sorted(nts, reverse=True), nts
            k += n
        if per_pass == 0:
            assert B.nt[order[0]] == max(B.nt) and passes[0] >= 2
            print(f"  first pass nt {[B.nt[j] for j in order[:passes[0]]]}", flush=True)
        outs.append((seis, grad, misfit))
    B.plan.set_tuning(0, False, 0)
    _release(B.plan)
    assert torch.equal(outs[0][0], outs[1][0])
    e = rel(outs[0][1].cpu().numpy(), outs[1][1].cpu().numpy())
    print(f"  gradient batched vs per-shot: {e:.2e}")
    assert e < 1e-5 and bool(torch.isfinite(outs[0][1]).all())
    assert outs[0][2] == pytest.approx(outs[1][2], rel=1e-9)
