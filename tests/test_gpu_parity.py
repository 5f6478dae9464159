"""GPU parity of the HIP propagator against the CPU oracle (through the C ABI).

Tolerances (FP32 kernels): records <= 1e-5 relative L2 against the oracle's
reference-precision instantiation (float storage, double intermediates);
gradients <= 1e-4 relative L2.  Errors against the FP64 oracle are asserted
at the same bounds where they hold and printed.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pyfwi_amd import workloads
from pyfwi_amd.wave_propagation import WavePropagator

from _helpers import oracle_setup, rel

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4


def _small(nt=400, **over):
    w = workloads.plain_small()
    w.inpa["t"] = workloads.t_for_nt(nt, w.inpa["dt"])
    w.inpa.update(over)
    return w


def test_forward_small_all_components(gpu):
    w = _small(nt=1000)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=4, chpr=0)
    d = W.forward_modeling(w.model)
    st = oracle_setup(w, chpr=0, components=4)
    seis, _ = O.forward(st, w.model, np.float32, keep_states=False)
    seis_d, _ = O.forward(st, w.model, np.float64, keep_states=False)
    ref, ref_d = O.section(seis, 4), O.section(seis_d, 4)
    for k in ref:
        e, e64 = rel(d[k], ref[k]), rel(d[k], ref_d[k])
        print(f"forward {k}: vs oracle {e:.2e}, vs FP64 {e64:.2e}")
        assert e < SEIS_TOL and e64 < SEIS_TOL, (k, e, e64)


def test_forward_order8_and_dt_scale(gpu):
    # order 8 stencil, dt above 0.6 ms so that two internal steps make one sample
    w = workloads.plain_small()
    w.inpa.update({"sdo": 8, "dh": 5.0, "dt": 0.0008, "rec_dis": 10.0})
    w.inpa["t"] = 0.16
    from pyfwi_amd.acquisition import Source
    src = Source(np.c_[w.src.i, w.src.j] * 5.0, 5.0, 0.0008)
    src.Ricker(15)
    rec = w.rec_loc * 5.0
    W = WavePropagator(w.inpa, src, rec, w.model_shape, components=4, chpr=0)
    assert W.dt_scale == 2 and W.nt != W.nt_ext
    d = W.forward_modeling(w.model)
    st = O.Setup(w.inpa, np.c_[src.i, src.j] * 5.0, rec, w.model_shape, src.w, src.component,
                 chpr=0, components=4)
    seis, _ = O.forward(st, w.model, np.float32, keep_states=False)
    ref = O.section(seis, 4)
    for k in ref:
        assert rel(d[k], ref[k]) < SEIS_TOL, k


@pytest.mark.parametrize("components,shape", [(0, "2d"), (2, "3d"), (1, "2d"), (3, "3d")])
def test_sections(gpu, components, shape):
    w = _small(nt=200, seimogram_shape=shape)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=components)
    d = W.forward_modeling(w.model)
    st = oracle_setup(w, chpr=0, components=components)
    st.seismo_shape = shape
    seis, _ = O.forward(st, w.model, np.float32, keep_states=False)
    ref = O.section(seis, components, shape)
    assert set(d) == set(ref)
    for k in ref:
        assert d[k].shape == ref[k].shape
        assert rel(d[k], ref[k]) < SEIS_TOL


def test_batched_equals_per_shot(gpu):
    w = _small(nt=300)
    a = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=4)
    da = a.forward_modeling(w.model)
    inpa = dict(w.inpa, per_shot=True)
    b = WavePropagator(inpa, w.src, w.rec_loc, w.model_shape, components=4)
    db = b.forward_modeling(w.model)
    for k in da:
        np.testing.assert_array_equal(da[k], db[k])


@pytest.mark.parametrize("chpr", [100, 0, 20])
def test_gradient_pyfwi_small(gpu, chpr):
    w = _small(nt=600)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=chpr)
    d_est = W.forward_modeling(w.model)
    obs = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2)
    d_obs = obs.forward_modeling(w.model_obs)
    res = {k: np.float32(d_est[k] - d_obs[k]) for k in d_est}
    g = W.gradient(res, parameterization="lmd")
    st = oracle_setup(w, chpr=chpr, components=2)
    seis, states = O.forward(st, w.model, np.float32)
    gr = O.gradient(st, w.model, res, states, parameterization="lmd")
    seis_d, states_d = O.forward(st, w.model, np.float64)
    gr_d = O.gradient(st, w.model, res, states_d, parameterization="lmd", dtype=np.float64)
    for k in gr:
        e, e64 = rel(g[k], gr[k]), rel(g[k], gr_d[k])
        print(f"pyfwi gradient chpr={chpr} {k}: vs oracle {e:.2e}, vs FP64 {e64:.2e}")
        assert e < GRAD_TOL and e64 < GRAD_TOL, (k, e, e64)
    # the 'dv' parameterisation applies the reference's chain rule
    gv = W.gradient(res, parameterization="dv")
    md = st.medium(w.model, np.float32)
    c = (slice(20, -20), slice(20, -20))
    ref = O.to_vd(gr["lam"], gr["mu"], gr["rho"], md["lam32"][c], md["mu32"][c], md["rho"][c])
    for k, r in zip(("vp", "vs", "rho"), ref):
        assert rel(gv[k], r) < GRAD_TOL


@pytest.mark.parametrize("chpr,tuning,balance", [(100, None, False), (0, None, False),
                                                 (100, (1, False, 0), False), (100, None, True)])
def test_gradient_order8(gpu, chpr, tuning, balance):
    """Order-8 reference gradient (kept states and time reversal; one shot a
    pass; energy balancing) against the oracle.  The balanced case runs 500
    steps: at 400 the wave has not reached every cell, and the balancing's
    0/0 there gives NaN in the reference's arithmetic as in the oracle."""
    print(f"order-8 gradient chpr={chpr} tuning={tuning} balance={balance}", flush=True)
    w = _small(nt=500 if balance else 400, sdo=8, energy_balancing=balance)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=chpr)
    if tuning:
        W._plan.set_tuning(*tuning)
    d = W.forward_modeling(w.model)
    res = {k: np.float32(0.2 * v) for k, v in d.items()}
    g = W.gradient(res, parameterization="lmd")
    st = oracle_setup(w, chpr=chpr, components=2)
    st.energy_balancing = balance
    seis, states = O.forward(st, w.model, np.float32)
    for k, r in O.section(seis, 2).items():
        assert rel(d[k], r) < SEIS_TOL, k
    gr = O.gradient(st, w.model, res, states, parameterization="lmd")
    for k in gr:
        e = rel(g[k], gr[k])
        print(f"  {k}: {e:.2e}")
        assert e < GRAD_TOL, (k, e)


@pytest.mark.parametrize("chpr,ns", [(100, 10), (0, 10), (100, 13)])
def test_gradient_many_shot_groups(gpu, chpr, ns):
    """Passes of 10 shots run in two blocks of five that each sum into their
    own accumulator set, 13 shots in three of four plus one; both against the
    oracle."""
    print(f"gradient chpr={chpr} shots={ns}", flush=True)
    from pyfwi_amd.acquisition import Source
    w = _small(nt=300)
    src_loc, _ = workloads._surface_geometry(ns, 100, 1.0, 50, 2, 4, 2, 1)
    w.src = Source(src_loc, 1.0, w.inpa["dt"])
    w.src.Ricker(30)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=chpr)
    d = W.forward_modeling(w.model)
    res = {k: np.float32(0.2 * v) for k, v in d.items()}
    g = W.gradient(res, parameterization="lmd")
    st = oracle_setup(w, chpr=chpr, components=2)
    seis, states = O.forward(st, w.model, np.float32)
    gr = O.gradient(st, w.model, res, states, parameterization="lmd")
    for k in gr:
        e = rel(g[k], gr[k])
        print(f"  {k}: {e:.2e}")
        assert e < GRAD_TOL, (k, e)


def test_gradient_energy_balancing(gpu):
    w = _small(nt=500, energy_balancing=True)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=100)
    d = W.forward_modeling(w.model)
    res = {k: np.float32(0.1 * d[k]) for k in d}
    g = W.gradient(res, parameterization="lmd")
    st = oracle_setup(w, chpr=100, components=2)
    st.energy_balancing = True
    seis, states = O.forward(st, w.model, np.float32)
    gr = O.gradient(st, w.model, res, states, parameterization="lmd")
    for k in gr:
        assert rel(g[k], gr[k]) < GRAD_TOL, k


@pytest.mark.parametrize("chpr,nt", [(5, 200), (100, 120)])
def test_snapshots_match_oracle(gpu, chpr, nt):
    """Kept states: copies at the checkpoint steps (chpr=5), or every step
    written slot to slot by the forward kernels with the source injection
    taken out of each slot again (chpr=100)."""
    w = _small(nt=nt)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=chpr)
    W.forward_modeling(w.model)
    st = oracle_setup(w, chpr=chpr, components=2)
    _, states = O.forward(st, w.model, np.float32)
    Wd = W.W
    for f, name in enumerate(O.FIELDS):
        for c in range(len(st.chp)):
            a = Wd[name][:, :, :, c].transpose(2, 0, 1)
            b = states[:, c, f]
            assert rel(a, b) < SEIS_TOL


def test_exact_gradient_vs_fp64(gpu):
    w = _small(nt=300)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=4, chpr=0)
    d = W.forward_modeling(w.model)
    rng = np.random.default_rng(7)
    res = {k: rng.standard_normal(v.shape).astype(np.float32) * np.abs(v).max() for k, v in d.items()}
    G = W.raw_gradient(res, mode="exact", ckpt_every=37)
    st = oracle_setup(w, chpr=0, components=4)
    rr = O.residual_records(st, res, np.float64)
    g3 = O.exact_gradient(st, w.model, rr, np.float64)
    md = st.medium(w.model, np.float64)
    grho = -(md["b"] ** 2) * g3[2]
    for name, a, b in (("lam", G[2], g3[0]), ("mu", G[0], g3[1]), ("rho", G[4], grho)):
        e = rel(a, b)
        print(f"exact gradient {name} vs FP64: {e:.2e}")
        assert e < GRAD_TOL, (name, e)


@pytest.mark.parametrize("dt_scale", [1, 2])
def test_exact_gradient_from_kept_states(gpu, dt_scale):
    """After a forward that kept every step (chpr = 100) the exact adjoint
    reads the kept states instead of recomputing segments from checkpoints:
    the same gradient as the recomputing path (same values, same shot-group
    sums) and within the FP64 oracle's tolerance."""
    w = _small(nt=300)
    if dt_scale == 2:  # two internal steps a sample
        from pyfwi_amd.acquisition import Source
        w.inpa.update({"dh": 5.0, "dt": 0.0008, "rec_dis": 10.0})
        w.inpa["t"] = workloads.t_for_nt(300, 0.0008)
        w.src = Source(np.c_[w.src.i, w.src.j] * 5.0, 5.0, 0.0008)
        w.src.Ricker(15)
        w.rec_loc = w.rec_loc * 5.0
    out = {}
    for chpr in (100, 0):
        W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=4, chpr=chpr)
        assert int(W.dt_scale) == dt_scale
        d = W.forward_modeling(w.model)
        rng = np.random.default_rng(7)
        res = {k: rng.standard_normal(v.shape).astype(np.float32) * np.abs(v).max()
               for k, v in d.items()}
        out[chpr] = W.raw_gradient(res, mode="exact", ckpt_every=0 if chpr else 37)
        if chpr:
            # the kept-state path ran: no checkpoint interval was used
            assert W._plan.last_passes(with_interval=True)[1] == 0
    st = oracle_setup(w, chpr=0, components=4)
    g3 = O.exact_gradient(st, w.model, O.residual_records(st, res, np.float64), np.float64)
    md = st.medium(w.model, np.float64)
    for k, ref in ((2, g3[0]), (0, g3[1]), (4, -(md["b"] ** 2) * g3[2])):
        e_path = rel(out[100][k], out[0][k])
        e64 = rel(out[100][k], ref)
        same = np.array_equal(out[100][k], out[0][k])
        print(f"exact from kept states, slot {k}: vs recomputed {e_path:.2e}"
              f"{' (bit-equal)' if same else ''}, vs FP64 {e64:.2e}", flush=True)
        assert e_path < 1e-6 and e64 < GRAD_TOL, (k, e_path, e64)


@pytest.mark.parametrize("tuning", [(2, False, 0), (3, True, 0), (1, True, 0), (0, False, 16),
                                    (2, True, 12)])
def test_execution_shapes_agree(gpu, tuning):
    """Shots per pass, fused or two-launch steps and the strip height change
    only the schedule: records and reference gradients match the oracle, and
    the exact gradient (recomputed per pass when the forward's checkpoints
    are not reused) matches the FP64 oracle."""
    w = _small(nt=300)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=4, chpr=0)
    W._plan.set_tuning(*tuning)
    d = W.forward_modeling(w.model)
    st = oracle_setup(w, chpr=0, components=4)
    seis, states = O.forward(st, w.model, np.float32)
    ref = O.section(seis, 4)
    for k in ref:
        assert rel(d[k], ref[k]) < SEIS_TOL, (k, tuning)
    res = {k: np.float32(0.3 * v) for k, v in d.items()}
    g = W.gradient(res, parameterization="lmd")
    gr = O.gradient(st, w.model, res, states, parameterization="lmd")
    for k in gr:
        assert rel(g[k], gr[k]) < GRAD_TOL, (k, tuning)
    G = W.raw_gradient(res, mode="exact", ckpt_every=41)
    g3 = O.exact_gradient(st, w.model, O.residual_records(st, res, np.float64), np.float64)
    md = st.medium(w.model, np.float64)
    for a, b in ((G[2], g3[0]), (G[0], g3[1]), (G[4], -(md["b"] ** 2) * g3[2])):
        assert rel(a, b) < GRAD_TOL, tuning


@pytest.mark.parametrize("chpr,comp,tuning", [(100, 0, (2, False, 0)), (0, 0, (0, True, 0)),
                                              (100, 4, (3, False, 0)), (100, 2, (0, True, 0))])
def test_misfit_gradient_matches_separate_calls(gpu, chpr, comp, tuning):
    """pyfwi_misfit_gradient (forward, L2 residual and reference adjoint per
    pass) equals forward_modeling -> CostFunction.l2 -> gradient."""
    from pyfwi_amd import _native
    from pyfwi_amd.fwi_tools import CostFunction
    w = _small(nt=300)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=comp, chpr=chpr)
    Wo = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=comp, chpr=chpr)
    d_obs = Wo.forward_modeling(w.model_obs)
    d_est = W.forward_modeling(w.model)
    rms, res = CostFunction("l2")(d_est, d_obs)
    G_ref = W.raw_gradient(res)
    # observed records in the plan's layout [ns, 5, nt, nr]
    obs = np.zeros((int(W.ns), 5, int(W.nt_ext), int(W.nr)), np.float32)
    for f, name in enumerate(("vx", "vz", "taux", "tauz", "tauxz")):
        if name in d_obs:
            obs[:, f] = np.reshape(d_obs[name], (int(W.nt_ext), int(W.ns), int(W.nr))).transpose(1, 0, 2)
    mask = 0 if comp == 0 else {2: 0b00011, 4: 0b11111}[comp]
    W._plan.set_tuning(*tuning)
    G = np.zeros_like(G_ref)
    seis = np.zeros_like(obs)
    misfit = W._plan.misfit_gradient(W._shots(list(range(int(W.ns)))), W._source_table(), obs,
                                     seis, G, W.chp, mask, 0, _native.MEM_HOST)
    assert abs(misfit - rms) < 1e-5 * abs(rms), (misfit, rms)
    for k in (0, 2, 4):
        assert rel(G[k], G_ref[k]) < 1e-5, (k, rel(G[k], G_ref[k]))


def test_sparse_terms_off_active_cells(gpu):
    """Without an absorbing layer a source or receiver can sit on the
    never-updated outer ring (row 1 < sdo here); the reference adds its terms
    in memory there, where they accumulate.  Records and the reference
    gradient must still match the oracle."""
    w = _small(nt=300, npml=0)
    w.src.j[:] = 1
    w.rec_loc[:, 1] = 1.0
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=100)
    d = W.forward_modeling(w.model)
    st = oracle_setup(w, chpr=100, components=2)
    seis, states = O.forward(st, w.model, np.float32)
    ref = O.section(seis, 2)
    for k in ref:
        assert rel(d[k], ref[k]) < SEIS_TOL, k
    res = {k: np.float32(0.3 * v) for k, v in d.items()}
    g = W.gradient(res, parameterization="lmd")
    gr = O.gradient(st, w.model, res, states, parameterization="lmd")
    for k in gr:
        assert rel(g[k], gr[k]) < GRAD_TOL, k


def test_loop_timing(gpu):
    """Timing mode 2 brackets only the time loops: one forward and one
    backward bracket per pass, counted in steps, and no launch brackets."""
    from pyfwi_amd import _native
    w = _small(nt=120)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=100)
    d_obs = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape,
                           components=2).forward_modeling(w.model_obs)
    W._plan.set_tuning(3, False, 0)   # passes of 3 and 2 shots
    W._plan.set_timing(_native.TIMING_LOOPS)
    W.misfit_gradient_device(w.model, d_obs, parameterization="lmd")
    fms, fn = W._plan.kernel_timing(_native.K_FWD_LOOP)
    bms, bn = W._plan.kernel_timing(_native.K_BWD_LOOP)
    assert (fn, bn) == (2 * int(W.nt), 2 * (int(W.nt) - 1))
    assert fms > 0 and bms > 0
    assert W._plan.kernel_timing(_native.K_ADJ_VEL)[1] == 0
    with pytest.raises(_native.NativeError):
        W._plan.set_timing(3)
    W._plan.set_timing(0)


def test_exact_schedule_sized_to_memory(gpu):
    """The exact adjoint with the automatic two-level schedule (forward with
    PYFWI_CKPT_AUTO, gradient with interval 0: checkpoints every
    ceil(sqrt(nt)) steps, kept by the forward, passes from the free memory)
    equals the one with checkpoints every 50 steps recomputed by the gradient
    -- the recomputed states are the forward's bit for bit, so the two differ
    only by the order of the shot groups' sums -- and reports its schedule."""
    import math
    import torch
    from pyfwi_amd import _native
    from pyfwi_amd.fwi_tools import expand_model
    w = _small(nt=400)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=4, chpr=0)
    m = {k: expand_model(v, int(W.tnz), int(W.tnx), W.npml) for k, v in w.model.items()}
    W.pml_preparation(m["vp"].max())
    W.elastic_buffers(m)
    plan = W._plan
    plan.set_stream(torch.cuda.current_stream().cuda_stream)
    dev = torch.device("cuda:0")
    shots = W._shots(list(range(int(W.ns))))
    src = torch.from_numpy(W._source_table()).to(dev)
    n = int(W.ns) * 5 * int(W.nt_ext) * int(W.nr)
    seis = torch.zeros(n, device=dev)
    grads = []
    for fwd_ck, grad_ck in ((_native.CKPT_AUTO, 0), (0, 50)):
        plan.forward(shots, src, seis, None, fwd_ck, _native.MEM_DEVICE)
        res = seis * 0.3
        g = torch.zeros(6 * int(W.tnz) * int(W.tnx), device=dev)
        plan.gradient(shots, src, res, g, _native.MODE_EXACT, grad_ck, 0, _native.MEM_DEVICE)
        torch.cuda.synchronize()
        passes, k = plan.last_passes(with_interval=True)
        print(f"exact schedule: forward ckpt {fwd_ck}, gradient ckpt {grad_ck}: passes {passes}, "
              f"interval {k}", flush=True)
        assert k == (math.ceil(math.sqrt(int(W.nt))) if grad_ck == 0 else grad_ck)
        assert sum(passes) == int(W.ns)
        grads.append(g.cpu().numpy())
    e = rel(grads[0], grads[1])
    print(f"  automatic vs every-50 schedule: {e:.2e}")
    assert np.isfinite(grads[0]).all() and e < 1e-6
