"""Uneven shot batches on the GPU: shots with their own nt, receiver aperture
and wavelet in one pyfwi_misfit_gradient call, against the oracle run one shot
at a time (each shot a reference propagator of its own geometry, chpr = 100).

Tolerances as the other parity tests: records <= 1e-5 and gradients <= 1e-4
relative L2 against the oracle; the misfit is compared with the L2 misfit of
the GPU's own records (<= 1e-6).
"""
import numpy as np
import pytest

from oracle import oracle as O
from pyfwi_amd import workloads
from pyfwi_amd.batch import UnevenBatch

from _helpers import rel

pytestmark = pytest.mark.gpu


def _batch():
    wl = workloads.uneven_small()
    B = UnevenBatch(wl.inpa, wl.specs, wl.model_shape, components=2)
    B.set_model(wl.model_obs)
    obs = B.forward()
    B.set_model(wl.model)
    return wl, B, obs


def test_uneven_batch_matches_per_shot_oracle(gpu):
    wl, B, obs = _batch()
    assert len(set(B.nt)) > 1 and len(set(B.nr)) > 1  # really uneven
    seis = np.zeros(B.n_records, np.float32)
    misfit, G = B.misfit_gradient(obs, seis=seis)
    G_ref = np.zeros_like(G)
    m_ref = 0.0
    for k, (sp, P) in enumerate(zip(B.specs, B.props)):
        inpa = dict(wl.inpa, t=sp.t, rec_dis=float(sp.rec_loc[1, 0] - sp.rec_loc[0, 0]))
        st = O.Setup(inpa, np.array([sp.src_loc]), sp.rec_loc, wl.model_shape, P.src.w,
                     P.src.component, chpr=100, components=2)
        rec_o, states = O.forward(st, wl.model, np.float32)
        mine = B.records(seis, k)
        e = rel(mine[:2], rec_o[0, :2])
        print(f"shot {k}: nt={B.nt[k]} nr={B.nr[k]} fdom={sp.fdom:.1f} records {e:.2e}")
        assert e < 1e-5, (k, e)
        ob = B.records(obs, k)
        res = {"vx": np.float32(mine[0] - ob[0]), "vz": np.float32(mine[1] - ob[1])}
        m_ref += sum(0.5 * float(r.astype(np.float64).ravel() @ r.astype(np.float64).ravel())
                     for r in res.values())
        G_ref += O.raw_gradient(st, wl.model, res, states, np.float32)
    assert abs(misfit - m_ref) <= 1e-6 * abs(m_ref), (misfit, m_ref)
    for k in (0, 2, 4):
        e = rel(G[k], G_ref[k])
        print(f"gradient slot {k}: {e:.2e}")
        assert e < 1e-4, (k, e)


@pytest.mark.parametrize("per_pass", [1, 2, 3])
def test_uneven_batch_pass_split(gpu, per_pass):
    """Passes change only the schedule: any split of the shots into passes
    gives the same records, bit for bit, and the same gradient up to float
    rounding (the shots that share a block add their imaging terms before
    they reach the accumulators, so the order of the sums follows how the
    passes group the shots)."""
    wl, B, obs = _batch()
    s1 = np.zeros(B.n_records, np.float32)
    m1, G1 = B.misfit_gradient(obs, seis=s1)
    B.plan.set_tuning(per_pass, False, 0)
    s2 = np.zeros(B.n_records, np.float32)
    m2, G2 = B.misfit_gradient(obs, seis=s2)
    B.plan.set_tuning(0, False, 0)
    np.testing.assert_array_equal(s1, s2)
    for k in (0, 2, 4):
        e = rel(G2[k], G1[k])
        print(f"per_pass={per_pass} gradient slot {k}: {e:.2e}")
        assert e < 1e-5, (k, e)
    assert abs(m1 - m2) <= 1e-12 * abs(m1)


def test_uneven_batch_device_buffers(gpu):
    """All buffers on the device (the source table included) give the same
    records and gradient as host buffers."""
    import torch
    from pyfwi_amd import _native
    wl, B, obs = _batch()
    seis_h = np.zeros(B.n_records, np.float32)
    m_h, G_h = B.misfit_gradient(obs, seis=seis_h)
    dev = torch.device("cuda:0")
    B.plan.set_stream(torch.cuda.current_stream().cuda_stream)
    src = torch.from_numpy(B.src_table).to(dev)
    obs_d = torch.from_numpy(obs).to(dev)
    fwd_d = torch.zeros(B.n_records, dtype=torch.float32, device=dev)
    B.forward(seis=fwd_d, mem=_native.MEM_DEVICE, src_table=src)
    seis_d = torch.zeros_like(fwd_d)
    grad_d = torch.zeros(6 * B.tnz * B.tnx, dtype=torch.float32, device=dev)
    m_d, _ = B.misfit_gradient(obs_d, grad=grad_d, seis=seis_d, mem=_native.MEM_DEVICE,
                               src_table=src)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(seis_d.cpu().numpy(), seis_h)
    np.testing.assert_array_equal(fwd_d.cpu().numpy(), seis_h)
    np.testing.assert_array_equal(grad_d.cpu().numpy().reshape(G_h.shape), G_h)
    assert abs(m_d - m_h) <= 1e-12 * abs(m_h)
