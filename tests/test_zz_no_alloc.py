"""pyfwi_forward_noalloc (allocation-free forward modelling through the C ABI)
against pyfwi_forward on the same inputs: records equal bit for bit, on two
calls in a row -- the second with fewer shots and fewer steps, so that it
runs on what the first left in the shared scratch array -- and an input
larger than the scratch array is refused before anything runs."""
import numpy as np
import pytest

from pyfwi_amd import _native, workloads
from pyfwi_amd.fwi_tools import expand_model
from pyfwi_amd.wave_propagation import WavePropagator

pytestmark = pytest.mark.gpu


def _plan(w, torch):
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=4, chpr=0)
    m = {k: expand_model(v, int(W.tnz), int(W.tnx), W.npml) for k, v in w.model.items()}
    W.pml_preparation(m["vp"].max())
    W.elastic_buffers(m)
    W._plan.set_stream(torch.cuda.current_stream().cuda_stream)
    return W


def _records(W, torch, ids, nt, noalloc):
    import copy
    dev = torch.device("cuda:0")
    shots = [copy.copy(s) for s in W._shots(ids)]
    for s in shots:
        s.nt = nt
        s.n_samples = -(-nt // s.dt_scale)
        s.rec_stride = s.n_rec
    # records packed shot after shot for the shorter rows
    off = 0
    for s in shots:
        s.rec_offset = off
        off += 5 * s.n_samples * s.rec_stride
    src = torch.from_numpy(W._source_table()).to(dev)
    seis = torch.full((off,), float("nan"), dtype=torch.float32, device=dev)
    print(f"no-alloc test: {len(ids)} shots, nt={nt}, noalloc={noalloc}", flush=True)
    if noalloc:
        W._plan.forward_noalloc(shots, src, seis)
    else:
        W._plan.forward(shots, src, seis, np.zeros(0, np.int32), 0, _native.MEM_DEVICE)
    torch.cuda.synchronize()
    return seis.cpu().numpy()


def test_noalloc_forward_equals_forward(gpu):
    import torch
    w = workloads.plain_small()
    W = _plan(w, torch)
    for ids, nt in (([0, 1, 2, 3, 4], 400), ([1, 3], 250)):
        want = _records(W, torch, ids, nt, noalloc=False)
        got = _records(W, torch, ids, nt, noalloc=True)
        assert np.isfinite(got).all()
        np.testing.assert_array_equal(got, want)
        assert np.abs(want).max() > 0


def test_noalloc_receivers_off_active_cells(gpu):
    """Without an absorbing layer, receivers on row 0 lie outside the active
    cells: they take the in-memory gather launch, in the no-alloc path too."""
    import torch
    w = workloads.plain_small()
    w.inpa["npml"] = 0
    w.rec_loc = w.rec_loc.copy()
    w.rec_loc[:, 1] = 0.0
    W = _plan(w, torch)
    want = _records(W, torch, [0, 4], 200, noalloc=False)
    got = _records(W, torch, [0, 4], 200, noalloc=True)
    np.testing.assert_array_equal(got, want)


def test_noalloc_refuses_more_than_the_scratch(gpu):
    import torch
    w = workloads.subset(workloads.plain_medium(), list(range(8)), nt=50)
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=0, chpr=0)
    dev = torch.device("cuda:0")
    shots = W._shots(list(range(8)))
    src = torch.from_numpy(W._source_table()).to(dev)
    seis = torch.zeros(8 * 5 * int(W.nt_ext) * int(W.nr), dtype=torch.float32, device=dev)
    with pytest.raises(_native.NativeError) as e:
        W._plan.forward_noalloc(shots, src, seis)
    assert e.value.code == _native.EINVAL and "scratch" in str(e.value)
