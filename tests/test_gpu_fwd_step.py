"""The forward step in one launch (k_fwd_step, PYFWI_TUNE_FUSED_STEPS) and two
forward steps in one launch (k_fwd_two, PYFWI_TUNE_FWD_TWO: the kept-state
forward at order 4) against the same steps as two launches each (k_vel then
k_str, the default).  The one-launch
kernel computes every value with the two kernels' arithmetic -- the new
velocities reach the stress update through registers and lane moves instead
of memory, the source term is added to the loaded values and the record
sample written by the owning lane as there, and the sums whose fused
multiply-adds the compiler could form either way are written out as fmaf in
both -- so records, kept states and the gradients built on them are
compared bit for bit.  (With chpr = 0 the gradient's time reversal also runs
in one launch a step, k_fused_rev.)  The misfit, a double sum whose terms
are added in no fixed order, is compared to 1e-12.

Each case runs with the one-step launch, the two-step launch (single steps
two launches) and both (single steps in one launch).

Cases: kept states (chpr = 100) and checkpoints (chpr = 0); passes of one
shot (the one-shot tile form), two shots and all five (blocks of four); ten
shots (blocks of five); dt_scale = 2 at orders 4 and 8; uneven shots whose
launches take the live shots only; well receivers (a receiver table, so a
gather launch records); sources off the active cells (the source term
injected in memory).
"""
import numpy as np
import pytest

from pyfwi_amd import _native, workloads
from pyfwi_amd.acquisition import Source
from pyfwi_amd.batch import UnevenBatch
from pyfwi_amd.wave_propagation import WavePropagator

from _helpers import rel

pytestmark = pytest.mark.gpu

FUSED = _native.TUNE_FUSED_STEPS
TWO = _native.TUNE_FWD_TWO
BITS = pytest.mark.parametrize("bits", [FUSED, TWO, TWO | FUSED])


def _both(make, run, per_pass=0, bits=FUSED):
    """run(W) with the forward steps launched as `bits` says, then with two
    launches a step."""
    out = []
    for b in (bits, 0):
        W = make()
        W._plan.set_tuning(per_pass, b, 0)
        out.append(run(W))
    return out


def _check(one, two, label):
    for k in two:
        a, b = np.asarray(one[k]), np.asarray(two[k])
        assert np.isfinite(a).all(), (label, k)
        same = np.array_equal(a, b)
        d = rel(a, b)
        print(f"{label} {k}: rel {d:.2e}{' (bit-equal)' if same else ''}", flush=True)
        if a.dtype == np.float64 and a.ndim == 0:  # the misfit
            assert d <= 1e-12, (label, k, d)
        else:
            assert same, (label, k, d)


def _records_and_gradient(model, grad=True):
    def run(W):
        d = W.forward_modeling(model)
        out = {f"rec_{k}": v for k, v in d.items()}
        if grad:
            res = {k: np.float32(0.3 * v) for k, v in d.items()}
            out.update({f"grad_{k}": v for k, v in W.gradient(res, parameterization="lmd").items()})
        return out
    return run


def _small(nt=400, **over):
    w = workloads.plain_small()
    w.inpa["t"] = workloads.t_for_nt(nt, w.inpa["dt"])
    w.inpa.update(over)
    return w


@BITS
@pytest.mark.parametrize("per_pass", [0, 1, 2])
@pytest.mark.parametrize("chpr", [100, 0])
def test_one_launch_forward_plain(gpu, chpr, per_pass, bits):
    w = _small(nt=401)  # an odd number of steps: the last one alone
    one, two = _both(lambda: WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape,
                                            components=4, chpr=chpr),
                     _records_and_gradient(w.model), per_pass, bits)
    _check(one, two, f"bits={bits} chpr={chpr} per_pass={per_pass}")


@BITS
@pytest.mark.parametrize("order", [4, 8])
def test_one_launch_forward_dt_scale(gpu, order, bits):
    w = workloads.plain_small()
    w.inpa.update({"sdo": order, "dh": 5.0, "dt": 0.0008, "rec_dis": 10.0, "t": 0.16})
    src = Source(np.c_[w.src.i, w.src.j] * 5.0, 5.0, 0.0008)
    src.Ricker(15)
    rec = w.rec_loc * 5.0

    def make():
        W = WavePropagator(w.inpa, src, rec, w.model_shape, components=4, chpr=100)
        assert W.dt_scale == 2
        return W
    one, two = _both(make, _records_and_gradient(w.model), bits=bits)
    _check(one, two, f"bits={bits} order {order} dt_scale=2")


@BITS
def test_one_launch_forward_blocks_of_five(gpu, bits):
    """Ten shots a pass: blocks of five shots."""
    w = workloads.subset(workloads.plain_medium(ns=10, nt=2000), list(range(10)), nt=240)
    one, two = _both(lambda: WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape,
                                            components=2, chpr=100),
                     _records_and_gradient(w.model), bits=bits)
    _check(one, two, f"bits={bits} 10 shots")


@BITS
def test_one_launch_forward_uneven_live_shots(gpu, bits):
    """Uneven nt: the kept states are time-major and each step's launch takes
    the shots live at that step.  The residual is the records themselves
    (zero observed data): a residual of two nearly equal records would scale
    the records' rounding differences up by their ratio."""
    wl = workloads.uneven_small()
    out = []
    for b in (bits, 0):
        B = UnevenBatch(wl.inpa, wl.specs, wl.model_shape, components=2)
        B.plan.set_tuning(0, b, 0)
        B.set_model(wl.model_obs)
        rec_obs = B.forward()
        B.set_model(wl.model)
        seis = np.zeros(B.n_records, np.float32)
        misfit, G = B.misfit_gradient(np.zeros_like(seis), seis=seis)
        out.append({"forward": rec_obs, "seis": seis, "misfit": np.float64(misfit), "grad": G})
    _check(out[0], out[1], f"bits={bits} uneven")


@BITS
@pytest.mark.parametrize("acq_type", [0, 2])
def test_one_launch_forward_wells(gpu, acq_type, bits):
    """Well receivers: records gathered through the receiver table."""
    from test_gpu_wells import wells
    _, _, model, _ = wells(acq_type)
    one, two = _both(lambda: wells(acq_type)[0], _records_and_gradient(model), bits=bits)
    _check(one, two, f"bits={bits} wells acq_type={acq_type}")


@BITS
@pytest.mark.parametrize("chpr", [100, 0])
def test_one_launch_forward_sources_off_active_cells(gpu, chpr, bits):
    """A source and receivers on the never-updated outer ring (no absorbing
    layer): the source term goes into memory before the step."""
    w = _small(nt=300, npml=0)
    w.src.j[:] = 1
    w.rec_loc[:, 1] = 1.0
    one, two = _both(lambda: WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape,
                                            components=2, chpr=chpr),
                     _records_and_gradient(w.model), bits=bits)
    _check(one, two, f"bits={bits} off active chpr={chpr}")
