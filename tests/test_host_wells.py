"""Host side of the well acquisitions (no GPU): the receiver table that
WavePropagator hands to the plan, the shot entries that point at it, and
the geometry errors."""
import numpy as np
import pytest

from pyfwi_amd import _native
from pyfwi_amd.acquisition import Source, acq_parameters
from pyfwi_amd.wave_propagation import WavePropagator

DH, DT, NZ, NX, NPML = 5.0, 5e-4, 24, 30, 6


def propagator(acq_type, ns=2, **kw):
    inpa = {"dh": DH, "dt": DT, "t": 0.05, "npml": NPML, "pmlR": 1e-5, "pml_dir": 2,
            "sdo": 4, "rec_dis": 2 * DH, "acq_type": acq_type}
    src_loc, rec_loc, _, nw = acq_parameters(ns, 2 * DH, NX * DH, NZ * DH, DH, 4, acq_type)
    src = Source(src_loc, DH, DT)
    src.Ricker(30)
    rec_loc = kw.pop("rec_loc", rec_loc)
    W = WavePropagator(inpa, src, rec_loc, (NZ, NX), components=4,
                       n_well_rec=kw.pop("n_well_rec", nw), chpr=100)
    return W, rec_loc, nw


def cells(rec_loc, tnx):
    return ((rec_loc[:, 1] / DH).astype(np.int64) + NPML) * tnx \
        + (rec_loc[:, 0] / DH).astype(np.int64) + NPML


@pytest.mark.parametrize("acq_type", [0, 2])
def test_receiver_table_rows(acq_type):
    W, rec_loc, nw = propagator(acq_type)
    tab = W.receiver_table()
    inj = cells(rec_loc, int(W.tnx))
    np.testing.assert_array_equal(tab[1], inj)
    rec = inj.copy()
    if acq_type == 2:
        rec[:nw] = inj[:nw][::-1]
    np.testing.assert_array_equal(tab[0], rec)


@pytest.mark.parametrize("acq_type", [0, 2])
def test_shots_point_at_the_table(acq_type):
    W, _, _ = propagator(acq_type)
    for sh in W._shots([0, 1]):
        assert sh.rec_layout == _native.REC_TABLE and sh.rec_table == 0
        assert sh.n_rec == sh.rec_stride == int(W.nr)
        assert sh.rec_flags == (_native.REC_VZ_FROM_VX if acq_type == 0 else 0)
    S, _, _ = propagator(1)
    assert S._shots([0])[0].rec_layout == _native.REC_LINE
    assert S._rec_table is None


def test_crosswell_source_column_sets_the_mask():
    W, _, _ = propagator(0)
    assert W.src_cts == W.src.i[0]
    G = np.ones((6, int(W.tnz), int(W.tnx)), np.float32)
    glam, _, _ = W.gradient_reading(G)
    sdo, c0 = int(W.sdo), int(W.sdo) + int(W.src_cts) + 2
    assert np.all(glam[sdo:NZ - sdo, c0:NX - sdo] == 1)
    assert glam[:, :c0].sum() == 0 and glam[:sdo].sum() == 0


def test_combined_without_well_receivers_raises():
    with pytest.raises(Exception, match="geophons"):
        propagator(2, n_well_rec=0)


def test_well_below_the_grid_raises():
    """The gather places receiver ir at rec_top_right_var + ir*dxr; from the
    model's last row, six receivers two cells apart leave the grid."""
    rec = np.c_[np.full(6, 100.0), (NZ - 1) * DH + 2 * DH * np.arange(6)]
    with pytest.raises(ValueError, match="outside the grid"):
        propagator(0, rec_loc=rec)
