"""Pinning the oracle's well acquisition (crosswell, acq_type 0, and wells +
surface, acq_type 2): the gathers of src/PyFWI/elastic_crosswell.cl:1-58 and
src/PyFWI/elastic.cl:43-98, 316-363 restated in oracle/elastic_oracle.c.

The expected receiver cells come straight from the receiver positions
(column int(x/dh) + npml, row int(z/dh) + npml), not from the reference's
macro-and-loop formulation that the oracle restates; the records are checked
against the kept states at those cells; the transpose of the gather by central
finite differences of the exact gradient."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from pyfwi_amd.acquisition import Source, acq_parameters

DH, DT, NZ, NX, NPML, ORDER = 5.0, 5e-4, 24, 30, 6, 4


def well_setup(acq_type, nt=120, ns=2, fdom=40, chpr=0):
    inpa = {"dh": DH, "dt": DT, "t": (nt - 1) * DT + DT / 2, "npml": NPML, "pmlR": 1e-5,
            "pml_dir": 2, "sdo": ORDER, "rec_dis": 2 * DH, "acq_type": acq_type}
    src_loc, rec_loc, _, n_well = acq_parameters(ns, 2 * DH, NX * DH, NZ * DH, DH, ORDER,
                                                 acq_type)
    src = Source(src_loc, DH, DT)
    src.Ricker(fdom)
    st = O.Setup(inpa, src_loc, rec_loc, (NZ, NX), src.w, src.component, chpr=chpr,
                 components=4, n_well_rec=n_well if acq_type != 1 else 0)
    rng = np.random.default_rng(3)
    vp = (2000 + 300 * rng.random((NZ, NX))).astype(np.float32)
    model = {"vp": vp, "vs": (vp / 1.8).astype(np.float32),
             "rho": (1800 + 200 * rng.random((NZ, NX))).astype(np.float32)}
    return st, model, rec_loc


def cells_from_positions(rec_loc, tnx):
    col = (rec_loc[:, 0] / DH).astype(np.int64) + NPML
    row = (rec_loc[:, 1] / DH).astype(np.int64) + NPML
    return row * tnx + col


@pytest.mark.parametrize("acq_type", [0, 1, 2])
def test_receiver_cells_follow_the_positions(acq_type):
    st, _, rec_loc = well_setup(acq_type)
    got = st.receiver_cells()
    assert None not in got
    np.testing.assert_array_equal(got, cells_from_positions(rec_loc, st.tnx))
    if acq_type == 2:
        # the left well's first receiver shares the surface line's first cell
        assert got[0] == got[st.n_well_rec]


@pytest.mark.parametrize("acq_type", [0, 2])
def test_records_sample_the_kept_states(acq_type):
    """Record row t+1 is the state after step t (kept at every step) at the
    receiver cells; the crosswell vz record holds vx."""
    st, model, rec_loc = well_setup(acq_type, nt=60)
    st.chp = np.arange(st.nt, dtype=np.int32)
    seis, W = O.forward(st, model, np.float32)
    cells = cells_from_positions(rec_loc, st.tnx)
    n = st.tnz * st.tnx
    for s in range(st.ns):
        states = W[s].reshape(st.nt, 5, n)[:, :, cells]          # [t, field, receiver]
        for f in range(5):
            field = 0 if (acq_type == 0 and f == 1) else f
            np.testing.assert_array_equal(seis[s, f, 1:st.nt], states[:st.nt - 1, field])
        assert np.all(seis[s, :, 0] == 0)
    assert np.abs(seis[:, 2]).max() > 0


@pytest.mark.parametrize("acq_type", [0, 2])
@pytest.mark.parametrize("exact", [0, 1])
def test_residual_injection_cells(acq_type, exact):
    """One residual row lands at the receiver cells, component f in field f
    (the reference's injection) or in the field it was recorded from (the
    transpose); receivers sharing a cell add up."""
    st, _, rec_loc = well_setup(acq_type)
    acq = st.acq()
    rng = np.random.default_rng(7)
    res = np.ascontiguousarray(rng.standard_normal((5, st.ns, st.nt_ext, st.nr)))
    n = st.tnz * st.tnx
    A = np.zeros((5, n))
    rc = O.lib().orc_inject_residual_d(ctypes.byref(st.grid), ctypes.byref(acq), O._p(A),
                                       O._p(res), ctypes.c_int(1), ctypes.c_int(5),
                                       ctypes.c_int(exact))
    assert rc == 0
    want = np.zeros((5, n))
    cells = cells_from_positions(rec_loc, st.tnx)
    for f in range(5):
        dst = 0 if (exact and acq_type == 0 and f == 1) else f
        np.add.at(want[dst], cells, res[f, 1, 5])
    np.testing.assert_allclose(A, want, rtol=1e-15, atol=1e-15)


def test_left_well_reordering():
    """forward_propagator reverses the left well's columns only."""
    st, _, _ = well_setup(2)
    seis = np.arange(2 * 5 * 3 * st.nr, dtype=np.float32).reshape(2, 5, 3, st.nr)
    out = st.flip_wells(seis)
    nw = st.n_well_rec
    np.testing.assert_array_equal(out[..., :nw], seis[..., nw - 1::-1])
    np.testing.assert_array_equal(out[..., nw:], seis[..., nw:])


@pytest.mark.parametrize("acq_type", [0, 2])
def test_exact_gradient_central_differences_wells(acq_type):
    """FP64 central differences of J(m) = <r, records(m)> along a smooth
    direction match the exact adjoint gradient: pins the transpose of the
    well gathers, the crosswell vz record included."""
    st, model, _ = well_setup(acq_type)
    md = st.medium(model, np.float64)
    rng = np.random.default_rng(4)
    res = rng.standard_normal((5, st.ns, st.nt_ext, st.nr))
    g3 = O.exact_gradient(st, model, res, np.float64, medium=md)
    z, x = np.mgrid[0:st.tnz, 0:st.tnx]
    for k, key in enumerate(("lam", "mu", "b")):
        dm = np.cos(0.2 * z + 0.3 * x + k) * md[key]
        eps = 1e-6

        def J(sign):
            m2 = dict(md)
            m2[key] = np.ascontiguousarray(md[key] + sign * eps * dm)
            seis, _ = O.forward(st, model, np.float64, keep_states=False, medium=m2)
            return np.sum(seis.transpose(1, 0, 2, 3) * res)
        fd = (J(1) - J(-1)) / (2 * eps)
        an = np.sum(g3[k] * dm)
        assert abs(fd - an) < 1e-6 * abs(an), (key, fd, an)


def test_crosswell_mask_keeps_columns_right_of_the_source():
    """gradient_reading's crosswell mask (wave_propagation.py:309-311): rows
    [sdo, nz - sdo), columns [sdo + src_i + 2, nx - sdo)."""
    st, _, _ = well_setup(0)
    G = np.ones((6, st.tnz, st.tnx), np.float32)
    glam, _, _ = O.read_gradient(st, G)
    sdo, c0 = st.sdo, st.sdo + st.src_cts + 2
    want = np.zeros((NZ, NX), np.float32)
    want[sdo:NZ - sdo, c0:NX - sdo] = 1
    np.testing.assert_array_equal(glam, want)
