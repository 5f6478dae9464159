"""Closed-form pins of the reference-adjoint terms of the CPU oracle
(oracle/elastic_oracle_impl.h) that the transposition and reversibility
checks of test_oracle.py leave open:

  adjoint damping     elastic.cl:395-408, 465-466
  reverse-step order  elastic.cl:175-182, 220-223, 262-268, 303-306
  imaging condition   elastic.cl:470-526 (Gmu, Glam, Grho, preconditioners,
                      the vz of "DzpVx")
  backward loop       wave_propagation.py:751-848

Each expected value is built from the reference's formulas with analytic
derivatives of constant, linear or quadratic fields -- on which the staggered
stencils are exact: a stencil with weights c_k on pairs at +-(k - 1/2) cells
maps a quadratic f to M1 f' at its half-cell point, M1 = sum c_k (2k - 1) --
not from the oracle's stencil code.  dx != dz and dx-damping != dz-damping so
that a swapped axis shows.  Every check is also run against builds of the
oracle with one planted mistake each (ORC_MUTANT in oracle/elastic_oracle.c)
and must fail there."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

from _helpers import rel

TOL = 1e-11       # FP64 closed forms
CAUGHT = 1e-6     # a planted mistake moves the checked values by more than this


def _grid(nz, nx, sdo, dt, dx=1.0, dz=2.0, npml=0):
    c = tuple(float("%f" % v) for v in O.coefficients(2 * sdo))
    return O.OGrid(nz, nx, sdo, npml, c[0], c[1], c[2], c[3], dt, dx, dz)


def _m1(g):
    c = (g.c1, g.c2, g.c3, g.c4)[:g.sdo]
    return float(sum(ck * (2 * k + 1) for k, ck in enumerate(c)))


def _medium(g, lam, mu, b, alpha, beta):
    shape = (g.nz, g.nx)
    full = lambda v: np.full(shape, v, np.float64)
    return dict(lam=full(lam), mu=full(mu), b=full(b), dpx=full(alpha), dpz=full(beta))


def _inner(g, margin):
    return (slice(margin, g.nz - margin), slice(margin, g.nx - margin))


# ------------------------------------------------------------------ checks
# Each returns the largest relative error of its closed form (per term).

def check_adjoint_damping(L, sdo):
    """Constant adjoint fields, constant lam, mu, b, constant damping alpha
    (x) and beta (z): every derivative vanishes away from the edges, so one
    Adj_update_tau + Adj_update_v step scales Atxx by 1 - dt alpha, Atzz by
    1 - dt beta, Atxz by 1 - dt (alpha + beta), Avx by 1 - dt (2 alpha + beta)
    and Avz by 1 - dt (alpha + 2 beta) (elastic.cl:395-396, 408, 465-466)."""
    g = _grid(40, 44, sdo, dt=0.05)
    alpha, beta = 1.7, 0.6
    md = _medium(g, 2.0, 0.7, 0.9, alpha, beta)
    c = np.array([0.8, -1.3, 2.1, 0.4, -0.9])  # avx, avz, atxx, atzz, atxz
    A = np.ascontiguousarray(np.broadcast_to(c[:, None, None], (5, g.nz, g.nx)), np.float64)
    O.step("adj", g, None, md, A=A, L=L)
    dt = g.dt
    keep = np.array([1 - dt * (2 * alpha + beta), 1 - dt * (alpha + 2 * beta),
                     1 - dt * alpha, 1 - dt * beta, 1 - dt * (alpha + beta)])
    win = _inner(g, 2 * g.sdo + 1)
    return {f: rel(A[k][win], np.full(A[k][win].shape, c[k] * keep[k]))
            for k, f in enumerate(("avx", "avz", "atxx", "atzz", "atxz"))}


def _linear_state(g):
    """Five linear fields f = f0 + f1 x + f2 z (x = j dx, z = i dz) and their
    coefficient rows [f0, f1, f2]."""
    co = np.array([[0.3, 0.7, -0.4],    # vx
                   [-0.2, 0.25, 0.9],   # vz
                   [1.1, -0.6, 0.35],   # txx
                   [0.5, 0.15, -0.8],   # tzz
                   [-0.7, 0.45, 0.55]])  # txz
    z, x = np.mgrid[0:g.nz, 0:g.nx].astype(np.float64)
    x, z = x * g.dx, z * g.dz
    F = np.stack([a + b * x + c * z for a, b, c in co])
    return np.ascontiguousarray(F), co, x, z


def check_step_order(L, sdo, direction):
    """Linear fields with constant damping k = alpha + beta.  Forward
    (elastic.cl:175-177, 218-219, 255-260, 299-300): each half step is
    (u + dt S)(1 - dt k), velocities first, the stresses then seeing the
    updated velocity slopes.  Time reversed in the reference's order txz,
    (txx, tzz), vz, vx (wave_propagation.py:780-812, elastic.cl:179-182,
    220-223, 262-268, 303-306): u (1 - dt k) - dt S, the velocities seeing the
    updated stress slopes.  S: b (Dx txx + Dz txz), b (Dz tzz + Dx txz),
    (lam + 2mu) Dx vx + lam Dz vz, (lam + 2mu) Dz vz + lam Dx vx,
    mu (Dx vz + Dz vx), each derivative M1 times the slope."""
    g = _grid(36, 40, sdo, dt=0.04)
    alpha, beta = 1.3, 0.45
    lam, mu, b = 1.8, 0.65, 0.75
    md = _medium(g, lam, mu, b, alpha, beta)
    F, co, x, z = _linear_state(g)
    M, dt, k = _m1(g), g.dt, alpha + beta
    q = 1 - dt * k
    want = np.empty_like(F)
    (p0, p1, p2), (q0, q1, q2), (r0, r1, r2), (s0, s1, s2), (u0, u1, u2) = co
    vx, vz, txx, tzz, txz = F.copy()  # F is updated in place
    if direction > 0:
        O.step("fwd", g, F, md, L=L)
        want[0] = (vx + dt * b * M * (r1 + u2)) * q
        want[1] = (vz + dt * b * M * (s2 + u1)) * q
        P1, P2, Q1, Q2 = p1 * q, p2 * q, q1 * q, q2 * q  # updated velocity slopes
        want[2] = (txx + dt * M * ((lam + 2 * mu) * P1 + lam * Q2)) * q
        want[3] = (tzz + dt * M * ((lam + 2 * mu) * Q2 + lam * P1)) * q
        want[4] = (txz + dt * mu * M * (Q1 + P2)) * q
    else:
        O.step("rev", g, F, md, L=L)
        want[4] = txz * q - dt * mu * M * (q1 + p2)
        want[2] = txx * q - dt * M * ((lam + 2 * mu) * p1 + lam * q2)
        want[3] = tzz * q - dt * M * ((lam + 2 * mu) * q2 + lam * p1)
        want[1] = vz * q - dt * b * M * (s2 * q + u1 * q)  # updated stress slopes
        want[0] = vx * q - dt * b * M * (r1 * q + u2 * q)
    win = _inner(g, 2 * g.sdo + 1)
    return {f: rel(F[kk][win], want[kk][win])
            for kk, f in enumerate(("vx", "vz", "txx", "tzz", "txz"))}


def check_imaging(L, sdo):
    """Quadratic forward fields f = a0 + a1 x + a2 z + a3 x^2 + a4 z^2 + a5 xz,
    arbitrary adjoint fields and buoyancy (read at the cell only).  With the
    analytic derivatives at the half-cell points (Dxp at x + dx/2, Dxm at
    x - dx/2, Dzp at z + dz/2, Dzm at z - dz/2, times M1) the six buffers of
    one Grad call are the reference's formulas (elastic.cl:485-524):
      DzpVx = Dzp vz (sic), DxmVz = Dxm vz, DxpVx = Dxp vx, DzmVz = Dzm vz,
      Gmu   += 2dt (Atxx DxpVx + Atzz DzmVz) + dt Atxz (DzpVx + DxmVz)
      Gmu_p += 4dt^2 (txx^2 DxpVx^2 + tzz^2 DzmVz^2) + dt^2 txz^2 (DzpVx^2 + DxmVz^2)
      Glam  += dt (Atxx + Atzz)(DxpVx + DzmVz)
      Glam_p+= dt^2 (txx^2 + tzz^2)(DxpVx^2 + DzmVz^2)
      Grho  += -dt b^2 (Avx (Dxm txx + Dzm txz) + Avz (Dzp tzz + Dxp txz))
      Grho_p+= dt^2 (vx^2 (Dxm txx^2 + Dzm txz^2) + vz^2 (Dzp tzz^2 + Dxp txz^2))
    on the interior rows and columns npml+sdo .. N-sdo-npml, zero elsewhere."""
    npml = 3
    g = _grid(34, 38, sdo, dt=0.07, npml=npml)
    rng = np.random.default_rng(7)
    co = rng.uniform(-1, 1, (5, 6))
    z, x = np.mgrid[0:g.nz, 0:g.nx].astype(np.float64)
    x, z = x * g.dx, z * g.dz

    def f(k, xx, zz):
        a = co[k]
        return a[0] + a[1] * xx + a[2] * zz + a[3] * xx ** 2 + a[4] * zz ** 2 + a[5] * xx * zz

    def fx(k, xx, zz):
        a = co[k]
        return a[1] + 2 * a[3] * xx + a[5] * zz

    def fz(k, xx, zz):
        a = co[k]
        return a[2] + 2 * a[4] * zz + a[5] * xx

    M, dt, hx, hz = _m1(g), g.dt, g.dx / 2, g.dz / 2
    F = np.ascontiguousarray(np.stack([f(k, x, z) for k in range(5)]))
    A = np.ascontiguousarray(rng.uniform(-1, 1, (5, g.nz, g.nx)))
    md = _medium(g, 1.0, 1.0, 1.0, 0.0, 0.0)
    md["b"] = np.ascontiguousarray(rng.uniform(0.5, 1.5, (g.nz, g.nx)))
    G = np.zeros((6, g.nz, g.nx))
    O.step("grad", g, F, md, A=A, G=G, L=L)
    VX, VZ, TXX, TZZ, TXZ = range(5)
    DzpVx = M * fz(VZ, x, z + hz)
    DxmVz = M * fx(VZ, x - hx, z)
    DxpVx = M * fx(VX, x + hx, z)
    DzmVz = M * fz(VZ, x, z - hz)
    DxmPx = M * fx(TXX, x - hx, z)
    DzmPxz = M * fz(TXZ, x, z - hz)
    DzpPz = M * fz(TZZ, x, z + hz)
    DxpPxz = M * fx(TXZ, x + hx, z)
    avx, avz, atxx, atzz, atxz = A
    vx, vz, txx, tzz, txz = F.copy()  # F is updated in place
    b = md["b"]
    want = np.stack([
        2 * dt * (atxx * DxpVx + atzz * DzmVz) + dt * atxz * (DzpVx + DxmVz),
        4 * dt ** 2 * (txx ** 2 * DxpVx ** 2 + tzz ** 2 * DzmVz ** 2)
        + dt ** 2 * txz ** 2 * (DzpVx ** 2 + DxmVz ** 2),
        dt * (atxx + atzz) * (DxpVx + DzmVz),
        dt ** 2 * (txx ** 2 + tzz ** 2) * (DxpVx ** 2 + DzmVz ** 2),
        -dt * b ** 2 * (avx * (DxmPx + DzmPxz) + avz * (DzpPz + DxpPxz)),
        dt ** 2 * (vx ** 2 * (DxmPx ** 2 + DzmPxz ** 2) + vz ** 2 * (DzpPz ** 2 + DxpPxz ** 2)),
    ])
    lo = npml + g.sdo
    mask = np.zeros((g.nz, g.nx), bool)
    mask[lo:g.nz - g.sdo - npml + 1, lo:g.nx - g.sdo - npml + 1] = True
    out = {}
    for k, name in enumerate(("gmu", "gmu_pre", "glam", "glam_pre", "grho", "grho_pre")):
        out[name] = max(rel(G[k][mask], want[k][mask]), float(np.abs(G[k][~mask]).max()))
    return out


def _tiny_setup(nt, npml=5, sdo=4):
    from pyfwi_amd.acquisition import Source
    dh, dt = 5.0, 5e-4
    nz, nx = 22, 26
    inpa = {"dh": dh, "dt": dt, "t": (nt - 1) * dt + dt / 2, "npml": npml, "pmlR": 1e-5,
            "pml_dir": 2, "sdo": sdo, "rec_dis": 2 * dh}
    rng = np.random.default_rng(9)
    vp = (2000 + 300 * rng.random((nz, nx))).astype(np.float32)
    model = {"vp": vp, "vs": (vp / 1.8).astype(np.float32),
             "rho": (1800 + 200 * rng.random((nz, nx))).astype(np.float32)}
    src_loc = np.array([[8 * dh, 4 * dh], [18 * dh, 4 * dh]])
    src = Source(src_loc, dh, dt)
    src.Ricker(60)
    rec = np.c_[(2 + 2 * np.arange(10)) * dh, np.full(10, 2 * dh)]
    st = O.Setup(inpa, src_loc, rec, (nz, nx), src.w, src.component, chpr=0, components=4)
    return st, model


def check_backward_loop(L):
    """The reference's backward loop (wave_propagation.py:751-848) composed
    in Python from the oracle's single-step primitives, for 2 shots of 7
    steps with checkpoints after steps 1 and 6 (chpr = 0): per t = 6 .. 1,
    restore the checkpoint at t = chp[c] or reverse one step, inject residual
    row t_src - 1 (t_src counting down from nt_ext), adjoint step, Grad.  It
    must equal the C loop (orc_gradient)."""
    st, model = _tiny_setup(nt=7)
    assert list(st.chp) == [1, 6]
    seis, W = O.forward(st, model, np.float64)
    rng = np.random.default_rng(10)
    res_rec = rng.standard_normal((5, st.ns, st.nt_ext, st.nr))
    res2d = {f: res_rec[k].transpose(1, 0, 2).reshape(st.nt_ext, st.ns * st.nr)
             for k, f in enumerate(O.FIELDS)}
    md = st.medium(model, np.float64)
    fn = getattr(L, "orc_gradient_d")
    G_c = np.zeros((6, st.tnz, st.tnx))
    chp = np.ascontiguousarray(st.chp, np.int32)
    res = np.ascontiguousarray(O.residual_records(st, res2d, np.float64))
    acq = st.acq()
    assert fn(ctypes.byref(st.grid), ctypes.byref(acq), O._p(md["lam"]), O._p(md["mu"]),
              O._p(md["b"]), O._p(md["dpx"]), O._p(md["dpz"]), O._p(chp), ctypes.c_int(len(chp)),
              O._p(np.ascontiguousarray(W)), O._p(res), O._p(G_c)) == 0
    inj = getattr(L, "orc_inject_residual_d")
    G_p = np.zeros_like(G_c)
    for s in range(st.ns):
        F = np.zeros((5, st.tnz, st.tnx))
        A = np.zeros((5, st.tnz, st.tnx))
        chpc, t_src = len(chp) - 1, st.nt_ext
        for t in range(st.nt - 1, 0, -1):
            if chpc >= 0 and t == chp[chpc]:
                F = np.ascontiguousarray(W[s, chpc])
                chpc -= 1
            else:
                O.step("rev", st.grid, F, md, L=L)
            if t % st.dt_scale == 0:
                assert inj(ctypes.byref(st.grid), ctypes.byref(acq), O._p(A), O._p(res), s,
                           t_src - 1, 0) == 0
                t_src -= 1
            O.step("adj", st.grid, None, md, A=A, L=L)
            O.step("grad", st.grid, F, md, A=A, G=G_p, L=L)
    return {"loop": rel(G_c, G_p)}


CHECKS = {
    "damping": lambda L, sdo: check_adjoint_damping(L, sdo),
    "forward_order": lambda L, sdo: check_step_order(L, sdo, +1),
    "reverse_order": lambda L, sdo: check_step_order(L, sdo, -1),
    "imaging": lambda L, sdo: check_imaging(L, sdo),
}


@pytest.mark.parametrize("sdo", [2, 4])
@pytest.mark.parametrize("check", list(CHECKS))
def test_closed_form(check, sdo):
    errs = CHECKS[check](O.lib(), sdo)
    assert max(errs.values()) < TOL, errs


def test_backward_loop():
    errs = check_backward_loop(O.lib())
    assert errs["loop"] < 1e-12, errs


# planted mistake -> (check, term that must move)
MUTANTS = {
    1: ("damping", "atxx"), 2: ("damping", "avx"), 3: ("damping", "avz"),
    4: ("reverse_order", "vx"), 5: ("imaging", "gmu"), 6: ("imaging", "gmu"),
    7: ("imaging", "grho"), 8: ("imaging", "gmu_pre"), 9: ("imaging", "glam_pre"),
    10: ("imaging", "grho_pre"), 11: ("damping", "atxz"), 12: ("loop", "loop"),
}


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_planted_mistake_is_caught(k):
    """Each check fails on the oracle built with the slip it is written for,
    and the other terms of that check still pass (the mistake is local)."""
    L = O.mutant(k)
    check, term = MUTANTS[k]
    errs = check_backward_loop(L) if check == "loop" else CHECKS[check](L, 2)
    assert errs[term] > CAUGHT, (k, errs)
    if check != "loop":
        # the unmutated library passes the same check
        assert max(CHECKS[check](O.lib(), 2).values()) < TOL


def test_oracle_under_sanitizers():
    """The oracle's entry points on tiny surface, crosswell and wells +
    surface cases, orders 4 and 8, built with AddressSanitizer and UBSan
    (oracle/asan_driver.c): no out-of-bounds access, undefined behaviour or
    leak."""
    import os
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    r = subprocess.run(["make", "-s", "-C", here, "asan-check"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "asan-check: clean" in r.stdout, r.stdout + r.stderr
