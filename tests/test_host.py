"""Host-side logic of the drop-in modules against closed forms and the
reference's own known-answer test (no GPU)."""
import numpy as np
import pytest

from pyfwi_amd import acquisition as acq
from pyfwi_amd import fwi_tools as tools
from pyfwi_amd import grad_switcher as gs
from pyfwi_amd.processing import prepare_residual
from pyfwi_amd.wave_propagation import WavePropagator, checkpoint_steps
from pyfwi_amd import workloads


def test_ricker_known_answer():
    """The reference's only unit test (test/test_ricker.py:16-22), verbatim in
    meaning: exact float32 equality with the closed form."""
    fdom, dt = 20.0, 0.004
    src = acq.Source(src_loc=np.array([[0, 0]]), dh=1.0, dt=dt)
    src.Ricker(fdom=fdom)
    t = np.arange(-1.0 / fdom, 1.0 / fdom + dt / 3, dt)
    assert np.all(src.w == np.float32((1.0 - 2.0 * (np.pi * fdom * t) ** 2)
                                      * np.exp(-(np.pi * fdom * t) ** 2)))


def test_source_samples_and_components():
    src = acq.Source(np.array([[10.0, 4.0], [30.0, 4.0]]), 1.0, 1e-3)
    src.Ricker(25)
    tab = src.samples(src.w.size + 5)
    assert tab.dtype == np.float32 and tab.shape == (src.w.size + 5, 5)
    np.testing.assert_array_equal(tab[:, 2], tab[:, 3])
    assert not tab[:, [0, 1, 4]].any() and not tab[src.w.size:].any()
    np.testing.assert_array_equal(tab[3], src(3))
    np.testing.assert_array_equal(src(src.w.size + 1), np.zeros(5, np.float32))
    assert list(src.i) == [10, 30] and list(src.j) == [4, 4]
    for kind, comp in ((1, 0), (2, 1)):
        s = acq.Source(np.array([[1.0, 1.0]]), 1.0, 1e-3, src_type=kind)
        assert s.component[comp] == 1 and s.component.sum() == 1


def test_surface_geometry():
    src, rec = acq.surface_seismic(5, 2.0, 100.0, 1.0, 4)
    np.testing.assert_allclose(src[:, 0], np.linspace(4, 91, 5))
    assert np.all(src[:, 1] == 4) and np.all(rec[:, 1] == 4)
    np.testing.assert_allclose(rec[:, 0], np.arange(4, 96, 2))
    s1, _ = acq.surface_seismic(1, 2.0, 100.0, 1.0, 4)
    assert s1[0, 0] == 50
    with pytest.raises(Exception):
        acq.acq_parameters(2, 0.5, 100, 100, 1.0, 4, 1)


def test_kernel_constants_are_printed_values():
    """The reference pastes constants into kernel source with %f
    (wave_propagation.py:434-490)."""
    assert tools.kernel_constant(-1 / 24) == -0.041667
    assert tools.kernel_constant(0.45 * 5 / 4000) == 0.000562
    assert [tools.kernel_constant(c) for c in tools.fd_coefficients(8)] == \
        [1.195955, -0.079498, 0.009763, -0.000697]


def test_damping_profile_closed_form():
    """d(r) = -(3/(2L)) (r/L)^2 log10(R), r = 0..npml-1 cells, mirrored on the
    left/top (fwi_tools.py:943-1063)."""
    dh, npml, R = 5.0, 20, 1e-5
    a = tools.pml_delta_calculation(np.float32(dh), npml, R)
    L = npml * dh
    for k in (0, 1, 7, 19):
        want = -(3 / (2 * L)) * ((k * dh / L) ** 2) * np.log10(R)
        assert a[k] == np.float32(want)
    dxg, dzg = tools.pml_counstruction(60, 70, np.float32(dh), npml, R, 2)
    assert dxg.shape == (60, 70) and dxg.dtype == np.float32
    np.testing.assert_array_equal(dxg[:, 70 - npml:], np.broadcast_to(a, (60, npml)))
    np.testing.assert_array_equal(dxg[:, :npml], np.broadcast_to(a[::-1], (60, npml)))
    assert not dxg[:, npml:70 - npml].any()
    np.testing.assert_array_equal(dzg[60 - npml:, 5], a)
    _, dz3 = tools.pml_counstruction(60, 70, np.float32(dh), npml, R, 3)
    assert not dz3[:npml].any() and dz3[-1, 0] == a[-1]
    dx0, _ = tools.pml_counstruction(60, 70, np.float32(dh), npml, R, 0)
    _, dz1 = tools.pml_counstruction(60, 70, np.float32(dh), npml, R, 1)
    assert not dx0.any() and not dz1.any()


def test_expand_model_replicates_edges():
    m = np.arange(12, dtype=np.float32).reshape(3, 4)
    e = tools.expand_model(m, 7, 8, 2)
    np.testing.assert_array_equal(e[2:5, 2:6], m)
    np.testing.assert_array_equal(e[0, 2:6], m[0])
    np.testing.assert_array_equal(e[6, 2:6], m[2])
    np.testing.assert_array_equal(e[3, :2], [m[1, 0]] * 2)
    assert e[0, 0] == m[0, 0] and e[6, 7] == m[2, 3] and e[0, 7] == m[0, 3]


def test_checkpoint_list():
    """wave_propagation.py:95-100."""
    np.testing.assert_array_equal(checkpoint_steps(1000, 0), [1, 999])
    np.testing.assert_array_equal(checkpoint_steps(1000, 100), np.arange(1000))
    c = checkpoint_steps(1000, 20)
    assert len(c) == 200 and c[0] == 0 and c[-1] == 999 and np.all(np.diff(c) > 0)
    np.testing.assert_array_equal(checkpoint_steps(50, 1), [1, 49])


def test_seismic_section_layout():
    rng = np.random.default_rng(0)
    seismo = {f: rng.standard_normal((7, 3, 2)).astype(np.float32)
              for f in ("vx", "vz", "taux", "tauz", "tauxz")}
    d = acq.seismic_section(seismo, 0, "2d")
    assert set(d) == {"taux", "tauz"} and d["taux"].shape == (7, 6)
    p = (seismo["taux"] + seismo["tauz"]) / 2
    for s in range(2):
        for r in range(3):
            np.testing.assert_array_equal(d["tauz"][:, s * 3 + r], p[:, r, s])
    d2 = acq.seismic_section(seismo, 2, "3d")
    assert set(d2) == {"vx", "vz"} and d2["vx"].shape == (7, 3, 2)
    assert set(acq.seismic_section(seismo, 3)) == {"taux", "tauz", "tauxz"}
    assert set(acq.seismic_section(seismo, 4)) == set(seismo)


def test_prepare_residual_and_l2():
    r = {"taux": np.ones((4, 6), np.float32)}
    d = prepare_residual(r, 2.0)
    assert set(d) == {"vx", "vz", "taux", "tauz", "tauxz"}
    assert np.all(d["taux"] == 2) and not d["vx"].any() and d["vx"].dtype == np.float32
    cf = tools.CostFunction("l2")
    a = {"vx": np.full((3, 2), 2.0, np.float32)}
    b = {"vx": np.full((3, 2), 0.5, np.float32)}
    rms, res = cf(a, b)
    assert rms == pytest.approx(0.5 * 6 * 1.5 ** 2) and np.all(res["vx"] == 1.5)


def _fd_jacobian(f, x, eps=1e-6):
    """Central-difference Jacobian of an elementwise map R^3 -> R^3 at each
    point: J[i][j] = d f_i / d x_j."""
    J = np.zeros((3, 3) + x[0].shape)
    for j in range(3):
        xp = [v.copy() for v in x]
        xm = [v.copy() for v in x]
        xp[j] = xp[j] + eps * np.abs(x[j])
        xm[j] = xm[j] - eps * np.abs(x[j])
        fp, fm = f(*xp), f(*xm)
        for i in range(3):
            J[i, j] = (fp[i] - fm[i]) / (2 * eps * np.abs(x[j]))
    return J


def test_chain_rules_against_finite_differences():
    """g_new = J^T g_old, with J the Jacobian of the parameter map, checked by
    central differences in float64 (grad_switcher.py:246-266, 331-448)."""
    rng = np.random.default_rng(1)
    vp = rng.uniform(1800, 4000, 50)
    vs = vp / rng.uniform(1.6, 2.0, 50)
    rho = rng.uniform(1800, 2600, 50)
    g = [rng.standard_normal(50) for _ in range(3)]
    # (vp, vs, rho) -> (lam, mu, rho)
    to_lmd = lambda a, b, r: (r * a ** 2 - 2 * r * b ** 2, r * b ** 2, r)
    J = _fd_jacobian(to_lmd, [vp, vs, rho])
    want = [sum(J[i, j] * g[i] for i in range(3)) for j in range(3)]
    lam, mu = rho * vp ** 2 - 2 * rho * vs ** 2, rho * vs ** 2
    got = gs.grad_lmd_to_vd(g[0], g[1], g[2], lam, mu, rho)
    for a, b in zip(got, want):
        np.testing.assert_allclose(a, b, rtol=2e-5)
    # (lam, mu, rho) -> (vp, vs, rho)
    to_vd = lambda l, m, r: (np.sqrt((l + 2 * m) / r), np.sqrt(m / r), r)
    J = _fd_jacobian(to_vd, [lam, mu, rho])
    want = [sum(J[i, j] * g[i] for i in range(3)) for j in range(3)]
    got = gs.grad_vd_to_lmd(g[0], g[1], g[2], vp, vs, rho)
    for a, b in zip(got, want):
        np.testing.assert_allclose(a, b, rtol=2e-5)
    # (K, mu, rho) -> (vp, vs, rho)
    k = lam + 2 * mu / 3
    to_vd_k = lambda kk, m, r: (np.sqrt((kk + 4 * m / 3) / r), np.sqrt(m / r), r)
    J = _fd_jacobian(to_vd_k, [k, mu, rho])
    want = [sum(J[i, j] * g[i] for i in range(3)) for j in range(3)]
    got = gs.grad_dv2kmd({"vp": g[0], "vs": g[1], "rho": g[2]}, vp, vs, rho)
    for a, b in zip(got, want):
        np.testing.assert_allclose(a, b, rtol=2e-5)


def test_chain_rule_on_torch_tensors():
    torch = pytest.importorskip("torch")
    a = [torch.rand(5, dtype=torch.float64) + 1 for _ in range(6)]
    out = gs.grad_lmd_to_vd(*a)
    ref = gs.grad_lmd_to_vd(*[x.numpy() for x in a])
    for t, r in zip(out, ref):
        assert t.dtype == torch.float32
        np.testing.assert_allclose(t.numpy(), r, rtol=1e-6)


def test_fdm_self_checks():
    for order in (4, 8):
        D = tools.Fdm(order)
        ex, ez = D.dot_product_test_derivatives(seed=3)
        assert abs(ex) < 1e-9 and abs(ez) < 1e-9
    D = tools.Fdm(4)
    c_sum = 9 / 8 + 1 / 24
    assert D.dt_computation(4000.0, 5.0) == pytest.approx(2 / 4000 / np.sqrt(2 * c_sum / 25 * 5))
    with pytest.raises(AssertionError):
        tools.Fdm(6)


def test_propagator_geometry_without_gpu():
    """The drop-in's timing and geometry (wave_propagation.py:28-144) are
    computed on the host; the device plan is created lazily."""
    w = workloads.plain_small()
    W = WavePropagator(w.inpa, w.src, w.rec_loc, w.model_shape, components=2, chpr=100)
    assert (W.nt, W.nt_ext, W.dt_scale) == (1000, 1000, 1)
    assert (int(W.tnz), int(W.tnx), int(W.dxr)) == (140, 140, 2)
    assert (int(W.rec_surface_const), int(W.rec_surface_var)) == (22, 21)
    assert list(W.srcx) == [24, 45, 67, 89, 111] and set(W.srcz) == {24}
    assert W.nchp == 1000 and W._n_live == 50
    assert W._grid.c[1] == -0.041667 and W._grid.dt == 1e-4
    m = workloads.plain_medium()
    Wm = WavePropagator(m.inpa, m.src, m.rec_loc, m.model_shape, components=0)
    assert Wm.nt == 2000 and list(Wm.chp) == [1, 1999]
    with pytest.raises(ValueError, match="acq_type"):
        WavePropagator(dict(w.inpa, acq_type=3), w.src, w.rec_loc, w.model_shape)


def test_workloads_are_seeded():
    a, b = workloads.plain_medium(), workloads.plain_medium()
    np.testing.assert_array_equal(a.model["vp"], b.model["vp"])
    assert a.model["vp"].shape == (500, 1000) and a.src.i.size == 20 and a.rec_loc.shape[0] == 400
    vp = a.model["vp"]
    assert vp.min() >= 1800 and vp.max() <= 4000
    assert a.inpa["dt"] == pytest.approx(0.45 * 5 / vp.max())


def test_fold_chain_and_section_vjp_against_fp64_differences():
    """Exact gradient w.r.t. the unpadded (vp, vs, rho): the oracle's FP64
    exact gradient over the padded grid, folded onto the model cells
    (transpose of expand_model) and chained (grad_lmd_to_vd), matches FP64
    central differences of the pressure-record loss; exercises the product's
    _fold_padding, _section_vjp and grad_lmd_to_vd."""
    from oracle import oracle as O
    from pyfwi_amd.wave_propagation import _fold_padding, _section_vjp
    from _helpers import oracle_setup
    w = workloads.plain_small()
    w.inpa["t"] = workloads.t_for_nt(150, w.inpa["dt"])
    w.inpa["npml"] = 8
    st = oracle_setup(w, chpr=0, components=0)
    m64 = {k: v.astype(np.float64) for k, v in w.model.items()}
    rng = np.random.default_rng(9)
    seis, _ = O.forward(st, m64, np.float64, keep_states=False)
    sec = O.section(seis, 0)
    obs = {k: v * 0.9 + 0.01 * rng.standard_normal(v.shape) * np.abs(v).max() for k, v in sec.items()}
    res = _section_vjp({k: sec[k] - obs[k] for k in sec}, 0)
    g3 = O.exact_gradient(st, m64, O.residual_records(st, res, np.float64), np.float64)
    md = st.medium(m64, np.float64)
    glam, gmu = _fold_padding(g3[0], 8), _fold_padding(g3[1], 8)
    grho = _fold_padding(-(md["b"] ** 2) * g3[2], 8)
    c = (slice(8, -8), slice(8, -8))
    gvp, gvs, gr = (np.asarray(a, np.float64) for a in
                    gs.grad_lmd_to_vd(glam, gmu, grho, md["lam"][c], md["mu"][c], md["rho"][c]))
    mask = (w.model["vp"] < w.model["vp"].max()).astype(np.float64)
    z, x = np.mgrid[0:100, 0:100]
    for key, g in (("vp", gvp), ("vs", gvs), ("rho", gr)):
        dm = 1e-3 * np.cos(0.07 * z + 0.11 * x) * mask * m64[key]

        def L(sign):
            m2 = dict(m64)
            m2[key] = m64[key] + sign * dm
            s2, _ = O.forward(st, m2, np.float64, keep_states=False)
            q = O.section(s2, 0)
            return sum(0.5 * np.sum((q[k] - obs[k]) ** 2) for k in q)

        fd = (L(1) - L(-1)) / 2
        an = np.sum(g * dm)
        # grad_lmd_to_vd returns float32 (as the reference); the directional
        # sum cancels heavily, so the error is measured against sum |g dm|
        scale = np.sum(np.abs(g * dm))
        assert abs(fd - an) < 1e-6 * scale, (key, fd, an, scale)


def test_pcs_parameterisations_match_reference_vectors():
    """PCS chain rules and the rock-physics relations they use, against
    input/output vectors written from the reference (tests/golden/
    make_pcs_vectors.py -> pcs_vectors.npz)."""
    import os
    from pyfwi_amd import rock_physics as rp
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "pcs_vectors.npz"))
    pcs = {k: g["in_" + k] for k in ("phi", "cc", "sw")}
    gdv = {k: g["in_g" + k] for k in ("vp", "vs", "rho")}
    rock = {"k_q": 37.0, "k_c": 21.0, "k_w": 2.25, "k_h": 0.8, "mu_q": 44.0, "mu_c": 7.0,
            "rho_q": 2.65, "rho_c": 2.55, "rho_w": 1.0, "rho_h": 0.7, "cs": 20.0}
    close = lambda a, b: np.testing.assert_allclose(np.asarray(a, np.float64), b, rtol=1e-6, atol=0)
    for method in ("VRH", "Voigt", "Reuss"):
        dv = rp.pcs2dv_gassmann(pcs, rock, method)
        for k in dv:
            close(dv[k], g[f"pcs2dv_gassmann_{method}_{k}"])
        out = gs.grad_dv2pcs_gassmann(gdv, rock, pcs, method)
        for k in out:
            close(out[k], g[f"grad_dv2pcs_gassmann_{method}_{k}"])
        close(gs.grad_vrh(rock["k_c"], rock["k_q"], pcs["cc"], method), g[f"grad_vrh_{method}"])
        close(rp.vrh(rock["k_c"], rock["k_q"], pcs["cc"], method), g[f"vrh_{method}"])
    for k, v in rp.pcs2dv_vrh(pcs, rock).items():
        close(v, g["pcs2dv_vrh_" + k])
    for k, v in rp.pcs2dv_han(pcs, rock).items():
        close(v, g["pcs2dv_han_" + k])
    for k, v in gs.grad_dv2pcs_han(gdv, rock, pcs).items():
        close(v, g["grad_dv2pcs_han_" + k])
    kd, gd = rp.drained_moduli(pcs["phi"], 30.0, 20.0, rock["cs"])
    close(kd, g["drained_kd"]); close(gd, g["drained_gd"])
    for a, b in zip(rp.biot_gassmann(pcs["phi"], 2.0, 30.0, kd), ("ku", "C", "M")):
        close(a, g["biot_" + b])
    assert int(g["grad_dv2pcs_vrh_raises"]) == 1
    with pytest.raises(TypeError):
        gs.grad_dv2pcs_vrh(gdv, rock, pcs)
    out = gs.grad_dv2pcs_vrh(gdv, rock, pcs, method="VRH")
    assert set(out) == {"phi", "cc", "sw"}
    par = gs.PcsParameterization(grad_args={"method": "VRH"}, model_args={"method": "VRH"})
    for k, v in par.grad_dv2pcs(gdv, rock, pcs).items():
        close(v, g[f"grad_dv2pcs_gassmann_VRH_{k}"])
