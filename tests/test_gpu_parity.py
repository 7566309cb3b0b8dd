"""GPU parity: the gfx950 kernels (through the C ABI) against the oracle and
the reference's golden vectors.

Tolerances:
  * FP64 pair kernels: every pair's table argument and cutoff decision are
    the reference's; values differ by a few ulps and rows are summed in
    another order, so |b_gpu - b_ref|_i <= 1e-12 * sum_j |A_ij a_j|.
  * Jacobi preconditioner and |b|: bit-identical.
  * FP32 pair kernel: |b32 - b_ref|_i <= 2e-5 * sum_j |A_ij a_j| plus, for
    each pair whose table argument lies within 1e-3 of the cutoff (where
    FP32 rounding may move it across), the jump var_s e^-8 |A-weight a_j|.
  * PCG: same success flag; iteration counts equal to +-1 on the
    synthetic problems, and within 10% on the reference fixture, whose
    weights reach 1.15e4 (condition ~1e7, where rounding alone moves CG's
    path); residual histories within 1e-6 relative over the first
    iterations; and the GPU solution at least as accurate as the
    reference's own: |x_gpu - x*| <= 2 |x_ref - x*| + 1e-12 |x*| against
    the dense solution x*.
  * Maps: smp_product and sigma_m_noise_dom to 1e-12 of sum_j |S_ij x_j|.
"""
import numpy as np
import pytest

from conftest import fixture_pixels, golden

pytestmark = pytest.mark.gpu

FORMS = {"w": 1, "sn": 0}


@pytest.fixture(scope="module")
def D():
    import dachshund_amd as D
    D.default_context()
    return D


def _problem(D, form, pix, params):
    sp = D.SignalCovarParams(*params)
    if form == "w":
        return D.WFVectorW(D.pixel_to_wpixel(pix), sp), sp
    return D.WFVectorSN(D.pixel_to_npixel(pix), sp), sp


def _cases():
    from dachshund_amd import synth
    rng = np.random.default_rng(12)
    cases = {"fixture": fixture_pixels(),
             "small": synth.make_pixels(synth.workload("small")),
             "tiny": synth.make_pixels(synth.workload("tiny"))}
    for n in (1, 2, 63, 64, 65, 129):
        p = np.empty((n, 5))
        p[:, :3] = rng.uniform(0, 6, (n, 3))
        p[:, 3] = rng.uniform(0.1, 0.3, n)
        p[:, 4] = rng.standard_normal(n)
        cases[f"n{n}"] = p
    # coincident points and a far-away loner
    p = np.repeat(cases["n65"][:10], 7, axis=0)
    p[-1, :3] = [1e4, -3e3, 5e5]
    cases["dups"] = p
    return cases


CASES = _cases()


@pytest.mark.parametrize("params", [(0.05, 2.5, 2.1), (0.05, 1.0, 1.0),
                                    (0.3, 0.7, 3.3)])
@pytest.mark.parametrize("form", ["w", "sn"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_matvec64_residual_precond(D, ora, case, form, params):
    pix = CASES[case]
    f = FORMS[form]
    fp = ora.to_form(pix, f)
    c = ora.covar(*params)
    wf, _ = _problem(D, form, pix, params)
    a = np.random.default_rng(len(pix)).standard_normal(len(pix))
    bound = ora.A_product(f, fp, c, np.abs(a))
    b = wf.A_product(a)
    assert np.all(np.abs(b - ora.A_product(f, fp, c, a)) <= 1e-12 * bound)
    r = wf.A_residual(a)
    rb = bound + np.abs(fp[:, 4])
    assert np.all(np.abs(r - ora.A_residual(f, fp, c, a)) <= 1e-12 * rb)
    assert np.array_equal(wf.preconditioner_solve(a), ora.precond(f, fp, c, a))
    assert wf.norm_b() == ora.norm_b(f, fp)
    b32 = wf.A_product(a, prec=D.DS_PREC_F32)
    assert np.all(np.abs(b32 - b) <= 2e-5 * bound +
                  _cutoff_allowance(fp, form, params, a))


F32_KINDS = {"valu": 0, "mfma": 1, "mfma_clamped": 2}


@pytest.mark.parametrize("kind", sorted(F32_KINDS))
@pytest.mark.parametrize("form", ["w", "sn"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_matvec32_kernels(D, ora, case, form, kind):
    """Both FP32 pair kernels (table argument expanded on the VALU or on the
    matrix cores) meet the FP32 row bound against the oracle."""
    pix = CASES[case]
    f = FORMS[form]
    params = (0.05, 2.5, 2.1)
    fp = ora.to_form(pix, f)
    c = ora.covar(*params)
    wf, _ = _problem(D, form, pix, params)
    wf.set_f32_kernel(F32_KINDS[kind])
    a = np.random.default_rng(len(pix) + 7).standard_normal(len(pix))
    bound = ora.A_product(f, fp, c, np.abs(a))
    b32 = wf.A_product(a, prec=D.DS_PREC_F32)
    ref = ora.A_product(f, fp, c, a)
    assert np.all(np.abs(b32 - ref) <= 2e-5 * bound +
                  _cutoff_allowance(fp, form, params, a))


@pytest.mark.parametrize("kind", sorted(F32_KINDS))
def test_matvec32_per_pair(D, ora, kind):
    """Per pair: A e_j is column j of A, so each entry is one pair's
    covariance.  FP32 covariance error <= 2e-6 var_s per pair (SN form, so
    the entry is S_ij itself) away from the cutoff, and the diagonal
    var_s + N_j."""
    pix = CASES["small"]
    params = (0.05, 2.5, 2.1)
    fp = ora.to_form(pix, 0)
    c = ora.covar(*params)
    wf, _ = _problem(D, "sn", pix, params)
    wf.set_f32_kernel(F32_KINDS[kind])
    n = len(pix)
    for j in np.random.default_rng(3).choice(n, 6, replace=False):
        e = np.zeros(n)
        e[j] = 1.0
        col = wf.A_product(e, prec=D.DS_PREC_F32)
        ref = ora.A_product(0, fp, c, e)
        allow = _cutoff_allowance(fp, "sn", params, e)
        assert np.all(np.abs(col - ref) <= 2e-6 * params[0] + allow), j


def _cutoff_allowance(fp, form, params, a):
    """Per row: var_s e^-8 * weight_i * sum over pairs within 1e-3 of the
    table cutoff (argument 199 in table units) of weight_j |a_j|."""
    var_s, lp, ll = params
    cp, cl = 1 / (2 * lp * lp), 1 / (2 * ll * ll)
    wj = fp[:, 3] if form == "w" else np.ones(len(fp))
    out = np.zeros(len(fp))
    for i0 in range(0, len(fp), 512):
        d = fp[i0:i0 + 512, None, :3] - fp[None, :, :3]
        xx = (cp * (d[..., 0] ** 2 + d[..., 1] ** 2)
              + cl * d[..., 2] ** 2) * 24.875
        near = np.abs(xx - 199.0) < 1e-3
        out[i0:i0 + 512] = near.astype(float) @ (wj * np.abs(a))
    return 1.01 * var_s * np.exp(-8.0) * wj * out


@pytest.mark.parametrize("lp,ll", [(1.0, 1.0), (2.0, 2.0), (2.5, 2.1)])
@pytest.mark.parametrize("form", ["w", "sn"])
def test_matvec64_golden(D, ora, form, lp, ll):
    g = golden("ops_fixture.npz")
    pix = fixture_pixels()
    f = FORMS[form]
    fp = ora.to_form(pix, f)
    c = ora.covar(0.05, lp, ll)
    wf, _ = _problem(D, form, pix, (0.05, lp, ll))
    vec = g["vec"]
    bound = ora.A_product(f, fp, c, np.abs(vec))
    key = f"{form}_{lp}_{ll}"
    assert np.all(np.abs(wf.A_product(vec) - g[key + "_prod"])
                  <= 1e-12 * bound)
    assert np.all(np.abs(wf.A_residual(vec) - g[key + "_res"])
                  <= 1e-12 * (bound + np.abs(fp[:, 4])))
    assert np.array_equal(wf.preconditioner_solve(vec), g[key + "_pc"])
    assert wf.norm_b() == g[key + "_normb"][0]


@pytest.mark.parametrize("waves", [1, 2, 4])
def test_launch_shape_changes_only_rounding(D, ora, waves):
    pix = CASES["small"]
    wf, _ = _problem(D, "w", pix, (0.05, 2.5, 2.1))
    a = np.random.default_rng(2).standard_normal(len(pix))
    fp = ora.to_form(pix, 1)
    c = ora.covar(0.05, 2.5, 2.1)
    bound = ora.A_product(1, fp, c, np.abs(a))
    wf.set_waves(waves)
    b = wf.A_product(a)
    assert np.all(np.abs(b - ora.A_product(1, fp, c, a)) <= 1e-12 * bound)
    assert np.array_equal(b, wf.A_product(a))   # deterministic


def _check_pcg(res, x, g_ok, g_it, g_norms, g_x, x_star, ill):
    assert res.success == g_ok
    if ill:
        assert abs(res.num_iters - g_it) <= max(1, 0.1 * g_it)
    else:
        assert abs(res.num_iters - g_it) <= 1
    m = min(len(res.residual_norms), len(g_norms), 10)
    np.testing.assert_allclose(res.residual_norms[:m], g_norms[:m],
                               rtol=1e-6)
    err_gpu = np.linalg.norm(x - x_star)
    err_ref = np.linalg.norm(g_x - x_star)
    assert err_gpu <= 2 * err_ref + 1e-12 * np.linalg.norm(x_star), \
        (err_gpu, err_ref)


def test_pcg_golden(D, ora):
    g = golden("pcg.npz")
    pix = fixture_pixels()
    small = g["small_pixels"]
    dense = {}
    for k in range(int(g["n_setups"])):
        form, vs, lp, ll, mi, sr, tol, src = g[f"s{k}_setup"]
        p = pix if src == 0 else small
        fname = "w" if int(form) == 1 else "sn"
        key = (int(form), vs, lp, ll, int(src))
        if key not in dense:
            dense[key] = ora.solve_cf(int(form), ora.to_form(p, int(form)),
                                      ora.covar(vs, lp, ll))
        wf, _ = _problem(D, fname, p, (vs, lp, ll))
        x = np.zeros(len(p))
        res = wf.solve_pcg(D.PCGParams(len(p), int(mi), int(sr), tol), x)
        _check_pcg(res, x, bool(g[f"s{k}_success"]), int(g[f"s{k}_iters"]),
                   g[f"s{k}_norms"], g[f"s{k}_x"], dense[key], src == 0)
    wf, _ = _problem(D, "w", small, (0.05, 2.5, 2.1))
    x = g["warm_x0"].copy()
    res = wf.solve_pcg(D.PCGParams(len(small), 1000, 50, 1e-6), x)
    _check_pcg(res, x, bool(g["warm_success"]), int(g["warm_iters"]),
               g["warm_norms"], g["warm_x"], dense[(1, 0.05, 2.5, 2.1, 1)],
               False)


def test_pcg_semantics(D, ora):
    pix = CASES["small"]
    wf, _ = _problem(D, "w", pix, (0.05, 2.5, 2.1))
    n = len(pix)
    fp = ora.to_form(pix, 1)
    c = ora.covar(0.05, 2.5, 2.1)
    for mi in (0, 1, 2, 3, 7):
        x = np.zeros(n)
        res = wf.solve_pcg(D.PCGParams(n, mi, 50, 1e-30), x)
        xo, ok, it, norms = ora.solve_pcg(1, fp, c, mi, 50, 1e-30)
        assert (res.success, res.num_iters, len(res.residual_norms)) == \
            (ok, it, len(norms)), mi
        np.testing.assert_allclose(res.residual_norms, norms, rtol=1e-10)
        np.testing.assert_allclose(x, xo, rtol=1e-9, atol=1e-12)
    # step_r = 1: a true residual every iteration
    x = np.zeros(n)
    res = wf.solve_pcg(D.PCGParams(n, 500, 1, 1e-8), x)
    xo, ok, it, norms = ora.solve_pcg(1, fp, c, 500, 1, 1e-8)
    assert res.success and ok and abs(res.num_iters - it) <= 1
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
    # |b| = 0: early return, success false, x untouched then x *= w
    z = pix.copy()
    z[:, 4] = 0.0
    wz, _ = _problem(D, "w", z, (0.05, 2.5, 2.1))
    x = np.ones(n)
    res = wz.solve_pcg(D.PCGParams(n, 100, 10, 1e-6), x)
    assert not res.success and res.num_iters == 0
    assert len(res.residual_norms) == 0
    np.testing.assert_array_equal(x, 1.0 / pix[:, 3])
    # bad parameters are errors, not exits
    from dachshund_amd import _lib
    with pytest.raises(_lib.DsError):
        wf.solve_pcg(D.PCGParams(n, 10, 0, 1e-6), np.zeros(n))
    with pytest.raises(_lib.DsError):
        wf.solve_pcg(D.PCGParams(n + 1, 10, 5, 1e-6), np.zeros(n))


@pytest.mark.parametrize("form", ["w", "sn"])
def test_mixed_solver_meets_fp64_tolerance(D, ora, form):
    pix = CASES["small"]
    f = FORMS[form]
    fp = ora.to_form(pix, f)
    c = ora.covar(0.05, 2.5, 2.1)
    wf, _ = _problem(D, form, pix, (0.05, 2.5, 2.1))
    n = len(pix)
    tol = 1e-8
    x = np.zeros(n)
    res = wf.solve_mixed(D.PCGParams(n, 2000, 50, tol), x, refine_drop=1e-3)
    assert res.success and res.n_true >= 1
    # the FP64 true residual of the returned solution, by the oracle
    y = x / fp[:, 3] if form == "w" else x
    r = ora.A_residual(f, fp, c, y)
    assert np.linalg.norm(r) <= tol * ora.norm_b(f, fp) * 1.0001
    xo, ok, _, _ = ora.solve_pcg(f, fp, c, 5000, 50, 1e-12)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("form", ["w", "sn"])
def test_mixed_solver_from_converged_start(D, ora, form):
    """The mixed solver's initial residual comes from the FP32 operator; a
    start that already meets the tolerance must still end on an FP64 true
    residual (at least one step, one true residual) and stay a solution."""
    pix = CASES["small"]
    f = FORMS[form]
    fp = ora.to_form(pix, f)
    c = ora.covar(0.05, 2.5, 2.1)
    wf, _ = _problem(D, form, pix, (0.05, 2.5, 2.1))
    n = len(pix)
    tol = 1e-6
    x0, ok, _, _ = ora.solve_pcg(f, fp, c, 5000, 50, 1e-12)
    assert ok
    # solve_pcg's x: W form y in, x = w .* y out (lib/wf_vector.cc:268-271)
    x = x0 / fp[:, 3] if form == "w" else x0.copy()
    res = wf.solve_mixed(D.PCGParams(n, 2000, 50, tol), x, refine_drop=1e-3)
    assert res.success and res.n_true >= 1 and res.num_iters >= 1
    assert res.n_matvec32 == res.num_iters + 1
    assert res.n_matvec64 == res.n_true
    y = x / fp[:, 3] if form == "w" else x
    r = ora.A_residual(f, fp, c, y)
    assert np.linalg.norm(r) <= tol * ora.norm_b(f, fp) * 1.0001
    assert np.linalg.norm(x - x0) <= 1e-5 * np.linalg.norm(x0)


@pytest.mark.parametrize("precond", ["jacobi", "block"])
@pytest.mark.parametrize("form", ["w", "sn"])
@pytest.mark.parametrize("case", ["n1", "n2", "n63", "n65", "n129", "dups"])
def test_mixed_solver_edge_cases(D, ora, case, form, precond):
    """Sizes around the 64-row block (padding rows in the last block), one
    and two pixels, and coincident pixels with a far-away loner (whose
    problem runs the direct-difference FP32 form): the mixed solver with
    either preconditioner meets the FP64 tolerance the oracle recomputes."""
    pix = CASES[case]
    f = FORMS[form]
    fp = ora.to_form(pix, f)
    c = ora.covar(0.05, 2.5, 2.1)
    wf, _ = _problem(D, form, pix, (0.05, 2.5, 2.1))
    wf.set_precond(D.DS_PRECOND_BLOCK if precond == "block"
                   else D.DS_PRECOND_JACOBI)
    n = len(pix)
    tol = 1e-8
    x = np.zeros(n)
    res = wf.solve_mixed(D.PCGParams(n, 2000, 50, tol), x, refine_drop=1e-3)
    assert res.success and res.n_true >= 1
    assert np.all(np.isfinite(x))
    y = x / fp[:, 3] if form == "w" else x
    r = ora.A_residual(f, fp, c, y)
    assert np.linalg.norm(r) <= tol * ora.norm_b(f, fp) * 1.0001


def test_resident_solution_and_map(D, ora):
    """app/dachshund.cc:218-264 end to end on the 8,192-pixel workload:
    W form from x = 0, then smp_product onto the 16x16x64 grid."""
    g = golden("app_c1.npz")
    pix = g["pixels"]
    cfg = dict(line.split(" = ") for line in str(g["config"]).splitlines())
    params = (float(cfg["corr_var_s"]), float(cfg["corr_l_perp"]),
              float(cfg["corr_l_para"]))
    wf, sp = _problem(D, "w", pix, params)
    n = len(pix)
    wf.set_x(np.zeros(n))
    res = wf.solve_resident(D.PCGParams(n, int(cfg["pcg_max_iter"]),
                                        int(cfg["pcg_step_r"]),
                                        float(cfg["pcg_tol"])))
    assert res.success
    it_ref = int(str(g["log_tail"]).split("iter ")[1].split()[0])
    assert abs(res.num_iters - it_ref) <= 1
    shape = (int(cfg["map_nx"]), int(cfg["map_ny"]), int(cfg["map_nz"]))
    box = (float(cfg["lx"]), float(cfg["ly"]), float(cfg["lz"]))
    m = wf.smp_grid(*box, *shape)
    ref = g["map"]
    x = wf.get_x()
    # accuracy against a tight FP64 solve: the GPU map is as close to it as
    # the reference's own map (both stop at |r| <= 1e-5 |b|)
    wf.set_x(np.zeros(n))
    wf.solve_resident(D.PCGParams(n, 5000, 50, 1e-12))
    m_star = wf.smp_grid(*box, *shape)
    err_gpu = np.linalg.norm(m - m_star)
    err_ref = np.linalg.norm(ref - m_star)
    assert err_gpu <= 2 * err_ref + 1e-12 * np.linalg.norm(m_star)
    assert np.linalg.norm(m - ref) <= 1e-4 * np.linalg.norm(ref)
    # same map from the oracle's own projection of the GPU solution
    mo = ora.smp_product(pix[:, :3], ora.map_coords(box, shape),
                         ora.covar(*params), x)
    bound = ora.smp_product(pix[:, :3], ora.map_coords(box, shape),
                            ora.covar(*params), np.abs(x))
    assert np.all(np.abs(m - mo) <= 1e-12 * bound)


@pytest.mark.parametrize("case", ["small", "n1", "n65", "dups", "ragged"])
def test_grid_lists_device_equal_host(D, ora, monkeypatch, case):
    """The map grid's tile lists are built on the device by a full scan;
    the host's binned builder (selected by DS_HOST_GRID_LISTS=1) must give
    the same lists, hence bit-identical maps, for ragged grids (sizes not a
    multiple of 4), a single pixel, coincident pixels and a far-away loner.
    The maps themselves are checked against the oracle to 1e-12."""
    from dachshund_amd import synth
    rng = np.random.default_rng(31)
    if case == "small":
        w = synth.workload("small")
        pix = synth.make_pixels(w)
        box, shape = w.box, w.map_shape
    else:
        pix = dict(_cases())[case] if case != "ragged" else None
        if pix is None:
            pix = np.empty((300, 5))
            pix[:, :3] = rng.uniform([0, 0, 0], [9, 7, 30], (300, 3))
            pix[:, 3] = rng.uniform(0.1, 0.3, 300)
            pix[:, 4] = rng.standard_normal(300)
        box, shape = (9.0, 7.0, 30.0), (13, 7, 45)
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    x = rng.standard_normal(len(pix))
    monkeypatch.setenv("DS_HOST_GRID_LISTS", "1")
    m_host = D.smp_product_grid(pix[:, :3], box, shape, sp, x)
    monkeypatch.delenv("DS_HOST_GRID_LISTS")
    m_dev = D.smp_product_grid(pix[:, :3], box, shape, sp, x)
    assert np.array_equal(m_dev, m_host)
    mc = ora.map_coords(box, shape)
    mo = ora.smp_product(pix[:, :3], mc, ora.covar(0.05, 2.5, 2.1), x)
    bound = ora.smp_product(pix[:, :3], mc, ora.covar(0.05, 2.5, 2.1),
                            np.abs(x))
    assert np.all(np.abs(m_dev - mo) <= 1e-12 * bound)


def test_map_side_golden(D, ora):
    g = golden("map_fixture.npz")
    pix = fixture_pixels()
    for lp, ll in [(2.0, 2.0), (2.5, 2.1)]:
        key = f"{lp}_{ll}"
        shape = tuple(int(v) for v in g["shape_" + key])
        sp = D.SignalCovarParams(0.05, lp, ll)
        c = ora.covar(0.05, lp, ll)
        mc = ora.map_coords((20.0, 20.0, 20.0), shape)
        bound = ora.smp_product(pix[:, :3], mc, c, np.abs(g["x"]))
        m = D.smp_product_grid(pix[:, :3], (20.0, 20.0, 20.0), shape, sp,
                               g["x"])
        assert np.all(np.abs(m - g["smp_" + key]) <= 1e-12 * bound)
        m2 = D.smp_product(pix[:, :3], mc, sp, g["x"])
        assert np.all(np.abs(m2 - g["smp_" + key]) <= 1e-12 * bound)
        s = D.sigma_m_noise_dom(D.pixel_to_npixel(pix), mc, sp)
        assert np.all(np.abs(s - g["nd_" + key]) <= 1e-12 * g["nd_" + key]
                      + 1e-300)


def test_map_covar_diag_golden(D, ora):
    """MapCovarDiag on the reference fixture against the reference's own
    sigma, flags and iteration counts (lib/map_covar.cc:73-113)."""
    g = golden("map_fixture.npz")
    pix = fixture_pixels()
    mc = ora.map_coords((10.0, 10.0, 10.0), (1, 1, 5))
    sp = D.SignalCovarParams(0.05, 2.0, 2.0)
    s, it, ok = D.map_covar_diag(D.pixel_to_npixel(pix), mc, sp,
                                 D.PCGParams(len(pix), 1000, 50, 1e-2))
    np.testing.assert_array_equal(ok, g["mcov_status"][:, 0] == 1)
    # the fixture's weights reach 1.15e4: iteration counts within 10%, and
    # where the paths are equally long sigma agrees to rounding
    g_it = g["mcov_status"][:, 1]
    assert np.all(np.abs(it - g_it) <= np.maximum(1, 0.1 * g_it))
    same = it == g_it
    assert same.sum() >= 3
    np.testing.assert_allclose(s[same], g["mcov_sigma"][same], rtol=1e-6)
    np.testing.assert_allclose(s, g["mcov_sigma"], rtol=2e-2)


def test_map_covar_diag_batches_equal_single_cells(D, ora):
    """Cells are solved as shared-geometry batches of 32: 70 cells (two
    full batches and a ragged one of 6, cells far from every pixel among
    them) give the same sigma, flags and iteration counts bit for bit as
    one call per cell, and agree with the oracle."""
    pix = CASES["small"]
    npx = D.pixel_to_npixel(pix)
    rng = np.random.default_rng(21)
    lo, hi = pix[:, :3].min(0), pix[:, :3].max(0)
    mc = rng.uniform(lo, hi, (70, 3))
    mc[[5, 40, 69]] = hi + 50.0          # |b| = 0: no solve, sigma 0
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    prm = D.PCGParams(len(pix), 200, 50, 1e-6)
    s, it, ok = D.map_covar_diag(npx, mc, sp, prm)
    assert not ok[[5, 40, 69]].any() and np.all(s[[5, 40, 69]] == 0.0)
    assert ok.sum() == 67
    for i in (0, 5, 31, 32, 63, 64, 68, 69):
        s1, it1, ok1 = D.map_covar_diag(npx, mc[i:i + 1], sp, prm)
        assert s1[0] == s[i] and it1[0] == it[i] and ok1[0] == ok[i]
    c = ora.covar(0.05, 2.5, 2.1)
    sel = np.r_[0:4, 30:34, 62:70]
    s_ref, it_ref, ok_ref = ora.map_covar_diag(
        ora.to_form(pix, ora.FORM_SN), mc[sel], c, 200, 50, 1e-6)
    np.testing.assert_array_equal(ok[sel], ok_ref)
    assert np.all(np.abs(it[sel] - it_ref) <= 1)
    np.testing.assert_allclose(s[sel], s_ref, rtol=1e-5)


@pytest.mark.parametrize("case", ["tiny", "n1", "n65", "dups"])
def test_map_points_edge_cases(D, ora, case):
    pix = CASES[case]
    rng = np.random.default_rng(5)
    lo, hi = pix[:, :3].min(0) - 3, pix[:, :3].max(0) + 3
    mc = rng.uniform(lo, hi, (333, 3))
    mc[:5] = pix[:5 if len(pix) >= 5 else len(pix), :3].repeat(
        5 // min(len(pix), 5) + 1, axis=0)[:5]
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    c = ora.covar(0.05, 2.5, 2.1)
    x = rng.standard_normal(len(pix))
    m = D.smp_product(pix[:, :3], mc, sp, x)
    bound = ora.smp_product(pix[:, :3], mc, c, np.abs(x))
    assert np.all(np.abs(m - ora.smp_product(pix[:, :3], mc, c, x))
                  <= 1e-12 * bound)
    npx = D.pixel_to_npixel(pix)
    s = D.sigma_m_noise_dom(npx, mc, sp)
    so = ora.sigma_m_noise_dom(npx, mc, c)
    assert np.all(np.abs(s - so) <= 1e-12 * so)


def _cutoff_pairs(lp, ll, span=40):
    """Pixel offsets (dy, dz) whose table argument, computed in the
    reference's operation order (lib/wf_vector.cc:62-66,
    lib/signal_covar.h:53-77; dx = 0), lies within 199 (1 +- span 2^-52):
    found by stepping one coordinate an ulp at a time along the pure
    perpendicular, pure parallel and a mixed direction."""
    cp, cl = 1 / (2 * lp * lp), 1 / (2 * ll * ll)
    dx_inv = 1.0 / (8.0 / 199)
    lo, hi = 199.0 * (1 - span * 2.0 ** -52), 199.0 * (1 + span * 2.0 ** -52)

    def xx_of(dy, dz):
        return (cp * (0.0 * 0.0 + dy * dy) + cl * (dz * dz)) * dx_inv

    out = set()
    for frac in (0.0, 0.37, 1.0):
        for sign in (1.0, -1.0):
            dy, dz = np.sqrt(frac * 8.0 / cp), np.sqrt((1 - frac) * 8.0 / cl)
            while True:
                x = xx_of(dy, dz)
                if (sign > 0 and x > hi) or (sign < 0 and x < lo):
                    break
                if lo <= x <= hi:
                    out.add((dy, dz, x))
                if frac == 1.0:
                    dy = np.nextafter(dy, sign * np.inf)
                else:
                    dz = np.nextafter(dz, sign * np.inf)
    return sorted(out, key=lambda t: t[2])


@pytest.mark.parametrize("form", ["sn", "w"])
@pytest.mark.parametrize("lp,ll", [(2.5, 2.1), (1.0, 1.0), (3.3, 1.7)])
def test_matvec64_cutoff_decision(D, ora, form, lp, ll):
    """Pairs whose table argument straddles the cutoff (argument 8, table
    index 199) within +-62 ulps: the FP64 kernel's off-diagonal entries
    (columns A e_j, read through one product since every pair is isolated
    25 Mpc/h from the next) have the oracle's zero pattern exactly and its
    values to 8 ulps.  The oracle's own decision is pinned by the
    reference's rule: zero exactly when the argument reaches 199."""
    pairs = _cutoff_pairs(lp, ll)
    m = len(pairs)
    pix = np.zeros((2 * m, 5))
    for p, (dy, dz, _) in enumerate(pairs):
        pix[2 * p, 0] = pix[2 * p + 1, 0] = 25.0 * p
        pix[2 * p + 1, 1] = dy
        pix[2 * p + 1, 2] = dz
    rng = np.random.default_rng(m)
    pix[:, 3] = rng.uniform(0.1, 0.3, 2 * m)
    pix[:, 4] = rng.standard_normal(2 * m)
    f = FORMS[form]
    fp = ora.to_form(pix, f)
    c = ora.covar(0.05, lp, ll)
    wf, _ = _problem(D, form, pix, (0.05, lp, ll))
    xs = np.array([t[2] for t in pairs])
    for side in (0, 1):
        a = np.zeros(2 * m)
        a[side::2] = 1.0
        rows = slice(1 - side, None, 2)
        ref = ora.A_product(f, fp, c, a)[rows]
        got = wf.A_product(a)[rows]
        assert np.array_equal(ref == 0.0, xs >= 199.0)
        assert np.array_equal(got == 0.0, ref == 0.0)
        assert np.all(np.abs(got - ref) <= 8 * np.spacing(np.abs(ref)))
    assert 0 < np.sum(xs >= 199.0) < m


@pytest.mark.parametrize("rows", [16, 32, 64])
@pytest.mark.parametrize("form", ["w", "sn"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_matvec64_tile_shapes(D, ora, case, form, rows):
    """The FP64 kernel's tile shapes (16, 32 or 64 rows per group x 16
    slots; the smaller ones skip the sub-blocks the cutoff box test
    culls): A_product and A_residual meet the 1e-12 row bound against the
    oracle at every shape, and the launch settings report the shape."""
    pix = CASES[case]
    f = FORMS[form]
    params = (0.05, 2.5, 2.1)
    fp = ora.to_form(pix, f)
    c = ora.covar(*params)
    wf, _ = _problem(D, form, pix, params)
    wf.set_tile(rows)
    assert wf.launch()[2] == rows
    a = np.random.default_rng(len(pix) + 3).standard_normal(len(pix))
    bound = ora.A_product(f, fp, c, np.abs(a))
    assert np.all(np.abs(wf.A_product(a) - ora.A_product(f, fp, c, a))
                  <= 1e-12 * bound)
    rb = bound + np.abs(fp[:, 4])
    assert np.all(np.abs(wf.A_residual(a) - ora.A_residual(f, fp, c, a))
                  <= 1e-12 * rb)


@pytest.mark.parametrize("rows", [16, 32])
def test_matvec64_cutoff_decision_tiles(D, ora, rows):
    """The cutoff-straddling pairs through the sub-block-culled shapes:
    the oracle's zero pattern exactly, values to 8 ulps."""
    pairs = _cutoff_pairs(2.5, 2.1)
    m = len(pairs)
    pix = np.zeros((2 * m, 5))
    for p, (dy, dz, _) in enumerate(pairs):
        pix[2 * p, 0] = pix[2 * p + 1, 0] = 25.0 * p
        pix[2 * p + 1, 1] = dy
        pix[2 * p + 1, 2] = dz
    pix[:, 3] = 0.2
    fp = ora.to_form(pix, 0)
    c = ora.covar(0.05, 2.5, 2.1)
    wf, _ = _problem(D, "sn", pix, (0.05, 2.5, 2.1))
    wf.set_tile(rows)
    a = np.zeros(2 * m)
    a[0::2] = 1.0
    ref = ora.A_product(0, fp, c, a)[1::2]
    got = wf.A_product(a)[1::2]
    assert np.array_equal(got == 0.0, ref == 0.0)
    assert np.all(np.abs(got - ref) <= 8 * np.spacing(np.abs(ref)))


def test_context_released_before_its_problem(D):
    """A context destroyed while a problem made on it is alive (the order a
    garbage collector may choose at interpreter exit) stays usable by that
    problem until the problem is destroyed: the library counts the
    context's holders and frees it with the last one."""
    from dachshund_amd import synth
    ctx = D.Context()
    w = synth.workload("small")
    pix = synth.make_pixels(w)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp, ctx=ctx)
    a = np.ones(len(pix))
    before = wf.A_product(a)
    ctx.close()
    assert np.array_equal(wf.A_product(a), before)
    wf.close()


def test_creation_rejects_non_finite_pixels(D):
    """Problem creation checks the records on several host threads while the
    upload runs beside them: a non-finite coordinate is reported before a
    non-finite weight or datum, each at its smallest record index, the
    upload's buffer goes back to the context, and the next problem builds."""
    from dachshund_amd import _lib
    from dachshund_amd import synth
    pix = synth.make_pixels(synth.workload("c1"))
    big = np.concatenate([pix] * 10)            # 81,920 records: three runs
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    bad = big.copy()
    bad[40000, 3] = np.nan
    bad[70000, 0] = np.nan
    bad[50000, 2] = np.inf
    with pytest.raises(_lib.DsError, match="coordinate 2 of record 50000"):
        D.WFVectorW(D.pixel_to_wpixel(bad), sp)
    bad[50000, 2] = big[50000, 2]
    bad[70000, 0] = big[70000, 0]
    with pytest.raises(_lib.DsError, match="pixel 40000: weight or data"):
        D.WFVectorW(D.pixel_to_wpixel(bad), sp)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    a = np.random.default_rng(3).standard_normal(len(pix))
    ref = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    assert np.array_equal(wf.A_product(a), ref.A_product(a))
    wf.close()
    ref.close()
