"""Batched solves of independent subvolumes (one launch per kernel for all
of them, hipGraph-captured iterations, autotuned launch shape) against the
same subvolumes solved one by one.

Each subvolume's arithmetic and reduction order in a batch are those of the
single-volume solve with the same launch shape, so results must be
bit-identical; graph replay must not change them; the autotuner may change
only the order of sums (solutions then agree to the solve tolerance).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _subvolumes(count=6, seed=4):
    from dachshund_amd import synth
    subs = synth.batch_workloads(seed=seed, count=count, n_min=500,
                                 n_max=6000)
    return subs, [synth.make_pixels(w) for w in subs]


@pytest.mark.parametrize("far", [False, True])
@pytest.mark.parametrize("mixed", [False, True])
def test_batch_bit_identical_to_single(mixed, far):
    """far: one subvolume gets a pixel 60 Mpc/h away from the others, so
    its FP32 kernel is the direct-difference form while the rest of the
    batch keeps the MFMA form; each subvolume must still match its own
    single-volume solve bit for bit."""
    import dachshund_amd as D
    subs, pix = _subvolumes()
    if far:
        pix[2] = pix[2].copy()
        pix[2][-1, 0] += 60.0
    sps = [D.SignalCovarParams(w.var_s, w.l_perp, w.l_para) for w in subs]
    wpix = [D.pixel_to_wpixel(p) for p in pix]
    batch = D.BatchWF(wpix, sps)
    batch.set_launch(4, False)
    n = batch.n
    prm = D.PCGParams(n, 1000, 50, 1e-6)
    batch.fill_x(0.0)
    res = batch.solve(prm, mixed=mixed, use_graph=False)
    xb = batch.split(batch.get_x())
    for k, (w, p, sp) in enumerate(zip(subs, wpix, sps)):
        wf = D.WFVectorW(p, sp)
        wf.set_launch(4, False)
        x = np.zeros(len(p))
        r = (wf.solve_mixed(D.PCGParams(len(p), 1000, 50, 1e-6), x)
             if mixed else wf.solve_pcg(D.PCGParams(len(p), 1000, 50, 1e-6),
                                        x))
        assert res[k].success and r.success
        assert res[k].num_iters == r.num_iters, k
        assert np.array_equal(xb[k], x), k
    if far:
        # the FP32 operator of the mixed batch against the FP64 one
        rng = np.random.default_rng(3)
        a = rng.standard_normal(n)
        b64 = batch.split(batch.A_product(a, D.DS_PREC_F64))
        b32 = batch.split(batch.A_product(a, D.DS_PREC_F32))
        for u, v in zip(b32, b64):
            assert np.linalg.norm(u - v) <= 1e-5 * np.linalg.norm(v)
    # graph replay of the same iterations
    if mixed:
        batch.fill_x(0.0)
        res_g = batch.solve(prm, mixed=True, use_graph=True)
        assert [r.num_iters for r in res_g] == [r.num_iters for r in res]
        for a, b in zip(batch.split(batch.get_x()), xb):
            assert np.array_equal(a, b)


def test_batch_meets_fp64_tolerance_per_subvolume(ora):
    import dachshund_amd as D
    subs, pix = _subvolumes(count=5, seed=9)
    sps = [D.SignalCovarParams(w.var_s, w.l_perp, w.l_para) for w in subs]
    batch = D.BatchWF([D.pixel_to_wpixel(p) for p in pix], sps)
    waves, order, ms = batch.autotune()
    assert waves in (1, 2, 4) and ms > 0
    batch.fill_x(0.0)
    res = batch.solve(D.PCGParams(batch.n, 1000, 50, 1e-6), mixed=True)
    for k, (w, p, x) in enumerate(zip(subs, pix, batch.split(batch.get_x()))):
        assert res[k].success
        fp = ora.to_form(p, ora.FORM_W)
        c = ora.covar(w.var_s, w.l_perp, w.l_para)
        r = ora.A_residual(ora.FORM_W, fp, c, x / fp[:, 3])
        assert np.linalg.norm(r) <= 1.0001e-6 * ora.norm_b(ora.FORM_W, fp), k


def test_batch_operator_is_block_diagonal(ora):
    import dachshund_amd as D
    subs, pix = _subvolumes(count=3, seed=2)
    sps = [D.SignalCovarParams(w.var_s, w.l_perp, w.l_para) for w in subs]
    batch = D.BatchWF([D.pixel_to_wpixel(p) for p in pix], sps)
    a = np.random.default_rng(0).standard_normal(batch.n)
    b = batch.split(batch.A_product(a))
    for k, (w, p, ak) in enumerate(zip(subs, pix, batch.split(a))):
        fp = ora.to_form(p, ora.FORM_W)
        c = ora.covar(w.var_s, w.l_perp, w.l_para)
        bound = ora.A_product(ora.FORM_W, fp, c, np.abs(ak))
        assert np.all(np.abs(b[k] - ora.A_product(ora.FORM_W, fp, c, ak))
                      <= 1e-12 * bound), k


def _realizations(R, seed=21):
    from dachshund_amd import synth
    w = synth.workload("small")
    pix, ds = synth.make_pixels(w, seed=seed, n_real=R)
    sets = []
    for k in range(R):
        p = pix.copy()
        p[:, 4] = ds[k]
        sets.append(p)
    return w, sets


def _multi_rhs(batch, mode):
    if mode == "explicit":
        batch.build_dense_bf16()
    else:
        batch.set_fused(True)


@pytest.mark.parametrize("mode", ["explicit", "fused"])
def test_dense_bf16_block_product(ora, mode):
    """S built in BF16 (relative rounding 2^-9) times BF16 directions, FP32
    accumulation: each row within 8e-3 of sum_j |S_ij u_j| of the FP64
    product; the identity part is exact."""
    import dachshund_amd as D
    w, sets = _realizations(5)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    batch = D.BatchWF([D.pixel_to_wpixel(p) for p in sets], [sp] * len(sets))
    _multi_rhs(batch, mode)
    a = np.random.default_rng(1).standard_normal(batch.n)
    b = batch.split(batch.dense_product(a))
    c = ora.covar(w.var_s, w.l_perp, w.l_para)
    errs = []
    for k, (p, ak) in enumerate(zip(sets, batch.split(a))):
        fp = ora.to_form(p, ora.FORM_W)
        ref = ora.A_product(ora.FORM_W, fp, c, ak)
        bound = ora.A_product(ora.FORM_W, fp, c, np.abs(ak)) - np.abs(ak)
        assert np.all(np.abs(b[k] - ref) <= 8e-3 * bound + 1e-12), k
        errs.append(np.linalg.norm(b[k] - ref) / np.linalg.norm(ref))
    assert max(errs) < 3e-3, errs


@pytest.mark.parametrize("mode", ["explicit", "fused"])
def test_dense_bf16_block_solve(ora, mode):
    """Multi-RHS mixed solve with the explicit BF16 block as inner operator:
    every right-hand side meets the FP64 tolerance (oracle residual) and
    matches its own FP64 solve."""
    import dachshund_amd as D
    w, sets = _realizations(6, seed=22)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    batch = D.BatchWF([D.pixel_to_wpixel(p) for p in sets], [sp] * len(sets))
    _multi_rhs(batch, mode)
    batch.fill_x(0.0)
    tol = 1e-6
    res = batch.solve(D.PCGParams(batch.n, 2000, 50, tol), mixed=True,
                      use_graph=True)
    c = ora.covar(w.var_s, w.l_perp, w.l_para)
    for k, (p, x) in enumerate(zip(sets, batch.split(batch.get_x()))):
        assert res[k].success and res[k].n_true >= 1, k
        fp = ora.to_form(p, ora.FORM_W)
        r = ora.A_residual(ora.FORM_W, fp, c, x / fp[:, 3])
        assert np.linalg.norm(r) <= 1.0001 * tol * ora.norm_b(ora.FORM_W, fp)
        xo, ok, _, _ = ora.solve_pcg(ora.FORM_W, fp, c, 5000, 50, 1e-12)
        assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo), k


@pytest.mark.parametrize("waves", [1, 4])
def test_multi_rhs_residual_equals_per_subvolume(waves):
    """The one-pass FP64 residual of a shared-geometry batch (the mixed
    solver's true residual when a multi-RHS operator is selected: the same
    FP64 pair values, multiplied into all right-hand sides on the FP64
    matrix cores) equals the per-subvolume parity kernel up to FP64
    rounding of the row sums: |difference| <= 1e-13 (sum_j |A_ij x_j| +
    |b_i|) per row (A is entrywise non-negative, so A |x| is that sum)."""
    import dachshund_amd as D
    w, sets = _realizations(7, seed=31)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    batch = D.BatchWF([D.pixel_to_wpixel(p) for p in sets], [sp] * len(sets))
    batch.set_launch(waves, True)
    batch.set_tile(64)
    x = np.random.default_rng(2).standard_normal(batch.n)
    r_sub = batch.A_residual(x)
    scale = batch.A_product(np.abs(x)) + np.abs(r_sub + batch.A_product(x))
    batch.set_fused(True)
    r_one = batch.A_residual(x)
    assert np.all(np.abs(r_one - r_sub) <= 1e-13 * scale)
    assert not np.array_equal(scale, 0 * scale)


@pytest.mark.parametrize("rows", [16, 32])
def test_batch_tile_shapes_equal_single(rows):
    """At the sub-block-culled FP64 tile shapes a batch's plain FP64 solve
    of each subvolume still equals that subvolume solved alone, bit for
    bit, and agrees with the 64-row shape to the solve tolerance."""
    import dachshund_amd as D
    subs, pix = _subvolumes(count=4)
    sps = [D.SignalCovarParams(w.var_s, w.l_perp, w.l_para) for w in subs]
    wpix = [D.pixel_to_wpixel(p) for p in pix]
    batch = D.BatchWF(wpix, sps)
    batch.set_launch(4, False)
    batch.set_tile(rows)
    prm = D.PCGParams(batch.n, 1000, 50, 1e-8)
    batch.fill_x(0.0)
    res = batch.solve(prm, mixed=False, use_graph=False)
    xb = batch.split(batch.get_x())
    batch.set_tile(64)
    batch.fill_x(0.0)
    batch.solve(prm, mixed=False, use_graph=False)
    x64 = batch.split(batch.get_x())
    for k, (p, sp) in enumerate(zip(wpix, sps)):
        wf = D.WFVectorW(p, sp)
        wf.set_launch(4, False)
        wf.set_tile(rows)
        x = np.zeros(len(p))
        r = wf.solve_pcg(D.PCGParams(len(p), 1000, 50, 1e-8), x)
        assert res[k].success and r.success
        assert np.array_equal(xb[k], x), k
        assert np.linalg.norm(xb[k] - x64[k]) <= 1e-6 * np.linalg.norm(x64[k])


def test_autotune_picks_a_shape_and_keeps_results():
    """The autotuner times every FP64 tile shape x waves x order and keeps
    the fastest; the solution at the tuned shape agrees with the default
    shape's to the solve tolerance (only the order of sums changes)."""
    import dachshund_amd as D
    subs, pix = _subvolumes(count=3)
    sps = [D.SignalCovarParams(w.var_s, w.l_perp, w.l_para) for w in subs]
    batch = D.BatchWF([D.pixel_to_wpixel(p) for p in pix], sps)
    prm = D.PCGParams(batch.n, 1000, 50, 1e-8)
    batch.fill_x(0.0)
    batch.solve(prm, mixed=False, use_graph=False)
    x0 = batch.get_x()
    waves, order, ms = batch.autotune(prec=D.DS_PREC_F64, reps=3)
    w, o, rows, _ = batch.launch()
    assert (w, o) == (waves, order) and rows in (16, 32, 64) and ms > 0
    batch.fill_x(0.0)
    batch.solve(prm, mixed=False, use_graph=False)
    x1 = batch.get_x()
    assert np.linalg.norm(x1 - x0) <= 1e-6 * np.linalg.norm(x0)
