"""Block-Jacobi preconditioner (exact inverses of the 64-row target blocks
of A; no reference counterpart, the reference preconditions with the
diagonal).

  * Apply: M^-1 r on the device equals the oracle's per-block
    inv(A_bb) r_b, with the blocks taken from the host planner's ordering
    (the same k-d blocks the problem uses): FP32-stored inverses, so
    |s - s_ref|_b <= 1e-5 |inv(A_bb)| |r_b| per block.
  * Solves: mixed and plain PCG with the block preconditioner meet the same
    FP64 tolerance (residual recomputed by the oracle) in fewer iterations
    than with Jacobi.
  * Batches: bit-identical to the single-volume solve.
"""
import numpy as np
import pytest

from conftest import fixture_pixels

pytestmark = pytest.mark.gpu


def _dense_A(ora, form, fp, c):
    n = len(fp)
    A = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = ora.A_product(form, fp, c, e)
    return A


def _cases():
    from dachshund_amd import synth
    rng = np.random.default_rng(5)
    p = np.empty((200, 5))
    p[:, :3] = rng.uniform(0, 8, (200, 3))
    p[:, 3] = rng.uniform(0.1, 0.3, 200)
    p[:, 4] = rng.standard_normal(200)
    return {"fixture": fixture_pixels(),
            "tiny": synth.make_pixels(synth.workload("tiny")),
            "n200": p}


@pytest.mark.parametrize("form", ["w", "sn"])
@pytest.mark.parametrize("case", ["fixture", "tiny", "n200"])
def test_block_precond_apply(ora, case, form):
    import dachshund_amd as D
    pix = _cases()[case]
    n = len(pix)
    if n > 700:   # dense oracle matrix: keep it small
        pix = pix[:700]
        n = 700
    params = (0.05, 2.5, 2.1)
    sp = D.SignalCovarParams(*params)
    f = 1 if form == "w" else 0
    fp = ora.to_form(pix, f)
    c = ora.covar(*params)
    wf = (D.WFVectorW(D.pixel_to_wpixel(pix), sp) if form == "w"
          else D.WFVectorSN(D.pixel_to_npixel(pix), sp))
    wf.set_precond(D.DS_PRECOND_BLOCK)
    r = np.random.default_rng(n).standard_normal(n)
    s = wf.preconditioner_solve(r)
    A = _dense_A(ora, f, fp, c)
    perm, _, _ = D.plan_tiles(pix[:, :3], sp)
    for b in range(len(perm) // 64):
        idx = perm[64 * b:64 * b + 64]
        idx = idx[:max(0, min(64, n - 64 * b))]
        if len(idx) == 0:
            continue
        Minv = np.linalg.inv(A[np.ix_(idx, idx)])
        ref = Minv @ r[idx]
        bound = np.abs(Minv) @ np.abs(r[idx])
        assert np.all(np.abs(s[idx] - ref) <= 1e-5 * bound + 1e-300), b
    # Jacobi is still the default of a fresh problem and bit-identical
    wf2 = (D.WFVectorW(D.pixel_to_wpixel(pix), sp) if form == "w"
           else D.WFVectorSN(D.pixel_to_npixel(pix), sp))
    assert np.array_equal(wf2.preconditioner_solve(r), ora.precond(f, fp, c, r))


@pytest.mark.parametrize("case", ["fixture", "n200"])
def test_block_precond_r32_apply_is_the_sequential_sum(case):
    """The mixed solver's block-Jacobi application (its cross-lane moves by
    whole-wave DPP rotations, the lower diagonals Horner-fashion) equals,
    bit for bit, the sequential form: for row i of a block, the FP64 sum over
    k = 0..63 in order of M_{i,(i+k) mod 64} r_j, j = (i+k) mod 64, with the
    stored FP32 inverse and r rounded to FP32.  (The product of two floats
    is exact in FP64, so each step rounds once, as an fma does.)"""
    import dachshund_amd as D
    pix = _cases()[case]
    n = len(pix)
    sp = D.SignalCovarParams(0.05, 2.5, 2.1)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    wf.set_precond(D.DS_PRECOND_BLOCK)
    rng = np.random.default_rng(3)
    r = rng.standard_normal(n) * np.exp(rng.uniform(-6, 6, n))
    s = wf.preconditioner_solve_r32(r)
    minv = wf.geometry("minv").reshape(-1, 33, 64).astype(np.float64)
    perm = wf.geometry("perm")
    real = perm >= 0
    rs = np.zeros(len(perm))
    rs[real] = r[perm[real]]
    rf = rs.astype(np.float32).astype(np.float64)
    i = np.arange(64)
    out = np.empty(len(perm))
    for b in range(len(perm) // 64):
        mt, rb = minv[b], rf[64 * b:64 * b + 64]
        acc = np.zeros(64)
        for k in range(64):
            j = (i + k) % 64
            m = mt[k, i] if k <= 32 else mt[64 - k, j]
            acc = acc + m * rb[j]
        out[64 * b:64 * b + 64] = acc
    ref = np.empty(n)
    ref[perm[real]] = out[real]
    assert np.array_equal(s, ref)
    # and it is the FP64-broadcast application to within the FP32 rounding
    # of r
    s64 = wf.preconditioner_solve(r)
    assert np.allclose(s, s64, rtol=1e-6, atol=1e-6 * np.abs(s64).max())


@pytest.mark.parametrize("mixed", [False, True])
def test_block_precond_solves(ora, mixed):
    import dachshund_amd as D
    from dachshund_amd import synth
    w = synth.workload("small")
    pix = synth.make_pixels(w)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    fp = ora.to_form(pix, 1)
    c = ora.covar(w.var_s, w.l_perp, w.l_para)
    tol = 1e-6
    iters = {}
    for kind in (D.DS_PRECOND_JACOBI, D.DS_PRECOND_BLOCK):
        wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
        wf.set_precond(kind)
        x = np.zeros(len(pix))
        prm = D.PCGParams(len(pix), 2000, 50, tol)
        res = wf.solve_mixed(prm, x) if mixed else wf.solve_pcg(prm, x)
        assert res.success
        r = ora.A_residual(1, fp, c, x / fp[:, 3])
        nb = ora.norm_b(1, fp)
        # plain PCG stops on the recurrence residual (lib/pcg.h:124-130),
        # the mixed solver on the FP64 true residual
        assert np.linalg.norm(r) <= (1.0001 if mixed else 1.1) * tol * nb
        iters[kind] = res.num_iters
    assert iters[D.DS_PRECOND_BLOCK] < iters[D.DS_PRECOND_JACOBI], iters


def test_block_precond_batch_equals_single():
    import dachshund_amd as D
    from dachshund_amd import synth
    subs = synth.batch_workloads(seed=9, count=4, n_min=500, n_max=4000)
    pix = [D.pixel_to_wpixel(synth.make_pixels(w)) for w in subs]
    sps = [D.SignalCovarParams(w.var_s, w.l_perp, w.l_para) for w in subs]
    batch = D.BatchWF(pix, sps)
    batch.set_launch(4, False)
    batch.set_precond(D.DS_PRECOND_BLOCK)
    batch.fill_x(0.0)
    res = batch.solve(D.PCGParams(batch.n, 1000, 50, 1e-6), mixed=True,
                      use_graph=True)
    xb = batch.split(batch.get_x())
    for k, (p, sp) in enumerate(zip(pix, sps)):
        wf = D.WFVectorW(p, sp)
        wf.set_launch(4, False)
        wf.set_precond(D.DS_PRECOND_BLOCK)
        x = np.zeros(len(p))
        r = wf.solve_mixed(D.PCGParams(len(p), 1000, 50, 1e-6), x)
        assert res[k].success and r.success
        assert res[k].num_iters == r.num_iters, k
        assert np.array_equal(xb[k], x), k


def test_block_precond_shared_geometry_batch_equals_single():
    """Realizations over one geometry share their block inverses (one set,
    applied to every subvolume from LDS): bit-identical to the
    single-volume solves."""
    import dachshund_amd as D
    from dachshund_amd import synth
    w = synth.workload("small")
    pix, ds = synth.make_pixels(w, n_real=4)
    sets = []
    for k in range(4):
        p = pix.copy()
        p[:, 4] = ds[k]
        sets.append(D.pixel_to_wpixel(p))
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    batch = D.BatchWF(sets, [sp] * 4)
    batch.set_launch(4, False)
    batch.set_precond(D.DS_PRECOND_BLOCK)
    batch.fill_x(0.0)
    prm = D.PCGParams(batch.n, 1000, 50, 1e-6)
    res = batch.solve(prm, mixed=True, use_graph=True)
    xb = batch.split(batch.get_x())
    for k, p in enumerate(sets):
        wf = D.WFVectorW(p, sp)
        wf.set_launch(4, False)
        wf.set_precond(D.DS_PRECOND_BLOCK)
        x = np.zeros(len(p))
        r = wf.solve_mixed(D.PCGParams(len(p), 1000, 50, 1e-6), x)
        assert res[k].success and r.success
        assert res[k].num_iters == r.num_iters, k
        assert np.array_equal(xb[k], x), k
