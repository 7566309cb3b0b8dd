"""Problems built from a context's buffer cache.

A destroyed problem's device buffers go back to its context, and the next
problems take them.  With DS_POOL_POISON=1 the context fills every buffer it
takes back with 0xFF bytes, so a kernel that reads a buffer before writing
it would see NaNs or wild indices.  Every result of problems built from such
buffers must equal, bit for bit, the same problem on a fresh context."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pixels(seed):
    from dachshund_amd import synth
    w = synth.workload("c1")
    return synth.make_pixels(w, seed=seed), w


def _run_all(D, ctx, seed):
    """Every path that reads a problem's buffers, on one problem."""
    from dachshund_amd.warmstart import WarmStart
    pix, w = _pixels(seed)
    n = len(pix)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp, ctx=ctx)
    a = np.random.default_rng(seed).standard_normal(n)
    out = {"A64": wf.A_product(a), "A32": wf.A_product(a, D.DS_PREC_F32)}
    prm = D.PCGParams(n, w.max_iter, w.step_r, w.tol)
    x = np.zeros(n)
    r = wf.solve_pcg(prm, x)
    out["pcg"] = x.copy()
    out["pcg_norms"] = np.asarray(r.residual_norms).copy()
    wf.set_precond(D.DS_PRECOND_BLOCK)
    out["bj"] = wf.preconditioner_solve(a)
    x = np.zeros(n)
    wf.solve_mixed(prm, x)
    out["mixed_bj"] = x.copy()
    ws = WarmStart(seed=0, ctx=ctx)
    ws.predict_into(wf)
    out["predicted"] = wf.get_x()
    r = wf.solve_resident(prm, mixed=True)
    out["learned"] = wf.get_x()
    out["learned_iters"] = np.array([r.num_iters])
    lx, ly, lz = w.box
    out["map"] = wf.smp_grid(lx, ly, lz, *w.map_shape)
    ws.close()
    wf.close()
    return out


def test_reused_poisoned_buffers_give_identical_results():
    import dachshund_amd as D
    os.environ["DS_POOL_POISON"] = "1"
    try:
        pctx = D.Context(0)
    finally:
        del os.environ["DS_POOL_POISON"]
    cctx = D.Context(0)
    # seed 11 fills the cache; 12 (another geometry, same pixel count) and
    # 11 again then run on poisoned buffers
    _run_all(D, pctx, 11)
    for seed in (12, 11):
        got = _run_all(D, pctx, seed)
        ref = _run_all(D, cctx, seed)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (seed, k)
        assert np.all(np.isfinite(got["map"]))
