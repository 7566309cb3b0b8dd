"""Learned warm start: the CPU oracle of the network (oracle/ws_oracle.py) is
pinned by finite differences; the device kernels (features, forward,
backward) are checked against it; and a trained network must start the
solve closer to the solution on held-out problems while the solve still
meets the same FP64 stopping rule (the iteration saving at the benchmark
size is reported by bench.py and dachshund_amd/train_warmstart.py).

Tolerances: features equal the oracle's after BF16 rounding (sets of
neighbours compared per pixel, since equidistant neighbours may be ordered
differently); predictions rel 1e-4 (1% of rows may differ where BF16
rounding flips); gradients rel 1e-3 in norm (FP32 accumulation, BF16
activations rounded identically).
"""
import numpy as np
import pytest

from oracle import ws_oracle as W


def _rand_problem(seed, n=300):
    rng = np.random.default_rng(seed)
    P = np.zeros(W.NPARAM, dtype=np.float32)
    P[W.OFF_W1:W.OFF_B1] = rng.uniform(-1, 1, W.FEAT * W.H) * np.sqrt(6 / W.FEAT)
    P[W.OFF_B1:W.OFF_W2] = rng.uniform(-0.1, 0.1, W.H)
    for o, b in ((W.OFF_W2, W.OFF_B2), (W.OFF_W3, W.OFF_B3)):
        P[o:b] = rng.uniform(-1, 1, W.H * W.H) * np.sqrt(6 / W.H)
        P[b:b + W.H] = rng.uniform(-0.1, 0.1, W.H)
    P[W.OFF_W4:W.OFF_B4] = rng.uniform(-0.3, 0.3, W.H)
    P[W.OFF_B4] = 0.05
    feat = W.bf16(rng.standard_normal((n, W.FEAT)).astype(np.float32))
    target = rng.standard_normal(n).astype(np.float32)
    return P, feat, target


def test_oracle_backward_matches_finite_differences():
    P, feat, target = _rand_problem(1, n=40)
    P = P.astype(np.float64)
    loss, g = W.loss_and_grad(P, feat, target, quant=False)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(W.NPARAM, 60, replace=False),
                          [W.OFF_B1, W.OFF_B2 + 3, W.OFF_B3, W.OFF_B4,
                           W.OFF_W4 + 5]])
    h = 1e-6
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (W.loss_and_grad(Pp, feat, target, False)[0]
              - W.loss_and_grad(Pm, feat, target, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-7, k


def test_oracle_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 3.0e38, -2.5, 0.0],
                 dtype=np.float32)
    r = W.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0   # halfway -> even
    assert r[2] == np.float32(1.0078125)
    assert r[4] == -2.5 and r[5] == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("quant", [True, False])
def test_gpu_forward_backward_vs_oracle(quant):
    from dachshund_amd.warmstart import WarmStart
    P, feat, target = _rand_problem(3, n=700)
    ws = WarmStart(seed=0, quantized=quant)
    ws.params = P
    pred, grad, loss = ws.eval_batch(feat, target)
    y, _ = W.forward(P, feat, quant)
    lo, go = W.loss_and_grad(P, feat, target, quant)
    close = np.abs(pred - y) <= 1e-4 * (1 + np.abs(y))
    # with BF16 activations a hidden value within FP32 rounding of a BF16
    # rounding boundary may round the other way: allow 1% of such rows
    assert close.mean() >= (0.99 if quant else 1.0), close.mean()
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.linalg.norm(grad - go) <= 1e-3 * np.linalg.norm(go)
    for a, b in ((W.OFF_W1, W.OFF_B1), (W.OFF_W2, W.OFF_B2),
                 (W.OFF_W3, W.OFF_B3), (W.OFF_W4, W.NPARAM)):
        assert (np.linalg.norm(grad[a:b] - go[a:b])
                <= 1e-3 * np.linalg.norm(go[a:b]) + 1e-7)


def _neighbours(f):
    nb = f[:5 * W.K].reshape(W.K, 5).astype(np.float64)
    nb = nb[np.any(nb != 0, axis=1)]
    d = 0.5 * np.sum(nb[:, :3] ** 2, axis=1)   # covariance-metric distance
    o = np.argsort(d, kind="stable")
    return d[o], nb[o]


@pytest.mark.gpu
def test_gpu_features_vs_oracle():
    """Equidistant neighbours (uniformly spaced skewer pixels) may take the
    last places in either order, so rows are compared as sorted distance
    lists and as sets of the neighbours strictly inside the 32nd distance."""
    import dachshund_amd as D
    from dachshund_amd import synth
    from dachshund_amd.warmstart import WarmStart
    w = synth.workload("small")
    pix = synth.make_pixels(w)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    ws = WarmStart(seed=0)
    fg = ws.features(wf)
    fo = W.features(pix, w.var_s, w.l_perp, w.l_para)
    np.testing.assert_array_equal(fg[:, 5 * W.K:], fo[:, 5 * W.K:])
    same_d = same_set = 0
    for a, b in zip(fg, fo):
        da, na = _neighbours(a)
        db, nbb = _neighbours(b)
        if len(da) == len(db) and np.allclose(da, db, rtol=2e-2, atol=1e-4):
            same_d += 1
        lim = 0.98 * db[-1] if len(db) else 0.0
        ia = sorted(map(tuple, np.round(na[da < lim], 2)))
        ib = sorted(map(tuple, np.round(nbb[db < lim], 2)))
        same_set += len(ia) == len(ib) and (len(ia) == 0 or np.allclose(
            np.array(ia), np.array(ib), atol=2e-2))
    n = len(pix)
    assert same_d >= 0.99 * n, same_d
    assert same_set >= 0.99 * n, same_set


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["w", "sn"])
def test_gpu_predict_equals_features_then_forward(form):
    """The one-pass prediction (features built in LDS, network on the matrix
    cores) gives the same initial guess, bit for bit, as the feature array
    followed by the separate forward kernel."""
    import dachshund_amd as D
    from dachshund_amd import synth
    from dachshund_amd.warmstart import WarmStart
    w = synth.workload("small")
    pix = synth.make_pixels(w)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    if form == "w":
        fp = D.pixel_to_wpixel(pix)
        wf = D.WFVectorW(fp, sp)
    else:
        fp = D.pixel_to_npixel(pix)
        wf = D.WFVectorSN(fp, sp)
    ws = WarmStart(seed=3)
    wf.fill_x(0.0)
    ws.predict_into(wf)
    y = wf.get_x()
    f = ws.features(wf)
    t, _, _ = ws.eval_batch(f, np.zeros(len(pix), dtype=np.float32))
    x0 = t.astype(np.float64) * (1.0 / np.sqrt(w.var_s))
    y_ref = x0 / fp[:, 3] if form == "w" else x0
    assert np.any(y != 0.0)
    assert np.array_equal(y, y_ref)


@pytest.mark.gpu
def test_gpu_predict_from_another_context():
    """The prediction does not wait on the host; a warm start on another
    context (another stream) is ordered against the problem's stream both
    ways, so it overwrites the x just filled and is in place for the read
    that follows: the same guess as from the problem's own context."""
    import dachshund_amd as D
    from dachshund_amd import synth
    from dachshund_amd.warmstart import WarmStart
    w = synth.workload("small")
    pix = synth.make_pixels(w)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    ctx2 = D.Context()
    ws_a = WarmStart(seed=5)
    ws_b = WarmStart(seed=5, ctx=ctx2)
    wf.fill_x(7.0)
    ws_a.predict_into(wf)
    ya = wf.get_x()
    for _ in range(3):
        wf.fill_x(7.0)
        ws_b.predict_into(wf)
        yb = wf.get_x()
        assert np.array_equal(ya, yb)
    assert not np.any(ya == 7.0)
    ws_b.close()
    ctx2.close()


@pytest.mark.gpu
def test_trained_warm_start_cuts_iterations():
    """Train on solutions of 12 seeded problems, evaluate on 4 held-out
    seeds against x0 = 0, same FP64 stopping rule."""
    import dachshund_amd as D
    from dachshund_amd import synth
    from dachshund_amd.warmstart import WarmStart
    from oracle import ora
    base = synth.workload("small")
    ws = WarmStart(seed=7)
    sp = D.SignalCovarParams(base.var_s, base.l_perp, base.l_para)
    for seed in range(100, 112):
        pix = synth.make_pixels(base, seed=seed)
        wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
        wf.fill_x(0.0)
        r = wf.solve_resident(D.PCGParams(wf.n, 1000, 50, 1e-8))
        assert r.success
        ws.add_sample(wf, seed=seed)
    loss = ws.train(epochs=40, batch=2048, lr=2e-3, lr_decay=0.95, seed=1)
    assert loss[-1] < 0.5 * loss[0]
    gains, ratios = [], []
    c = ora.covar(base.var_s, base.l_perp, base.l_para)
    for seed in range(900, 904):
        pix = synth.make_pixels(base, seed=seed)
        wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
        prm = D.PCGParams(wf.n, 1000, 50, 1e-6)
        x = np.zeros(wf.n)
        r0 = wf.solve_pcg(prm, x)
        ws.predict_into(wf)
        y0 = wf.get_x()
        x = y0.copy()
        r1 = wf.solve_pcg(prm, x)
        assert r0.success and r1.success
        fp = ora.to_form(pix, ora.FORM_W)
        rt = ora.A_residual(ora.FORM_W, fp, c, x / fp[:, 3])
        assert np.linalg.norm(rt) <= 1e-6 * ora.norm_b(ora.FORM_W, fp) * 1.5
        gains.append(r0.num_iters - r1.num_iters)
        # error of the initial guess in the norm CG minimises (A-norm, W
        # space), relative to that of x0 = 0
        y_star = x / fp[:, 3]
        e = y0 - y_star
        ratios.append(np.sqrt(e @ ora.A_product(ora.FORM_W, fp, c, e)
                              / (y_star @ ora.A_product(ora.FORM_W, fp, c,
                                                        y_star))))
    # the predicted guess is closer in the A-norm on every held-out problem,
    # and costs no iterations on average
    assert max(ratios) < 0.9, ratios
    assert np.mean(gains) >= 0, gains


@pytest.mark.gpu
def test_committed_network_gain_on_held_out_300k():
    """The committed network (dachshund_amd/warmstart_c5.npz, trained on
    seeds 1000-1199 of the 300,000-pixel workload) on held-out seeds 2001-
    2003, each solve to the FP64 tolerance 1e-6 from x0 = 0 and from the
    prediction, on a fresh problem (so the prediction's time includes the
    neighbour search):

      * plain FP64 PCG with Jacobi (the reference's solver): at least 7 %
        fewer iterations, and less time including the neighbour search and
        the prediction;
      * the mixed solver with block-Jacobi (the benchmark's solver): at
        least 2 fewer iterations on average (about 6 %).  Here the saved
        iterations
        save about as much time as a fresh geometry's neighbour search
        and prediction cost; the gain in time needs the geometry's
        neighbour lists to be reused (several data vectors over the same
        pixels)."""
    import os
    import time
    import dachshund_amd as D
    from dachshund_amd import synth
    from dachshund_amd.warmstart import WarmStart
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    ws = WarmStart()
    ws.load(os.path.join(root, "dachshund_amd", "warmstart_c5.npz"))
    w = synth.workload("c5")
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    ctx = D.default_context()
    it0, it1, t0s, t1s, mit0, mit1 = [], [], [], [], [], []
    for seed in (2001, 2002, 2003):
        pw = D.pixel_to_wpixel(synth.make_pixels(w, seed=seed))
        prm = D.PCGParams(len(pw), w.max_iter, w.step_r, 1e-6)
        for learned in (False, True):
            wf = D.WFVectorW(pw, sp)
            ctx.synchronize()
            t = time.perf_counter()
            if learned:
                ws.predict_into(wf)
            else:
                wf.fill_x(0.0)
            r = wf.solve_resident(prm)
            dt = time.perf_counter() - t
            assert r.success
            (it1 if learned else it0).append(r.num_iters)
            (t1s if learned else t0s).append(dt)
            wf.set_precond(D.DS_PRECOND_BLOCK)
            if learned:
                ws.predict_into(wf)
            else:
                wf.fill_x(0.0)
            rm = wf.solve_resident(prm, mixed=True)
            assert rm.success and rm.true_norm <= 1e-6 * wf.norm_b() * 1.0001
            (mit1 if learned else mit0).append(rm.num_iters)
            wf.close()
    print(f"FP64 Jacobi iterations {it0} -> {it1}, s {t0s} -> {t1s}; "
          f"mixed block-Jacobi {mit0} -> {mit1}")
    assert np.sum(it1) <= 0.93 * np.sum(it0), (it0, it1)
    assert np.sum(t1s) < np.sum(t0s), (t0s, t1s)
    assert np.sum(mit1) <= np.sum(mit0) - 2 * len(mit0), (mit0, mit1)
