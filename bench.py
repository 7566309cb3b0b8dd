"""Benchmark: Wiener-filter PCG time to relative residual 1e-6 on one GPU.

Workload: 300,000 pixels (1000 skewers x 300), box 64 x 64 x 600 Mpc/h, a
Gaussian random field (var_s 0.05, l_perp 2.5, l_para 2.1) drawn from a seed,
noise U[0.1, 0.3], W form (the reference app's system, app/dachshund.cc:234),
x0 = 0 (or the learned warm start when a trained network is given).

A step is one complete solve to |r| <= 1e-6 |b| with the inputs resident
in HBM.  The headline solver is the mixed FP32/FP64 one (FP32 pair
evaluation, FP64 vectors and true residuals) started from the learned
warm start (neighbour search + network inference, inside the timed step);
plain FP64 PCG from x0 = 0 (the reference's algorithm on the GPU) is timed
on the same problem for the speed-up.  The benchmark realization (seed
2000) is held out from the network's training seeds (1000-1199).  `value`
is evaluated pixel-pair covariances per second over the timed solves.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                  [--warmstart dachshund_amd/warmstart_c5.npz | none]
  python bench.py --workload c3    # 32-RHS explicit BF16 block (MFMA)
  python bench.py --workload c4    # 512 subvolumes, batched
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Ceiling of the pair kernels: one evaluated pair needs at least 12 VALU
# lane-operations with pre-scaled coordinates (3 differences, 3 for the
# quadratic form, index, fraction, cutoff select, 2 for the interpolation, 1
# multiply-accumulate) and one transcendental or table fetch, so
# P_peak = min(R_VALU / 12, R_trans) with, on MI355X's 256 CUs at 2.4 GHz,
# R_VALU = 256 x 64 lanes x 2.4e9 (non-packed) and R_trans = 256 x 16 x
# 2.4e9.
VALU_OPS_PER_PAIR = 12
CLOCK_HZ = 2.4e9
R_VALU = 256 * 64 * CLOCK_HZ
R_TRANS = 256 * 16 * CLOCK_HZ
PAIR_PEAK = min(R_VALU / VALU_OPS_PER_PAIR, R_TRANS)   # 3.28e12 pairs/s
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one FP32 operator launch per pixel: its record as
# FP32 (16 B), its FP32 direction (4 B) and its FP64 output (8 B); of one
# FP64 launch: the FP64 record (32 B), FP64 input (8 B) and output (8 B).
ALGO_BYTES_PER_PIXEL = {"f32": 28, "f64": 48}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_ms(wf, prec):
    """Average launch duration (ms, HIP events on the problem's stream) of
    the pair kernel of precision `prec`: 200 launches to reach the steady
    clock of a running solve, then the median of 5 averages of 25."""
    wf.time_matvec(prec, reps=200)
    ts = sorted(wf.time_matvec(prec, reps=25) for _ in range(5))
    return ts[2]


def pmc_traffic(*kernels):
    """HBM bytes per launch of the kernels together on the 300k-pixel
    problem, from the calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    summarised in profiles/ (null if that summary lacks one of them)."""
    path = os.path.join(ROOT, "profiles", "pair_kernel_traffic_c5_300k.txt")
    if not os.path.exists(path):
        return None
    found = {}
    for line in open(path):
        f = line.split()
        if f and f[0] in kernels:
            found[f[0]] = float(f[1]) + float(f[2])
    return sum(found.values()) if len(found) == len(kernels) else None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def pairs_inside_cutoff(pix, l_perp, l_para):
    """Ordered pairs i != j whose table argument is inside the cutoff
    (argument 8, lib/signal_covar.h:53-65): the nonzero off-diagonal
    entries of A, counted with a k-d tree in the scaled metric."""
    from scipy.spatial import cKDTree
    dx_inv = 199.0 / 8.0
    s = np.array([np.sqrt(dx_inv / (2 * l_perp * l_perp))] * 2 +
                 [np.sqrt(dx_inv / (2 * l_para * l_para))])
    tree = cKDTree(pix[:, :3] * s)
    return int(tree.count_neighbors(tree, np.sqrt(199.0))) - len(pix)


def cpu_baseline(pix_w, var_s, l_perp, l_para, n_full, n_mv_full,
                 budget_s=12.0):
    """The reference's own WFVectorW::A_product (lib/wf_vector.cc:196-220),
    compiled from its sources into oracle/_ref/libdsref.so with its
    platforms/edison-gnu flags, timed on this host's cores on a bounded
    subset of the pixels; returns the dense pair rate and the extrapolated
    full-size solve time."""
    import ctypes as C
    lib_path = os.path.join(ROOT, "oracle", "_ref", "libdsref.so")
    if not os.path.exists(lib_path):
        return None
    L = C.CDLL(lib_path)
    dp = C.POINTER(C.c_double)
    L.dsref_w_A_product.argtypes = [C.c_int, dp, C.c_double, C.c_double,
                                    C.c_double, dp, dp]
    L.dsref_num_threads.restype = C.c_int
    threads = L.dsref_num_threads()
    n = 8000
    rate = None
    while True:
        sub = np.ascontiguousarray(pix_w[:n])
        a = np.random.default_rng(0).standard_normal(n)
        b = np.empty(n)
        t0 = time.perf_counter()
        L.dsref_w_A_product(n, sub.ctypes.data_as(dp), var_s, l_perp, l_para,
                            a.ctypes.data_as(dp), b.ctypes.data_as(dp))
        dt = time.perf_counter() - t0
        rate = n * n / dt
        log(f"cpu reference A_product n={n}: {dt:.2f} s, "
            f"{rate:.3e} pairs/s on {threads} threads")
        nxt = int(n * min(4.0, max(1.2, (budget_s / max(dt, 1e-3)) ** 0.5)))
        if dt > budget_s / 4 or nxt >= len(pix_w) or n >= 120000:
            break
        n = min(nxt, len(pix_w))
    return {
        "value": rate, "unit": "pair-evals/s", "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "reference",
        "sample": (f"one WFVectorW::A_product of the reference (all n^2 "
                   f"pairs, OpenMP) on the first {n} pixels of the same "
                   f"workload"),
        "extrapolated_solve_s": n_full * n_full * n_mv_full / rate,
    }


def bench_block(args):
    """100,000 pixels, 32 noise realizations of one field: multi-RHS mixed
    solve whose inner operator is the explicit BF16 S on MFMA (one GEMM
    per iteration for all right-hand sides), FP64 true residuals; against
    the matrix-free FP32 inner operator and plain FP64 PCG on the same
    batch.  Roofline of the GEMM: HBM bytes per launch (S + U + partials)
    over its time, and MFMA FLOP/s."""
    import dachshund_amd as D
    from dachshund_amd import synth
    w = synth.workload("c3")
    R = 32
    t0 = time.perf_counter()
    pix, ds = synth.make_pixels(w, n_real=R)
    sets = []
    for k in range(R):
        p = pix.copy()
        p[:, 4] = ds[k]
        sets.append(D.pixel_to_wpixel(p))
    n = len(pix)
    log(f"c3: {R} realizations of {n} pixels in {time.perf_counter() - t0:.1f}"
        f" s")
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    batch = D.BatchWF(sets, [sp] * R)
    kind = (D.DS_PRECOND_BLOCK if args.precond == "block"
            else D.DS_PRECOND_JACOBI)
    t = time.perf_counter()
    batch.build_dense_bf16()
    t_build = time.perf_counter() - t
    log(f"explicit BF16 S ({2 * n * n / 1e9:.1f} GB) built in {t_build:.2f} s")
    prm = D.PCGParams(batch.n, 2000, w.step_r, args.tol)

    def run(mode, k):
        ts, res = [], None
        for _ in range(k):
            batch.fill_x(0.0)
            t = time.perf_counter()
            if mode == "dense":
                batch.set_dense(True)
                batch.set_precond(kind)
                res = batch.solve(prm, mixed=True, use_graph=True)
            elif mode == "fused":
                batch.set_fused(True)
                batch.set_precond(kind)
                res = batch.solve(prm, mixed=True, use_graph=True)
            elif mode == "mixed":
                batch.set_dense(False)
                batch.set_fused(False)
                batch.set_precond(kind)
                res = batch.solve(prm, mixed=True, use_graph=True)
            else:
                batch.set_precond(D.DS_PRECOND_JACOBI)
                res = batch.solve(prm, mixed=False, use_graph=False)
            ts.append(time.perf_counter() - t)
            assert all(r.success for r in res), mode
        return float(np.mean(ts)), res

    out = {}
    xs = {}
    for mode in ("dense", "fused", "mixed", "fp64"):
        run(mode, args.warmup)
        tm, res = run(mode, args.steps)
        out[mode] = {"s": tm, "iters_max": max(r.num_iters for r in res),
                     "fp64_true_residuals": max(r.n_true for r in res)}
        log(f"{mode}: {tm * 1e3:.1f} ms for {R} RHS, iterations <= "
            f"{out[mode]['iters_max']}")
        xs[mode] = batch.get_x()
    # accuracy against tight FP64 solves of every right-hand side
    batch.fill_x(0.0)
    batch.solve(D.PCGParams(batch.n, 4000, w.step_r, 1e-10), mixed=False,
                use_graph=False)
    x_tight = batch.get_x()
    errs = {m: max(float(np.linalg.norm(a - b) / np.linalg.norm(b))
                   for a, b in zip(batch.split(xs[m]), batch.split(x_tight)))
            for m in ("dense", "fused", "mixed")}
    batch.set_dense(True)
    ms = batch.time_dense(reps=10)
    batch.set_fused(True)
    ms_fused = batch.time_dense(reps=10)
    batch.set_fused(False)
    st = batch.stats()
    pairs1 = st.pairs_per_matvec / R     # one geometry's 64 x 16 pairs
    log(f"fused multi-RHS kernel {ms_fused:.3f} ms: {pairs1:.3e} S tiles "
        f"pairs x {R} RHS ({2 * pairs1 * R / (ms_fused * 1e-3) / 1e12:.0f} "
        f"MFMA TFLOP/s, {pairs1 / (ms_fused * 1e-3) / 1e9:.0f} Gpair/s)")
    # algorithmic bytes of one block product: S once (BF16), the R
    # directions (BF16) and the R results (FP32)
    bytes_launch = 2.0 * n * n + 2.0 * R * n + 4.0 * R * n
    flops = 2.0 * n * n * R
    gbs = bytes_launch / (ms * 1e-3) / 1e9
    tfs = flops / (ms * 1e-3) / 1e12
    log(f"GEMM {ms:.3f} ms: {gbs:.0f} GB/s, {tfs:.0f} TFLOP/s")
    print(json.dumps({
        "metric": "multi-RHS PCG time-to-rel-residual 1e-6 (32 RHS, explicit "
                  "BF16 S on MFMA, FP64 refinement)",
        "value": R * n * n * (out["dense"]["iters_max"]) / out["dense"]["s"],
        "unit": "dense-equivalent pair-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": out["dense"]["s"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 S, fp32 accumulate, fp64 vectors",
        "data": "synthetic (seeded GRF, 32 noise realizations)",
        "config": {"workload": f"{n} pixels (400 skewers x 250), box "
                               f"{w.box}, 32 RHS, seed {w.seed}",
                   "tol": args.tol,
                   "precond": {"dense": args.precond, "fused": args.precond,
                               "mixed": args.precond, "fp64": "jacobi"}},
        "time_to_tol_s": {k: v["s"] for k, v in out.items()},
        "iterations_max": {k: v["iters_max"] for k, v in out.items()},
        "s_build_explicit_S": t_build,
        "x_rel_l2_vs_fp64_tight_max": errs,
        "fused_kernel": {"ms_per_launch": ms_fused,
                         "pairs_per_launch": pairs1,
                         "pair_evals_per_s": pairs1 / (ms_fused * 1e-3),
                         "mfma_tflops": 2 * pairs1 * R / (ms_fused * 1e-3)
                         / 1e12},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": None, "kernel": "k_gemm_bf16",
                     "ms_per_launch": ms, "bytes_per_launch": bytes_launch,
                     "mfma_tflops": tfs, "mfma_peak_tflops": 2500.0},
    }), flush=True)


def bench_chunk(args):
    """76,800 pixels (240 skewers x 320, lognormal noise) and the 60 x 48 x
    880 map grid (2.53M cells): plain FP64 PCG and the mixed solve to
    1e-6, then the S_mp x projection onto the grid (lib/map.cc:20-43), the
    step that dominates this configuration in the reference (1.95e11 pairs
    dense)."""
    import dachshund_amd as D
    from dachshund_amd import synth
    w = synth.workload("c2")
    pix = synth.make_pixels(w)
    n = len(pix)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    st = wf.stats()
    prm = D.PCGParams(n, w.max_iter, w.step_r, w.tol)
    log(f"c2: n={n}, setup {st.setup_seconds:.2f} s, {st.pairs_per_matvec:.3e}"
        f" pairs per FP64 matvec")

    def solve(mixed, block):
        wf.set_precond(D.DS_PRECOND_BLOCK if block else D.DS_PRECOND_JACOBI)
        wf.fill_x(0.0)
        return (wf.solve_resident(prm, mixed=True) if mixed
                else wf.solve_resident(prm))

    out = {}
    for name, mixed, block in (("plain_fp64", False, False),
                               ("mixed_block_jacobi", True, True)):
        for _ in range(args.warmup):
            solve(mixed, block)
        ts = []
        for _ in range(args.steps):
            t = time.perf_counter()
            r = solve(mixed, block)
            ts.append(time.perf_counter() - t)
            assert r.success, name
        out[name] = {"s": float(np.mean(ts)), "iters": r.num_iters}
        log(f"{name}: {np.mean(ts) * 1e3:.1f} ms, {r.num_iters} iters")
    ms64 = kernel_ms(wf, D.DS_PREC_F64)
    ms32 = kernel_ms(wf, D.DS_PREC_F32)
    log(f"matvec: FP64 {ms64:.3f} ms "
        f"({st.f64_pairs_per_matvec / ms64 * 1e3:.3e} "
        f"pairs/s), FP32 {ms32:.3f} ms "
        f"({st.sub_pairs_per_matvec / ms32 * 1e3:.3e} pairs/s)")
    box, shape = w.box, w.map_shape
    cells = int(np.prod(shape))
    wf.smp_grid(*box, *shape)
    ts = []
    for _ in range(max(args.steps, 1)):
        t = time.perf_counter()
        m = wf.smp_grid(*box, *shape)
        ts.append(time.perf_counter() - t)
    t_map = float(np.mean(ts))
    log(f"S_mp x onto {shape} ({cells} cells): {t_map * 1e3:.1f} ms "
        f"(host copy of the map included), {n * cells / t_map:.3e} "
        f"dense-equivalent pairs/s")
    print(json.dumps({
        "metric": "survey-chunk solve + map projection (76,800 px, 2.53M "
                  "cells)",
        "value": n * cells / t_map,
        "unit": "dense-equivalent map pair-evals/s (S_mp x)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_map * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded GRF skewers, lognormal noise)",
        "config": {"workload": f"{n} pixels (240 skewers x 320), box "
                               f"{w.box}, map {shape}", "tol": w.tol},
        "time_to_tol_s": {k: v["s"] for k, v in out.items()},
        "iterations": {k: v["iters"] for k, v in out.items()},
        "matvec_ms": {"f64": ms64, "f32": ms32},
        "pairs_per_matvec": {"f64": st.f64_pairs_per_matvec,
                             "f32": st.sub_pairs_per_matvec},
        "map_s": t_map, "map_norm": float(np.linalg.norm(m)),
    }), flush=True)


def bench_kernels(args):
    """Pair kernels alone on the 300,000-pixel problem, for hardware-counter
    runs: first a 1 GiB device-to-device copy (twice; the runtime's copy
    kernel, whose FETCH_SIZE / WRITE_SIZE calibrate the byte counters), then
    `--steps` launches each of the VALU FP32, the MFMA FP32 and the FP64
    pair kernel at its three tile shapes (4 waves per block, longest list
    first; the 16-row shape, the default, last), then once each the
    per-geometry kernels of the headline step: the block-Jacobi build, the
    neighbour search and the learned prediction."""
    import ctypes
    import dachshund_amd as D
    from dachshund_amd import synth
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    nbytes = ctypes.c_size_t(1 << 30)
    assert hip.hipMalloc(ctypes.byref(a), nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(b), nbytes) == 0
    assert hip.hipMemset(a, 1, nbytes) == 0
    for _ in range(2):
        assert hip.hipMemcpy(b, a, nbytes, 3) == 0      # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    w = synth.workload("c5")
    pix = synth.make_pixels(w)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    wf = D.WFVectorW(D.pixel_to_wpixel(pix), sp)
    wf.set_launch(4, True)
    st = wf.stats()
    out = {}
    for name, kind in (("k_matvec32", D.DS_F32_VALU),
                       ("k_matvec32m", D.DS_F32_MFMA)):
        wf.set_f32_kernel(kind)
        out[name] = wf.time_matvec(D.DS_PREC_F32, reps=args.steps)
    for rows in (64, 32, 16):   # FP64 tile shapes; 16 last (the default)
        wf.set_tile(rows)
        out[f"k_matvec64_r{rows}"] = wf.time_matvec(D.DS_PREC_F64,
                                                   reps=args.steps)
    out["k_matvec64"] = out["k_matvec64_r16"]
    for k, ms in out.items():
        log(f"{k}: {ms * 1e3:.1f} us per launch")
    # the headline step's per-geometry kernels, once each
    from dachshund_amd.warmstart import WarmStart
    ctx = D.default_context()
    ctx.synchronize()
    t = time.perf_counter()
    wf.set_precond(D.DS_PRECOND_BLOCK)
    ctx.synchronize()
    t_bj = time.perf_counter() - t
    ws = WarmStart()
    ws.load(args.warmstart)
    t = time.perf_counter()
    ws.predict_into(wf)
    ctx.synchronize()
    t_pred = time.perf_counter() - t
    log(f"block-Jacobi build {t_bj * 1e3:.2f} ms, neighbour search + "
        f"prediction {t_pred * 1e3:.2f} ms")
    print(json.dumps({
        "metric": "pair kernel launch time (300,000 pixels)",
        "value": out["k_matvec32m"] * 1e3, "unit": "us per launch",
        "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": out["k_matvec32m"], "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 / f64",
        "data": "synthetic (seeded GRF skewers)",
        "config": {"workload": "300000 pixels, W form, pair kernels alone"},
        "us_per_launch": {k: v * 1e3 for k, v in out.items()},
        "pairs_per_launch": {"k_matvec32": st.pairs_per_matvec,
                             "k_matvec32m": st.sub_pairs_per_matvec,
                             "k_matvec64": st.f64_pairs_per_matvec,
                             "k_matvec64_r64": st.pairs_per_matvec},
        "setup_s": {"block_jacobi_build": t_bj,
                    "neighbour_search_and_prediction": t_pred},
    }), flush=True)


def bench_batch(args):
    """512 subvolumes with log-uniform pixel counts in [2000, 40000] and
    per-subvolume correlation lengths: one batched mixed solve (every
    kernel launched once for all subvolumes, hipGraph-captured iterations,
    autotuned launch shape) against solving the subvolumes one by one."""
    import dachshund_amd as D
    from dachshund_amd import synth
    subs = synth.batch_workloads(seed=4, count=args.count)
    t0 = time.perf_counter()
    pix = []
    for k, w in enumerate(subs):
        pix.append(D.pixel_to_wpixel(synth.make_pixels(w)))
        if k % 64 == 0:
            log(f"subvolume {k}: n={len(pix[-1])}, "
                f"{time.perf_counter() - t0:.1f} s")
    sps = [D.SignalCovarParams(w.var_s, w.l_perp, w.l_para) for w in subs]
    batch = D.BatchWF(pix, sps)
    kind = (D.DS_PRECOND_BLOCK if args.precond == "block"
            else D.DS_PRECOND_JACOBI)
    batch.set_precond(kind)
    st = batch.stats()
    log(f"batch: {batch.n} pixels, setup {st.setup_seconds:.2f} s, "
        f"{st.pairs_per_matvec:.3e} pairs per batched matvec")
    waves, order, ms_tuned = batch.autotune()
    log(f"autotuner: waves {waves}, longest-first {order}, {ms_tuned:.3f} ms")
    prm = D.PCGParams(batch.n, 2000, 50, args.tol)
    for _ in range(args.warmup):
        batch.fill_x(0.0)
        batch.solve(prm, mixed=True, use_graph=True)
    ts = []
    for _ in range(args.steps):
        batch.fill_x(0.0)
        t = time.perf_counter()
        res = batch.solve(prm, mixed=True, use_graph=True)
        ts.append(time.perf_counter() - t)
        assert all(r.success for r in res)
    t_batch = float(np.mean(ts))
    log(f"batched: {t_batch * 1e3:.1f} ms, iterations <= "
        f"{max(r.num_iters for r in res)}")
    # the same subvolumes one after another, each with the library's
    # default launch shape (4 waves per block, longest list first; the
    # batch's tuned shape suits a 6.7M-pixel launch, not a 2k-pixel one)
    t_seq = 0.0
    for p, sp in zip(pix, sps):
        wf = D.WFVectorW(p, sp)
        wf.set_precond(kind)
        wf.fill_x(0.0)
        t = time.perf_counter()
        r = wf.solve_resident(D.PCGParams(wf.n, 2000, 50, args.tol),
                              mixed=True)
        t_seq += time.perf_counter() - t
        wf.close()
    log(f"one by one: {t_seq * 1e3:.1f} ms")
    p64, p32 = batch.sub_pairs()
    evals = float(sum(int(a) * r.n_matvec64 + int(b) * r.n_matvec32
                      for a, b, r in zip(p64, p32, res)))
    print(json.dumps({
        "metric": "batched PCG time-to-rel-residual 1e-6 (512 subvolumes, "
                  "one launch per kernel, hipGraph)",
        "value": evals / t_batch,
        "unit": "evaluated pair-evals/s (each subvolume's own pair count "
                "x its own FP32 and FP64 matvecs)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pairs + f64 vectors",
        "data": "synthetic (seeded GRF subvolumes)",
        "config": {"workload": f"{len(subs)} subvolumes, {batch.n} pixels",
                   "tol": args.tol, "precond": args.precond},
        "time_to_tol_s": {"batched": t_batch, "one_by_one": t_seq},
        "speedup_vs_one_by_one": t_seq / t_batch,
        "autotuned": {"waves_per_block": waves, "longest_first": order,
                      "matvec_ms": ms_tuned},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5",
                    choices=("c5", "c1", "c2", "c3", "c4", "kernels"),
                    help="c5: the 300k-pixel headline (default; c1 runs the "
                         "same comparison on 8,192 pixels); c2: 76,800 "
                         "pixels and the 2.53M-cell map; c3: 100k pixels x "
                         "32 realizations; c4: 512 subvolumes; kernels: the "
                         "pair kernels alone (for counter runs)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--drop", type=float, default=1e-3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--warmstart", default=os.path.join(
        ROOT, "dachshund_amd", "warmstart_c5.npz"))
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--precond", choices=("jacobi", "block"), default="block",
                    help="preconditioner of the mixed solves of the c3 and "
                         "c4 workloads (plain FP64 keeps Jacobi)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the "
              "solve is not sharded across GPUs in this build.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "c3":
        return bench_block(args)
    if args.workload == "c4":
        return bench_batch(args)
    if args.workload == "c2":
        return bench_chunk(args)
    if args.workload == "kernels":
        return bench_kernels(args)

    import dachshund_amd as D
    from dachshund_amd import synth

    w = synth.workload(args.workload)
    t0 = time.perf_counter()
    pix = synth.make_pixels(w)
    n = len(pix)
    log(f"workload {w.name}: n={n}, generated in "
        f"{time.perf_counter() - t0:.1f} s")
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    pw = D.pixel_to_wpixel(pix)
    wf = D.WFVectorW(pw, sp)
    st = wf.stats()
    log(f"setup {st.setup_seconds:.2f} s; evaluated pairs per matvec: FP64 "
        f"{st.pairs_per_matvec:.3e} (64 x 16 culling), FP32 "
        f"{st.sub_pairs_per_matvec:.3e} (16 x 16); "
        f"({st.pairs_per_matvec / n / n:.2e} of "
        f"n^2)")
    prm = D.PCGParams(n, w.max_iter, w.step_r, args.tol)

    ws = None
    if args.warmstart != "none" and os.path.exists(args.warmstart):
        from dachshund_amd.warmstart import WarmStart
        ws = WarmStart()
        ws.load(args.warmstart)
        log(f"learned warm start: {args.warmstart}")
    else:
        log("no trained warm start: headline solver starts from x0 = 0")

    def solve(mixed, learned=False, block=False):
        wf.set_precond(D.DS_PRECOND_BLOCK if block else D.DS_PRECOND_JACOBI)
        if learned:
            ws.predict_into(wf)
        else:
            wf.fill_x(0.0)
        if mixed:
            return wf.solve_resident(prm, mixed=True, refine_drop=args.drop)
        return wf.solve_resident(prm)

    def timed(mixed, k, learned=False, block=False):
        D.default_context()
        ts, res = [], None
        for _ in range(k):
            t = time.perf_counter()
            res = solve(mixed, learned, block)
            ts.append(time.perf_counter() - t)
            if not res.success:
                raise RuntimeError(f"solve (mixed={mixed}) did not converge")
        return ts, res

    def report(name, ts, r):
        log(f"{name}: {np.mean(ts) * 1e3:.1f} ms, {r.num_iters} iters, "
            f"{r.n_matvec32} FP32 + {r.n_matvec64} FP64 matvecs")

    learned = ws is not None
    ctx = D.default_context()
    creates = []   # problem creation (ordering + culling on the device)

    def fresh_step():
        """The headline step on a freshly created problem: block-Jacobi
        inverses, neighbour search, prediction and the mixed solve -- all
        the per-geometry work only the headline solver needs -- inside the
        timed region.  Ordering and culling (shared with plain FP64, which
        needs the same tile lists) happen at problem creation, outside."""
        ctx.synchronize()
        tc = time.perf_counter()
        wf2 = D.WFVectorW(pw, sp)
        ctx.synchronize()
        creates.append(time.perf_counter() - tc)
        t = time.perf_counter()
        wf2.set_precond(D.DS_PRECOND_BLOCK)
        if learned:
            ws.predict_into(wf2)
        else:
            wf2.fill_x(0.0)
        r = wf2.solve_resident(prm, mixed=True, refine_drop=args.drop)
        dt = time.perf_counter() - t
        if not r.success:
            raise RuntimeError("headline solve did not converge")
        x = wf2.get_x()
        wf2.close()
        return dt, r, x

    for _ in range(args.warmup):
        solve(False)
        solve(True)
        solve(False, block=True)
        solve(True, block=True)
        solve(True, learned, True)
        if learned:
            solve(True, True)
        fresh_step()
    t64, r64 = timed(False, args.steps)
    x64 = wf.get_x()
    report("plain FP64 PCG (Jacobi, the reference's algorithm)", t64, r64)
    tm0, rm0 = timed(True, args.steps)
    report("mixed, x0 = 0, Jacobi", tm0, rm0)
    tml, rml = (timed(True, args.steps, learned=True) if learned
                else (tm0, rm0))
    report("mixed + learned x0, Jacobi", tml, rml)
    tb64, rb64 = timed(False, args.steps, block=True)
    report("plain FP64 PCG, block-Jacobi", tb64, rb64)
    tbm0, rbm0 = timed(True, args.steps, block=True)
    report("mixed, x0 = 0, block-Jacobi", tbm0, rbm0)
    tmc, rmc = timed(True, args.steps, learned=learned, block=True)
    report("mixed + learned x0 + block-Jacobi, per-geometry data cached",
           tmc, rmc)
    steps = [fresh_step() for _ in range(args.steps)]
    tmx = [s_[0] for s_ in steps]
    rmx, xmx = steps[-1][1], steps[-1][2]
    report("headline: mixed + learned x0 + block-Jacobi, fresh geometry "
           "(inverses, neighbour search, prediction, solve)", tmx, rmx)

    # breakdown of the per-geometry work (not timed as the step)
    t_pred = t_knn = t_bj = None
    wf3 = D.WFVectorW(pw, sp)
    ctx.synchronize()
    t = time.perf_counter()
    wf3.set_precond(D.DS_PRECOND_BLOCK)
    ctx.synchronize()
    t_bj = time.perf_counter() - t
    if ws:
        t = time.perf_counter()
        ws.predict_into(wf3)
        ctx.synchronize()
        t_first = time.perf_counter() - t
        tp = []
        for _ in range(args.steps):
            ctx.synchronize()
            t = time.perf_counter()
            ws.predict_into(wf3)
            ctx.synchronize()
            tp.append(time.perf_counter() - t)
        t_pred = float(np.mean(tp))
        t_knn = t_first - t_pred
        log(f"learned prediction (features + network) {t_pred * 1e3:.2f} ms;"
            f" neighbour search {t_knn * 1e3:.2f} ms")
    wf3.close()
    log(f"block-Jacobi inverses {t_bj * 1e3:.2f} ms")

    # accuracy of the mixed map against a tight FP64 solve
    tight = D.PCGParams(n, 4 * w.max_iter, w.step_r, 1e-10)
    wf.set_precond(D.DS_PRECOND_JACOBI)
    wf.fill_x(0.0)
    rt = wf.solve_resident(tight)
    box, shape = w.box, w.map_shape
    m_ref = wf.smp_grid(*box, *shape)
    wf.set_x(xmx)
    m_mx = wf.smp_grid(*box, *shape)
    wf.set_x(x64)
    m_64 = wf.smp_grid(*box, *shape)
    rel_mx = float(np.linalg.norm(m_mx - m_ref) / np.linalg.norm(m_ref))
    rel_hp = float(np.linalg.norm(m_mx - m_64) / np.linalg.norm(m_64))
    rel_64 = float(np.linalg.norm(m_64 - m_ref) / np.linalg.norm(m_ref))
    log(f"map rel-L2 vs FP64 tol 1e-10 solve ({rt.num_iters} iters): mixed "
        f"{rel_mx:.2e}, plain FP64 {rel_64:.2e}")

    # dominant kernel: the FP32 operator (symmetric pair kernel + gather)
    ms32 = kernel_ms(wf, D.DS_PREC_F32)
    ms64 = kernel_ms(wf, D.DS_PREC_F64)
    pairs = st.f64_pairs_per_matvec      # FP64 kernel at its tile shape
    prod32 = st.sub_pairs_per_matvec     # FP32 pair products (16 x 16)
    eval32 = st.sub_pairs_evaluated      # FP32 evaluations (each once)
    rate32 = eval32 / (ms32 * 1e-3)
    rate64 = pairs / (ms64 * 1e-3)
    inside = pairs_inside_cutoff(pix, w.l_perp, w.l_para)
    log(f"pair kernels: FP32 {ms32:.3f} ms ({rate32:.3e} evaluations/s, "
        f"{rate32 / PAIR_PEAK:.2f} of {PAIR_PEAK:.3e}), FP64 {ms64:.3f} ms "
        f"({rate64:.3e}/s, {rate64 / PAIR_PEAK:.2f}); {inside:.3e} ordered "
        f"pairs inside the cutoff")

    t_mx = float(np.mean(tmx))
    t_64 = float(np.mean(t64))
    t_create = float(np.mean(creates[-args.steps:]))
    log(f"problem creation from pixels (device ordering and culling) "
        f"{t_create * 1e3:.2f} ms; from pixels: headline "
        f"{(t_create + t_mx) * 1e3:.2f} ms, plain FP64 "
        f"{(t_create + t_64) * 1e3:.2f} ms")
    evals = eval32 * rmx.n_matvec32 + pairs * rmx.n_matvec64
    traffic = pmc_traffic("k_matvec32m", "k_sym_epi")
    out = {
        "metric": "PCG time-to-rel-residual 1e-6 (mixed+learned vs plain "
                  "FP64, 1 GPU); pair-evals/s",
        "value": evals / t_mx,
        "unit": "pair-evals/s (covariance evaluations performed per second "
                "over the headline step, setup included)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mx * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pairs + f64 vectors",
        "data": "synthetic (seeded Gaussian random field skewers)",
        "config": {
            "workload": (f"{n} pixels ({w.n_skewers} skewers x "
                         f"{w.n_per_skewer}), box {w.box}, GRF var_s "
                         f"{w.var_s} l_perp {w.l_perp} l_para {w.l_para}, "
                         f"noise U[0.1,0.3], seed {w.seed} (held out), W "
                         f"form, headline x0 = "
                         f"{'learned' if ws else 'zero'}, block-Jacobi; "
                         f"step = one solve on a fresh geometry incl. its "
                         f"block-Jacobi build, neighbour search and "
                         f"prediction"),
            "tol": args.tol, "step_r": w.step_r,
            "refine_drop": args.drop,
        },
        "time_to_tol_s": {"headline": t_mx,
                          "headline_setup_cached": float(np.mean(tmc)),
                          "mixed_learned_jacobi": float(np.mean(tml)),
                          "mixed_x0_zero_jacobi": float(np.mean(tm0)),
                          "plain_fp64_block_jacobi": float(np.mean(tb64)),
                          "mixed_x0_zero_block_jacobi": float(np.mean(tbm0)),
                          "plain_fp64": t_64},
        "speedup_vs_plain_fp64": t_64 / t_mx,
        "speedup_vs_plain_fp64_setup_cached": t_64 / float(np.mean(tmc)),
        "iterations": {"headline": rmx.num_iters, "headline_fp64_residuals":
                       rmx.n_true, "mixed_learned_jacobi": rml.num_iters,
                       "mixed_x0_zero_jacobi": rm0.num_iters,
                       "plain_fp64_block_jacobi": rb64.num_iters,
                       "mixed_x0_zero_block_jacobi": rbm0.num_iters,
                       "plain_fp64": r64.num_iters},
        "learned_warm_start": bool(ws),
        "per_geometry_s": {"block_jacobi_build": t_bj,
                           "neighbour_search": t_knn,
                           "prediction": t_pred},
        "setup_s": st.setup_seconds,
        "problem_create_s": t_create,
        "time_from_pixels_s": {"headline": t_create + t_mx,
                               "plain_fp64": t_create + t_64},
        "speedup_from_pixels": (t_create + t_64) / (t_create + t_mx),
        "map_rel_l2_vs_fp64_tight": {"headline": rel_mx,
                                     "plain_fp64": rel_64},
        "map_rel_l2_headline_vs_plain_fp64": rel_hp,
        "pairs_per_matvec": {"fp64_evaluated": pairs,
                             "fp64_64x16_tiles": st.pairs_per_matvec,
                             "fp32_evaluated": eval32,
                             "fp32_pair_products": prod32,
                             "inside_cutoff": inside,
                             "dense": n * n},
        "useful_pair_fraction": {"fp32_products": inside / prod32,
                                 "fp64": inside / pairs},
        "pair_products_per_s": (prod32 * rmx.n_matvec32 +
                                pairs * rmx.n_matvec64) / t_mx,
        "dense_equiv_pair_evals_per_s": n * n * (rmx.n_matvec32 +
                                                 rmx.n_matvec64) / t_mx,
        "roofline": {
            "bound": "valu", "achieved": rate32, "peak": PAIR_PEAK,
            "unit": "pair-evals/s", "frac": rate32 / PAIR_PEAK,
            "peak_basis": "min(R_VALU/12, R_trans), 256 CU x 64 lanes x "
                          "2.4 GHz / 12 VALU ops per pair, non-packed",
            "traffic": traffic,
            "traffic_source": "profiles/pair_kernel_traffic_c5_300k.txt "
                              "(rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                              "launch, copy-calibrated)",
            "algorithmic_bytes": ALGO_BYTES_PER_PIXEL["f32"] * n,
            "kernel": "k_matvec32m + k_sym_epi (symmetric FP32 operator: "
                      "each 16 x 16 tile pair evaluated once and used for "
                      "both products, then the per-tile gather)",
            "ms_per_launch": ms32,
            "pairs_evaluated_per_launch": eval32,
            "pair_products_per_launch": prod32,
        },
        "fp64_kernel": {"kernel": f"k_matvec64 ({wf.launch()[2]} x 16 "
                                  f"tiles)", "achieved": rate64,
                        "peak": PAIR_PEAK, "unit": "pair-evals/s",
                        "frac": rate64 / PAIR_PEAK, "ms_per_launch": ms64,
                        "pairs_evaluated_per_launch": pairs,
                        "traffic": pmc_traffic("k_matvec64"),
                        "algorithmic_bytes": ALGO_BYTES_PER_PIXEL["f64"] * n},
    }
    if not args.no_cpu:
        cb = cpu_baseline(pw, w.var_s, w.l_perp, w.l_para, n, r64.n_matvec64)
        if cb:
            out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
