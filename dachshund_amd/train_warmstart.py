"""Train the learned warm start on solutions the solver itself produces.

    python -m dachshund_amd.train_warmstart [--workload c5] [--train 200]
        [--test 20] [--epochs 30] [--out dachshund_amd/warmstart_c5.npz]

For each training seed a realization of the workload is drawn, solved by
plain FP64 PCG on the GPU to the benchmark tolerance, and its features and
solution are kept in HBM as training samples; the network is then trained
by the library's own forward/backward/Adam kernels.  Held-out seeds compare
the iterations of the solve from x0 = 0 and from the predicted x0, with the
same stopping rule.  The parameters and a JSON summary are written to --out.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import dachshund_amd as D  # noqa: E402
from dachshund_amd import synth  # noqa: E402
from dachshund_amd.warmstart import WarmStart  # noqa: E402

TRAIN_SEED0 = 1000
TEST_SEED0 = 2000


def log(msg):
    print(f"[train] {msg}", flush=True)


def problem(w, seed):
    pix = synth.make_pixels(w, seed=seed)
    sp = D.SignalCovarParams(w.var_s, w.l_perp, w.l_para)
    return D.WFVectorW(D.pixel_to_wpixel(pix), sp)


def evaluate(ws, w, seeds, tol):
    rows = []
    for seed in seeds:
        wf = problem(w, seed)
        prm = D.PCGParams(wf.n, w.max_iter, w.step_r, tol)
        out = {"seed": seed}
        for mode in ("fp64", "mixed"):
            mixed = mode == "mixed"
            wf.fill_x(0.0)
            r0 = wf.solve_resident(prm, mixed=mixed)
            t = time.perf_counter()
            ws.predict_into(wf)
            r1 = wf.solve_resident(prm, mixed=mixed)
            t1 = time.perf_counter() - t
            log(f"held-out seed {seed} {mode}: {r0.num_iters} -> "
                f"{r1.num_iters} iterations")
            out[mode] = {"iters_x0_zero": r0.num_iters,
                         "iters_learned": r1.num_iters,
                         "ok": bool(r0.success and r1.success),
                         "s_x0_zero": r0.seconds,
                         "s_learned_incl_predict": t1}
        rows.append(out)
        wf.close()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--train", type=int, default=200)
    ap.add_argument("--test", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-decay", type=float, default=0.9)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    w = synth.workload(args.workload)
    out = args.out or os.path.join(ROOT, "dachshund_amd",
                                   f"warmstart_{args.workload}.npz")
    ws = WarmStart(seed=args.seed)
    t0 = time.perf_counter()
    iters = []
    for k in range(args.train):
        seed = TRAIN_SEED0 + k
        wf = problem(w, seed)
        wf.fill_x(0.0)
        r = wf.solve_resident(D.PCGParams(wf.n, w.max_iter, w.step_r,
                                          args.tol))
        if not r.success:
            raise RuntimeError(f"training solve {seed} did not converge")
        ws.add_sample(wf, seed=seed)
        iters.append(r.num_iters)
        wf.close()
        if k % 20 == 0:
            log(f"sample {k}: seed {seed}, {r.num_iters} iters, "
                f"{time.perf_counter() - t0:.1f} s")
    log(f"{ws.num_samples()} training pixels from {args.train} solves in "
        f"{time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    loss = []
    for e in range(args.epochs):
        lr = args.lr * args.lr_decay ** e
        loss.append(float(ws.train(1, batch=args.batch, lr=lr, lr_decay=1.0,
                                   seed=args.seed + e)[0]))
        log(f"epoch {e}: lr {lr:.2e}, loss {loss[-1]:.4e}, "
            f"{time.perf_counter() - t1:.1f} s")
    log(f"trained {args.epochs} epochs in {time.perf_counter() - t1:.1f} s; "
        f"loss {loss[0]:.4e} -> {loss[-1]:.4e}")
    ws.clear_samples()
    rows = evaluate(ws, w, range(TEST_SEED0, TEST_SEED0 + args.test),
                    args.tol)
    summary = {
        "workload": args.workload, "train_seeds": [TRAIN_SEED0,
                                                   TRAIN_SEED0 + args.train],
        "test_seeds": [TEST_SEED0, TEST_SEED0 + args.test],
        "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
        "lr_decay": args.lr_decay, "tol": args.tol,
        "loss": [float(v) for v in loss],
        "train_iters_mean": float(np.mean(iters)),
        "test": rows,
    }
    for mode in ("fp64", "mixed"):
        z = np.mean([r[mode]["iters_x0_zero"] for r in rows])
        lrn = np.mean([r[mode]["iters_learned"] for r in rows])
        log(f"held-out {mode}: iterations {z:.1f} (x0 = 0) -> {lrn:.1f} "
            f"(learned)")
    np.savez(out, params=ws.params,
             summary=np.array(json.dumps(summary)))
    log(f"wrote {out}")
    print(json.dumps({k: summary[k] for k in ("loss", "train_iters_mean")}))
    return summary


if __name__ == "__main__":
    main()
