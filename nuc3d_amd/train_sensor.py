"""Training of the hybrid reconstruction's shock sensor, on the GPU.

    python -m nuc3d_amd.train_sensor [--out nuc3d_amd/sensor_params.npz]

1. Solver snapshots with the plain schemes (every run on the GPU):
     Taylor-Green vortex 64^3, Navier-Stokes, WENO5-JS, 200 steps;
     Shu-Osher 512x64x64, Euler, WENO5-Z, inflow/outflow, 400 steps;
     supersonic flat plate over a bump, 8 blocks of uneven sizes (0.72 M
     cells, curvilinear tanh-stretched mesh, isothermal wall, inflow/outflow),
     Navier-Stokes, WENO5-JS, 300 steps;
     decaying isotropic turbulence n^3 (256 by default), Navier-Stokes,
     WENO5-JS, from the seeded spectrum (seed 1: the benchmark uses seed 0).
   Every few steps the faces of all three directions are sampled on the GPU
   (nuc3d_sensor_sample): 30 features and the teacher's label per face, at a
   stride that gives about 1e7 training stencils, and at another offset a
   held-out set.
2. Minibatch Adam on the GPU (nuc3d_sensor_train), two learning rates.
3. The logit threshold: the smallest that keeps the recall of the teacher's
   WENO5 faces on the held-out set at or above `recall` (faces the network is
   unsure about go to WENO5-JS).
The parameters, threshold and training record are written to an .npz file
(nuc3d_amd/sensor_params.npz is the one the benchmark loads).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import cases, solver

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "sensor_params.npz")
TAU = 1e-3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def init_params(seed):
    """He-initialised FP32 parameters with zero biases (padding inputs zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(solver.SENSOR_NPARAM, dtype=np.float32)
    W1 = rng.normal(0.0, np.sqrt(2.0 / 30), (32, 32)).astype(np.float32)
    W1[:, 30:] = 0.0
    p[0:1024] = W1.ravel()
    p[1056:2080] = rng.normal(0.0, np.sqrt(2.0 / 32), 1024).astype(np.float32)
    p[2112:2144] = rng.normal(0.0, np.sqrt(1.0 / 32), 32).astype(np.float32)
    return p


def snapshot_runs(n_turb):
    """(name, params, blocks, snapshot steps, sample stride, setup) of the runs;
    setup(case, blocks) sets metrics and state where the blocks do not carry
    them (the flat plate: metrics from the mesh nodes on the GPU)."""
    runs = [("tgv64", solver.Params(), cases.taylor_green(64), (0, 40, 80, 120, 160, 200), 4, None)]
    runs.append(("shu_osher", solver.Params(model=solver.EULER3D, mach=1.0, scheme=(solver.WENO5Z,) * 3),
                 cases.shu_osher(512, 64, 64, seed=1), (0, 100, 200, 300, 400), 8, None))
    runs.append(("flatplate", solver.Params(mach=2.9, reynolds=2.0e4, t_wall=2.5),
                 cases.flat_plate([44, 36, 40, 36], [32, 40], 64), (0, 100, 200, 300), 2, cases.setup_flat_plate))
    blocks, phys = cases.isotropic_turbulence(n=n_turb, seed=1)
    stride = max(1, int(round(40 * (n_turb / 256.0) ** 3)))
    runs.append(("turbulence%d" % n_turb, solver.Params(mach=phys.mach, reynolds=phys.reynolds), blocks,
                 (0, 100, 200, 300), stride, None))
    return runs


def collect(train, held, n_turb, tau=TAU):
    record = []
    for name, params, blocks, snaps, stride, setup in snapshot_runs(n_turb):
        case = solver.Case(params, blocks)
        if setup:
            setup(case, blocks)
        step = 0
        for s in snaps:
            case.steps_graph(s - step, 0.5)
            step = s
            a = train.sample(case, tau, stride=stride, offset=0)
            b = held.sample(case, tau, stride=stride * 4, offset=stride // 2 if stride > 1 else 0)
            record.append({"run": name, "step": s, "train": a, "held": b})
            log("train_sensor: %s step %d: %d + %d held-out stencils" % (name, s, a, b))
        case.close()
    return record


def choose_threshold(z, y, recall):
    """Smallest logit threshold whose recall of y == 1 is at least `recall`."""
    zp = np.sort(z[y > 0.5])
    if zp.size == 0:
        return 0.0
    k = int(np.floor((1.0 - recall) * zp.size))
    return float(zp[k])


def evaluate(z, y, thr):
    v = y >= 0
    z, y = z[v], y[v] > 0.5
    pred = z >= thr
    return {
        "samples": int(v.sum()),
        "positive_fraction": float(y.mean()),
        "accuracy": float((pred == y).mean()),
        "recall": float(pred[y].mean()) if y.any() else 1.0,
        "central_fraction": float((~pred).mean()),
        "missed_fraction": float((~pred & y).mean()),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--n-turb", type=int, default=256)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--recall", type=float, default=0.995)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)

    t0 = time.time()
    train = solver.Sensor(18_000_000)
    held = solver.Sensor(5_000_000)
    record = collect(train, held, args.n_turb)
    n, npos = train.count()
    log("train_sensor: %d training stencils, %d labelled WENO5 (%.1f%%)" % (n, npos, 100.0 * npos / max(n, 1)))
    train.set_params(init_params(args.seed))
    held.set_params(train.params())
    _, yh = held.download()
    z0 = held.infer()
    t1 = time.time()
    losses = [train.train(int(args.steps * 2 / 3), batch=args.batch, lr=2e-3, seed=args.seed + 1)]
    losses.append(train.train(args.steps - int(args.steps * 2 / 3), batch=args.batch, lr=4e-4, seed=args.seed + 2))
    losses = np.concatenate(losses)
    t2 = time.time()
    params = train.params()
    held.set_params(params)
    z = held.infer()
    thr = choose_threshold(z, yh, args.recall)
    ev = evaluate(z, yh, thr)
    ev0 = evaluate(z0, yh, choose_threshold(z0, yh, args.recall))
    log("train_sensor: %d Adam steps in %.2f s, loss %.4f -> %.4f" % (args.steps, t2 - t1, losses[:50].mean(),
                                                                     losses[-50:].mean()))
    log("train_sensor: held-out %s" % json.dumps(ev))
    log("train_sensor: untrained network at the same recall: %s" % json.dumps(ev0))
    np.savez(args.out, params=params, threshold=np.float32(thr), tau=np.float64(TAU),
             losses=losses.astype(np.float32), record=json.dumps(record), heldout=json.dumps(ev),
             heldout_untrained=json.dumps(ev0), recall=np.float64(args.recall))
    log("train_sensor: wrote %s (%.1f s total)" % (args.out, time.time() - t0))
    train.close()
    held.close()
    return params, thr, ev


def load(path=DEFAULT_OUT):
    """(params, threshold) of a trained sensor file."""
    d = np.load(path, allow_pickle=False)
    return d["params"].astype(np.float32), float(d["threshold"])


if __name__ == "__main__":
    main()
