#!/usr/bin/env python3
"""Throughput of the per-block RHS + TVD-RK3 path on one MI355X: plain WENO5-JS
against the learned hybrid (central-8 / WENO5-JS per face, chosen by the
BF16-MFMA sensor) on the same workload, with the hybrid's L2 density deviation
from plain WENO5-JS as the accuracy measure.

Metric: cell-updates/s per RK3 step in FP64 (one cell-update = one cell
advanced one full TVD-RK3 step: three RHS evaluations and three stage
updates).  Workload: decaying compressible isotropic turbulence on a 256^3
periodic block (Navier-Stokes, local Lax-Friedrichs splitting, 6th-order
central viscous terms, CFL 0.5), synthetic initial field of the prescribed
spectrum (seed 0), developed for --develop steps with plain WENO5-JS; both
methods then start from that state, resident in HBM, and run the same steps.
Timed steps are launched as a captured HIP graph of two RK3 steps.

`value` is the throughput of the faster method among those whose L2 density
deviation from plain WENO5-JS is within 1e-3 of the density rms (plain
WENO5-JS always is); both methods' figures are in the line.  The line also
carries the live roofline figure of the kernel class that takes the most time
per step (HIP events recorded on the stream the kernels are launched on) and a
CPU baseline: the reference compiled from its own sources (oracle/_ref) timed
on the host cores on a bounded sample of the same workload, or the oracle
where that build is absent.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s per RK3 step FP64: plain WENO5 vs learned hybrid at equal L2 error"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell-update of one fused TVD-RK3 step: stage 0
# reads Q (40) and J + nine metric terms (80) and writes Q1 (40); stages 1
# and 2 also read Q^n (40): 160 + 200 + 200 = 560.  Blocks on the uniform-mesh
# path (J and the metric terms one value each, taken from the block's
# descriptor) skip the 80 B of metrics per stage: 320.
ALG_BYTES_PER_STEP = 560.0
ALG_BYTES_PER_STEP_UNIFORM = 320.0


def kernel_alg_bytes(cls, ns, uniform=False):
    """Compulsory HBM bytes per cell of one launch of a kernel class, averaged
    over the three RK stages (every input field read once, every output
    written once; ghost layers and halo re-reads not counted); with
    `uniform` the J and metric bytes are left out (the uniform-mesh path
    does not load them):
      x sweep:      Q 40 + J 8 + xi metrics 24 + RHS write 40                 = 112
      y sweep:      Q 40 + J 8 + eta metrics 24 + RHS read and write 80        = 152
      z sweep:      Q 40 + J 8 + zeta metrics 24 + RHS read 40, Q^n 40 at
                    stages 1 and 2, Q_out write 40, xi/eta metrics 48 at
                    stage 2 for the eigenvalues                               = 194.7
                    (Navier-Stokes: + the twelve viscous fluxes fv 96, whose
                     divergence the z sweep takes                            = 290.7)
      viscous flux: Q 40 + J 8 + nine metrics 72 read (u v w T computed in
                    the pass), fv write 96                                    = 216
      viscous div:  fv 96 + RHS 40 + Q 40 + J 8 read, Q_out 40 write,
                    Q^n 40 at stages 1 and 2, nine metrics 72 at stage 2      = 274.7
                    (single RHS evaluations, and CRWENO5 z lines)
    uniform: x sweep 80, y sweep 120, z sweep 146.7 (Navier-Stokes 242.7),
    viscous flux 136, viscous div 242.7.
    """
    m = 0.0 if uniform else 1.0
    if cls == "x_sweep":
        return 80.0 + m * 32.0
    if cls == "y_sweep":
        return 120.0 + m * 32.0
    if cls == "z_sweep":
        return 80.0 + m * 32.0 + (0 + 40 + 40) / 3.0 + 40.0 + m * 48.0 / 3.0 + (96.0 if ns else 0.0)
    if cls == "viscous_flux":
        return 136.0 + m * 80.0
    return 96.0 + 40.0 + 40.0 + m * 8.0 + 40.0 + (0 + 40 + 40) / 3.0 + m * 72.0 / 3.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """(physical cores of the host from lscpu, CPUs this process may run on)."""
    phys = None
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        cores = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
        phys = len(cores) or None
    except Exception:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return phys, avail


def cpu_baseline(n, warm=10, timed=2, timeout=600):
    """The reference (oracle/_ref: its own sources built twice, the makefile's
    -O0 and -O2 -march=x86-64-v3) on a bounded sample of the workload: a
    periodic slab of the same developed-free turbulence field split along x
    into one block per MPI rank, one rank per core this process may use (at
    most the host's physical cores).  Timed as block::solve times a run
    (block.cpp:576-584, MPI_Wtime around each RK3 step): `warm` warm-up
    steps, then the mean wall time of the next `timed` steps.  `value` is the
    -O2 build's cell-updates/s; the -O0 build's is beside it."""
    import numpy as np
    from nuc3d_amd import cases
    from oracle import oracle as O

    phys_cores, avail = host_cores()
    ranks = min(avail, phys_cores or avail)
    blocks, phys = cases.isotropic_turbulence(n=n, seed=0)
    spec = blocks[0]
    p = O.Params(model=O.NS, mach=phys.mach, reynolds=phys.reynolds)

    def slab(nx, ny, nz):
        sub = cases.BlockSpec((nx, ny, nz), cases.periodic_faces(), spec.jac[:nz, :ny, :nx].copy(),
                              spec.met[:, :nz, :ny, :nx].copy(), spec.dxyz[:, :nz, :ny, :nx].copy(),
                              spec.q[:, :nz, :ny, :nx].copy())
        r = ranks
        while nx // r < 8:
            r //= 2
        return cases.split_block_x(sub, r), r

    if not (O.ref_available(optimised=True) and O.ref_available()):
        sub = cases.BlockSpec((64, 64, 32), cases.periodic_faces(), spec.jac[:32, :64, :64].copy(),
                              spec.met[:, :32, :64, :64].copy(), spec.dxyz[:, :32, :64, :64].copy(),
                              spec.q[:, :32, :64, :64].copy())
        oc = O.OracleCase(p, [sub])
        t0 = time.perf_counter()
        oc.steps(1, 0.5)
        t = time.perf_counter() - t0
        return {"value": 64 * 64 * 32 / t, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                "sample": "C restatement (oracle/), 1 core, 64x64x32 periodic slab of the 256^3 field, 1 RK3 step "
                          "(the reference build oracle/_ref is absent)"}
    res = {}
    # the -O0 build is about 4x slower: half the slab keeps both runs near 10-20 s
    for name, opt, dims in (("O2", True, (256, 128, 32)), ("O0", False, (256, 64, 32))):
        parts, r = slab(*dims)
        d = tempfile.mkdtemp(prefix="nuc3d_cpu_")
        O.write_ref_case(d, p, parts)
        t0 = time.perf_counter()
        out = O.run_ref(d, r, "steps", warm + timed, 0.5, optimised=opt, timeout=timeout)
        walls = np.asarray(out[0]["walls"], dtype=np.float64)
        t = float(np.mean(walls[warm:warm + timed]))
        cells = dims[0] * dims[1] * dims[2]
        res[name] = {"value": cells / t, "ms_per_step": 1e3 * t, "ranks": r, "slab": "%dx%dx%d" % dims,
                     "run_s": time.perf_counter() - t0}
        log("cpu baseline: -%s build, %d ranks, %s slab: %.4g cell-updates/s" % (name, r, res[name]["slab"],
                                                                                   res[name]["value"]))
    return {"value": res["O2"]["value"], "unit": "cell-updates/s", "cores": res["O2"]["ranks"], "kind": "reference",
            "host_physical_cores": phys_cores, "cpus_available": avail,
            "O0_makefile": res["O0"], "O2": res["O2"],
            "sample": "the reference built from its own sources (oracle/_ref): -O2 -march=x86-64-v3 (value) and the "
                      "makefile's -O0 (O0_makefile), %d MPI ranks (one per available core; the host has %s physical "
                      "cores, %d available to this job), periodic slab of the 256^3 turbulence field (%s for -O2, "
                      "%s for -O0) in x-blocks; MPI_Wtime per RK3 step as block.cpp:576-584, %d warm-up steps, "
                      "mean of steps %d..%d" % (res["O2"]["ranks"], phys_cores, avail, res["O2"]["slab"],
                                               res["O0"]["slab"], warm, warm + 1, warm + timed)}


# timing class -> kernel (rocprofv3 names: k_sweep_<variant><DIR,...>, ...)
KERNEL_OF_CLASS = {"x_sweep": "k_sweep_x", "y_sweep": "k_sweep_y", "z_sweep": "k_sweep_z",
                   "viscous_flux": "k_visc_flux", "viscous_div_update": "k_visc_div"}


def read_traffic(cls, n, variant):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy
    kernel calibration), when present for this workload and variant."""
    path = os.path.join(ROOT, "profiles", "%s_iso%d_%s_traffic.txt" % (KERNEL_OF_CLASS[cls], n, variant))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            try:
                return float(line.split()[1])
            except (IndexError, ValueError):
                return None
    return None


def read_valu(kernel, n, variant):
    """FP64 VALU figures of a kernel from the committed rocprofv3 counter
    summary (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 and SQ_INSTS_VALU per
    dispatch over the kernel-trace duration), when present for this workload
    and variant: its FP64 rate against the 78.6 TFLOP/s vector peak, the
    fraction of cycles in which a SIMD could issue a VALU instruction that it
    did, and the step's FP64 instructions per cell-update."""
    path = os.path.join(ROOT, "profiles", "valu_iso%d_%s.txt" % (n, variant))
    if not os.path.exists(path):
        return None
    out = {"source": os.path.relpath(path, ROOT)}
    for line in open(path):
        if line.startswith(kernel + " "):
            f = line[len(kernel):].split()
            try:
                out.update({"kernel": kernel, "kernel_fp64_tflops": float(f[2]), "fp64_peak_tflops": 78.6,
                            "kernel_fp64_frac": float(f[3].rstrip("%")) / 100.0,
                            "kernel_valu_issue_frac": float(f[5].rstrip("%")) / 100.0,
                            "kernel_fp64_lane_ops_per_cell": float(f[7])})
            except (IndexError, ValueError):
                return None
        elif line.startswith("plain step: FP64 instructions executed per cell-update"):
            try:
                out["step_fp64_lane_ops_per_cell_update"] = float(line.split("=")[1].split()[0])
            except (IndexError, ValueError):
                return None
    return out if "kernel" in out else None


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters: reads and writes `nbytes` each."""
    import torch
    n = nbytes // 8
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    assert torch.equal(x, y)
    print(json.dumps({"copy_bytes_each_way": n * 8, "launches": 3}), flush=True)


def time_method(case, steps, warmup, cells, cfl=0.5):
    """(cell-updates/s, ms/step, per-class kernel times) of graph-launched steps."""
    case.steps_graph(max(warmup, 2), cfl)
    case.sync()
    t0 = time.perf_counter()
    case.steps_graph(steps, cfl)
    case.sync()
    elapsed = time.perf_counter() - t0
    # per-kernel times: two more steps launched one kernel at a time, each
    # bracketed by HIP events on the stream it runs on
    case.timing(True)
    case.timing_read()
    case.steps(2, cfl)
    tm = case.timing_read()
    case.timing(False)
    return cells * steps / elapsed, 1e3 * elapsed / steps, tm


def bench_multiblock(args):
    """Multi-block workloads, plain WENO5-JS and the learned hybrid from the
    same developed states:
      batch48    48 independent cases (one block group each) at CFL 0.2: with
                 the reference's local Lax-Friedrichs splitting the blasts need
                 e<0 fix-ups from the first steps at CFL 0.5, and the
                 Navier-Stokes blast of 136^3 diverges there (on the plain path,
                 which is the reference's scheme);
      flatplate  the 16-block supersonic flat plate over a bump (60.5 M cells,
                 curvilinear metrics computed on the GPU from the mesh) at CFL
                 0.5."""
    import numpy as np
    from nuc3d_amd import cases, solver, train_sensor

    exact = 0 if args.fast else 1
    if args.workload == "batch48":
        members = cases.batch_cases(n_cases=48, lo=32, hi=160, seed=0)
        blocks = [m.block for m in members]
        phys, cfl = {}, 0.2
        workload = ("batch of 48 independent cases (24 TGV Re 400-3200, 24 spherical blasts), cubes of seeded "
                    "sizes 32^3..160^3, one block group each, Navier-Stokes, CFL 0.2")
    else:
        members = None
        blocks = cases.flat_plate_blocks()
        phys, cfl = dict(mach=2.9, reynolds=2.0e4, t_wall=2.5), 0.5
        workload = ("supersonic flat plate over a Gaussian bump, Ma 2.9, Re_delta 2e4, isothermal wall T_w 2.5, "
                    "16 blocks of seeded uneven sizes (60.5 M cells), curvilinear tanh-stretched mesh, "
                    "Navier-Stokes, CFL 0.5")
    cells = int(sum(np.prod(b.dims) for b in blocks))
    log("bench: %s: %d blocks, %d cells" % (args.workload, len(blocks), cells))

    def make(params, bl, state=None):
        c = solver.Case(params, bl)
        c.set_uniform_mesh(not args.no_uniform_mesh)
        c.set_overlap(not args.no_overlap)
        c.set_persistent(args.persistent)
        if members:
            for i, m in enumerate(members):
                c.set_group(i, m.group, m.mach, m.reynolds)
        else:
            cases.setup_flat_plate(c, bl)
        if state is not None:
            for i, q in enumerate(state):
                c.set_q(i, q)
        return c

    plain = make(solver.Params(exact=exact, **phys), blocks)
    kseg = plain.autotune()
    log("bench: autotuned k segments per block: %s" % kseg)
    if args.develop:
        plain.steps_graph(args.develop, cfl)
        plain.sync()
    start = [plain.get_q(i) for i in range(len(blocks))]
    v_plain, ms_plain, _ = time_method(plain, args.steps, args.warmup, cells, cfl)
    log("bench: %s plain %.3f ms/step, %.4g cell-updates/s" % (args.workload, ms_plain, v_plain))
    params, thr = train_sensor.load()
    hyb = make(solver.Params(exact=exact, scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), **phys), blocks, start)
    hyb.set_sensor(params, thr)
    hyb.set_sensor_interval(args.sensor_interval)
    hyb.autotune()
    v_h, ms_h, _ = time_method(hyb, args.steps, args.warmup, cells, cfl)
    cen, net, fb = hyb.hybrid_counts()
    num = den = 0.0
    finite = True
    rps = []
    for i in range(len(blocks)):
        jac = plain.get_metrics(i)[0] if members is None else blocks[i].jac
        rp = plain.get_q(i)[0] * jac
        rh = hyb.get_q(i)[0] * jac
        finite = finite and bool(np.isfinite(rp).all() and np.isfinite(rh).all())
        num += float(np.sum((rh - rp) ** 2))
        den += float(np.sum(rp ** 2))
        rps.append(rp)
    # the density fluctuation: about each case's own mean (the batch), about
    # the mean of the whole flow (the flat plate)
    if members:
        fden = sum(float(np.sum((rp - rp.mean()) ** 2)) for rp in rps)
    else:
        mean = sum(float(rp.sum()) for rp in rps) / cells
        fden = sum(float(np.sum((rp - mean) ** 2)) for rp in rps)
    dev = float(np.sqrt(num / den))
    fdev = float(np.sqrt(num / fden)) if fden > 0 else float("inf")
    log("bench: %s hybrid %.3f ms/step, %.4g cell-updates/s (x%.3f), central %.3f, L2 density deviation %.2e "
        "of the rms, %.2e of the fluctuation rms" % (args.workload, ms_h, v_h, v_h / v_plain,
                                                     cen / max(cen + net + fb, 1), dev, fdev))
    plain.close()
    hyb.close()
    ok = finite and fdev <= 1e-3
    value, method, ms = (v_h, "hybrid", ms_h) if (ok and v_h > v_plain) else (v_plain, "plain", ms_plain)
    print(json.dumps({
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload + ", %d plain steps developed: plain WENO5-JS vs learned hybrid" % args.develop,
                   "cells": cells, "blocks": len(blocks), "variant": "fast" if args.fast else "exact",
                   "value_method": method, "column_tile_k_segments": kseg, "persistent_sweeps": args.persistent},
        "plain": {"value": v_plain, "ms_per_step": ms_plain},
        "hybrid": {"value": v_h, "ms_per_step": ms_h, "speedup_vs_plain": v_h / v_plain,
                   "decision_interval_steps": args.sensor_interval,
                   "central_fraction": cen / max(cen + net + fb, 1), "l2_over_density_rms": dev,
                   "l2_over_fluctuation_rms": fdev, "finite": finite,
                   "gate": "L2 density deviation from plain WENO5-JS <= 1e-3 of the rms of the density "
                           "fluctuation (about each case's mean for the batch, the flow's mean for the plate)"},
    }), flush=True)


def bench_tgv(args):
    """The reference's own CPU-runnable case: the Taylor-Green vortex on one
    64^3 periodic block (Navier-Stokes, Ma 0.1, Re 1600, WENO5-JS with local
    Lax-Friedrichs splitting, cd6 viscous terms, TVD-RK3 at CFL 0.5), plain
    method, column segments autotuned."""
    from nuc3d_amd import cases, solver

    n = 64
    blocks = cases.taylor_green(n)
    cells = n ** 3
    c = solver.Case(solver.Params(exact=0 if args.fast else 1), blocks)
    c.set_uniform_mesh(not args.no_uniform_mesh)
    c.set_overlap(not args.no_overlap)
    c.set_persistent(args.persistent)
    kseg = c.autotune()
    v, ms, tm = time_method(c, args.steps, args.warmup, cells)
    c.close()
    log("bench: tgv64 %.3f ms/step, %.4g cell-updates/s" % (ms, v))
    print(json.dumps({
        "metric": "cell-updates/s per RK3 step FP64, Taylor-Green vortex 64^3, plain WENO5-JS", "value": v,
        "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Taylor-Green vortex 64^3 periodic, Navier-Stokes, Ma 0.1, Re 1600, WENO5-JS, cd6, "
                               "TVD-RK3 CFL 0.5", "cells": cells, "variant": "fast" if args.fast else "exact",
                   "column_tile_k_segment": kseg[0]},
        "kernel_ms_per_step": {k: t / 2.0 for k, (t, n_) in tm.items() if n_},
    }), flush=True)


def bench_shock(args):
    """The shock workload of the compact scheme: the 3-D Shu-Osher
    shock/entropy-wave problem (1024 x 128 x 128 cells by default; Euler, Mach
    1, inflow/outflow in x, periodic in y and z, CFL 0.4), CRWENO5 (a
    tridiagonal system per line, component and upwind side) against WENO5-Z
    from the same initial state; `value` is CRWENO5's throughput.  CFL 0.4:
    at 0.5 the reference's CRWENO5 (crweno5.cpp with TVD-RK3) loses
    positivity on this problem within 30 steps and ends in NaN (its own build,
    512 x 64 x 64 in 8 ranks, 40 steps); at 0.4 it runs the 400 steps."""
    from nuc3d_amd import cases, solver

    exact = 0 if args.fast else 1
    nx, ny, nz = args.shock_dims
    blocks = cases.shu_osher(nx, ny, nz, seed=0)
    cells = nx * ny * nz
    res = {}
    for name, sch in (("weno5z", solver.WENO5Z), ("crweno5", solver.CRWENO5)):
        c = solver.Case(solver.Params(model=solver.EULER3D, mach=1.0, scheme=(sch,) * 3, exact=exact), blocks)
        c.set_uniform_mesh(not args.no_uniform_mesh)
        c.set_overlap(not args.no_overlap)
        c.set_persistent(args.persistent)
        v, ms, tm = time_method(c, args.steps, args.warmup, cells, cfl=0.4)
        c.sync()
        fx, pos = c.fixups(), c.positivity()
        c.close()
        if pos != (0, 0):
            raise RuntimeError("shu_osher %s: positivity failed during the timed steps %s" % (name, pos))
        res[name] = {"value": v, "ms_per_step": ms, "fixups": fx,
                     "kernel_ms_per_step": {k: t / 2.0 for k, (t, n) in tm.items() if n}}
        log("bench: shu_osher %-8s %.3f ms/step, %.4g cell-updates/s  %s" % (
            name, ms, v, " ".join("%s %.3f" % (k, t / 2.0) for k, (t, n) in tm.items() if n)))
    print(json.dumps({
        "metric": "cell-updates/s per RK3 step FP64, Shu-Osher: CRWENO5 (compact) vs WENO5-Z", "value":
        res["crweno5"]["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["crweno5"]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "3-D Shu-Osher shock/entropy wave %dx%dx%d, Euler, Mach 1, inflow/outflow in x, "
                               "periodic y and z, CFL 0.4" % (nx, ny, nz), "cells": cells,
                   "variant": "fast" if args.fast else "exact"},
        "crweno5": res["crweno5"], "weno5z": res["weno5z"],
        "crweno5_over_weno5z": res["crweno5"]["value"] / res["weno5z"]["value"],
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--develop", type=int, default=100,
                    help="plain WENO5-JS steps from the initial field before both methods start")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--fast", action="store_true",
                    help="contracted WENO arithmetic (within 1e-12 of the reference on well-conditioned fields only)")
    ap.add_argument("--no-hybrid", action="store_true")
    ap.add_argument("--no-other-variant", action="store_true",
                    help="skip timing the other kernel variant (fast beside exact, or exact beside fast)")
    ap.add_argument("--sensor-interval", type=int, default=1, choices=(1, 2),
                    help="multi-block workloads: steps between the hybrid network's decisions (the turbulence "
                         "workload measures both)")
    ap.add_argument("--shock-dims", type=int, nargs=3, default=(1024, 128, 128), metavar=("NX", "NY", "NZ"),
                    help="cells of the shu_osher workload")
    ap.add_argument("--workload", choices=("turbulence", "batch48", "flatplate", "shu_osher", "tgv64"),
                    default="turbulence",
                    help="batch48: 48 independent TGV and blast cases of seeded sizes 32^3..160^3 in one case "
                         "(block groups); flatplate: the 16-block supersonic flat plate over a bump (60.5 M "
                         "cells); both plain WENO5-JS and the learned hybrid; shu_osher: CRWENO5 against "
                         "WENO5-Z on the Shu-Osher problem (--shock-dims); tgv64: the Taylor-Green vortex "
                         "on one 64^3 block, plain WENO5-JS")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-uniform-mesh", action="store_true",
                    help="load J and the metric fields even where they are one value each")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run the viscous-flux pass in order with the sweeps (for profiles of single kernels)")
    ap.add_argument("--persistent", action="store_true",
                    help="launch every direction sweep as one persistent grid over the (block, tile) work queue")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (every block of a case lives on one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import numpy as np
    from nuc3d_amd import cases, solver, train_sensor

    if args.workload == "shu_osher":
        return bench_shock(args)
    if args.workload == "tgv64":
        return bench_tgv(args)
    if args.workload != "turbulence":
        return bench_multiblock(args)
    n = args.n
    cells = n ** 3
    log("bench: generating %d^3 isotropic turbulence" % n)
    blocks, phys = cases.isotropic_turbulence(n=n, seed=0)
    common = dict(model=solver.NAVIERSTOKES3D, mach=phys.mach, reynolds=phys.reynolds, exact=0 if args.fast else 1)
    plain = solver.Case(solver.Params(**common), blocks)
    plain.set_uniform_mesh(not args.no_uniform_mesh)
    plain.set_overlap(not args.no_overlap)
    plain.set_persistent(args.persistent)
    kseg = plain.autotune()
    log("bench: autotuned k segments of the column tiles: %s" % kseg)
    if args.develop:
        log("bench: developing the field, %d plain steps" % args.develop)
        plain.steps_graph(args.develop, 0.5)
        plain.sync()
    spec = blocks[0]
    start = [cases.BlockSpec(spec.dims, spec.faces, spec.jac, spec.met, spec.dxyz, plain.get_q(0))]

    log("bench: plain WENO5-JS, %d timed steps" % args.steps)
    v_plain, ms_plain, tm = time_method(plain, args.steps, args.warmup, cells)
    ns = True
    for k, (ms, cnt) in tm.items():
        if cnt:
            log("bench: plain  %-19s %8.3f ms avg over %d launches" % (k, ms / cnt, cnt))
    uniform = plain.uniform_mesh(0)
    alg_step = ALG_BYTES_PER_STEP_UNIFORM if uniform else ALG_BYTES_PER_STEP
    dom = max(tm, key=lambda k: tm[k][0])
    dom_ms, dom_n = tm[dom]
    avg_ms = dom_ms / max(dom_n, 1)
    # the algorithmic bytes of one stage-kernel launch: a third of the step's
    # compulsory bytes per cell-update (320 B on the uniform-mesh path, 560 B
    # with the metric fields) for every cell of the case
    alg_launch = alg_step / 3.0 * cells
    achieved = alg_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    model_gbs = kernel_alg_bytes(dom, ns, uniform) * cells / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    variant = "fast" if args.fast else "exact"
    traffic = read_traffic(dom, n, variant)
    log("bench: plain %.3f ms/step, %.4g cell-updates/s" % (ms_plain, v_plain))
    sweep_tiles = plain.sweep_tiles(0)
    # the dominant kernel's FP64 VALU figures, from the committed counters
    nf_dom = {"x_sweep": sweep_tiles[0], "y_sweep": sweep_tiles[1], "z_sweep": sweep_tiles[2]}.get(dom)
    dom_kernel = ("k_sweep_%s<%d, 0, true, %d, %d, false, false>" % (variant, {"x_sweep": 0, "y_sweep": 1,
                                                                            "z_sweep": 2}[dom],
                                                                     1 if dom == "z_sweep" else 0, nf_dom)
                  if nf_dom else "k_visc_flux_%s<3, true>" % variant)
    valu = read_valu(dom_kernel, n, variant)

    # the other kernel variant on the same developed state, for the record
    other = None
    if not args.no_other_variant:
        oc = solver.Case(solver.Params(**dict(common, exact=1 if args.fast else 0)), start)
        oc.set_uniform_mesh(not args.no_uniform_mesh)
        oc.set_overlap(not args.no_overlap)
        oc.set_persistent(args.persistent)
        oc.autotune()
        v_o, ms_o, _ = time_method(oc, args.steps, args.warmup, cells)
        oc.close()
        other = {"variant": "exact" if args.fast else "fast", "value": v_o, "ms_per_step": ms_o,
                 "note": "the fast variant contracts the WENO arithmetic into FMA with one division per weight "
                         "set; its parity bar on ill-conditioned components is 8x the reference's own "
                         "-O0/-O2 spread (the exact variant's is 1x)"}
        log("bench: plain, %s variant: %.3f ms/step, %.4g cell-updates/s" % (other["variant"], ms_o, v_o))

    hybrid = None
    if not args.no_hybrid:
        params, thr = train_sensor.load()
        rho_p = plain.get_q(0)[0] * spec.jac
        rms = float(np.sqrt(np.mean(rho_p ** 2)))
        rms_f = float(np.sqrt(np.mean((rho_p - rho_p.mean()) ** 2)))
        cadences = {}
        # the network decides at stage 0 of every step (interval 1) or of every
        # other step (interval 2); the positivity fallback is evaluated at every
        # stage either way
        for interval in (1, 2):
            hcase = solver.Case(solver.Params(scheme=(solver.HYBRID,) * 3, width=(4, 4, 4), **common), start)
            hcase.set_uniform_mesh(not args.no_uniform_mesh)
            hcase.set_overlap(not args.no_overlap)
            hcase.set_persistent(args.persistent)
            hcase.set_sensor(params, thr)
            hcase.set_sensor_interval(interval)
            hcase.autotune()
            log("bench: learned hybrid, decisions every %d step(s), %d timed steps" % (interval, args.steps))
            v_h, ms_h, tmh = time_method(hcase, args.steps, args.warmup, cells)
            for k, (ms, cnt) in tmh.items():
                if cnt:
                    log("bench: hybrid %-19s %8.3f ms avg over %d launches" % (k, ms / cnt, cnt))
            cen, net, fb = hcase.hybrid_counts()
            tot = max(cen + net + fb, 1)
            # both cases have now run the same steps from the developed state
            rho_h = hcase.get_q(0)[0] * spec.jac
            l2 = float(np.sqrt(np.mean((rho_h - rho_p) ** 2)))
            cadences[interval] = {
                "value": v_h, "ms_per_step": ms_h, "speedup_vs_plain": v_h / v_plain,
                "central_fraction": cen / tot, "weno_by_network_fraction": net / tot, "fallback_fraction": fb / tot,
                "l2_density_deviation": l2, "l2_over_density_rms": l2 / rms, "l2_over_fluctuation_rms": l2 / rms_f,
                "max_density_deviation": float(np.abs(rho_h - rho_p).max()),
                "kernel_ms_per_step": {k: v[0] / 2 for k, v in tmh.items()},
            }
            log("bench: hybrid (interval %d) %.3f ms/step, %.4g cell-updates/s (x%.3f), central faces %.3f, "
                "L2 density deviation %.3e of rms, %.3e of the fluctuation rms" % (
                    interval, ms_h, v_h, v_h / v_plain, cen / tot, l2 / rms, l2 / rms_f))
            hcase.close()
        hybrid = {"logit_threshold": thr, "decisions_every_step": cadences[1],
                  "decisions_every_other_step": cadences[2],
                  "gate": "L2 density deviation from plain WENO5-JS <= 1e-3 of the rms of the density "
                          "fluctuation (the stricter reading; the deviation over the rms of the density itself "
                          "is reported beside it)"}
    plain.close()

    value, method, ms_step = v_plain, "plain", ms_plain
    if hybrid:
        for interval, key in ((1, "decisions_every_step"), (2, "decisions_every_other_step")):
            h = hybrid[key]
            if h["l2_over_fluctuation_rms"] <= 1e-3 and h["value"] > value:
                value, method, ms_step = h["value"], "hybrid (decision interval %d)" % interval, h["ms_per_step"]

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        try:
            cpu = cpu_baseline(n)
        except Exception as e:
            log("cpu baseline failed: %s" % e)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "decaying isotropic turbulence %d^3 periodic (seed 0, %d plain steps developed), "
                        "Navier-Stokes, local LF splitting, cd6 viscous terms, TVD-RK3 CFL 0.5: plain WENO5-JS "
                        "vs learned hybrid (central-8 / WENO5-JS per face)" % (n, args.develop),
            "cells": cells,
            "variant": variant,
            "value_method": method,
            "column_tile_k_segment": kseg[0],
            "persistent_sweeps": args.persistent,
        },
        "plain": {"value": v_plain, "ms_per_step": ms_plain, "kernel_ms_per_step": {k: v[0] / 2 for k, v in tm.items()},
                  "sweep_faces_per_tile_line": sweep_tiles},
        "plain_other_variant": other,
        "hybrid": hybrid,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_over_alg": traffic / alg_launch if traffic else None,
            "kernel": KERNEL_OF_CLASS[dom],
            "avg_launch_ms": avg_ms,
            "alg_bytes_per_launch": alg_launch,
            "alg_bytes_basis": "%g B per cell-update / 3 stages x %d cells per stage-kernel launch" % (alg_step, cells),
            "uniform_mesh": uniform,
            "step_alg_bytes_per_cell_update": alg_step,
            "step_level_gbs": alg_step * v_plain / 1e9,
            "step_frac": alg_step * v_plain / 1e9 / HBM_PEAK_GBS,
            "kernel_model_bytes_per_cell": kernel_alg_bytes(dom, ns, uniform),
            "kernel_model_gbs": model_gbs,
            "fp64_valu": valu,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
