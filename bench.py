#!/usr/bin/env python3
"""Benchmark: 3-D heat diffusion 512^3 FP64 (T_eff, fraction of MI355X HBM peak), with the
triad A = B + C over three 512^3 arrays as the bandwidth calibration and the plain
(non-memopt) kernel beside the memopt (z-streaming) kernel.

One JSON line on stdout. T_eff follows the reference's definition
(A_eff = 2*D_u + D_k; T, T2 written/read once, Cp read once: 3 * 8 B * n^3 per step).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallelstencil.jl_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E peak


def log(msg):
    print(msg, file=sys.stderr, flush=True)


SETTLE = 150  # untimed back-to-back launches before each kernel's warm-up


def time_kernel(ps, launch, steps, warmup, settle=SETTLE):
    """Runs `launch(i)` settle + warmup + steps times asynchronously on the library stream;
    returns (wall ms per step, kernel ms per launch from HIP events on that stream), timed over
    the last `steps` launches only. The `settle` launches come first because a sustained load
    makes the GPU's power management cut the clock a few launches in and restore it over the
    next few tens: at 512^3 the exact diffusion kernel runs 0.63 ms, then up to 0.94 ms, and is
    back at a steady 0.65 ms by launch 60. The timed steps then measure the sustained rate."""
    for i in range(settle):
        launch(i)
    for i in range(warmup):
        launch(i)
    ps.synchronize()
    e0, e1 = ps.Event(), ps.Event()
    t0 = time.perf_counter()
    e0.record()
    for i in range(steps):
        launch(i)
    e1.record()
    ps.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps, e0.elapsed_ms(e1) / steps


HEAD_ARITHMETIC = {
    "memopt_fastmath": "FP64 with FMA contraction and division through the reciprocal of uniform divisors "
                       "(fastmath=true); within 1e-12 relative per field of the exact oracle after 200 steps",
    "memopt": "FP64, bit-exact against the oracle (IEEE division, no contraction)",
    "plain": "FP64, bit-exact against the oracle (IEEE division, no contraction)",
}


def _traffic(n, head):
    """(HBM bytes per launch of the head kernel, where the figure comes from): read from the
    committed rocprofv3 counter summary of the same kernel (FETCH_SIZE scaled by the triad
    calibration + WRITE_SIZE), not measured in this run; (None, None) when there is none."""
    rel = os.path.join("profiles", f"diffusion3d_{n}_bench.md")
    path = os.path.join(ROOT, rel)
    if head == "triad" or not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 3 and p[0] == "hbm_bytes_per_launch" and p[1].startswith("psk_diffusion3D") and \
                    (("_memopt" in p[1]) == head.startswith("memopt")) and (("_fastmath" in p[1]) == head.endswith("_fastmath")):
                return int(p[2]), f"{rel} (rocprofv3 FETCH_SIZE/WRITE_SIZE of {p[1]}, committed profile, not this run)"
    return None, None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def per_launch_ms(ps, launch, n=20):
    """Event-timed durations of n single launches (after the timed loop): min and median, to
    show the spread behind the mean."""
    ev = [(ps.Event(), ps.Event()) for _ in range(n)]
    for i, (a, b) in enumerate(ev):
        a.record()
        launch(i)
        b.record()
    ps.synchronize()
    t = sorted(a.elapsed_ms(b) for a, b in ev)
    return t[0], t[len(t) // 2]


def cpu_baseline(n, budget_s=12.0):
    """The oracle's diffusion step (Threads-backend loop nest, OpenMP over z) on the host cores."""
    from oracle import ps_oracle as O
    import numpy as np
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    O.set_threads(cores)
    T0, Cp, p = O.diffusion_inputs(n, seed=1)
    T2 = np.asfortranarray(T0.copy())
    args = (p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
    O.diffusion3d_flux(T2, T0, Cp, *args, fast=True)  # warm-up
    steps, t0 = 0, time.perf_counter()
    a, b = T0, T2
    while steps < 3 or time.perf_counter() - t0 < budget_s:
        O.diffusion3d_flux(b, a, Cp, *args, fast=True)
        a, b = b, a
        steps += 1
    dt = (time.perf_counter() - t0) / steps
    return {"value": round(3 * 8 * n ** 3 / dt / 1e9, 2), "unit": "GB/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{steps} steps of the diffusion3D {n}^3 FP64 step (oracle, contraction allowed, OpenMP)"}


def _sig(v):
    """Per-sample record values to 4 significant digits (residuals of 1e-9 stay visible)."""
    if isinstance(v, float):
        return float(f"{v:.4g}")
    if isinstance(v, dict):
        return {k: _sig(x) for k, x in v.items()}
    return v


def stokes_warm_start(args):
    """Stokes2D time to tolerance (residual 1e-8, FP64 on the GPU), warm start from the U-Net
    (inference included) against the plain PT iteration from rest, on held-out samples at
    --stokes-n and at --stokes-n2. The network is trained here, on the GPU with the library's own
    kernels, on samples the solver generates at --stokes-train-n (default --stokes-n); at other
    resolutions its inputs and outputs are resampled. The solver is the two-launch fused
    iteration (bit-identical to the miniapp's six launches); with --stokes-ts maxV the velocity
    pseudo-time steps use the 3x3 maximum viscosity, which keeps the iteration stable at the
    contrasts of --stokes-contrast."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import unet as U
    from parallelstencil.stokes_fused import Stokes2DBatch, Stokes2DFused

    mod = "bench.stokes"
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=mod)
    n, contrast = args.stokes_n, args.stokes_contrast
    nt = args.stokes_train_n or n
    tsolver = Stokes2DFused(mod, nt, nt, eps=1e-8, nout=200, eta_ts=args.stokes_ts, history=True)
    t0 = time.perf_counter()
    data, stats = [], []
    if args.stokes_batch > 1 and not args.stokes_load and args.train:
        # the training solves S at a time (Stokes2DBatch's slot queue): the same iterations, residuals
        # and fields per seed as the single-sample solver
        bsolver = Stokes2DBatch(mod, nt, nt, args.stokes_batch, eps=1e-8, nout=200, eta_ts=args.stokes_ts)
        data, stats = U.generate_samples_batch(bsolver, range(1, args.train + 1), contrast, log=log)
        bsolver.close()
        del bsolver
    else:
        for i0 in range(1, (0 if args.stokes_load else args.train) + 1, 64):
            d, st = U.generate_samples(tsolver, range(i0, min(args.train, i0 + 63) + 1), contrast)
            data += d
            stats += st
            log(f"stokes {nt}^2: {len(stats)} training solves, {time.perf_counter() - t0:.1f} s, last {st[-1][0]} iterations")
    t1 = time.perf_counter()
    if args.stokes_load:
        ws = U.WarmStart.load(args.stokes_load)
        if (ws.nx, ws.ny) != (nt, nt):
            raise SystemExit(f"{args.stokes_load} holds a warm start for {ws.nx}x{ws.ny}, not the training resolution {nt}")
        m = ws.meta
        ok, its, hist, t2, n_train = [True] * m.get("train_solves_converged", 0), [0], [float("nan")], t1, m.get("train_samples", 0)
        n_solves = m.get("train_solves", 0)
        args.epochs, args.stokes_mirror = m.get("epochs"), m.get("mirrored", args.stokes_mirror)
        log(f"warm start loaded from {args.stokes_load}: {m}")
    else:
        n_solves = len(stats)
        ok = [s_[2] <= tsolver.eps for s_ in stats]
        data = [d for d, k in zip(data, ok) if k]  # samples on which the plain iteration did not converge are dropped
        its = [s_[0] for s_, k in zip(stats, ok) if k]
        log(f"stokes {nt}^2: {len(stats)} training solves in {t1 - t0:.1f} s, {sum(ok)} converged, mean {np.mean(its):.0f} PT iterations, "
            f"{1e6 * (t1 - t0) / max(1, sum(s_[0] for s_ in stats)):.2f} us per sample-iteration"
            f"{f' ({args.stokes_batch} samples per launch pair)' if args.stokes_batch > 1 else ''}")
        if args.stokes_mirror:  # the problem's mirror symmetry in x doubles the training set
            data = data + [U.mirror_x(d) for d in data]
        n_train = len(data)
        ws = U.WarmStart(nt, nt, base=32, seed=0)
        X, T, Mk = U.dataset(ws, data)
        ws.set_normalisation(T, Mk)
        del data
        hist = ws.train(X, T, Mk, epochs=args.epochs, batch=8, lr=3e-3, log=log, cosine=True)
        ps.synchronize()
        t2 = time.perf_counter()
        log(f"training: {n_train} samples, {args.epochs} epochs in {t2 - t1:.1f} s, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
        if args.stokes_save:
            ws.save(args.stokes_save, dict(train_samples=n_train, train_solves_converged=sum(ok), train_solves=len(stats), epochs=args.epochs,
                                           mirrored=args.stokes_mirror, contrast_x10=round(10 * contrast)))
            log(f"warm start saved to {args.stokes_save}")
    evals = {}
    sizes = []
    todo = [(n, args.eval), (args.stokes_n2, args.eval2)]
    if args.eval_train and nt not in (n, args.stokes_n2):  # the training resolution itself
        todo.insert(0, (nt, args.eval_train))
    for m, ne in todo:
        if not m or not ne:
            continue
        sv = tsolver if m == nt else Stokes2DFused(mod, m, m, eps=1e-8, nout=200, eta_ts=args.stokes_ts, history=True)
        sizes.append((m, ne, sv, None if m == nt else U.resampled_predictor(ws, nt, m)))
    for m, ne, sv, pred in sizes:
        ev_all = U.evaluate(ws, sv, range(100001, 100001 + ne), contrast, predict=pred, log=log)
        ev = [e for e in ev_all if e["err_cold"] <= sv.eps]  # held-out samples the plain iteration solves
        cold_s, warm_s = sum(e["cold_s"] for e in ev), sum(e["warm_s"] for e in ev)
        cold_it, warm_it = np.mean([e["cold_iters"] for e in ev]), np.mean([e["warm_iters"] for e in ev])
        perr = {f: float(np.median([e["pred_err"][f] for e in ev])) for f in ("Vx", "Vy", "P")}
        log(f"eval {m}^2: cold {cold_it:.0f} iterations / {cold_s:.3f} s, warm {warm_it:.0f} iterations / {warm_s:.3f} s, "
            f"gain {cold_s / warm_s:.3f}x; median relative error of the prediction {perr}")
        evals[m] = dict(n=m, samples=len(ev), dropped_not_converged=len(ev_all) - len(ev), cold_iters_mean=float(cold_it),
                        warm_iters_mean=float(warm_it), cold_s=cold_s, warm_s=warm_s, gain=cold_s / warm_s,
                        warm_fallbacks=sum(e["fallback"] for e in ev), warm_converged=all(e["err_warm"] <= sv.eps for e in ev),
                        pred_err_median=perr, per_sample=[{k: _sig(v) for k, v in e.items()} for e in ev])
    head = evals[n]
    out = {
        "metric": "T_eff GB/s, %HBM peak (diffusion3D 512³); Stokes2D time-to-tol: warm vs plain",
        "value": round(head["gain"], 3),
        "unit": "x (plain PT time / warm-started time, inference included)",
        "n_gpus": 1, "steps": head["samples"], "warmup": 0, "ms_per_step": round(1e3 * head["warm_s"] / max(1, head["samples"]), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 (solver), bf16/fp32 (network)",
        "data": "synthetic",
        "config": {"workload": f"Stokes2D {n}x{n} pseudo-transient to residual 1e-8, 1-8 inclusions, log10 contrast in "
                               f"[-{contrast:g}, {contrast:g}], pseudo-time steps {args.stokes_ts}; U-Net 4 levels, 32 base channels, BF16, "
                               f"inputs log10 viscosity and body force, trained at {nt}x{nt}",
                   "train_samples": n_train, "train_solves_converged": sum(ok), "train_mirrored": bool(args.stokes_mirror),
                   "train_n": nt, "eval_samples": {str(k): v["samples"] for k, v in evals.items()}, "epochs": args.epochs,
                   "warm_start_loaded_from": args.stokes_load or None},
        "stokes": {"datagen_s": t1 - t0, "datagen_samples_per_launch": args.stokes_batch if stats and args.stokes_batch > 1 else 1,
                   "train_s": t2 - t1, "loss_first": hist[0], "loss_last": hist[-1],
                   "train_dropped_not_converged": n_solves - sum(ok), "train_iters_mean": float(np.mean(its)) if stats else None,
                   "eval": {str(k): {kk: vv for kk, vv in v.items() if kk != "per_sample"} for k, v in evals.items()}},
    }
    if args.stokes_out:
        with open(args.stokes_out, "w") as f:
            json.dump(dict(out, per_sample={str(k): v["per_sample"] for k, v in evals.items()}, loss_history=hist), f)
    print(json.dumps(out), flush=True)


def acoustic_study(args):
    """acoustic3D n^3 (miniapps/acoustic3D.jl), `steps` time steps from the miniapp's initial
    state in FP64, FP32 and BFloat16 storage with FP32 arithmetic, plain and memopt kernels, exact
    and (unless --acoustic-exact-only) fastmath: T_eff per step (reference A_eff = 8 field sweeps,
    miniapps/acoustic3D.jl:62, and the 12 sweeps the two unfused kernels need) and the per-field
    error of each run against the exact FP64 plain run."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import workloads as W

    n, steps = args.acoustic_n, args.acoustic_steps
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module="bench.a64")
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 3, module="bench.a32")
    P0, prm = W.acoustic3d_inputs(n)
    runs, ref = {}, None
    variants = [(prec, mod, elt, size, memopt, fast) for fast in (False, True)
                for prec, mod, elt, size in (("f64", "bench.a64", ps.Float64, 8), ("f32", "bench.a32", ps.Float32, 4),
                                             ("bf16", "bench.a32", ps.BFloat16, 2))
                for memopt in (False, True) if not (fast and args.acoustic_exact_only)]
    for prec, mod, elt, size, memopt, fast in variants:
        kV, kP = W.define_acoustic3d(mod, memopt, fastmath=fast)
        P, Vx, Vy, Vz = (ps.zeros(n, n, n, eltype=elt), ps.zeros(n + 1, n, n, eltype=elt), ps.zeros(n, n + 1, n, eltype=elt),
                         ps.zeros(n, n, n + 1, eltype=elt))
        a = (prm["dt"], prm["rho"], prm["dx"], prm["dy"], prm["dz"])
        b = (prm["dt"], prm["k"], prm["dx"], prm["dy"], prm["dz"])
        for _ in range(3):  # compile and warm up on the zero state
            kV(Vx, Vy, Vz, P, *a, async_=True)
            kP(P, Vx, Vy, Vz, *b, async_=True)
        for A in (Vx, Vy, Vz):
            A.fill_(0.0)
        P.copy_from(W._to_storage(P0, elt))
        ps.synchronize()
        e0, e1 = ps.Event(), ps.Event()
        e0.record()
        for _ in range(steps):
            kV(Vx, Vy, Vz, P, *a, async_=True)
            kP(P, Vx, Vy, Vz, *b, async_=True)
        e1.record()
        ps.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        fields = tuple(W._from_storage(A.to_host(), elt) for A in (P, Vx, Vy, Vz))
        if prec == "f64" and not memopt and not fast:
            ref = fields
        errs = {name: W.field_errors(f, r) for name, f, r in zip(("P", "Vx", "Vy", "Vz"), fields, ref)}
        key = f"{prec}_{'memopt' if memopt else 'plain'}" + ("_fastmath" if fast else "")
        runs[key] = {"ms_per_step": round(ms, 4), "T_eff_gbs_ref": round(8 * size * n ** 3 / ms / 1e6, 1),
                     "T_eff_gbs_12sweeps": round(12 * size * n ** 3 / ms / 1e6, 1),
                     "err_vs_f64 (max, l2)": {k: [float(f"{v[0]:.3e}"), float(f"{v[1]:.3e}")] for k, v in errs.items()}}
        log(f"acoustic3D {n}^3 {key}: {ms:.3f} ms/step, T_eff {runs[key]['T_eff_gbs_ref']} GB/s (ref A_eff), "
            f"P err max {errs['P'][0]:.2e} l2 {errs['P'][1]:.2e}")
        del P, Vx, Vy, Vz
    best = min(runs.values(), key=lambda r: r["ms_per_step"])
    f64 = min((runs["f64_plain"], runs["f64_memopt"]), key=lambda r: r["ms_per_step"])  # the faster exact FP64 path
    out = {"metric": "T_eff GB/s (acoustic3D, reference A_eff = 8 field sweeps)", "value": f64["T_eff_gbs_ref"],
           "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": 3, "ms_per_step": f64["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 / f32 / bf16-storage", "data": "synthetic",
           "config": {"workload": f"acoustic3D {n}^3, {steps} steps from the miniapp initial state (P0 = exp(-r^2), V0 = 0)"},
           "runs": runs, "fastest_ms_per_step": best["ms_per_step"]}
    print(json.dumps(out), flush=True)


def xgroups_study(args):
    """acoustic3D n^3 plain kernels (miniapps/acoustic3D.jl), exact, in BFloat16, FP32 and FP64
    storage: kernel time per step (V + P, HIP events) over x-group launch shapes: cells per
    thread (cells_x), block width along x (nthreads_x_groups) and the odd x grid (xcd_spread),
    with the library's default shape beside them. Under sustained load the GPU's clock drops over
    tens of seconds (by up to 30% for instruction-bound kernels), so every shape of every
    precision is measured once per round, rounds alternate the order, and each shape reports the
    median of its rounds."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import workloads as W

    n, burst, rounds = args.acoustic_n, 100, 5
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 3, module="bench.x32")
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module="bench.x64")
    P0, prm = W.acoustic3d_inputs(n)
    a = (prm["dt"], prm["rho"], prm["dx"], prm["dy"], prm["dz"])
    b = (prm["dt"], prm["k"], prm["dx"], prm["dy"], prm["dz"])
    fields = {}
    for prec, elt in (("bf16", ps.BFloat16), ("f32", ps.Float32), ("f64", ps.Float64)):
        F = (ps.zeros(n, n, n, eltype=elt), ps.zeros(n + 1, n, n, eltype=elt), ps.zeros(n, n + 1, n, eltype=elt), ps.zeros(n, n, n + 1, eltype=elt))
        F[0].copy_from(W._to_storage(P0, elt))
        fields[prec] = F
    cfgs = [("f64", 0, 0, True)] + [("f64", cx, w, True) for cx, w in ((2, 64), (2, 32), (4, 32), (4, 16))]
    cfgs += [(p, cx, w, True, cz) for p, cx, w in (("f64", 1, 0), ("f32", 4, 32), ("bf16", 4, 32)) for cz in (2, 4)]
    cfgs += [(p, 2, w, True, 2) for p in ("f32", "bf16", "f64") for w in (32, 64)]
    cfgs += [("bf16", 0, 0, True), ("bf16", 1, 0, True)] + [("bf16", cx, w, sp) for cx in (2, 4, 8) for w in (16, 32, 64) for sp in (False, True)]
    cfgs += [("f32", 0, 0, True), ("f32", 1, 0, True)] + [("f32", cx, w, sp) for cx in (2, 4) for w in (16, 32, 64) for sp in (False, True)]
    kern = {c: W.define_acoustic3d("bench.x64" if c[0] == "f64" else "bench.x32", False, cells_x=c[1], nthreads_x_groups=c[2], xcd_spread=c[3],
                                   cells_z=c[4] if len(c) > 4 else 0 if c[1] == 0 else 1)
            for c in cfgs}
    res = {c: [] for c in cfgs}
    for r in range(rounds):
        for c in (cfgs if r % 2 == 0 else cfgs[::-1]):
            kV, kP = kern[c]
            P, Vx, Vy, Vz = fields[c[0]]
            for _ in range(20):
                kV(Vx, Vy, Vz, P, *a, async_=True)
                kP(P, Vx, Vy, Vz, *b, async_=True)
            e = [ps.Event() for _ in range(3)]
            e[0].record()
            for _ in range(burst):
                kV(Vx, Vy, Vz, P, *a, async_=True)
            e[1].record()
            for _ in range(burst):
                kP(P, Vx, Vy, Vz, *b, async_=True)
            e[2].record()
            ps.synchronize()
            res[c].append((e[0].elapsed_ms(e[1]) / burst, e[1].elapsed_ms(e[2]) / burst))
        log(f"round {r + 1} of {rounds} done")
    out_shapes = {}
    for c in cfgs:
        prec, cx, w, sp = c[:4]
        cz = c[4] if len(c) > 4 else 0
        key = f"{prec} " + ("default" if cx == 0 else f"cells_x={cx}" + (f" width={w} odd_grid={int(sp)}" if cx > 1 else ""))
        if cz:
            key = f"{prec} cells_x={cx} cells_z={cz}" + (f" width={w}" if cx > 1 else "")
        v = float(np.median([x[0] for x in res[c]]))
        p = float(np.median([x[1] for x in res[c]]))
        tot = [x[0] + x[1] for x in res[c]]
        out_shapes[key] = {"V_ms": round(v, 4), "P_ms": round(p, 4), "V+P_ms": round(float(np.median(tot)), 4),
                           "V+P_min_max": [round(min(tot), 4), round(max(tot), 4)]}
        log(f"acoustic3D {n}^3 {key}: V {v:.4f} ms, P {p:.4f} ms, V+P median {np.median(tot):.4f} ms (rounds {min(tot):.4f} - {max(tot):.4f})")
    # 2-D: acoustic2D 4096^2 (miniapps/acoustic2D.jl), same rounds
    n2 = 4096
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 2, module="bench.x2d")
    P2, prm2 = W.acoustic2d_inputs(n2)
    a2 = (prm2["dt"], prm2["rho"], prm2["dx"], prm2["dy"])
    b2 = (prm2["dt"], prm2["k"], prm2["dx"], prm2["dy"])
    f2 = {}
    for prec, elt in (("bf16", ps.BFloat16), ("f32", ps.Float32)):
        F = (ps.zeros(n2, n2, eltype=elt), ps.zeros(n2 + 1, n2, eltype=elt), ps.zeros(n2, n2 + 1, eltype=elt))
        F[0].copy_from(W._to_storage(P2, elt))
        f2[prec] = F
    c2 = [(prec, cx, w) for prec in ("bf16", "f32") for cx, w in ((0, 0), (1, 0), (2, 32), (2, 64), (4, 16), (4, 32), (4, 64))]
    k2 = {c: W.define_acoustic2d("bench.x2d", False, cells_x=c[1], nthreads_x_groups=c[2]) for c in c2}
    r2 = {c: [] for c in c2}
    for r in range(rounds):
        for c in (c2 if r % 2 == 0 else c2[::-1]):
            kV, kP = k2[c]
            P, Vx, Vy = f2[c[0]]
            for _ in range(20):
                kV(Vx, Vy, P, *a2, async_=True)
                kP(P, Vx, Vy, *b2, async_=True)
            e = [ps.Event() for _ in range(3)]
            e[0].record()
            for _ in range(burst):
                kV(Vx, Vy, P, *a2, async_=True)
            e[1].record()
            for _ in range(burst):
                kP(P, Vx, Vy, *b2, async_=True)
            e[2].record()
            ps.synchronize()
            r2[c].append((e[0].elapsed_ms(e[1]) / burst, e[1].elapsed_ms(e[2]) / burst))
    for c in c2:
        prec, cx, w = c
        key = f"2d {prec} " + ("default" if cx == 0 else f"cells_x={cx}" + (f" width={w}" if cx > 1 else ""))
        tot = [x[0] + x[1] for x in r2[c]]
        out_shapes[key] = {"V_ms": round(float(np.median([x[0] for x in r2[c]])), 4), "P_ms": round(float(np.median([x[1] for x in r2[c]])), 4),
                           "V+P_ms": round(float(np.median(tot)), 4), "V+P_min_max": [round(min(tot), 4), round(max(tot), 4)]}
        log(f"acoustic2D {n2}^2 {key}: V+P median {np.median(tot):.4f} ms (rounds {min(tot):.4f} - {max(tot):.4f})")
    ms16, ms32, ms64 = (out_shapes[f"{p} default"]["V+P_ms"] for p in ("bf16", "f32", "f64"))
    log(f"default shapes: BF16 {ms16:.4f} ms, FP32 {ms32:.4f} ms, FP64 {ms64:.4f} ms per step; BF16 / FP32 speed-up {ms32 / ms16:.2f}x")
    out = {"metric": "acoustic3D plain kernel time per step (V + P) by x-group launch shape", "value": ms16,
           "unit": "ms per step (BFloat16, default shape)", "n_gpus": 1, "steps": burst * rounds, "warmup": 20, "ms_per_step": ms16,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16-storage / f32 / f64", "data": "synthetic",
           "config": {"workload": f"acoustic3D {n}^3 plain kernels, exact, miniapp initial state; {rounds} interleaved rounds of {burst} steps"},
           "f32_default_ms_per_step": ms32, "f64_default_ms_per_step": ms64, "bf16_speedup_over_f32": round(ms32 / ms16, 3),
           "shapes": out_shapes}
    print(json.dumps(out), flush=True)


def adjoint_study(args):
    """Acoustic 2-D inversion gradient at n^2 over `steps` steps (sqrt(N) checkpointing): time of
    the forward run and of the checkpointed gradient in FP32 and FP64, the FP32 gradient against
    the FP64 one, and the FP64 gradient against central finite differences along 8 random
    directions."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import acoustic_adjoint as AA

    n, nt = args.adjoint_n, args.adjoint_steps
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float32, 2, module="bench.adj32")
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module="bench.adj64")
    s = AA.Setup(n, n, nt, seed=1, nrec=64, sigma=100.0)
    res = {}
    grads = {}
    for prec, mod in (("f32", "bench.adj32"), ("f64", "bench.adj64")):
        inv = AA.Acoustic2DInversion(mod, s)
        Dobs = inv.forward(s.K0)  # observed data from the background model
        inv.set_observed(Dobs)
        inv.forward(s.K)  # warm-up: compiles the forward and adjoint kernels outside the timing
        inv.gradient(s.K)
        t0 = time.perf_counter()
        inv.forward(s.K)
        t1 = time.perf_counter()
        J, g = inv.gradient(s.K)
        t2 = time.perf_counter()
        grads[prec] = g
        res[prec] = {"forward_s": round(t1 - t0, 4), "gradient_s": round(t2 - t1, 4), "J": J, "checkpoints": len(inv.ck),
                     "segment": inv.seg}
        log(f"adjoint {n}^2 x {nt} {prec}: forward {t1 - t0:.3f} s, gradient {t2 - t1:.3f} s ({len(inv.ck)} checkpoints), J {J:.6e}")
        if prec == "f64":
            rng = np.random.default_rng(11)
            fd = []
            for _ in range(8):
                d = rng.standard_normal(s.K.shape) * np.exp(-((np.arange(n)[:, None] - n / 2) ** 2 + (np.arange(n)[None, :] - n / 2) ** 2) / (2 * (n / 4) ** 2))
                h = 1e-4
                f = (inv.misfit(s.K + h * d) - inv.misfit(s.K - h * d)) / (2 * h)
                a = float(np.sum(g * d))
                fd.append(abs(f - a) / max(abs(f), 1e-300))
            res[prec]["fd_rel_err_max"] = float(max(fd))
            log(f"finite differences (8 directions): max relative error {max(fd):.2e}")
    rel = float(np.linalg.norm(grads["f32"] - grads["f64"]) / np.linalg.norm(grads["f64"]))
    out = {"metric": "acoustic2D inversion gradient time (sqrt(N) checkpointed adjoint)", "value": res["f32"]["gradient_s"],
           "unit": "s per gradient", "n_gpus": 1, "steps": nt, "warmup": 1, "ms_per_step": round(1e3 * res["f32"]["gradient_s"] / nt, 4),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 check)", "data": "synthetic",
           "config": {"workload": f"acoustic2D adjoint {n}x{n}, {nt} steps, 64 receivers, Ricker source, K = 1 + 0.1 Gaussian (sigma 100 cells)"},
           "runs": res, "f32_vs_f64_gradient_rel_l2": rel}
    log(f"FP32 vs FP64 gradient: relative L2 {rel:.2e}")
    print(json.dumps(out), flush=True)


def gemm_study(args):
    """BF16 MFMA GEMM of the warm-start network, HIP-event timed: the three products of a 3x3
    convolution layer of the U-Net at batch 8 x 128^2 (forward, weight gradient, data gradient),
    with the im2col operand gathered inside the GEMM (ps_nn_gemm_im2col, what the network runs)
    and through a stored column matrix (ps_nn_im2col / ps_nn_col2im + ps_nn_gemm), and a square
    4096^3 product, as TFLOP/s against the BF16 dense MFMA peak."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import unet as U

    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module="bench.gemm")
    peak = 2500.0  # TFLOP/s, MI355X BF16 dense MFMA
    rng = np.random.default_rng(0)
    Bn, H, W, cin, cout = 8, 128, 128, 32, 32
    P, K = Bn * H * W, 9 * cin

    def buf(shape, dt="bf16"):
        return U.Buf(shape, dt).upload(rng.standard_normal(shape).astype(np.float32))

    x, col, w, dy = buf((P, cin)), buf((P, K)), buf((K, cout)), buf((P, cout))
    y, dcol, gw, bias = U.Buf((P, cout)), U.Buf((P, K)), U.Buf((K, cout), "f32"), U.Buf((cout,), "f32")
    wf, dx = U.Buf((9 * cout, cin)), U.Buf((P, cin))

    def dgrad_implicit():
        ps.core.check(U.lib.ps_nn_conv_flip(w.p, cin, cout, wf.p, None))
        U.gemm_im2col(0, P, cin, 9 * cout, dy, H, W, cout, wf, cin, 1, dx, cin)

    a4, b4, c4 = buf((4096, 4096)), buf((4096, 4096)), U.Buf((4096, 4096), "f32")
    sk = U._splitk(K, cout, P)
    cases = [
        ("conv 32->32 forward, im2col inside the GEMM (x . W)", P, cout, K,
         lambda: U.gemm_im2col(0, P, cout, K, x, H, W, cin, w, cout, 1, y, cout, bias, U.EPI_BIAS_RELU)),
        ("conv 32->32 weight gradient, im2col inside the GEMM (col(x)^T . dY)", K, cout, P,
         lambda: U.gemm_im2col(1, K, cout, P, x, H, W, cin, dy, cout, 1, gw, cout, epi=U.EPI_F32_ADD, splitk=sk)),
        ("conv 32->32 data gradient, im2col of dY inside the GEMM (ps_nn_conv_flip + col(dY) . Wf)", P, cin, 9 * cout, dgrad_implicit),
        ("im2col of x (the column matrix, 9x the activations)", 0, 0, 0,
         lambda: ps.core.check(U.lib.ps_nn_im2col(x.p, Bn, H, W, cin, col.p, None))),
        ("conv 32->32 forward on the column matrix (col . W)", P, cout, K,
         lambda: U.gemm(P, cout, K, col, K, 1, w, cout, 1, y, cout, bias=bias, epi=U.EPI_BIAS_RELU)),
        ("conv 32->32 weight gradient on the column matrix (col^T . dY)", K, cout, P,
         lambda: U.gemm(K, cout, P, col, 1, K, dy, cout, 1, gw, cout, epi=U.EPI_F32_ADD, splitk=sk)),
        ("conv 32->32 data gradient into the column matrix (dY . W^T)", P, K, cout,
         lambda: U.gemm(P, K, cout, dy, cout, 1, w, 1, cout, dcol, K, epi=U.EPI_BF16)),
        ("col2im of the column gradient", 0, 0, 0, lambda: ps.core.check(U.lib.ps_nn_col2im(dcol.p, Bn, H, W, cin, dx.p, None))),
        ("square 4096^3", 4096, 4096, 4096, lambda: U.gemm(4096, 4096, 4096, a4, 4096, 1, b4, 4096, 1, c4, 4096, epi=U.EPI_F32)),
    ]
    res = []
    for name, M, N, Kd, fn in cases:
        for _ in range(3):
            fn()
        ps.synchronize()
        reps = 20
        e0, e1 = ps.Event(), ps.Event()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        ps.synchronize()
        ms = e0.elapsed_ms(e1) / reps
        if M == 0:  # a data-movement kernel: HBM bytes, the activations once and the column matrix
            gbs = (1 + 9) * P * cin * 2 / (ms * 1e-3) / 1e9
            res.append({"case": name, "ms": round(ms, 4), "GB/s": round(gbs, 1)})
            log(f"{name}: {ms:.4f} ms, {gbs:.0f} GB/s")
            continue
        tf = 2.0 * M * N * Kd / (ms * 1e-3) / 1e12
        res.append({"case": name, "M": M, "N": N, "K": Kd, "ms": round(ms, 4), "tflops": round(tf, 2), "frac_of_peak": round(tf / peak, 4)})
        log(f"{name}: M={M} N={N} K={Kd}: {ms:.4f} ms, {tf:.1f} TFLOP/s ({100 * tf / peak:.2f}% of BF16 MFMA peak)")
    ms = [r["ms"] for r in res]
    implicit, explicit = ms[0] + ms[1] + ms[2], sum(ms[3:8])
    log(f"forward + weight gradient + data gradient of the layer: {implicit:.4f} ms with im2col inside the GEMM, "
        f"{explicit:.4f} ms through the column matrix (im2col, col2im and the products)")
    out = {"metric": "BF16 MFMA GEMM TFLOP/s (U-Net convolution products)", "value": res[0]["tflops"], "unit": "TFLOP/s",
           "n_gpus": 1, "steps": 20, "warmup": 3, "ms_per_step": res[0]["ms"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
           "config": {"workload": "ps_nn_gemm_im2col / ps_nn_gemm on the U-Net's 3x3 convolution products (batch 8 x 128^2, 32->32 channels) "
                                  "and a square 4096^3 product"},
           "roofline": {"bound": "mfma", "achieved": res[0]["tflops"], "peak": peak, "unit": "TFLOP/s",
                        "frac": res[0]["frac_of_peak"], "traffic": None},
           "cases": res}
    print(json.dumps(out), flush=True)


def unet_study(args):
    """The warm-start U-Net on the GPU (4 levels, 32 base channels, inputs 2 channels, BF16
    storage with FP32 accumulation): the forward pass at 512^2 (batch 1 and 8) and a training
    step at batch 8 x 128^2 (forward, backward, Adam; the backward counted as twice the forward
    products, less the first layer's data gradient), HIP-event timed. FLOPs are those of the
    convolution products, 2 B H W K Cout per layer at its level's resolution; reported against
    the BF16 dense MFMA peak."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import unet as U

    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module="bench.unet")
    peak = 2500.0
    net = U.UNet(base=32, cin=2, seed=0)
    rng = np.random.default_rng(0)

    def conv_flops(B, H, W):
        f = 0
        for l in range(4):
            for cv in net.enc[l]:
                f += 2 * B * (H >> l) * (W >> l) * cv.K * cv.cout
        for l in range(3):
            for cv in net.dec[l]:
                f += 2 * B * (H >> l) * (W >> l) * cv.K * cv.cout
        return f + 2 * B * H * W * net.head.K * net.head.cout

    def timed(fn, reps):
        for _ in range(3):
            fn()
        ps.synchronize()
        e0, e1 = ps.Event(), ps.Event()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        ps.synchronize()
        return e0.elapsed_ms(e1) / reps

    runs = []
    for B, H in ((1, 512), (8, 512)):
        x = U.Buf((B, H, H, 2)).upload(rng.standard_normal((B, H, H, 2)).astype(np.float32))
        ms = timed(lambda: net.forward(x, B, H, H), 10)
        fl = conv_flops(B, H, H)
        tf = fl / (ms * 1e-3) / 1e12
        runs.append({"case": f"forward {B} x {H}^2", "ms": round(ms, 3), "GFLOP": round(fl / 1e9, 2), "tflops": round(tf, 1), "frac_of_peak": round(tf / peak, 4)})
        log(f"U-Net forward, batch {B} x {H}^2: {ms:.3f} ms, {fl / 1e9:.1f} GFLOP, {tf:.1f} TFLOP/s ({100 * tf / peak:.2f}% of BF16 MFMA peak)")
    B, H = 8, 128
    x = U.Buf((B, H, H, 2)).upload(rng.standard_normal((B, H, H, 2)).astype(np.float32))
    dout = U.Buf((B, H, H, 3)).upload(0.01 * rng.standard_normal((B, H, H, 3)).astype(np.float32))

    def step():
        net.zero_grad()
        net.forward(x, B, H, H)
        net.backward(dout)
        net.adam(1e-4)

    ms = timed(step, 10)
    f1 = conv_flops(B, H, H)
    fl = 3 * f1 - 2 * B * H * H * net.enc[0][0].K * net.enc[0][0].cout
    tf = fl / (ms * 1e-3) / 1e12
    runs.append({"case": f"training step {B} x {H}^2", "ms": round(ms, 3), "GFLOP": round(fl / 1e9, 2), "tflops": round(tf, 1), "frac_of_peak": round(tf / peak, 4)})
    log(f"U-Net training step, batch {B} x {H}^2: {ms:.3f} ms, {fl / 1e9:.1f} GFLOP, {tf:.1f} TFLOP/s ({100 * tf / peak:.2f}% of BF16 MFMA peak)")
    head = runs[0]
    out = {"metric": "U-Net forward TFLOP/s (warm-start network, 512^2 sample)", "value": head["tflops"], "unit": "TFLOP/s", "n_gpus": 1,
           "steps": 10, "warmup": 3, "ms_per_step": head["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
           "config": {"workload": "U-Net 4 levels, 32 base channels, 3x3 convolutions as implicit-GEMM MFMA products, inputs 2 x 512^2"},
           "roofline": {"bound": "mfma", "achieved": head["tflops"], "peak": peak, "unit": "TFLOP/s", "frac": head["frac_of_peak"], "traffic": None},
           "runs": runs}
    print(json.dumps(out), flush=True)


def stokes_batch_study(args):
    """Stokes2D PT iterations of S samples at once (Stokes2DBatch: the two fused launches over
    (n, n, S) arrays) against one sample at a time (Stokes2DFused), at 127^2 (the training
    resolution of the warm start) and 511^2: graph replays of 200-iteration blocks with their
    per-sample residual reductions, HIP-event timed over 1000 iterations. Reported per batch
    iteration and per sample-iteration; the effective rate counts the 27 field sweeps of n^2
    doubles that one fused iteration makes (10 by compute_Pτ!, 17 by compute_dVV!). Before the
    timing, each batch's first plane is checked bit for bit against the single-sample solver
    after the same iterations."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import stokes_fused as F
    from parallelstencil import workloads as W

    mod = "bench.stokesbatch"
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=mod)
    nblk = 5
    runs = []
    for n, sizes in ((127, (1, 4, 16, 64)), (511, (1, 2, 4, 8, 16))):
        mats = [W.stokes2d_inclusions(n, n, 7000 + s, log10_contrast=1.0) for s in range(max(sizes))]
        one = F.Stokes2DFused(mod, n, n)
        one.set_material(*mats[0])
        one.reset()
        one.iterate_block()
        ps.synchronize()
        ref_vx = one.Vx.to_host()
        e0, e1 = ps.Event(), ps.Event()
        e0.record()
        for _ in range(nblk):
            one.iterate_block()
        e1.record()
        ps.synchronize()
        t1 = e0.elapsed_ms(e1) / (nblk * one.nout) * 1e3
        one.close()
        runs.append({"n": n, "samples": 0, "solver": "Stokes2DFused", "us_per_iteration": round(t1, 3), "us_per_sample_iteration": round(t1, 3),
                     "eff_TBs": round(27 * 8 * n * n / (t1 * 1e-6) / 1e12, 3)})
        log(f"{n}^2 one sample (Stokes2DFused): {t1:.2f} us per iteration, {27 * 8 * n * n / (t1 * 1e-6) / 1e12:.2f} TB/s effective")
        for S in sizes:
            B = F.Stokes2DBatch(mod, n, n, S)
            B.set_material(np.asfortranarray(np.stack([m for m, _ in mats[:S]], axis=2)),
                           np.asfortranarray(np.stack([r for _, r in mats[:S]], axis=2)))
            B.reset()
            B.iterate_block()
            ps.synchronize()
            same = bool(np.array_equal(np.asarray(B.Vx.to_host())[:, :, 0].view(np.uint64), np.asarray(ref_vx).view(np.uint64)))
            e0.record()
            for _ in range(nblk):
                B.iterate_block()
            e1.record()
            ps.synchronize()
            t = e0.elapsed_ms(e1) / (nblk * B.nout) * 1e3
            B.close()
            del B
            eff = 27 * 8 * n * n * S / (t * 1e-6) / 1e12
            runs.append({"n": n, "samples": S, "solver": "Stokes2DBatch", "us_per_iteration": round(t, 3), "us_per_sample_iteration": round(t / S, 3),
                         "speedup_per_sample": round(t1 / (t / S), 2), "eff_TBs": round(eff, 3), "plane0_bit_identical": same})
            log(f"{n}^2 batch of {S}: {t:.2f} us per iteration, {t / S:.3f} us per sample-iteration ({t1 / (t / S):.2f}x one sample at a time), "
                f"{eff:.2f} TB/s effective; first plane bit-identical to the single solver: {same}")
            if not same:
                raise SystemExit("batched Stokes2D differs from the single-sample solver")
    best = max((r for r in runs if r["samples"] and r["n"] == 127), key=lambda r: r["speedup_per_sample"])
    out = {"metric": "Stokes2D PT sample-iterations per second (batched samples)", "value": round(1e6 / best["us_per_sample_iteration"], 1),
           "unit": "sample-iterations/s", "n_gpus": 1, "steps": nblk * 200, "warmup": 200, "ms_per_step": round(best["us_per_iteration"] / 1e3, 5),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (inclusion samples, contrast 10^[-1,1])",
           "config": {"workload": f"Stokes2D PT iteration, {best['samples']} samples of 127^2 per launch pair, graph replays"},
           "roofline": {"bound": "hbm", "achieved": round(best["eff_TBs"] * 1e3, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(best["eff_TBs"] * 1e3 / HBM_PEAK_GBS, 4), "traffic": None},
           "runs": runs}
    print(json.dumps(out), flush=True)


def shmem_study(args):
    """diffusion2D written with the kernel language (examples/diffusion2D_shmem_novis.jl: an
    @sharedMem tile with halo and @sync_threads) at the example's 16384^2, against the same update
    written with FiniteDifferences2D macros (plain and memopt) and a 2-D triad over the same three
    arrays. T_eff per step uses the example's A_eff = 3 arrays x 8 B x n^2; the fields of the
    kernel-language and macro kernels are compared bit for bit."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import workloads as W

    n, K, Wm = args.shmem_n, args.steps, args.warmup
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module="bench.shmem")
    mod = "bench.shmem"
    kl, kli = W.define_diffusion2d_shmem(mod), W.define_diffusion2d_shmem(mod, inbounds=True)
    fd, fdm = W.define_diffusion2d_fd(mod), W.define_diffusion2d_fd(mod, memopt=True)
    tri = ps.core._define("parallel", "@parallel function copy2D!(T2, T, Ci)\n    @all(T2) = @all(T) + @all(Ci)\n    return\nend", {}, mod)
    lx = ly = 1.0
    dx, dy = lx / (n - 1), ly / (n - 1)
    lam, c0 = 1.0, 2.0
    _dx, _dy = 1.0 / dx, 1.0 / dy
    dt = min(dx * dx, dy * dy) / lam / (1.0 / c0) / 4.1
    bytes_step = 3 * 8 * n * n
    rng = np.random.default_rng(1)
    T0 = np.asfortranarray(1.7 + 0.1 * rng.random((n, n)))
    Ci = ps.fill(1.0 / c0, n, n)
    runs = {}
    finals = {}
    cases = [("kernel language, threads (32,8) (the example)", "kl", (32, 8)), ("kernel language, threads (64,4)", "kl", (64, 4)),
             ("kernel language, threads (128,2)", "kl", (128, 2)), ("kernel language inbounds=true, threads (32,8)", "kli", (32, 8)),
             ("kernel language inbounds=true, threads (128,2)", "kli", (128, 2)), ("FiniteDifferences2D", "fd", None),
             ("FiniteDifferences2D memopt", "fdm", None), ("triad", "tri", None)]
    for name, kind, th in cases:
        T, T2 = ps.zeros(n, n), ps.zeros(n, n)
        T.copy_from(T0)
        T2.copy_from(T0)
        st = [T, T2]
        if kind in ("kl", "kli"):
            kw = dict(nblocks=(n // th[0], n // th[1]), nthreads=th, shmem=(th[0] + 2) * (th[1] + 2) * 8)
            k = kl if kind == "kl" else kli
            launch = lambda i, k=k, kw=kw: (k(st[1], st[0], Ci, lam, dt, _dx, _dy, async_=True, **kw), st.reverse())
        elif kind in ("fd", "fdm"):
            k = fd if kind == "fd" else fdm
            launch = lambda i, k=k: (k(st[1], st[0], Ci, lam, dt, _dx, _dy, async_=True), st.reverse())
        else:
            launch = lambda i: tri(st[1], st[0], Ci, async_=True)
        wall, ms = time_kernel(ps, launch, K, Wm)
        runs[name] = {"ms_per_step": round(ms, 4), "T_eff_gbs": round(bytes_step / ms / 1e6, 1)}
        log(f"diffusion2D {n}^2 {name}: {ms:.4f} ms/step, T_eff {runs[name]['T_eff_gbs']} GB/s")
        if kind != "tri":
            finals[name] = st[0].to_host()
        del T, T2, st
    names = [c[0] for c in cases if c[1] != "tri"]
    same = all(np.array_equal(finals[names[0]], finals[m]) for m in names[1:])
    log(f"kernel-language and FiniteDifferences2D fields after {K + Wm} steps bit-identical: {same}")
    ex = runs[cases[0][0]]
    out = {"metric": "T_eff GB/s (diffusion2D, kernel language: @sharedMem + @sync_threads)", "value": ex["T_eff_gbs"], "unit": "GB/s",
           "n_gpus": 1, "steps": K, "warmup": Wm, "ms_per_step": ex["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"examples/diffusion2D_shmem_novis.jl at {n}^2, threads (32,8), shmem 34*10*8 B"},
           "roofline": {"bound": "hbm", "achieved": ex["T_eff_gbs"], "peak": 8000.0, "unit": "GB/s",
                        "frac": round(ex["T_eff_gbs"] / 8000.0, 4), "traffic": None},
           "fields_bit_identical": bool(same), "runs": runs}
    print(json.dumps(out), flush=True)


def cellarray_study(args):
    """examples/memcopyCellArray3D.jl: T2 = T + Ci over CellArrays of 4x4 FP64 cells at
    128 x 128 x 1024 cells, in the example's three kernel forms (FiniteDifferences3D @all,
    @parallel_indices over whole cells, per-field loops with field(A, i, j)), with each cell field
    stored contiguously (blocklength 0, the GPU default) and with cells stored whole (blocklength
    1); against a triad over plain fields of the same byte volume. T_eff uses the example's
    A_eff = 3 x 8 B x cells x 16."""
    import parallelstencil as ps

    nx, ny, nz = 128, 128, 1024
    K, Wm = args.steps, args.warmup
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module="bench.cells")
    mod = "bench.cells"
    d = lambda kind, text: ps.core._define(kind, text, {}, mod)
    kfd = d("parallel", "@parallel function copy3D!(T2, T, Ci)\n    @all(T2) = @all(T) + @all(Ci)\n    return\nend")
    kex = d("parallel_indices", "@parallel_indices (ix,iy,iz) function copy3D_explicit!(T2, T, Ci)\n"
                                "    T2[ix,iy,iz] = T[ix,iy,iz] + Ci[ix,iy,iz]\n    return\nend")
    klg = d("parallel_indices", "@parallel_indices (ix,iy,iz) function copy3D_large_cells!(T2, T, Ci)\n"
                                "    for cy = 1:cellsize(T2,2)\n        for cx = 1:cellsize(T2,1)\n"
                                "            field(T2,cx,cy)[ix,iy,iz] = field(T,cx,cy)[ix,iy,iz] + field(Ci,cx,cy)[ix,iy,iz]\n"
                                "        end\n    end\n    return\nend")
    tri = d("parallel", "@parallel function copy3D!(T2, T, Ci)\n    @all(T2) = @all(T) + @all(Ci)\n    return\nend")
    bytes_step = 3 * 8 * nx * ny * nz * 16
    runs = {}
    for bl in (0, 1):
        T = ps.fill(1.7, nx, ny, nz, celldims=(4, 4), blocklength=bl)
        T2 = ps.fill(1.7, nx, ny, nz, celldims=(4, 4), blocklength=bl)
        Ci = ps.fill(0.5, nx, ny, nz, celldims=(4, 4), blocklength=bl)
        for name, k in (("@all (FiniteDifferences3D)", kfd), ("whole cells (@parallel_indices)", kex), ("field(A,i,j) loops", klg)):
            st = [T, T2]
            launch = lambda i, k=k: (k(st[1], st[0], Ci, async_=True), st.reverse())
            wall, ms = time_kernel(ps, launch, K, Wm)
            key = f"blocklength {bl}: {name}"
            runs[key] = {"ms_per_step": round(ms, 4), "T_eff_gbs": round(bytes_step / ms / 1e6, 1)}
            log(f"memcopyCellArray3D {nx}x{ny}x{nz} 4x4 cells, {key}: {ms:.4f} ms/step, T_eff {runs[key]['T_eff_gbs']} GB/s")
        del T, T2, Ci
    A, B, C = (ps.fill(1.0, nx, ny, nz * 16) for _ in range(3))
    st = [A, B]
    wall, ms = time_kernel(ps, lambda i: tri(st[1], st[0], C, async_=True), K, Wm)
    runs["triad over plain fields of the same bytes"] = {"ms_per_step": round(ms, 4), "T_eff_gbs": round(bytes_step / ms / 1e6, 1)}
    log(f"triad {nx}x{ny}x{nz * 16} plain: {ms:.4f} ms/step, T_eff {runs['triad over plain fields of the same bytes']['T_eff_gbs']} GB/s")
    main_key = "blocklength 0: @all (FiniteDifferences3D)"
    out = {"metric": "T_eff GB/s (memcopyCellArray3D, 4x4 FP64 cells)", "value": runs[main_key]["T_eff_gbs"], "unit": "GB/s", "n_gpus": 1,
           "steps": K, "warmup": Wm, "ms_per_step": runs[main_key]["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"examples/memcopyCellArray3D.jl at {nx}x{ny}x{nz} cells of 4x4"},
           "roofline": {"bound": "hbm", "achieved": runs[main_key]["T_eff_gbs"], "peak": 8000.0, "unit": "GB/s",
                        "frac": round(runs[main_key]["T_eff_gbs"] / 8000.0, 4), "traffic": None}, "runs": runs}
    print(json.dumps(out), flush=True)


def ad_study(args):
    """Derivative kernels of the diffusion step at n^3 FP64 (@parallel ∇=(T2->T̄2, T->T̄)):
    the primal (memopt) beside the forward-mode (tangent) and reverse-mode (adjoint) kernels,
    launched directly. Algorithmic bytes per interior point: primal and tangent read two fields
    and write one (24 B); the adjoint reads the seed T̄2 and Cp, clears the seed and adds into
    T̄ (read + write), 40 B. The adjoint runs in its gather form (each point sums the adjoints
    of its 7 stencil neighbours, then a second kernel clears the seeds) and in the atomic
    scatter form (each point adds into its 7 neighbours)."""
    import parallelstencil as ps
    from parallelstencil import workloads as W

    n, K, Wm = args.n, args.steps, args.warmup
    mod = "bench.ad"
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=mod)
    k = W.define_diffusion3d(True, mod)
    hT, hCp, c = W.diffusion_inputs(n)
    T, T2, Cp = (ps.zeros(n, n, n).copy_from(h) for h in (hT, hT, hCp))
    del hT, hCp
    sc = (c["lam"], c["dt"], c["dx"], c["dy"], c["dz"])
    dup = {"T2": "DuplicatedNoNeed", "T": "DuplicatedNoNeed"}
    tan, adj = k._derivative("Forward", dup), k._derivative("Reverse", dup)
    S1, S2 = ps.fill(0.0, n, n, n), ps.fill(1e-3, n, n, n)
    runs = {}
    pts = (n - 2) ** 3

    def scatter(i):
        adj.use_gather = False
        adj(T2, T, Cp, *sc, S2, S1, async_=True)
        adj.use_gather = True

    for name, launch, bpp in (
            ("primal (memopt)", lambda i: k(T2, T, Cp, *sc, async_=True), 24),
            ("forward mode (tangent)", lambda i: tan(T2, T, Cp, *sc, S1, S2, async_=True), 24),
            ("reverse mode (adjoint, gather + clear)", lambda i: adj(T2, T, Cp, *sc, S2, S1, async_=True), 40),
            ("reverse mode (adjoint, atomic scatter)", scatter, 40)):
        wall, ms = time_kernel(ps, launch, K, Wm)
        runs[name] = {"ms_per_launch": round(ms, 4), "bytes_per_point": bpp, "gbs": round(bpp * pts / ms / 1e6, 1)}
        log(f"diffusion3D {n}^3 {name}: {ms:.4f} ms/launch, {runs[name]['gbs']} GB/s algorithmic")
    head = runs["reverse mode (adjoint, gather + clear)"]
    del T, T2, Cp, S1, S2
    # acoustic2D (miniapps/acoustic2D.jl with a bulk-modulus field, as in the inversion): one
    # adjoint step by the hand-written transposed-stencil kernels and by the derivative kernels
    # of the forward kernels (compute_PK! then compute_V! in reverse)
    from parallelstencil import acoustic_adjoint as AA
    m2 = "bench.ad2"
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=m2)
    d2 = lambda text: ps.core._define("parallel", text, {}, m2)
    kv, kp, kgk, kav, kap = d2(AA.FWD_V), d2(AA.FWD_P), d2(AA.GRAD_K), d2(AA.ADJ_V), d2(AA.ADJ_P)
    na = 2048
    P, Kf, lP, gK = (ps.fill(v, na, na) for v in (0.1, 1.0, 1e-3, 0.0))
    Vx, lVx = ps.fill(0.2, na + 1, na), ps.fill(1e-3, na + 1, na)
    Vy, lVy = ps.fill(0.3, na, na + 1), ps.fill(1e-3, na, na + 1)
    sa = (1e-3, 0.5, 0.5)
    dp = kp._derivative("Reverse", {"P": "DuplicatedNoNeed", "Vx": "DuplicatedNoNeed", "Vy": "DuplicatedNoNeed", "K": "DuplicatedNoNeed"})
    dv = kv._derivative("Reverse", {"Vx": "DuplicatedNoNeed", "Vy": "DuplicatedNoNeed", "P": "DuplicatedNoNeed"})

    def hand(i):
        kgk(gK, lP, Vx, Vy, *sa, async_=True)
        kav(lVx, lVy, lP, Kf, *sa, async_=True)
        kap(lP, lVx, lVy, sa[0], 1.0, sa[1], sa[2], async_=True)

    def generated(gather):
        def run(i):
            dp.use_gather = dv.use_gather = gather
            dp(P, Vx, Vy, Kf, *sa, lP, lVx, lVy, gK, async_=True)
            dv(Vx, Vy, P, sa[0], 1.0, sa[1], sa[2], lVx, lVy, lP, async_=True)
            dp.use_gather = dv.use_gather = True
        return run

    for name, launch in (("acoustic2D 2048^2 adjoint step, hand-written (grad_K!, adjoint_V!, adjoint_P!)", hand),
                         ("acoustic2D 2048^2 adjoint step, derivative kernels (gather form)", generated(True)),
                         ("acoustic2D 2048^2 adjoint step, derivative kernels (atomic scatter form)", generated(False))):
        wall, ms = time_kernel(ps, launch, args.steps, Wm)
        runs[name] = {"ms_per_step": round(ms, 4)}
        log(f"{name}: {ms:.4f} ms/step")
    out = {"metric": "GB/s algorithmic (reverse-mode derivative of diffusion3D)", "value": head["gbs"], "unit": "GB/s", "n_gpus": 1,
           "steps": K, "warmup": Wm, "ms_per_step": head["ms_per_launch"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"@parallel ∇=(T2->T̄2, T->T̄) diffusion3D_step! at {n}^3"},
           "roofline": {"bound": "hbm", "achieved": head["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(head["gbs"] / HBM_PEAK_GBS, 4), "traffic": None}, "runs": runs}
    print(json.dumps(out), flush=True)


def stokes3d_study(args):
    """miniapps/Stokes3D.jl at its own configuration (127^3, inclusion viscosity 0.1, tolerance
    1e-6 checked every 500 iterations): time per PT iteration, eager and as graph replays, T_eff
    with the miniapp's A_eff = 8 n^3 x 8 B per iteration (:160), and the time to tolerance."""
    import parallelstencil as ps
    from parallelstencil import workloads as W
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3)
    n = args.stokes3d_n
    Mus, Rog = W.stokes3d_material(n, n, n)
    a_eff = 8 * 8 * n ** 3
    runs = {}
    for graph in [m == "graph" for m in args.stokes3d_modes.split(",")]:
        S = W.Stokes3D(__name__, n, n, n, graph=graph)
        S.set_material(Mus, Rog)
        S.reset()
        S.iterate_block()
        ps.synchronize()
        t0 = time.perf_counter()
        S.iterate_block()
        ps.synchronize()
        ms_it = (time.perf_counter() - t0) * 1e3 / S.nout
        it, err, t_tol = None, None, None
        if args.stokes3d_solve:
            S.reset()
            ps.synchronize()
            t0 = time.perf_counter()
            it, err = S.solve()
            ps.synchronize()
            t_tol = round(time.perf_counter() - t0, 4)
        key = "graph" if graph else "eager"
        runs[key] = {"ms_per_iteration": round(ms_it, 5), "T_eff_gbs": round(a_eff / (ms_it * 1e-3) / 1e9, 2),
                     "iterations": it, "error": err, "s_to_tol": t_tol}
        log(f"stokes3d {n}^3 {key}: {ms_it:.4f} ms/iteration, T_eff {a_eff / ms_it / 1e6:.1f} GB/s, "
            f"{it} iterations to err {err} in {t_tol} s")
        del S
    head = runs.get("graph", runs.get("eager"))
    out = {"metric": "Stokes3D PT iteration T_eff GB/s (miniapp A_eff = 8 n^3 x 8 B per iteration)", "value": head["T_eff_gbs"],
           "unit": "GB/s", "n_gpus": 1, "steps": 500, "warmup": 500, "ms_per_step": head["ms_per_iteration"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"Stokes3D {n}^3 FP64, spherical inclusion (viscosity 0.1, density 1), eps 1e-6, nout 500"},
           "roofline": {"bound": "hbm", "achieved": head["T_eff_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(head["T_eff_gbs"] / HBM_PEAK_GBS, 4), "traffic": None}, "runs": runs}
    print(json.dumps(out), flush=True)


def thermal_study(args):
    """miniapps/ThermalConvection2D.jl at its own configuration ((96 ar - 1) x 95, Ra 1e7, PT
    tolerance 1e-4 checked every 100 iterations): PT iterations per second (graph replays of 100
    iterations, 10 launches each) and the wall time of the first time steps."""
    import parallelstencil as ps
    from parallelstencil import workloads as W
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2)
    S = W.ThermalConvection2D(__name__, graph=bool(args.graph))
    S._block()
    ps.synchronize()
    nb = 20
    t0 = time.perf_counter()
    for _ in range(nb):
        S._block()
    ps.synchronize()
    ms_it = (time.perf_counter() - t0) * 1e3 / (nb * S.nerr)
    S.reset()
    ps.synchronize()
    steps = []
    for _ in range(args.thermal_steps):
        t0 = time.perf_counter()
        niter, errV, errP, dt = S.time_step()
        ps.synchronize()
        steps.append({"pt_iterations": niter, "errV": errV, "errP": errP, "dt": dt, "s": round(time.perf_counter() - t0, 4)})
        log(f"thermal step: {niter} PT iterations, errV {errV:.2e} errP {errP:.2e}, dt {dt:.3e}, {steps[-1]['s']} s")
    out = {"metric": "ThermalConvection2D PT iterations per second", "value": round(1e3 / ms_it, 1), "unit": "iterations/s",
           "n_gpus": 1, "steps": nb * S.nerr, "warmup": S.nerr, "ms_per_step": round(ms_it, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"ThermalConvection2D {S.nx}x{S.ny} FP64, Ra 1e7, eps 1e-4, nerr 100, " + ("graph replays" if args.graph else "eager launches")},
           "time_steps": steps}
    print(json.dumps(out), flush=True)


def porowaves_study(args):
    """miniapps/scalar_porowaves2D.jl at its own configuration (256 x 512, blob setup): time per
    time step (four launches; the time step is refreshed from a GPU maximum every 100 steps)."""
    import parallelstencil as ps
    from parallelstencil import workloads as W
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2)
    S = W.Porowaves2D(__name__)
    S.step(10)
    ps.synchronize()
    K = args.steps * 10
    t0 = time.perf_counter()
    S.step(K)
    ps.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / K
    log(f"porowaves {S.nx}x{S.ny}: {ms:.4f} ms per time step, t = {S.t:.4e}")
    out = {"metric": "scalar porosity waves time steps per second", "value": round(1e3 / ms, 1), "unit": "steps/s", "n_gpus": 1,
           "steps": K, "warmup": 10, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic", "config": {"workload": f"scalar_porowaves2D {S.nx}x{S.ny} FP64, blob setup"}}
    print(json.dumps(out), flush=True)


def hydromech_study(args):
    """miniapps/HydroMech2D.jl at its own configuration (res 128: 127 x 255): time per PT
    iteration (graph replays of 200 iterations, 11 launches each), T_eff against the miniapp's
    A_eff = 16 n_x n_y x 8 B per iteration (:186), and the first time steps to tolerance."""
    import parallelstencil as ps
    from parallelstencil import workloads as W
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2)
    S = W.HydroMech2D(__name__, 128, graph=bool(args.graph))
    S._block()
    ps.synchronize()
    nb = 10
    t0 = time.perf_counter()
    for _ in range(nb):
        S._block()
    ps.synchronize()
    ms_it = (time.perf_counter() - t0) * 1e3 / (nb * S.nout)
    a_eff = 16 * S.nx * S.ny * 8
    S.reset()
    ps.synchronize()
    steps = []
    for _ in range(args.thermal_steps):
        t0 = time.perf_counter()
        niter, err, dt = S.time_step()
        ps.synchronize()
        steps.append({"pt_iterations": niter, "error": err, "dt": dt, "s": round(time.perf_counter() - t0, 4)})
        log(f"hydromech step: {niter} PT iterations, err {err:.2e}, dt {dt:.3e}, {steps[-1]['s']} s")
    out = {"metric": "HydroMech2D PT iteration T_eff GB/s (miniapp A_eff = 16 n_x n_y x 8 B per iteration)",
           "value": round(a_eff / (ms_it * 1e-3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1, "steps": nb * S.nout, "warmup": S.nout,
           "ms_per_step": round(ms_it, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "config": {"workload": f"HydroMech2D {S.nx}x{S.ny} FP64, eps 1e-5, nout 200, " + ("graph replays" if args.graph else "eager launches")},
           "time_steps": steps}
    print(json.dumps(out), flush=True)


def hidecomm_study(args):
    """examples/diffusion3D_multigpucpu_hidecomm_novis.jl on one GPU at n^3 FP64 (periodic in
    every dimension, so that update_halo! exchanges with the process itself): the example's loop
    (@hide_communication around the step and update_halo!) beside the same step and halo update
    in sequence, and the step alone. T_eff is the example's (A_eff = 3 x 8 B x n^3 per step, 100
    steps of which the first 10 are not timed)."""
    import numpy as np
    import parallelstencil as ps
    from parallelstencil import igg

    n, nt = args.n, 100
    mod = "bench.hidecomm"
    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3, module=mod)
    text = ("@parallel function diffusion3D_step!(T2, T, Ci, lam, dt, _dx, _dy, _dz)\n"
            "    @inn(T2) = @inn(T) + dt*(lam*@inn(Ci)*(@d2_xi(T)*_dx^2 + @d2_yi(T)*_dy^2 + @d2_zi(T)*_dz^2));\n"
            "    return\nend")
    igg.init_global_grid(n, n, n, periodx=1, periody=1, periodz=1)
    dx = dy = dz = 1.0 / (igg.nx_g() - 1)
    lam, c0 = 1.0, 2.0
    dt = min(dx ** 2, dy ** 2, dz ** 2) / lam / (1 / c0) / 8.1
    sc = (lam, dt, 1 / dx, 1 / dy, 1 / dz)
    a_eff = 3 * 8 * n ** 3
    runs = {}
    for memopt in (False, True):
        k = ps.core._define("parallel", text, {"memopt": memopt}, mod)
        T, T2, Ci = ps.fill(1.7, n, n, n), ps.fill(1.7, n, n, n), ps.fill(1 / c0, n, n, n)
        variants = (
            ("step only", lambda: k(T2, T, Ci, *sc, async_=True)),
            ("step, then update_halo!", lambda: (k(T2, T, Ci, *sc, async_=True), igg.update_halo_(T2, async_=True))),
            ("@hide_communication (16,4,4) step + update_halo!",
             lambda: ps.hide_communication((16, 4, 4), [(k, (T2, T, Ci) + sc)], lambda: igg.update_halo_(T2))))
        for name, step in variants:
            key = f"{'memopt' if memopt else 'plain'}: {name}"
            try:
                for it in range(nt):
                    if it == 10:
                        ps.synchronize()
                        t0 = time.perf_counter()
                    step()
                    T, T2 = T2, T
            except ps.ArgumentError as e:  # memopt launches need ranges divisible by the block (as in the reference)
                ps.synchronize()
                runs[key] = {"error": str(e)}
                log(f"hidecomm example {n}^3 {key}: {e}")
                continue
            ps.synchronize()
            t_it = (time.perf_counter() - t0) / (nt - 10)
            runs[key] = {"ms_per_step": round(1e3 * t_it, 4), "T_eff_gbs": round(a_eff / t_it / 1e9, 1)}
            log(f"hidecomm example {n}^3 {key}: {1e3 * t_it:.4f} ms/step, T_eff {runs[key]['T_eff_gbs']} GB/s")
        del T, T2, Ci
    igg.finalize_global_grid()
    head = runs["plain: @hide_communication (16,4,4) step + update_halo!"]
    out = {"metric": "T_eff GB/s (diffusion3D_multigpucpu_hidecomm_novis, one GPU, periodic halos)", "value": head["T_eff_gbs"],
           "unit": "GB/s", "n_gpus": 1, "steps": nt - 10, "warmup": 10, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"examples/diffusion3D_multigpucpu_hidecomm_novis.jl at {n}^3, periodic in x, y, z"},
           "roofline": {"bound": "hbm", "achieved": head["T_eff_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(head["T_eff_gbs"] / HBM_PEAK_GBS, 4), "traffic": None}, "runs": runs}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="", help="memopt | plain | triad (profiling)")
    ap.add_argument("--workload", default="diffusion", help="diffusion (default) | stokes (learned warm start) | acoustic | xgroups | adjoint | gemm | shmem | cellarray | ad | hidecomm | stokes-batch | unet | stokes3d | thermal | porowaves | hydromech")
    ap.add_argument("--acoustic-n", type=int, default=384)
    ap.add_argument("--acoustic-steps", type=int, default=1000)
    ap.add_argument("--acoustic-exact-only", action="store_true", help="acoustic: skip the fastmath variants")
    ap.add_argument("--adjoint-n", type=int, default=2048)
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--stokes-n", type=int, default=511)
    ap.add_argument("--stokes-n2", type=int, default=1023, help="stokes: second evaluation resolution (0: none)")
    ap.add_argument("--stokes-train-n", type=int, default=0, help="stokes: training resolution (0: --stokes-n)")
    ap.add_argument("--stokes-contrast", type=float, default=2.0, help="stokes: log10 viscosity contrast range [-c, c]")
    ap.add_argument("--stokes-ts", default="maxV", help="stokes: pseudo-time steps 'av' (miniapp) or 'maxV'")
    ap.add_argument("--stokes-out", default="", help="stokes: JSON file for per-sample results and residual histories")
    ap.add_argument("--stokes-save", default="", help="stokes: write the trained warm start (.npz)")
    ap.add_argument("--stokes-load", default="", help="stokes: evaluate the warm start in this .npz instead of training one")
    ap.add_argument("--stokes-mirror", type=int, default=1, help="stokes: 1 adds the x-mirrored training samples")
    ap.add_argument("--stokes-batch", type=int, default=64,
                    help="stokes: training solves this many samples per launch pair (Stokes2DBatch); 1 solves them one at a time")
    ap.add_argument("--train", type=int, default=256, help="stokes: training samples")
    ap.add_argument("--eval", type=int, default=24, help="stokes: held-out samples at --stokes-n")
    ap.add_argument("--eval2", type=int, default=4, help="stokes: held-out samples at --stokes-n2")
    ap.add_argument("--eval-train", type=int, default=0, help="stokes: held-out samples at --stokes-train-n as well")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--shmem-n", type=int, default=16384)
    ap.add_argument("--stokes3d-n", type=int, default=127)
    ap.add_argument("--stokes3d-solve", type=int, default=1, help="stokes3d: 0 skips the solve to tolerance")
    ap.add_argument("--stokes3d-modes", default="eager,graph", help="stokes3d: launch modes measured (eager, graph)")
    ap.add_argument("--graph", type=int, default=1, help="thermal, hydromech: 0 launches each PT iteration eagerly")
    ap.add_argument("--thermal-steps", type=int, default=3, help="thermal, hydromech: time steps run to PT tolerance")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "stokes":
        stokes_warm_start(args)
        return
    if args.workload == "acoustic":
        acoustic_study(args)
        return
    if args.workload == "adjoint":
        adjoint_study(args)
        return
    if args.workload == "xgroups":
        xgroups_study(args)
        return
    if args.workload == "unet":
        unet_study(args)
        return
    if args.workload == "stokes-batch":
        stokes_batch_study(args)
        return
    if args.workload == "gemm":
        gemm_study(args)
        return
    if args.workload == "shmem":
        shmem_study(args)
        return
    if args.workload == "cellarray":
        cellarray_study(args)
        return
    if args.workload == "ad":
        ad_study(args)
        return
    if args.workload == "hidecomm":
        hidecomm_study(args)
        return
    if args.workload == "stokes3d":
        stokes3d_study(args)
        return
    if args.workload == "thermal":
        thermal_study(args)
        return
    if args.workload == "porowaves":
        porowaves_study(args)
        return
    if args.workload == "hydromech":
        hydromech_study(args)
        return

    import parallelstencil as ps
    from parallelstencil import workloads as W

    ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 3)
    n, K, Wm = args.n, args.steps, args.warmup
    dev = ps.device_info()
    log(f"device {dev['name']} {dev['arch']} CUs={dev['cu_count']} reported-HBM={dev['hbm_peak_gbs']:.0f} GB/s")
    bytes_step = 3 * 8 * n ** 3
    res = {}

    if args.only in ("", "triad"):
        k = W.define_triad3d(__name__)
        A, B, C = ps.zeros(n, n, n), ps.fill(1.7, n, n, n), ps.fill(0.5, n, n, n)
        state = [A, B]

        def triad(i):
            k(state[0], state[1], C, async_=True)
            state[0], state[1] = state[1], state[0]
        ms, kms = time_kernel(ps, triad, K, Wm)
        res["triad"] = (ms, kms)
        log(f"triad {n}^3: {ms:.4f} ms/rep wall, {kms:.4f} ms kernel, T_eff {bytes_step / kms / 1e6:.1f} GB/s")
        del A, B, C, state

    T0, Cp, p = W.diffusion_inputs(n, seed=1)
    pa = (p["lam"], p["dt"], p["dx"], p["dy"], p["dz"])
    T, T2, Cpd = ps.zeros(n, n, n), ps.zeros(n, n, n), ps.zeros(n, n, n)
    Cpd.copy_from(Cp)
    for name, memopt, fast in (("plain", False, False), ("memopt", True, False), ("memopt_fastmath", True, True)):
        if args.only and args.only != name:
            continue
        k = W.define_diffusion3d(memopt, __name__, fastmath=fast)
        T.copy_from(T0)
        T2.copy_from(T0)
        st = [T, T2]

        def step(i):
            k(st[1], st[0], Cpd, *pa, async_=True)
            st[0], st[1] = st[1], st[0]
        ms, kms = time_kernel(ps, step, K, Wm)
        lo, med = per_launch_ms(ps, step)
        res[name] = (ms, kms, lo, med)
        log(f"diffusion {name} {n}^3: {ms:.4f} ms/step wall, {kms:.4f} ms kernel (single launches: min {lo:.4f}, median {med:.4f}), "
            f"T_eff {bytes_step / kms / 1e6:.1f} GB/s")
    del T0, Cp

    # the headline kernel: memopt with fastmath=true (FMA contraction, reciprocal division), held
    # to 1e-12 relative per field of the exact oracle over the 200 steps at 512^3 in the GPU tests;
    # the bit-exact memopt kernel is reported beside it in "kernels"
    head = next((h for h in ("memopt_fastmath", "memopt", "plain", "triad") if h in res), None)
    ms, kms = res[head][:2]
    teff_wall = bytes_step / (ms * 1e-3) / 1e9
    achieved = bytes_step / (kms * 1e-3) / 1e9
    traffic, traffic_src = _traffic(n, head)
    out = {
        "metric": "T_eff GB/s, %HBM peak (diffusion3D 512³); Stokes2D time-to-tol: warm vs plain",
        "value": round(teff_wall, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"diffusion3D {n}^3 FP64, flux form with on-the-fly fluxes, {head} kernel",
                   "n": n, "fields": "T, T2, Cp (3 x 8 B per cell per step)", "hbm_peak_gbs": HBM_PEAK_GBS,
                   "arithmetic": HEAD_ARITHMETIC.get(head, "exact"), "settle_launches": SETTLE},
        "hbm_frac": round(teff_wall / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src},
        "kernels": {k_: dict({"ms_wall": round(v[0], 5), "ms_kernel": round(v[1], 5),
                              "T_eff_gbs": round(bytes_step / (v[1] * 1e-3) / 1e9, 2)},
                             **({"ms_single_min": round(v[2], 5), "ms_single_median": round(v[3], 5)} if len(v) > 2 else {}))
                    for k_, v in res.items()},
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
