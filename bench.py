#!/usr/bin/env python3
"""Benchmark of the RPY grand-mobility hot path on one MI355X.

Headline line (`value`): one full grand-mobility apply (all four reference
blocks: seg<-seg 6x6, seg<-blob, blob<-seg, blob<-blob) over the
46,722-particle swimmer (1024 filaments x 30 segments + 16,002 blobs on a
sphere of radius 40a), forces and torques i.i.d. N(0,1), inputs resident in
HBM.  It reports pair-interactions/s of the FP32-Kahan kernel, the FP64
kernel beside it, the FP32 error against FP64, the pair kernel's achieved
FLOP/s against the vector roofline, and the CPU FP64 oracle timed on the host
cores on a bounded sample of the same workload.

Further legs in the same JSON line:
  sim_steps   implicit time steps of the 256-filament x 20-segment + 4002-blob
              swimmer after two driving periods: plain FP64 against the
              learned warm start, mixed precision, the fitted GMRES guess and
              the exact body-frame preconditioner, alone and combined, with
              Broyden and GMRES iteration counts and the CPU oracle's rate;
  ensemble    64 uneven swimmers stepped as one batch;
  rollout     1000 steps of the 962-particle swimmer with and without the
              learned warm start;
  cilia       prescribed-cilia saddle-point steps;
  carpet      a carpet of cilia above a no-slip wall (Swan-Brady kernel).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from generic_mobility_interface_class_amd import capi, geometry  # noqa: E402

FLOPS = {"ss": 92, "sb": 56, "bs": 56, "bb": 40,      # include/gmi_flops.h
         "ss_wall_tt": 137, "ss_wall": 324}
# Launch shape (target workgroups, minimum source chunk) of the time-stepping
# legs.  The shape sets the order of the pair kernels' partial sums, so a
# timed pick would change trajectories, and the networks trained on them, from
# run to run.  Of the shapes timed on the 9,122-particle swimmer's steps
# (target workgroups 1024-8192, minimum source chunks 64-256) this one gave
# the shortest step.
STEP_LAUNCH_SHAPE = (4096, 64)
# The 962-particle swimmer's tiles are too few for that shape to fill the
# chip (16 target tiles; 64-source chunks leave each lane a long serial loop):
# of (2048-8192, 16-64) timed over its body-inverse rollout, 16-source chunks
# gave the shortest step (1.49 against 1.71 ms).
ROLLOUT_LAUNCH_SHAPE = (4096, 16)
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def measured_traffic():
    """HBM bytes per FP32 grand-apply launch at 46,722 particles from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated on a copy of known
    size), recorded next to the profiles; None if absent."""
    path = os.path.join(ROOT, "profiles", "traffic_pair_apply_46722.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def finite_json(v):
    """Non-finite floats become null (strict JSON)."""
    if isinstance(v, dict):
        return {k: finite_json(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [finite_json(x) for x in v]
    if isinstance(v, float) and not np.isfinite(v):
        return None
    return v


def grand_pairs_flops(ns, nb):
    pairs = ns * ns + 2 * ns * nb + nb * nb
    flops = FLOPS["ss"] * ns * ns + (FLOPS["sb"] + FLOPS["bs"]) * ns * nb + FLOPS["bb"] * nb * nb
    return pairs, flops


def time_apply(ctx, steps, warmup):
    for _ in range(warmup):
        ctx.call("gmi_evaluate_grand_mobility")
    ctx.call("gmi_wait_for_device")
    ctx.call("gmi_profile_begin")
    ctx.call("gmi_timer_start")
    for _ in range(steps):
        ctx.call("gmi_evaluate_grand_mobility")
    ms = capi.C.c_double(0.0)
    ctx.call("gmi_timer_stop", capi.C.byref(ms))
    pair_ms = capi.C.c_double(0.0)
    n = capi.C.c_int(0)
    ctx.call("gmi_profile_end", capi.C.byref(pair_ms), capi.C.byref(n))
    ctx.call("gmi_wait_for_device")
    return ms.value, pair_ms.value / max(1, n.value)


def host_cpu():
    """CPU model and counts of the host the CPU legs run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu": model, "nproc": os.cpu_count() or 1, "usable": usable}


def cpu_baseline(xs, xb, fs, fb, seconds=12.0, seconds_one_core=6.0):
    """Oracle (CPU FP64 restatement) on a bounded sample of target rows
    against all sources, on up to 16 host threads (`value`) and on one
    (`one_core`); pairs/s."""
    from oracle import binding as orc
    threads = max(1, min(16, os.cpu_count() or 1))
    ns, nb = len(xs), len(xb)

    def run(rs, rb, nt):
        t0 = time.perf_counter()
        v = orc.Mss(xs, fs, 1.0, rows=(0, rs), nthreads=nt)
        orc.Msb(xs, xb, fb, 1.0, 1.0, V=v, rows=(0, rs), nthreads=nt)
        orc.Mbs(xb, xs, fs, 1.0, 1.0, rows=(0, rb), nthreads=nt)
        orc.Mbb(xb, fb, 1.0, rows=(0, rb), nthreads=nt)
        return time.perf_counter() - t0, (rs + rb) * (ns + nb)

    def sized(nt, target):
        dt, _ = run(64, 32, nt)
        scale = max(1.0, target / max(dt, 1e-6))
        rs = min(ns, int(64 * scale))
        rb = min(nb, int(32 * scale))
        dt, pairs = run(rs, rb, nt)
        return pairs / dt, (f"{rs} segment + {rb} blob target rows x all {ns + nb} sources "
                            f"(all four blocks, FP64, {dt:.1f} s)")

    value, sample = sized(threads, seconds)
    one, one_sample = sized(1, seconds_one_core)
    return {
        "value": value,
        "unit": "pair-interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": sample,
        "one_core": {"value": one, "sample": one_sample},
        "host": host_cpu(),
    }


def bench_step(args):
    """Implicit time steps of the 256 x 20-segment + 4002-blob swimmer.

    Identical runs take `step_warmup` FP64 steps (default two driving periods,
    600 steps); the learned ones record every 5th step's (features, converged
    - guess) rows and each trains a network on its own rows.  Then each run steps over the
    same `step_steps` steps in its own mode:
      fp64                  plain FP64 (the reference method: extrapolated
                            Broyden iterate, previous solution as the GMRES
                            guess)
      learned_mixed         network Broyden warm start + GMRES guess fitted on
                            the last 16 right-hand sides + mixed precision
      learned_fp64          network + fitted GMRES guess, FP64
      fitted_fp64           fitted GMRES guess only, FP64
      fitted_mixed          fitted GMRES guess + mixed precision
      mixed                 mixed precision only
      inverse_fp64          exact body-frame GMRES preconditioner, FP64
      inverse_mixed         the preconditioner + mixed precision
      learned_inverse_fp64  network + the preconditioner, FP64
      learned_inverse_mixed network + the preconditioner + mixed precision
    The fitted guess is on from the first step in the runs that use it.
    """
    cfg = geometry.sim_workload(args.step_workload, seed=1)
    modes = (("fp64", capi.GMI_FP64, False, False, False),
             ("learned_mixed", capi.GMI_FP32_KAHAN, True, True, False),
             ("learned_fp64", capi.GMI_FP64, True, True, False),
             ("fitted_fp64", capi.GMI_FP64, False, True, False),
             ("fitted_mixed", capi.GMI_FP32_KAHAN, False, True, False),
             ("mixed", capi.GMI_FP32_KAHAN, False, False, False),
             ("inverse_fp64", capi.GMI_FP64, False, False, True),
             ("inverse_mixed", capi.GMI_FP32_KAHAN, False, False, True),
             ("learned_inverse_fp64", capi.GMI_FP64, True, False, True),
             ("learned_inverse_mixed", capi.GMI_FP32_KAHAN, True, False, True))
    if args.step_modes:
        keep = args.step_modes.split(",")
        modes = tuple(m for m in modes if m[0] in keep)
    stride = 5
    sims = {}
    setup = {}
    for name, _, learned, fitted, inverse in modes:
        sim = capi.Sim(cfg, precision=capi.GMI_FP64)
        sim.set_sim_precision(capi.GMI_FP64)
        sim.set_graphs(not args.no_graphs)
        sim.set_gram_schmidt(args.gs_passes)
        if fitted:
            sim.set_gmres_history(16)
        if inverse:
            t0 = time.perf_counter()
            sim.set_body_preconditioner(capi.GMI_PRECOND_BODY_INVERSE)
            setup[name] = time.perf_counter() - t0
        if learned:
            sim.warmstart_init(hidden=128, seed=7)
            sim.warmstart_configure(apply=False, collect=(args.step_warmup // stride) * sim.Ns,
                                    stride=stride)
        shape = tuple(args.step_shape) if args.step_shape else STEP_LAUNCH_SHAPE
        sim.call("gmi_set_launch_shape", *shape)
        sim.make_body_reference_matrices()
        sims[name] = sim
    t0 = time.perf_counter()
    for nt in range(args.step_warmup):
        for name, sim in sims.items():
            rc, bi, gi = sim.step(nt)
            assert rc == 0, (name, nt, rc, sim.last_error())
    warm_s = time.perf_counter() - t0
    rows = next((s.warmstart_dataset_size() for (n, _, lr, _, _), s in
                 zip(modes, sims.values()) if lr), 0)
    print(f"# {args.step_warmup} warm-up steps x {len(sims)} runs in {warm_s:.1f} s, "
          f"{rows} training rows", file=sys.stderr, flush=True)
    train = {}
    for name, prec, learned, _, _ in modes:
        sim = sims[name]
        if learned:
            t0 = time.perf_counter()
            loss = sim.warmstart_train(iters=3000, batch=1024, lr=1e-3, seed=1)
            train = {"training_rows": rows, "train_iters": 3000,
                     "train_s": time.perf_counter() - t0, "final_loss_standardised": loss}
            sim.warmstart_configure(apply=True, collect=0)
        sim.set_sim_precision(prec)
    res = {}
    nt0 = args.step_warmup
    for name, prec, learned, _, _ in modes:
        sim = sims[name]
        a0, f0 = sim.warmstart_stats() if learned else (0, 0)
        r = _timed_steps(sim, nt0, args.step_steps, f"step {name}")
        if learned:
            a1, f1 = sim.warmstart_stats()
            r["learned_steps"], r["fallbacks"] = a1 - a0, f1 - f0
        res[name] = r
    for name in setup:
        res[name]["preconditioner_setup_s"] = setup[name]
    if "fp64" in sims:
        ref = sims["fp64"].state()["x"]
        for name in sims:
            if name != "fp64":
                res[name]["max_position_diff_vs_fp64"] = float(
                    np.abs(sims[name].state()["x"] - ref).max())
        res["fp64"]["gram_schmidt"] = _gram_schmidt(sims["fp64"], nt0 + args.step_steps, 2,
                                                    "step fp64")
    for sim in sims.values():
        sim.close()
    out = {"workload": args.step_workload, "warmup_steps": args.step_warmup,
           "steps_per_window": args.step_steps, "window": f"{nt0}-{nt0 + args.step_steps - 1}",
           "launch_shape": list(shape), **res, "warm_start": train}
    if "fp64" in res:
        for name in res:
            if name != "fp64":
                out[f"speedup_{name}_vs_fp64"] = res[name]["steps_per_s"] / res["fp64"]["steps_per_s"]
    return out


# Name of the one timed window whose bounds --profile-window prints, for
# --summarize-trace to cut a whole-run kernel trace to it (None: no window).
PROFILE_WINDOW = None


def _timed_steps(sim, nt0, k, name):
    """k steps from nt0 timed with the context's HIP events.  A step that
    fails (non-convergence, divergence) ends the window and is reported."""
    window = name == PROFILE_WINDOW
    sim.call("gmi_wait_for_device")
    if window:
        print(f"# {name}: profiled window opens at {time.monotonic_ns()} ns (CLOCK_MONOTONIC)",
              file=sys.stderr, flush=True)
    sim.call("gmi_timer_start")
    tb = tg = 0
    failed = None
    for nt in range(nt0, nt0 + k):
        rc, bi, gi = sim.step(nt)
        tb += bi
        tg += gi
        if rc != 0:
            failed = {"step": nt, "status": rc, "message": sim.last_error()}
            print(f"# {name}: step {nt} failed: {failed['message']}", file=sys.stderr, flush=True)
            k = nt - nt0 + 1
            break
    ms = capi.C.c_double(0.0)
    sim.call("gmi_timer_stop", capi.C.byref(ms))
    if window:
        sim.call("gmi_wait_for_device")
        print(f"# {name}: profiled window closes at {time.monotonic_ns()} ns (CLOCK_MONOTONIC)",
              file=sys.stderr, flush=True)
    if failed:
        return {"failed": failed, "steps": f"{nt0}-{nt0 + k - 1}", "steps_per_s": 0.0}
    r = dict(steps=f"{nt0}-{nt0 + k - 1}", steps_per_s=k / (ms.value * 1e-3),
             ms_per_step=ms.value / k, broyden_iters_per_step=tb / k,
             gmres_iters_per_step=tg / k)
    print(f"# {name}: {r['ms_per_step']:.2f} ms/step, {tb / k:.2f} Broyden, "
          f"{tg / k:.1f} GMRES it/step", file=sys.stderr, flush=True)
    return r


def _gram_schmidt(sim, nt0, k, name):
    """The Gram-Schmidt kernels of the fused Arnoldi step (k_kprod_dots and
    k_gs_chunk<2>, with k_gs_chunk<1> when it takes two passes and
    k_colsum_wide beyond 256 chunks) over k steps from nt0, launched eagerly
    so that every Arnoldi step brackets them with HIP events.  Algorithmic
    bytes per Arnoldi step with j previous columns of length n: per
    Gram-Schmidt pass (j + 2) n doubles (the columns and w read, w written),
    against 8 TB/s; the K products and split sums that k_kprod_dots also does
    are not counted."""
    sim.set_graphs(False)
    sim.call("gmi_profile_begin")
    for nt in range(nt0, nt0 + k):
        assert sim.step(nt)[0] == 0
    ms, nbytes, steps = capi.C.c_double(0.0), capi.C.c_double(0.0), capi.C.c_int(0)
    sim.call("gmi_profile_krylov", capi.C.byref(ms), capi.C.byref(nbytes), capi.C.byref(steps))
    pair_ms, n = capi.C.c_double(0.0), capi.C.c_int(0)
    sim.call("gmi_profile_end", capi.C.byref(pair_ms), capi.C.byref(n))
    gbs = nbytes.value / (ms.value * 1e-3) / 1e9 if ms.value > 0 else 0.0
    r = {"kernels": "k_kprod_dots + k_gs_chunk<2> (+ k_colsum_wide): K products and the "
                    "one Gram-Schmidt pass of the fused Arnoldi step (the reference's one pass)",
         "steps": f"{nt0}-{nt0 + k - 1}", "arnoldi_steps": steps.value, "ms": ms.value,
         "us_per_arnoldi_step": 1e3 * ms.value / max(1, steps.value),
         "algorithmic_bytes": nbytes.value, "achieved_gbs": gbs,
         "frac_hbm_peak": gbs / PEAK_HBM_GBS}
    print(f"# {name} Gram-Schmidt: {steps.value} Arnoldi steps, "
          f"{r['us_per_arnoldi_step']:.1f} us each, {gbs:.0f} GB/s algorithmic",
          file=sys.stderr, flush=True)
    return r


def bench_ensemble(args):
    """64 independent swimmers of uneven size (17..474 filaments x 20
    segments, 642..7292 blobs; 306,518 particles) stepped as one batch:
    every pair launch covers all swimmers, Arnoldi chunks replay as
    captured graphs.  Two steps untimed, then FP64, mixed-precision and
    mixed-precision with the exact body-frame preconditioner (18.4 GB of
    packed per-swimmer inverses in HBM) windows from the same state (three
    contexts); after the FP64 window one more step measures the
    Gram-Schmidt kernels at the batch's GMRES length."""
    cfg = geometry.ensemble_workload("ensemble_64", seed=4)
    out = {"workload": "ensemble_64", "swimmers": len(cfg.swimmers),
           "filaments": int(sum(g.nfil for g in cfg.swimmers)),
           "segments": int(sum(g.ns for g in cfg.swimmers)),
           "blobs": int(sum(g.nblob for g in cfg.swimmers))}
    for name, prec, k, inverse in (("fp64", capi.GMI_FP64, args.ens_steps // 2, False),
                                   ("mixed", capi.GMI_FP32_KAHAN, args.ens_steps, False),
                                   ("inverse_mixed", capi.GMI_FP32_KAHAN, args.ens_steps, True)):
        sim = capi.Sim(cfg, precision=prec)
        sim.set_sim_precision(prec)
        sim.set_graphs(not args.no_graphs)
        t0 = time.perf_counter()
        sim.make_body_reference_matrices()
        sim.autotune_blocks(capi.GMI_BLOCK_BB, reps=5)
        if inverse:
            sim.set_body_preconditioner(capi.GMI_PRECOND_BODY_INVERSE)
        setup_s = time.perf_counter() - t0
        for nt in range(2):
            assert sim.step(nt)[0] == 0
        out[name] = _timed_steps(sim, 2, k, f"ensemble {name}")
        out[name]["setup_s"] = setup_s
        if name == "fp64":
            out[name]["gram_schmidt"] = _gram_schmidt(sim, 2 + k, 1, f"ensemble {name}")
        sim.close()
    out["speedup_mixed_vs_fp64"] = out["mixed"]["steps_per_s"] / out["fp64"]["steps_per_s"]
    out["speedup_inverse_mixed_vs_fp64"] = (out["inverse_mixed"]["steps_per_s"] /
                                            out["fp64"]["steps_per_s"])
    return out


def bench_rollout(args):
    """Learned warm start over a long rollout of the 16 x 20-segment +
    642-blob swimmer.  Two identical runs take one driving period (300
    steps) of plain FP64 steps; one records every 5th step (18,880 rows of
    (features, converged - guess)) and trains the network; then both run
    `rollout` more steps, one plain, one with the network.  The same pair
    then repeats with the exact body-frame preconditioner in both runs, so
    the network's gain is measured with and without GMRES dominating."""
    cfg = geometry.sim_workload("step_962", seed=1)
    plain, learned = capi.Sim(cfg), capi.Sim(cfg)
    plain_inv, learned_inv = capi.Sim(cfg), capi.Sim(cfg)
    for s in (plain, learned, plain_inv, learned_inv):
        s.set_graphs(not args.no_graphs)
    for s in (learned, learned_inv):
        s.warmstart_init(hidden=128, seed=7)
    for s in (learned, learned_inv):
        s.warmstart_configure(apply=False, collect=20000, stride=5)
    for s in (plain, learned, plain_inv, learned_inv):
        s.call("gmi_set_launch_shape", *ROLLOUT_LAUNCH_SHAPE)
        s.make_body_reference_matrices()
    for s in (plain_inv, learned_inv):
        s.set_body_preconditioner(capi.GMI_PRECOND_BODY_INVERSE)
    for nt in range(300):
        for s in (plain, learned, plain_inv, learned_inv):
            assert s.step(nt)[0] == 0
    rows = learned.warmstart_dataset_size()
    t0 = time.perf_counter()
    loss = learned.warmstart_train(iters=3000, batch=1024, lr=1e-3, seed=1)
    train_s = time.perf_counter() - t0
    learned.warmstart_configure(apply=True, collect=0)
    # the preconditioned learned run trains on its own rows, same schedule
    learned_inv.warmstart_train(iters=3000, batch=1024, lr=1e-3, seed=1)
    learned_inv.warmstart_configure(apply=True, collect=0)
    k = args.rollout
    rp = _timed_steps(plain, 300, k, "rollout plain")
    rl = _timed_steps(learned, 300, k, "rollout learned")
    rpi = _timed_steps(plain_inv, 300, k, "rollout plain + body inverse")
    rli = _timed_steps(learned_inv, 300, k, "rollout learned + body inverse")
    dx = float(np.abs(plain.state()["x"] - learned.state()["x"]).max())
    for s in (plain, learned, plain_inv, learned_inv):
        s.close()
    return {"workload": "step_962", "warmup_steps": 300, "rollout_steps": k,
            "launch_shape": list(ROLLOUT_LAUNCH_SHAPE),
            "plain": rp, "learned": rl,
            "speedup_learned_vs_plain": rl["steps_per_s"] / rp["steps_per_s"],
            "plain_inverse": rpi, "learned_inverse": rli,
            "speedup_learned_inverse_vs_plain_inverse": rli["steps_per_s"] / rpi["steps_per_s"],
            "speedup_learned_inverse_vs_plain": rli["steps_per_s"] / rp["steps_per_s"],
            "final_max_position_diff": dx,
            "warm_start": {"training_rows": rows, "train_iters": 3000, "train_s": train_s,
                           "final_loss_standardised": loss}}


def _cilia_table(cfg, steps, use_graphs=True):
    """Reference generalised forces of one filament of `cfg`'s body, computed
    by the device as the reference computes them (fixed phase speed 2 pi from
    phase 0 over one period, vel_dir . f after each solve)."""
    one = dataclasses.replace(cfg, nfil=np.array([1]),
                              fil_refs=np.asarray(cfg.fil_refs).reshape(-1, 3)[:1],
                              phase=np.zeros(1), dynamic_phase=False, dt=1.0 / steps)
    g = capi.Cilia(one)
    g.set_graphs(use_graphs)
    out = np.zeros(steps)
    for nt in range(steps):
        g.initial_guess(nt)
        assert g.compute_velocities(nt)[0] == 0
        out[nt] = g.generalised_forces()[0]
        g.integrate_bodies()
        g.end_of_step(nt)
    g.close()
    return out


def _cilia_window(g, nt0, k, name):
    g.call("gmi_wait_for_device")
    g.call("gmi_timer_start")
    its = 0
    for nt in range(nt0, nt0 + k):
        rc, it = g.step(nt)
        assert rc == 0, g.last_error()
        its += it
    ms = capi.C.c_double(0.0)
    g.call("gmi_timer_stop", capi.C.byref(ms))
    r = dict(steps=f"{nt0}-{nt0 + k - 1}", steps_per_s=k / (ms.value * 1e-3),
             ms_per_step=ms.value / k, gmres_iters_per_step=its / k)
    print(f"# {name}: {r['ms_per_step']:.2f} ms/step, {its / k:.1f} GMRES it/step",
          file=sys.stderr, flush=True)
    return r


def bench_cilia(args):
    """Prescribed-cilia steps (coral-larva beat, dynamic phases, free body,
    the reference's defaults): the 32-cilium swimmer (1,282 particles) and
    the 512-cilium swimmer (14,242 particles).  Per workload: the reference
    generalised-force table (one period, 300 steps) from the device, 3
    untimed steps, then FP64 and mixed-precision windows from fresh starts."""
    out = {}
    names = [args.cilia_workload] if args.cilia_workload else ["cilia_1282", "cilia_14242"]
    for name in names:
        k = args.cilia_steps
        base = geometry.cilia_workload(name)
        t0 = time.perf_counter()
        table = _cilia_table(base, args.cilia_table_steps, not args.no_graphs)
        table_s = time.perf_counter() - t0
        cfg = dataclasses.replace(base, gen_force_refs=table)
        r = {"segments": cfg.ns, "blobs": cfg.nb, "filaments": cfg.nf,
             "system_size": cfg.system_size(), "table_s": table_s}
        for prec_name, prec in (("fp64", capi.GMI_FP64), ("mixed", capi.GMI_FP32_KAHAN)):
            g = capi.Cilia(cfg, precision=prec)
            g.set_graphs(not args.no_graphs)
            g.set_gram_schmidt(args.gs_passes)
            for nt in range(3):
                assert g.step(nt)[0] == 0
            r[prec_name] = _cilia_window(g, 3, k, f"{name} {prec_name}")
            g.close()
        r["speedup_mixed_vs_fp64"] = r["mixed"]["steps_per_s"] / r["fp64"]["steps_per_s"]
        out[name] = r
    if not args.no_cpu and "cilia_1282" in out:
        out["cilia_1282"]["cpu_baseline"] = cpu_cilia_baseline()
    return out


def bench_carpet(args):
    """Prescribed-cilia carpets above a no-slip wall (the reference's
    INFINITE_PLANE_WALL build with its defaults: coral-larva beat, dynamic
    phases, rectangular seeding): 64 and 1,024 cilia of 20 segments.  Per
    workload: the reference generalised-force table of one filament from the
    device (one period), 3 untimed steps, an FP64 window, then the wall pair
    kernel alone (M_ss of the last solution on the final geometry, 20
    launches timed with HIP events on the context stream), and the full 6x6
    wall apply (forces and torques, N(0,1), seed 5) on the same positions."""
    out = {}
    names = [args.carpet_workload] if args.carpet_workload else list(geometry.CARPET_WORKLOADS)
    for name in names:
        k = args.cilia_steps
        base = geometry.carpet_workload(name)
        t0 = time.perf_counter()
        table = _cilia_table(base, args.cilia_table_steps, not args.no_graphs)
        table_s = time.perf_counter() - t0
        cfg = dataclasses.replace(base, dynamic_phase=True, gen_force_refs=table)
        r = {"segments": cfg.ns, "filaments": cfg.nf, "system_size": cfg.system_size(),
             "table_s": table_s}
        g = capi.Cilia(cfg)
        g.set_graphs(not args.no_graphs)
        g.set_gram_schmidt(args.gs_passes)
        for nt in range(3):
            assert g.step(nt)[0] == 0
        r["fp64"] = _cilia_window(g, 3, k, f"{name} fp64")
        g.segment_velocities()
        g.call("gmi_profile_begin")
        for _ in range(20):
            g.segment_velocities()
        pair_ms = capi.C.c_double(0.0)
        n = capi.C.c_int(0)
        g.call("gmi_profile_end", capi.C.byref(pair_ms), capi.C.byref(n))
        xs_final = g.state()["x"]
        g.close()
        ms = pair_ms.value / max(1, n.value)
        pairs = cfg.ns * cfg.ns
        tflops = FLOPS["ss_wall_tt"] * pairs / (ms * 1e-3) / 1e12
        r["pair_kernel"] = {"kernel": "k_pair_wall<true>", "ms": ms, "launches": n.value,
                            "pairs_per_s": pairs / (ms * 1e-3), "achieved_tflops": tflops,
                            "frac_fp64_peak": tflops / PEAK_FP64_TFLOPS}
        print(f"# {name} wall pairs: {ms:.3f} ms, {tflops:.1f} TFLOP/s", file=sys.stderr,
              flush=True)
        # the full 6x6 wall mobility (forces and torques) on the same positions
        ctx = capi.Context([cfg.ns], [0], cfg.nseg)
        ctx.set_wall(True)
        ctx.upload(xs_final, np.random.default_rng(5).standard_normal((cfg.ns, 6)))
        total_ms, kern_ms = time_apply(ctx, 10, 2)
        ctx.close()
        tf6 = FLOPS["ss_wall"] * pairs / (kern_ms * 1e-3) / 1e12
        r["wall_apply_6x6"] = {"kernel": "k_pair_wall<false>", "ms": kern_ms,
                               "ms_per_apply": total_ms / 10, "pairs_per_s": pairs / (kern_ms * 1e-3),
                               "achieved_tflops": tf6, "frac_fp64_peak": tf6 / PEAK_FP64_TFLOPS}
        print(f"# {name} 6x6 wall apply: {kern_ms:.3f} ms, {tf6:.1f} TFLOP/s", file=sys.stderr,
              flush=True)
        out[name] = r
    if not args.no_cpu and "carpet_1280" in out:
        out["carpet_1280"]["cpu_baseline"] = cpu_carpet_baseline()
    return out


def cpu_carpet_baseline():
    """Oracle (CPU FP64) steps of the 64-cilium carpet (bounded sample: ten
    steps after two, table from the oracle itself)."""
    from oracle import binding as orc
    threads = max(1, min(16, os.cpu_count() or 1))
    base = geometry.carpet_workload("carpet_1280")
    one = dataclasses.replace(base, nfil=np.array([1]),
                              fil_refs=np.asarray(base.fil_refs).reshape(-1, 3)[:1],
                              phase=np.zeros(1))
    table = orc.cilia_reference_gen_forces(one, 60, nthreads=threads)
    o = orc.Cilia(dataclasses.replace(base, dynamic_phase=True, gen_force_refs=table),
                  nthreads=threads)
    for nt in range(2):
        o.step(nt)
    t0 = time.perf_counter()
    its = 0
    for nt in range(2, 12):
        rc, it = o.step(nt)
        assert rc == 0
        its += it
    dt = time.perf_counter() - t0
    return {"value": 10.0 / dt, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"steps 2-11 of carpet_1280 ({its / 10:.1f} GMRES it/step, {dt:.2f} s; "
                      f"pair sums on {threads} threads, the rest single-threaded; 60-entry table)"}


def cpu_cilia_baseline():
    """Oracle (CPU FP64) prescribed-cilia step of the 1,282-particle swimmer
    (bounded sample: one step after two, table from the oracle itself)."""
    from oracle import binding as orc
    threads = max(1, min(16, os.cpu_count() or 1))
    base = geometry.cilia_workload("cilia_1282")
    one = dataclasses.replace(base, nfil=np.array([1]),
                              fil_refs=np.asarray(base.fil_refs).reshape(-1, 3)[:1],
                              phase=np.zeros(1), dynamic_phase=False)
    table = orc.cilia_reference_gen_forces(one, 60, nthreads=threads)
    o = orc.Cilia(dataclasses.replace(base, gen_force_refs=table), nthreads=threads)
    for nt in range(2):
        o.step(nt)
    t0 = time.perf_counter()
    rc, it = o.step(2)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"step 2 of cilia_1282 ({it} GMRES iterations, {dt:.2f} s; pair sums on "
                      f"{threads} threads, the rest single-threaded; 60-entry table)"}


def cpu_step_baseline(args):
    """Oracle (CPU FP64) time step on the same swimmer: one step after the
    body reference matrices (bounded sample)."""
    from oracle import binding as orc
    threads = max(1, min(16, os.cpu_count() or 1))
    cfg = geometry.sim_workload(args.step_workload, seed=1)
    sim = orc.Sim(cfg, nthreads=threads)
    sim.make_body_reference_matrices()
    t0 = time.perf_counter()
    rc, bi, gi = sim.step(0)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"step 0 of {args.step_workload} ({bi} Broyden, {gi} GMRES iterations, "
                      f"{dt:.1f} s; pair sums on {threads} threads, the rest single-threaded)"}


def summarize_trace(trace_csv, bench_log, steps):
    """Markdown table of the kernels of a rocprofv3 kernel trace (CSV) that
    ran inside the window a --profile-window run printed (CLOCK_MONOTONIC ns,
    the trace's clock): calls, time and share per kernel, per step."""
    import csv
    import re
    steps = int(steps)
    bounds = {}
    for line in open(bench_log):
        m = re.search(r"profiled window (opens|closes) at (\d+) ns", line)
        if m:
            bounds[m.group(1)] = int(m.group(2))
    if "opens" not in bounds or "closes" not in bounds:
        sys.exit(f"{bench_log}: no profiled window (run bench.py --profile-window MODE)")
    t0, t1 = bounds["opens"], bounds["closes"]
    per = {}
    busy = 0
    for r in csv.DictReader(open(trace_csv)):
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if s < t0 or e > t1:
            continue
        name = r["Kernel_Name"].replace("void ", "").replace("gmi::(anonymous namespace)::", "")
        name = re.sub(r"\(.*", "", name)
        c = per.setdefault(name, [0, 0])
        c[0] += 1
        c[1] += e - s
        busy += e - s
    wall = t1 - t0
    print(f"window {wall / 1e6:.2f} ms over {steps} steps: {wall / 1e6 / steps:.3f} ms/step wall, "
          f"{busy / 1e6 / steps:.3f} ms/step in kernels ({100.0 * busy / wall:.0f}% busy), "
          f"{sum(c for c, _ in per.values()) / steps:.1f} kernels/step\n")
    print("| kernel | calls/step | us/step | avg us | share of kernel time |")
    print("|---|---:|---:|---:|---:|")
    for name, (c, d) in sorted(per.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{name}` | {c / steps:.2f} | {d / 1e3 / steps:.1f} | {d / 1e3 / c:.2f} | "
              f"{100.0 * d / busy:.1f}% |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="apply_46722")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--step-workload", default="step_9122")
    ap.add_argument("--step-steps", type=int, default=40)
    ap.add_argument("--step-warmup", type=int, default=600,
                    help="FP64 steps before the timed window (two driving periods: the "
                         "learned runs train on every 5th step's rows)")
    ap.add_argument("--no-step", action="store_true")
    ap.add_argument("--step-modes", default=None,
                    help="comma list of step-leg modes (default: all nine)")
    ap.add_argument("--gs-passes", type=int, default=1, choices=(1, 2),
                    help="Gram-Schmidt passes per Arnoldi step in the step, cilia and carpet "
                         "legs (1: one, as the reference)")
    ap.add_argument("--step-shape", type=int, nargs=2, default=None,
                    metavar=("TARGET_WORKGROUPS", "MIN_CHUNK"),
                    help="pair launch shape of the step leg (default: STEP_LAUNCH_SHAPE)")
    ap.add_argument("--ens-steps", type=int, default=40)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--rollout", type=int, default=1000)
    ap.add_argument("--no-rollout", action="store_true")
    ap.add_argument("--cilia-steps", type=int, default=20)
    ap.add_argument("--cilia-workload", default=None,
                    help="one of geometry.CILIA_WORKLOADS (default: both)")
    ap.add_argument("--cilia-table-steps", type=int, default=300,
                    help="entries of the reference generalised-force table (steps per period)")
    ap.add_argument("--no-cilia", action="store_true")
    ap.add_argument("--carpet-workload", default=None,
                    help="one of geometry.CARPET_WORKLOADS (default: both)")
    ap.add_argument("--no-carpet", action="store_true")
    ap.add_argument("--carpet-only", action="store_true",
                    help="only the carpets above the wall (diagnostic line)")
    ap.add_argument("--cilia-only", action="store_true",
                    help="only the prescribed-cilia steps (diagnostic line, e.g. for profiling)")
    ap.add_argument("--apply-only", action="store_true",
                    help="only the 46,722-particle apply (no CPU baseline, no time steps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device (a transfer of known size for "
                         "calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the Arnoldi kernels eagerly (for kernel-trace profiling)")
    ap.add_argument("--summarize-trace", nargs=3, default=None,
                    metavar=("KERNEL_TRACE_CSV", "BENCH_LOG", "STEPS"),
                    help="no GPU work: summarise a rocprofv3 kernel trace over the window "
                         "that a --profile-window run printed to BENCH_LOG (per step over "
                         "STEPS steps), as a markdown table")
    ap.add_argument("--profile-window", default=None, metavar="MODE",
                    help="run only the time-step leg in MODE (e.g. learned_inverse_mixed), "
                         "eagerly, and print the bounds of its timed window, so that "
                         "--summarize-trace can cut a kernel trace of the run to it")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(*args.summarize_trace)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one MI355X; "
              "multi-GPU sharding is design-only (see the sharding notes).", file=sys.stderr)
        sys.exit(2)

    if args.cilia_only:
        print(json.dumps(finite_json({"cilia": bench_cilia(args)}), allow_nan=False), flush=True)
        return
    if args.carpet_only:
        print(json.dumps(finite_json({"carpet": bench_carpet(args)}), allow_nan=False), flush=True)
        return
    if args.apply_only:
        args.no_cpu = args.no_step = args.no_ensemble = args.no_rollout = args.no_cilia = True
        args.no_carpet = True
    if args.profile_window:
        global PROFILE_WINDOW
        PROFILE_WINDOW = f"step {args.profile_window}"
        args.step_modes = args.profile_window
        args.no_cpu = args.no_ensemble = args.no_rollout = args.no_cilia = True
        args.no_carpet = args.no_graphs = True
    if args.copy_calibration:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = src.clone()
        torch.cuda.synchronize()
        print(f"# calibration copy: {src.numel() * 4} bytes read and written",
              file=sys.stderr, flush=True)
        del src, dst

    g = geometry.workload_geometry(args.workload)
    xs, xb = g.positions()
    fs, fb = geometry.random_forces(g.ns, g.nblob, seed=3)
    pairs, flops = grand_pairs_flops(g.ns, g.nblob)

    res = {}
    vel = {}
    for name, prec in (("fp32_kahan", capi.GMI_FP32_KAHAN), ("fp64", capi.GMI_FP64)):
        ctx = capi.Context([g.ns], [g.nblob], g.nseg, precision=prec)
        ctx.upload(xs, fs, xb, fb)
        # five launches per candidate: with two, box noise sometimes picked a
        # 4096-workgroup shape 4% slower than the 8192 one
        shape = ctx.autotune(reps=5)
        steps = args.steps if name == "fp32_kahan" else max(3, args.steps // 2)
        total_ms, kern_ms = time_apply(ctx, steps, args.warmup)
        vel[name] = ctx.velocities()
        res[name] = dict(steps=steps, ms_per_step=total_ms / steps, kernel_ms=kern_ms,
                         launch_shape=shape)
        ctx.close()
        print(f"# {name}: {total_ms / steps:.3f} ms/apply, pair kernel {kern_ms:.3f} ms",
              file=sys.stderr, flush=True)

    v32s, v32b = vel["fp32_kahan"]
    v64s, v64b = vel["fp64"]
    num = np.sqrt(np.sum((v32s - v64s) ** 2) + np.sum((v32b - v64b) ** 2))
    den = np.sqrt(np.sum(v64s ** 2) + np.sum(v64b ** 2))
    err_norm = float(num / den)
    err_max = float(max(np.abs(v32s - v64s).max() / np.abs(v64s).max(),
                        np.abs(v32b - v64b).max() / np.abs(v64b).max()))

    r32, r64 = res["fp32_kahan"], res["fp64"]
    achieved32 = flops / (r32["kernel_ms"] * 1e-3) / 1e12
    achieved64 = flops / (r64["kernel_ms"] * 1e-3) / 1e12
    line = {
        "metric": "RPY pair-interactions/s (% FP32 peak)",
        "value": pairs / (r32["ms_per_step"] * 1e-3),
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": r32["steps"],
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (geodesic sphere + Fibonacci filaments, N(0,1) forces, seed 3)",
        "config": {
            "workload": args.workload,
            "segments": g.ns, "blobs": g.nblob,
            "pairs_per_apply": pairs, "flop_per_apply": flops,
            "blocks": "ss+sb+bs+bb", "precision": "FP32 pairs, Kahan cross-tile, FP64 reduce",
        },
        "roofline": {
            # compute-bound on the vector ALUs (a pairwise N-body sum is not a
            # contraction, so no MFMA); priced against the published FP32 peak
            "bound": "valu",
            "units": "VALU FP32 (pair sums are not a contraction; no MFMA)",
            "achieved": round(achieved32, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved32 / PEAK_FP32_TFLOPS, 4),
            "traffic": measured_traffic(),
            "kernel": "k_pair<float>",
            "kernel_ms": r32["kernel_ms"],
            "launch_shape": r32["launch_shape"],
        },
        "fp64": {
            "value": pairs / (r64["ms_per_step"] * 1e-3),
            "ms_per_step": r64["ms_per_step"],
            "kernel_ms": r64["kernel_ms"],
            "achieved_tflops": round(achieved64, 3),
            "frac_fp64_peak": round(achieved64 / PEAK_FP64_TFLOPS, 4),
        },
        "fp32_error_vs_fp64": {"norm_rel": err_norm, "max_rel": err_max},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(xs, xb, fs, fb, args.cpu_seconds)
    if not args.no_step and args.step_modes:
        line["sim_steps"] = bench_step(args)  # a subset of modes (diagnostics)
    elif not args.no_step:
        st = line["sim_steps"] = bench_step(args)
        # the second half of the metric: learned + mixed vs plain FP64 steps/s
        fp = st["fp64"]["steps_per_s"]
        line["steps_learned_mixed_vs_fp64"] = {
            "workload": st["workload"],
            "fp64_steps_per_s": fp,
            # learned + mixed with the reference's GMRES preconditioner
            "learned_mixed_steps_per_s": st["learned_mixed"]["steps_per_s"],
            "learned_mixed_speedup": st["learned_mixed"]["steps_per_s"] / fp,
            # learned + mixed with the exact body-frame preconditioner (fastest mode)
            "learned_inverse_mixed_steps_per_s": st["learned_inverse_mixed"]["steps_per_s"],
            "speedup": st["learned_inverse_mixed"]["steps_per_s"] / fp,
            "path": "speedup: network Broyden warm start + FP32-Kahan segment pair blocks + "
                    "exact body-frame GMRES preconditioner (FP64 Krylov products, one Arnoldi "
                    "step per solve); learned_mixed_speedup: the same without the body-frame "
                    "preconditioner",
            "attribution_speedups_vs_fp64": {
                k: st[k]["steps_per_s"] / fp for k in
                ("mixed", "fitted_fp64", "fitted_mixed", "learned_fp64", "learned_mixed",
                 "inverse_fp64",
                 "inverse_mixed", "learned_inverse_fp64", "learned_inverse_mixed")},
        }
        if not args.no_cpu:
            line["sim_steps"]["cpu_baseline"] = cpu_step_baseline(args)
    if not args.no_ensemble:
        line["ensemble"] = bench_ensemble(args)
    if not args.no_rollout:
        line["rollout"] = bench_rollout(args)
    if not args.no_cilia:
        line["cilia"] = bench_cilia(args)
    if not args.no_carpet:
        line["carpet"] = bench_carpet(args)
    print(json.dumps(finite_json(line), allow_nan=False), flush=True)


if __name__ == "__main__":
    main()
