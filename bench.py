#!/usr/bin/env python3
"""Throughput of the transient path tracer on one MI355X.

A step is one full render pass of the workload (every path of every pixel,
every bounce, and the compensated film accumulation) with the scene and film
already resident in HBM. Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload icosphere6|box|media|batch64] [--guided on|only]

The CPU baseline (the oracle, oracle/oracle.c, same RNG scheme and workload)
runs first, on a bounded sample of rows, before anything touches the GPU.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from jarabotransientrendering_amd import scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# algorithmic HBM bytes per unit of work (wavefront formulation, path state round-trips
# through HBM between stages):
BYTES_EXTEND_PER_RAY = 32 + 16          # ray in, hit out
BYTES_SHADOW_PER_RAY = 32 + 1           # shadow ray in, visibility byte out
BYTES_FILM_PER_RECORD = 1               # visibility byte of every shadow record
BYTES_FILM_PER_CONTRIB = 20 + 48        # record in (RGB, pixel, bin) + FP32 RGB sum and compensation RMW
BYTES_SHADE_PER_SEGMENT = 64 + 64 + 20  # path state in + out, NEE record out
BYTES_PATH_PER_SEGMENT = 300            # whole segment: extension + shadow + state + record + film


def workload(name: str):
    if name == "icosphere6":
        s = scenes.icosphere_box(level=6, width=256, height=256, time_bins=1024, exposure=0.01, sqrt_spp=32,
                                 n_incoming=1, max_bounces=8, seed=0)
        desc = ("81,920-triangle displaced icosphere (subdivision 6, value-noise amplitude 0.05) in the open "
                "unit box, 0.3x0.3 area light; 256x256 px x 1024 time bins of 0.01; 1024 spp "
                "(32x32 camera samples x 1 path); max 8 bounces; FP32 film with compensated per-bin sums")
    elif name == "box":
        s = scenes.cornell_box(width=64, height=64, time_bins=256, exposure=0.01, sqrt_spp=32, n_incoming=1,
                               max_bounces=4, seed=0)
        desc = "22-triangle box, point light; 64x64 px x 256 bins of 0.01; 1024 spp; max 4 bounces"
    elif name == "media":
        s = scenes.media_box(width=512, height=512, scanline=256, time_bins=1024, exposure=0.01, sqrt_spp=64,
                             n_incoming=1, max_bounces=16, seed=0)
        desc = ("box filled with a homogeneous isotropic medium (sigma_s 0.5, sigma_a 0.05), point light; "
                "scanline 256 of 512x512, 1024 bins of 0.01; 4096 spp; max 16 events")
    elif name.startswith("batch64:"):
        k = int(name.split(":")[1])
        s = scenes.batch_scenes(1, k)[0]
        desc = f"job {k} of the 64-scene batched workload"
    else:
        raise SystemExit(f"unknown workload {name}")
    return s, desc


# ---------------------------------------------------------------------------------------------
# CPU baseline: the oracle on host cores, same workload scene and RNG scheme, bounded row sample

def _cpu_worker(args):
    name, rows, sqrt_spp, seed = args
    from oracle import pyoracle as po
    s, _ = workload(name)
    s.render.sqrt_spp = sqrt_spp
    _, _, st = po.render(s, po.RNG_PER_PATH, rows=rows, seed=seed)
    return st["paths"], st["render_seconds"]


def host_cores():
    """(processes the CPU baseline may run, description of the host's cores). The baseline uses
    every core this process may run on, up to the host's CPU share for one GPU when the
    environment states one (OMP_NUM_THREADS; a shared GPU host gives each GPU's jobs that many)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    procs = max(1, min(affinity, share) if share > 0 else affinity)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return procs, {"os_cpu_count": os.cpu_count(), "affinity": affinity, "cpu_share": share or None,
                   "processes": procs, "model": model}


def cpu_baseline(name: str, seconds: float):
    """The oracle on `procs` host cores for about `seconds`: each process renders a disjoint row
    range (or, for a scanline workload, the scanline with its own seed) of the same scene."""
    from oracle import pyoracle as po
    if not os.path.exists(po.LIB_PATH):
        po.build()
    s, _ = workload(name)
    procs, host = host_cores()
    H, W, n_inc = s.film.height, s.film.width, s.render.n_incoming
    scan = s.film.scanline
    y_mid = scan if scan >= 0 else H // 2
    paths1, sec1 = _cpu_worker((name, (y_mid, y_mid + 1), 2, 99))   # calibration
    rate1 = paths1 / max(sec1, 1e-9)
    rows_per_proc = 1 if scan >= 0 else max(1, H // procs)
    spp = max(1, int(round((rate1 * seconds / (rows_per_proc * W * n_inc)) ** 0.5)))
    jobs = []
    for p in range(procs):
        r0 = scan if scan >= 0 else p * rows_per_proc
        jobs.append((name, (r0, r0 + rows_per_proc), spp, p))
    ctx = mp.get_context("fork")  # no GPU state exists yet in this process
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    paths = sum(r[0] for r in res)
    per_core = paths / sum(r[1] for r in res)
    sample = (f"{procs} processes, each {rows_per_proc} row(s) x {W} px x {spp * spp} camera samples x {n_inc} "
              f"path(s) of the same scene (per-path RNG streams, as on the GPU): {paths} paths in {wall:.1f} s "
              f"wall; per core {per_core:.4g} paths/s of render loop")
    return {"value": paths / wall, "unit": "paths/s", "cores": procs, "kind": "port", "sample": sample,
            "value_per_core": per_core, "host": host, "calibration": CPU_CALIBRATION}


# The oracle's render loop against the reference binary (Embree BVH8 + AVX2, single-threaded), per
# core on the same machine (an 8-core 2.7 GHz Xeon), best of three runs each, render time only.
# Not calibrated on the 81,920-triangle scene: the reference builds only with Embree there (its
# non-Embree intersection tests every triangle), and Embree's CMake build is out of scope here.
CPU_CALIBRATION = {
    "box (22 triangles, 64x64 px, 16 camera samples x 64 paths, 4 bounces)":
        {"reference_paths_per_s": 1.89e6, "oracle_paths_per_s": 2.2e6, "ratio": 1.16},
    "media scanline (22 triangles + medium, 512 px, 16 x 16 paths, 16 events)":
        {"reference_paths_per_s": 0.63e6, "oracle_paths_per_s": 0.64e6, "ratio": 1.02},
    "icosphere6 (81,930 triangles)": "not calibrated",
}


# ---------------------------------------------------------------------------------------------
# Guided versus plain: time to reach a target relative MSE on the guided workload (the icosphere
# geometry filled with the homogeneous medium). Both estimators sample Lambertian directions in an
# orthonormal frame, so they converge to the same film; the reference film is a long plain render.

TRIM = 1e-3  # fraction of per-element errors left out of the trimmed relMSE


def guided_scene(width: int, height: int):
    s = scenes.icosphere_box(level=6, width=width, height=height, time_bins=1024, exposure=0.01, sqrt_spp=8,
                             n_incoming=1, max_bounces=16, seed=0)
    s.medium = [0.05, 0.05, 0.05, 0.5, 0.5, 0.5]
    return s


def rel_mse(x, r, eps_scale=1e-2, trim=0.0):
    """mean over (pixel, bin, channel) of (x - r)^2 / (r^2 + eps), eps = eps_scale * mean(r)^2;
    with trim > 0 the largest `trim` fraction of the per-element errors is left out."""
    eps = eps_scale * float(np.mean(r)) ** 2
    e = ((x.astype(np.float64) - r) ** 2 / (r.astype(np.float64) ** 2 + eps)).ravel()
    if trim > 0:
        k = int(len(e) * (1.0 - trim))
        e = np.partition(e, k)[:k]
    return float(np.mean(e))


def bins_sum(x, group=8):
    """Film [rows][W][T][3] with every `group` consecutive time bins summed."""
    r, w, t, c = x.shape
    return x[:, :, : t // group * group].reshape(r, w, t // group, group, c).sum(axis=3)


GUIDED_WORKLOADS = {
    # icosphere geometry + the medium, 128x128 px
    "icosphere_media": dict(ref_sqrt=32, ref_spp=262144, pass_sqrt=16, checkpoints=(4, 8, 16, 32, 64),
                            train_passes=8, train_sqrt=8, steps_per_wave=96, train_paths=1 << 18),
    # the 1024-bin media streak film: one 512-pixel scanline of the medium-filled box, point light
    "media_scanline": dict(ref_sqrt=64, ref_spp=1 << 22, pass_sqrt=64, checkpoints=(4, 8, 16, 32, 64),
                           train_passes=5, train_sqrt=32, steps_per_wave=64, train_paths=1 << 19, bsdf_fraction=0.5),
}
# Training schedules, swept with eight repeats each (median time-to-target speedup on the trimmed
# per-bin relMSE), with training steps that read the batch in pass order: media streak film 4 x 32
# Adam steps 1.25x, 4 x 48 1.42x (one repeat diverged), 4 x 64 1.39x, 4 x 96 1.48x (worst repeat
# 1.28x), 5 x 64 1.42-1.47x (worst 1.35x), 6 x 64 1.48x (worst 1.27x), 8 x 48 1.36x; icosphere
# 8 x 64 1.23x, 8 x 96 1.20x, 8 x 128 1.20x, 12 x 64 1.17x, 16 x 32 1.16x. Batches of 65,536
# records instead of 16,384 lost (0.64-0.78x at 4 x 8..32). With steps that read the batch through
# a bijection (spread over the pass): media 4 x 64 1.45x, 5 x 64 1.49-1.56x (worst 1.38x), 6 x 64
# 1.46x (one repeat 0.79x), 5 x 96 1.53x (worst 1.29x); icosphere 8 x 64 1.24x (two repeats below
# 1x), 8 x 96 1.33x (worst 1.20x), 6 x 64 1.28x.
# BSDF / phase-function fraction of the one-sample MIS mixture, swept over 0.3 / 0.5 / 0.7 with three
# training repeats each: 0.3 gives the icosphere workload its best trimmed per-bin relMSE, 0.5 the
# media streak film its best on every metric (a larger fraction also bounds each vertex's weight by
# albedo / fraction, so fewer rare high-weight paths reach the time-integrated image).


def guided_comparison(name="icosphere_media", ref_sqrt=32, ref_spp=262144, pass_sqrt=16, checkpoints=(4, 8, 16, 32, 64),
                      train_passes=8, train_sqrt=8, steps_per_wave=64, train_paths=1 << 18, bsdf_fraction=0.3,
                      learning_rate=3e-3, batch_size=1 << 14, repeats=3):
    """Guided versus plain rendering of the same film, timed. The guide first trains on
    `train_passes` passes of train_sqrt^2 spp (one wavefront each, the first `train_paths` paths of
    which record training data, `steps_per_wave` Adam steps after
    each); those passes are timed and their film discarded (an untrained network's samples would stay
    in the estimate). The network is then frozen and both estimators render passes of pass_sqrt^2 spp.
    Training is repeated `repeats` times with independent initialisations and training samples
    (a network's quality varies from one training run to the next); the reported speedups and error
    ratios are the medians over the repeats, each repeat's figures listed beside them.

    Metrics against a long plain render (ref_spp, FP32 film with compensated sums):
      rel_mse_trim   per (pixel, bin, channel) relMSE without the largest 0.1 % of errors (the target
                     metric: NEE at vertices near the light has unbounded variance, so the untrimmed
                     value is set by single samples, for both estimators alike);
      rel_mse        the same, untrimmed;
      rel_mse_bins8  trimmed, on the film with 8 consecutive time bins summed (less sensitive to
                     per-bin sparsity at low sample counts);
      rel_mse_steady untrimmed, on the time-integrated image."""
    from jarabotransientrendering_amd import renderer
    s = guided_scene(128, 128) if name == "icosphere_media" else workload("media")[0]
    world, cam, sampler, film, integ = renderer.setup_scene(s)
    eng = renderer.TransientRenderer()
    eng.orthonormal_frame = True
    # reference: plain passes of ref_sqrt^2 camera samples until ref_spp
    sampler.sqrt_spp = ref_sqrt
    t0 = time.perf_counter()
    n_ref = max(1, ref_spp // ref_sqrt ** 2)
    for k in range(n_ref):
        eng.render(None, world, integ, cam, film, sampler, seed=100000 + k)
    ref = film.read().astype(np.float64)
    ref_steady, ref_b8 = ref.sum(axis=2), bins_sum(ref)
    ref_sec = time.perf_counter() - t0
    sampler.sqrt_spp = pass_sqrt
    pass_paths = s.n_paths() // (s.render.sqrt_spp ** 2) * pass_sqrt ** 2

    def run(guide, rep=0):
        t_render, out, done = 0.0, [], 0
        if guide is not None:
            sampler.sqrt_spp = train_sqrt
            eng.wave_paths = pass_paths // pass_sqrt ** 2 * train_sqrt ** 2
            for k in range(train_passes):
                st = eng.render(None, world, integ, cam, film, sampler, seed=1000000 + 1000 * rep + k, guide=guide)
                t_render += st["render_ms"] * 1e-3
            guide.set_steps(0)
            sampler.sqrt_spp = pass_sqrt
        film.clean()
        eng.wave_paths = pass_paths if guide is not None else 0
        for cp in checkpoints:
            while done < cp:
                st = eng.render(None, world, integ, cam, film, sampler, seed=done, guide=guide)
                t_render += st["render_ms"] * 1e-3
                done += 1
            img = film.read()
            out.append({"passes": cp, "spp": cp * pass_sqrt ** 2, "seconds": t_render,
                        "rel_mse_trim": rel_mse(img, ref, trim=TRIM), "rel_mse": rel_mse(img, ref),
                        "rel_mse_bins8": rel_mse(bins_sum(img), ref_b8, trim=TRIM),
                        "rel_mse_steady": rel_mse(img.sum(axis=2), ref_steady)})
        return out

    # Untimed: one training pass and one frozen pass of a throw-away guide, so that the guided
    # kernels' first launches in this process (code loading, first-use set-up) fall outside the timed
    # region as the plain kernels' do (the reference render ran them). Each timed guide still
    # allocates its own training buffers inside its timed first pass.
    warm = renderer.PathGuide(steps_per_wave=1, bsdf_fraction=bsdf_fraction, learning_rate=learning_rate,
                              train_paths=train_paths, batch_size=batch_size, init_seed=999)
    sampler.sqrt_spp = train_sqrt
    eng.wave_paths = pass_paths // pass_sqrt ** 2 * train_sqrt ** 2
    eng.render(None, world, integ, cam, film, sampler, seed=999999, guide=warm)
    warm.set_steps(0)
    sampler.sqrt_spp = pass_sqrt
    eng.wave_paths = pass_paths
    eng.render(None, world, integ, cam, film, sampler, seed=999998, guide=warm)
    del warm

    plain = run(None)
    guided_runs, infos = [], []
    for rep in range(repeats):
        guide = renderer.PathGuide(steps_per_wave=steps_per_wave, bsdf_fraction=bsdf_fraction,
                                   learning_rate=learning_rate, train_paths=train_paths, batch_size=batch_size,
                                   init_seed=1 + rep)
        guided_runs.append(run(guide, rep))
        infos.append(guide.info())

    def time_to(curve, key, target):
        prev = None
        for c in curve:
            if c[key] <= target:
                if prev is None:
                    return c["seconds"] * c[key] / target  # reached at the first checkpoint: relMSE ~ 1/time
                # log-log interpolation between checkpoints
                a = (np.log(target) - np.log(prev[key])) / (np.log(c[key]) - np.log(prev[key]))
                return float(np.exp(np.log(prev["seconds"]) + a * (np.log(c["seconds"]) - np.log(prev["seconds"]))))
            prev = c
        # not reached: extrapolate with relMSE ~ 1/time from the last checkpoint
        return float(curve[-1]["seconds"] * curve[-1][key] / target)

    keys = ("rel_mse_trim", "rel_mse_bins8", "rel_mse_steady")
    per_rep = [{k: time_to(plain, k, plain[-1][k]) / time_to(g, k, plain[-1][k]) for k in keys} for g in guided_runs]
    ratio_rep = [{k: g[-1][k] / plain[-1][k] for k in keys} for g in guided_runs]
    speedups = {k: float(np.median([r[k] for r in per_rep])) for k in keys}
    ratios = {k: float(np.median([r[k] for r in ratio_rep])) for k in keys}
    key = "rel_mse_trim"
    # the repeat with the median trimmed-relMSE speedup
    mid = int(np.argsort([r[key] for r in per_rep])[len(per_rep) // 2])
    guided = guided_runs[mid]
    t_plain, t_guided = time_to(plain, key, plain[-1][key]), time_to(guided, key, plain[-1][key])
    if name == "icosphere_media":
        wdesc = (f"{s.n_tris}-triangle displaced icosphere in the open unit box filled with a homogeneous medium "
                 f"(sigma_a 0.05, sigma_s 0.5), 0.3x0.3 area light; 128x128 px x 1024 bins of 0.01; max 16 events")
    else:
        wdesc = ("box filled with a homogeneous isotropic medium (sigma_s 0.5, sigma_a 0.05), point light; "
                 "scanline 256 of 512x512, 1024 bins of 0.01; max 16 events")
    return {
        "workload": f"{name}: {wdesc}; passes of {pass_sqrt ** 2} spp ({pass_paths} paths)",
        "reference": {"spp": n_ref * ref_sqrt ** 2, "seconds": ref_sec, "estimator": "plain, orthonormal frame"},
        "rel_mse_eps": "1e-2 * mean(reference)^2",
        "guide_settings": {"bsdf_fraction": bsdf_fraction, "learning_rate": learning_rate,
                           "steps_per_wave": steps_per_wave, "batch_size": batch_size, "train_passes": train_passes,
                           "train_paths": train_paths,
                           "train_pass_spp": train_sqrt ** 2},
        "repeats": repeats,
        "error_ratio_at_equal_spp": ratios,
        "error_ratio_by_repeat": ratio_rep,
        "speedup_by_repeat": per_rep,
        "target_metric": "rel_mse_trim (largest 0.1 % of per-element errors left out)",
        "target_rel_mse": plain[-1][key],
        "seconds_to_target": {"plain": t_plain, "guided": t_guided},
        "speedup": t_plain / t_guided,
        "speedup_by_metric": speedups,
        "plain": plain,
        "guided": guided,
        "guide": infos[mid],
    }


# ---------------------------------------------------------------------------------------------
# Batched workload: 64 uneven scenes (seeds 0..63) rendered together by one launch sequence.

def cpu_baseline_batch(n_jobs: int, seconds: float):
    """The oracle on host cores over a bounded sample of the batched workload: process p renders
    the middle row of job p (of the first min(cores, n_jobs) jobs) at a reduced sample count."""
    from oracle import pyoracle as po
    if not os.path.exists(po.LIB_PATH):
        po.build()
    cores, host = host_cores()
    procs = max(1, min(cores, n_jobs))
    jobs = []
    for p in range(procs):
        s, _ = workload(f"batch64:{p}")
        H, W = s.film.height, s.film.width
        paths1, sec1 = _cpu_worker((f"batch64:{p}", (H // 2, H // 2 + 1), 1, 99))  # calibration
        rate = paths1 / max(sec1, 1e-9)
        spp = max(1, int(round((rate * seconds / W) ** 0.5)))
        jobs.append((f"batch64:{p}", (H // 2, H // 2 + 1), spp, p))
    ctx = mp.get_context("fork")  # no GPU state exists yet in this process
    t0 = time.perf_counter()
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    paths = sum(r[0] for r in res)
    per_core = paths / sum(r[1] for r in res)
    sample = (f"{procs} processes; process p renders the middle row of batched job p at a reduced sample count "
              f"(per-path RNG streams, as on the GPU): {paths} paths in {wall:.1f} s wall; per core "
              f"{per_core:.4g} paths/s of render loop")
    return {"value": paths / wall, "unit": "paths/s", "cores": procs, "kind": "port", "sample": sample,
            "value_per_core": per_core, "host": host, "calibration": CPU_CALIBRATION}


def progress(msg: str) -> None:
    """A progress line on stderr (the JSON result stays the only line on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def bench_batch(args) -> dict:
    n_jobs = 64
    cpu = None if args.no_cpu_baseline else cpu_baseline_batch(n_jobs, args.cpu_seconds)
    from jarabotransientrendering_amd import renderer
    t0 = time.perf_counter()
    scs = scenes.batch_scenes(n_jobs)
    setups = [renderer.setup_scene(s) for s in scs]
    build_s = time.perf_counter() - t0
    progress(f"batch: {n_jobs} scenes built in {build_s:.1f} s")
    b = renderer.Batch()
    for (w, cam, sampler, film, integ), s in zip(setups, scs):
        b.add(w, integ, cam, film, sampler, seed=s.render.seed)
    tuned, tune_ms = b.autotune()
    launch = dict(tuned)
    progress(f"batch: autotuned {launch}")
    for k in range(args.warmup):
        b.render(launch)
    for (_, _, _, film, _) in setups:
        film.clean()
    steps = []
    t1 = time.perf_counter()
    for k in range(args.steps):
        steps.append(b.render(launch))
    wall = time.perf_counter() - t1
    paths = sum(s["paths"] for s in steps)
    value = paths / wall
    # one profiled batch pass without graph replay for the per-kernel split
    prof = b.render({**launch, "flags": launch["flags"] & ~8 | 1})
    kern_ms, kern_n = prof["kernel_ms"], prof["kernel_launches"]
    alg = {"extend": BYTES_EXTEND_PER_RAY * prof["extension_rays"],
           "shadow": BYTES_SHADOW_PER_RAY * prof["shadow_rays"],
           "shade": BYTES_SHADE_PER_SEGMENT * prof["extension_rays"],
           "film": BYTES_FILM_PER_RECORD * prof["shadow_rays"] + BYTES_FILM_PER_CONTRIB * prof["contributions"],
           "camera": 48 * prof["paths"]}
    dom = max(kern_ms, key=lambda n: kern_ms[n])
    avg_ms = kern_ms[dom] / max(kern_n[dom], 1)
    bpl = alg[dom] / max(kern_n[dom], 1)
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    # the other batch mode (shared wavefronts <-> one job after another) with the same shape
    other = {**launch, "flags": launch["flags"] ^ 128}  # TR_RENDER_PER_JOB
    t3 = time.perf_counter()
    other_paths = b.render(other)["paths"]
    other_rate = other_paths / (time.perf_counter() - t3)
    progress("batch: timed and profiled renders done")
    small = small_jobs_comparison(renderer)
    progress("batch: small-jobs comparison done")
    # the same jobs one render call at a time
    eng = renderer.TransientRenderer()
    t2 = time.perf_counter()
    seq_paths = 0
    for (w, cam, sampler, film, integ), s in zip(setups, scs):
        seq_paths += eng.render(None, w, integ, cam, film, sampler, seed=s.render.seed)["paths"]
    seq_wall = time.perf_counter() - t2
    ext = sum(s["extension_rays"] for s in steps)
    path_gbs = BYTES_PATH_PER_SEGMENT * ext / wall / 1e9
    params = [scenes.batch_scene_params(k) for k in range(n_jobs)]
    return {
        "metric": "time-resolved paths/s on 1 MI355X",
        "value": value, "unit": "paths/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": ("batch64: 64 icosphere-in-box scenes (seeds 0..63; triangle target, film size, "
                                "time bins and spp drawn log-uniform per seed; nearest icosphere level), "
                                "max 8 bounces, rendered as one batched launch sequence"),
                   "paths_per_step": paths // args.steps,
                   "triangles_total": int(sum(s.n_tris for s in scs)),
                   "film_bins_total": int(sum(p["size"] ** 2 * p["time_bins"] for p in params)),
                   "parallelism": "single GPU", "launch": launch, "autotune_trial_ms": tune_ms,
                   "scene_build_s": build_s},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": avg_ms,
                     "launches": kern_n[dom], "algorithmic_bytes_per_launch": bpl},
        "path_roofline": {"bytes_per_segment": BYTES_PATH_PER_SEGMENT, "segments_per_path": ext / paths,
                          "achieved": path_gbs, "frac": path_gbs / HBM_PEAK_GBS},
        "kernels_ms": {k: round(v, 3) for k, v in kern_ms.items()},
        "sequential": {"paths_per_s": seq_paths / seq_wall, "seconds": seq_wall,
                       "note": "the same 64 jobs rendered one tr_render_transient call at a time"},
        "other_batch_mode": {"per_job": bool(other["flags"] & 128), "paths_per_s": other_rate},
        "small_jobs": small,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        # against one core of the same CPU baseline (its render-loop rate per process)
        "gpu_over_cpu_core": (value / cpu["value_per_core"]) if cpu else None,
    }


def small_jobs_comparison(renderer, n_jobs: int = 256, reps: int = 3) -> dict:
    """Many small jobs (icosphere level 3-4 in the box, 32x32 px, 128 bins, 16 spp, max 8 bounces):
    one shared sequence of wavefronts against one render call per job (the per-job batch mode
    shares wavefronts among jobs this small, so it is not listed)."""
    scs = []
    for k in range(n_jobs):
        s = scenes.icosphere_box(level=3 + k % 2, width=32, height=32, time_bins=128, exposure=0.08, sqrt_spp=4,
                                 n_incoming=1, max_bounces=8, seed=1000 + k)
        scs.append(s)
    setups = [renderer.setup_scene(s) for s in scs]
    b = renderer.Batch()
    for (w, cam, sampler, film, integ), s in zip(setups, scs):
        b.add(w, integ, cam, film, sampler, seed=s.render.seed)
    paths = sum(s.n_paths() for s in scs)

    def rate(fn):
        fn()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        return paths * reps / (time.perf_counter() - t)

    eng = renderer.TransientRenderer()
    shared = rate(lambda: b.render({"flags": 0}))
    shared_graph = rate(lambda: b.render({"flags": 8}))
    calls = rate(lambda: [eng.render(None, w, integ, cam, film, sampler, seed=s.render.seed)
                          for (w, cam, sampler, film, integ), s in zip(setups, scs)])
    return {"jobs": n_jobs, "paths_per_render": paths,
            "paths_per_s": {"shared_wavefronts": shared, "shared_wavefronts_graph": shared_graph,
                            "render_calls": calls}}


FLOP_PER_EVAL = 35840      # guiding network forward pass per sample (48-64-64-64-64-40, multiply-adds x 2)
FLOP_PER_TRAIN = 107520    # forward + backward (activation and weight gradients) per training record
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix-core peak


def guide_mlp_bench(n_eval: int = 1 << 22, reps: int = 5, train_sizes=(1 << 14, 1 << 18), train_reps: int = 8,
                    render_passes: int = 4) -> dict:
    """Throughput of the guiding network's matrix-core kernels: batched evaluation on device
    buffers (HIP events around the kernel), Adam steps on host-built record batches (wall time per
    step, record upload included), and guided render passes of the guided workload with a frozen
    network (the network evaluated inside the shading kernel)."""
    import ctypes
    import torch
    from jarabotransientrendering_amd import renderer, _native as nat
    rng = np.random.default_rng(0)
    g = renderer.PathGuide()
    raw = torch.rand((n_eval, 12), device="cuda", dtype=torch.float32)
    out = torch.empty((n_eval, 40), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    ms = ctypes.c_float()
    times = []
    for r in range(reps + 1):
        nat.check(nat.lib().tr_guide_eval_device(g.handle, ctypes.c_void_p(raw.data_ptr()), n_eval,
                                                 ctypes.c_void_p(out.data_ptr()), ctypes.byref(ms)))
        if r:
            times.append(ms.value)
    eval_ms = float(np.median(times))
    eval_tflops = n_eval * FLOP_PER_EVAL / (eval_ms * 1e-3) / 1e12
    res = {"eval": {"samples": n_eval, "kernel_ms": eval_ms, "tflops": eval_tflops,
                    "frac_bf16_peak": eval_tflops / BF16_PEAK_TFLOPS}}
    train = {}
    for b in train_sizes:
        gt = renderer.PathGuide(batch_size=b)
        rec = np.zeros((b, 16), np.float32)
        rec[:, 0:3] = rng.random((b, 3))
        rec[:, 3] = rng.random(b)
        om = rng.normal(size=(b, 3))
        rec[:, 4:7] = om / np.linalg.norm(om, axis=1, keepdims=True)
        rec[:, 8:11] = rec[:, 4:7]
        rec[:, 11] = rng.exponential(1.0, b)
        om = rng.normal(size=(b, 3))
        rec[:, 12:15] = om / np.linalg.norm(om, axis=1, keepdims=True)
        gt.train_batch(rec)
        t0 = time.perf_counter()
        for _ in range(train_reps):
            gt.train_batch(rec)
        dt = (time.perf_counter() - t0) / train_reps
        train[str(b)] = {"records": b, "step_ms_wall": dt * 1e3,
                         "tflops_wall": b * FLOP_PER_TRAIN / dt / 1e12}
    res["train"] = train
    # guided passes (frozen network) of the guided workload
    s = guided_scene(128, 128)
    world, cam, sampler, film, integ = renderer.setup_scene(s)
    eng = renderer.TransientRenderer()
    eng.orthonormal_frame = True
    eng.profile = True
    sampler.sqrt_spp = 16
    g.set_steps(0)
    kms = {}
    for k in range(render_passes):
        st = eng.render(None, world, integ, cam, film, sampler, seed=k, guide=g)
        for n, v in st["kernel_ms"].items():
            kms[n] = kms.get(n, 0.0) + v
    res["guided_render"] = {"passes": render_passes, "paths_per_pass": st["paths"],
                            "kernel_ms": {k: round(v, 3) for k, v in kms.items()}}
    return res


def copy_probe(nbytes: int, reps: int = 3) -> None:
    """Device-to-device copies of a known size (the HIP runtime's copy kernel), the calibration
    workload for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise SystemExit("copy probe: hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise SystemExit("copy probe: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="icosphere6")
    ap.add_argument("--wave", type=int, default=0, help="paths per wavefront (0 = library default)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--binary-bvh", action="store_true", help="traverse the binary BVH (default: 4-wide)")
    ap.add_argument("--launch-flags", default="",
                    help="extra TR_RENDER_* launch flags, comma-separated names (e.g. POSTPONE_LEAVES,STACK32)")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"))
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (counter calibration)")
    ap.add_argument("--guide-mlp", action="store_true",
                    help="only measure the guiding network's kernels (evaluation, training, guided shading)")
    ap.add_argument("--guided", choices=["quick", "on", "off", "only"], default="quick",
                    help="guided-versus-plain time-to-relMSE comparison in the line: quick (default; one "
                         "trained guide per workload, shorter reference renders), on (three guides, full "
                         "references), off, or only (print just that, full)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures a single MI355X; --gpus must be 1 (the multi-GPU split is a "
              "design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: run as a single process", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.guide_mlp:
        print(json.dumps({"metric": "guiding network throughput", "unit": "TFLOP/s", "dtype": "bf16 x bf16 -> f32",
                          "peak_tflops": BF16_PEAK_TFLOPS, "data": "synthetic", **guide_mlp_bench()}))
        return
    if args.guided == "only":
        res = {k: guided_comparison(k, **v) for k, v in GUIDED_WORKLOADS.items()}
        print(json.dumps({"metric": "seconds to target relative MSE, guided vs plain",
                          "value": res["icosphere_media"]["speedup"],
                          "unit": "x (plain time / guided time)", "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f32 tracer, bf16 network", "data": "synthetic", "guided": res["icosphere_media"],
                          "guided_media_scanline": res["media_scanline"]}))
        return
    if args.workload == "batch64":
        print(json.dumps(bench_batch(args)))
        return
    scene, desc = workload(args.workload)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.workload, args.cpu_seconds)
        progress(f"cpu baseline: {cpu['value']:.4g} paths/s on {cpu['cores']} processes")

    from jarabotransientrendering_amd import renderer, _native as nat
    world, cam, sampler, film, integ = renderer.setup_scene(scene)
    eng = renderer.TransientRenderer()
    eng.wave_paths = args.wave
    # the timed steps carry HIP events around the dominant kernel (extend) only: an event pair
    # between two launches drains the GPU, and timing every kernel would add 2-14 % to a step
    launch_flags = 0
    for name in filter(None, args.launch_flags.split(",")):
        launch_flags |= getattr(nat, "TR_RENDER_" + name.strip().upper())
    eng.profile = False
    eng.extra_flags = nat.TR_RENDER_PROFILE_EXTEND | launch_flags
    eng.binary_bvh = args.binary_bvh
    for k in range(args.warmup):
        eng.render(None, world, integ, cam, film, sampler, seed=1000 + k)
    film.clean()
    steps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        steps.append(eng.render(None, world, integ, cam, film, sampler, seed=k))
    wall = time.perf_counter() - t0
    paths = sum(s["paths"] for s in steps)
    value = paths / wall
    dom = "extend"
    dom_ms = sum(s["kernel_ms"][dom] for s in steps)
    dom_n = sum(s["kernel_launches"][dom] for s in steps)
    # per-kernel split: the same steps again, every kernel timed (after the timed region)
    eng.profile = True
    eng.extra_flags = launch_flags
    t1 = time.perf_counter()
    psteps = [eng.render(None, world, integ, cam, film, sampler, seed=k) for k in range(args.steps)]
    pwall = time.perf_counter() - t1
    kern_ms = {n: sum(s["kernel_ms"][n] for s in psteps) for n in psteps[0]["kernel_ms"]}
    kern_n = {n: sum(s["kernel_launches"][n] for s in psteps) for n in psteps[0]["kernel_launches"]}
    ext = sum(s["extension_rays"] for s in steps)
    shd = sum(s["shadow_rays"] for s in steps)
    con = sum(s["contributions"] for s in steps)
    alg_bytes = {
        "extend": BYTES_EXTEND_PER_RAY * ext,
        "shadow": BYTES_SHADOW_PER_RAY * shd,
        "film": BYTES_FILM_PER_RECORD * shd + BYTES_FILM_PER_CONTRIB * con,
        "shade": BYTES_SHADE_PER_SEGMENT * ext,
        "camera": 48 * paths,  # direction, path id + stream state, light stream + pixel + free flight
    }
    avg_ms = dom_ms / max(dom_n, 1)
    bytes_per_launch = alg_bytes[dom] / max(dom_n, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    limiter = None
    try:
        with open(args.traffic_json) as fh:
            tj = json.load(fh)
        key = f"{args.workload}:{dom}"
        if key in tj:
            traffic = tj[key]["bytes_per_launch"]
            # what bounds the kernel, from the same profile's SQ counters: the share of SIMD cycles
            # that issue a VALU instruction, and the active lanes per instruction (of 64)
            if "valu_issue_frac" in tj[key]:
                limiter = {"kind": "valu_issue", "valu_issue_frac": round(tj[key]["valu_issue_frac"], 3),
                           "active_lanes_per_valu": round(tj[key]["active_lanes_per_valu"], 1), "lanes": 64,
                           "source": "profiles/icosphere6_valu_counters.md"}
    except (OSError, ValueError, KeyError):
        traffic = None
    gpu_time_ms = sum(kern_ms.values())
    path_gbs = BYTES_PATH_PER_SEGMENT * ext / (wall) / 1e9
    out = {
        "metric": "time-resolved paths/s on 1 MI355X",
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "paths_per_step": paths // args.steps,
                   "triangles": scene.n_tris, "bvh": "binary" if args.binary_bvh else "4-wide",
                   "launch_flags": args.launch_flags or None, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": avg_ms, "launches": dom_n,
                     "algorithmic_bytes_per_launch": bytes_per_launch, "limiter": limiter},
        "path_roofline": {"bytes_per_segment": BYTES_PATH_PER_SEGMENT, "segments_per_path": ext / paths,
                          "achieved": path_gbs, "frac": path_gbs / HBM_PEAK_GBS},
        # per-kernel split from the same number of further steps with every kernel timed
        "kernels_ms": {k: round(v, 3) for k, v in kern_ms.items()},
        "kernels_ms_step_ms": pwall / args.steps * 1e3,
        # algorithmic bytes / kernel time for every stage (the film accumulation among them)
        "kernels_gbs": {k: {"achieved": round(alg_bytes[k] / (kern_ms[k] * 1e-3) / 1e9, 1),
                            "frac": round(alg_bytes[k] / (kern_ms[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                        for k in alg_bytes if kern_ms.get(k)},
        "kernel_time_share": {k: round(v / gpu_time_ms, 4) for k, v in kern_ms.items()} if gpu_time_ms else None,
        "cpu_baseline": cpu,
        "gpu_over_cpu": (value / cpu["value"]) if cpu else None,
        # against one core of the same CPU baseline (its render-loop rate per process)
        "gpu_over_cpu_core": (value / cpu["value_per_core"]) if cpu else None,
    }
    if args.guided in ("on", "quick"):
        quick = args.guided == "quick"
        t_g = time.perf_counter()
        res = {}
        for k, v in GUIDED_WORKLOADS.items():
            progress(f"guided comparison: {k}")
            kw = dict(v)
            if quick:
                # three trained guides, median reported: one guide in eight trains a weak network
                kw["repeats"] = 3
                kw["ref_spp"] = kw["ref_spp"] // 4
            res[k] = guided_comparison(k, **kw)
        # the headline half of the metric: time to the plain tracer's relMSE, guided vs plain, on the
        # 1024-bin media streak film and on the icosphere geometry in the medium
        out["guided"] = {
            "metric": "seconds to the plain tracer's relMSE after 64 passes, plain / guided (trimmed per-bin relMSE)",
            "mode": args.guided, "seconds": time.perf_counter() - t_g,
            "speedup": {k: r["speedup"] for k, r in res.items()},
            "speedup_by_metric": {k: r["speedup_by_metric"] for k, r in res.items()},
            "error_ratio_at_equal_spp": {k: r["error_ratio_at_equal_spp"] for k, r in res.items()},
            "target": ">= 2x (guided reaches plain's relMSE in <= 0.5x the wall time)",
            "detail": res,
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
