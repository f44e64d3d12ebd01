"""Benchmark of the per-frame synthesis hot path on one MI355X.

One step = one stylized 1024x1024 still frame through the full reference
schedule (device box pyramids of every guide, random init, 5 levels of
jump-flood propagation + random search with 7x7 guided patches, 8 iterations
per level, upsampling between levels), inputs already resident in HBM.
Beside it, on the same GPU in the same run:
  animation -- 24 frames 960x540, exemplar 512^2, FMBM_RM with forward and
             backward passes, temporal coherence 0.5, 4 levels, 3 iterations;
  batched -- 64 uneven single-frame jobs (128^2..1024^2, aspect 0.5..2,
             5x5/7x7 patches, 3-5 levels, 6 iterations) planned as one
             multi-job launch sequence, shapes autotuned, replayed as a HIP
             graph, then run as one persistent launch; compared per job,
             bit for bit, with the plain path;
  extended_terms -- a 512^2 still with the default patch distance, with
             object IDs, coherence or the LAB colour space alone and with
             all three; a temporal sequence with and without LAB and
             coherence;
  still256 -- the CPU-runnable still (256^2, 5x5 patches, 3 levels, 4
             iterations, no orientation): GPU frame time, the oracle on all
             host cores (median of 5) and on 1 core, offsets compared;
  served  -- 32 streamed sequences (256..1920 x 256..1080, exemplars
             512..1024, 7x7 patches, 4 levels, 3 iterations, temporal
             coherence 0.5) under Poisson frame arrivals, the server loop
             with 2 frames in flight, p50/p99 latency and frames/s against
             the plain per-frame path (a FIFO over service times measured
             before and after the rates).
Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-levels L]
                  [--no-batch] [--no-serve] [--serve-rates 4,20,100,400]
                  [--serve-in-flight 2]
"""
import argparse
import ctypes as C
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "stylized frames/sec + patch-dist evals/s at 1024²; p50/p99 latency, served vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_TAP = 102          # canonical per-tap FLOP count of the patch distance
NNF_BYTES_PER_PX = 96       # algorithmic HBM bytes per pixel per propagate step


def measured_traffic(kernel="propagate"):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain copy), or
    None when absent."""
    path = os.path.join(ROOT, "profiles", "traffic_still1024.md")
    try:
        for line in open(path):
            if line.startswith(kernel + "_traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def workload():
    from painttween_amd import workloads as W
    style = W.style_pair(1024, 1024, seed=2, guides=True)
    frame = W.still_frame(1024, 1024, seed=2, guides=True)
    params = dict(residual_window_size=3, levels=5, dist_trans_weight=0.5, input_analogy_weight=1.0,
                  orientation_enabled=1, canvas_output_weight=1.0)
    return style, frame, params, 8


def cpu_baseline(style, frame, params, iterations, cpu_levels, threads):
    """The oracle on the host cores over the coarsest `cpu_levels` levels of the
    same schedule; frames/s extrapolated by patch-distance count."""
    from tests.harness import Session
    from tests.oracle_ctx import OracleContext
    levels = params["levels"]
    ops = []
    for lv in range(levels - 1, levels - 1 - cpu_levels, -1):
        ops += [("style", lv), ("reset_adv",), ("frame", lv, 0)]
        ops += [("randomize",)] if lv == levels - 1 else [("upsample",), ("sync", 0)]
        ops += [("sync", 0)]
        for it in range(iterations):
            ops += [("propagate", j) for j in (128, 64, 32, 16, 8, 4, 2, 1, 1)]
            if it < iterations - 1:
                ops += [("scatter",)]
    with OracleContext(threads) as o:
        s = Session(o, style, [frame], o.default_params(**params))
        p0 = o.pde_count()
        t0 = time.perf_counter()
        s.run(ops)
        dt = time.perf_counter() - t0
        pdes = o.pde_count() - p0
    return dt, pdes


def _pool_map(fn, items):
    """Host-side input generation in worker processes (before any GPU use)."""
    n = max(1, min(16, os.cpu_count() or 1, len(items)))
    with mp.get_context("spawn").Pool(n) as pool:
        return pool.map(fn, items)


def _gen_batch(spec):
    from painttween_amd import workloads as W
    return W.batch_job(spec)


def _gen_seq(arg):
    from painttween_amd import workloads as W
    spec, frames = arg
    return W.serving_sequence(spec, frames)


BANDWIDTH_KINDS = ("canvas", "upsample", "random", "pyramid", "advect", "merge")


def host_cpu():
    """The host's logical CPU count and model name (reported beside the CPU
    timings)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"logical_cpus": os.cpu_count(), "model": model}


def profile_kinds(ctx, kinds):
    """HIP-event totals and algorithmic bytes of the recorded launches of
    each kernel class; achieved GB/s and the fraction of the HBM peak."""
    out = {}
    for name in kinds:
        n, ms, px, by = C.c_uint64(), C.c_double(), C.c_uint64(), C.c_uint64()
        ctx._check(ctx.lib.pt_profile_get(ctx.handle, name.encode(), C.byref(n), C.byref(ms), C.byref(px)))
        ctx._check(ctx.lib.pt_profile_bytes(ctx.handle, name.encode(), C.byref(by)))
        if n.value == 0:
            continue
        gbs = by.value / (ms.value * 1e-3) / 1e9 if ms.value > 0 else 0.0
        out[name] = {"launches": n.value, "total_ms": ms.value, "avg_us": ms.value * 1e3 / n.value,
                     "pixels": px.value, "bytes_per_launch": by.value / n.value, "achieved_gbs": gbs,
                     "hbm_frac": gbs / HBM_PEAK_GBS}
        # the launches at the largest level size alone (small levels are
        # bound by launch latency, not bandwidth)
        cnt = C.c_int()
        m = n.value
        pa, ba, ta = (C.c_uint64 * m)(), (C.c_uint64 * m)(), (C.c_float * m)()
        ctx._check(ctx.lib.pt_profile_records(ctx.handle, name.encode(), m, pa, ba, ta, C.byref(cnt)))
        top = max(pa[i] for i in range(m))
        sel = [i for i in range(m) if pa[i] == top]
        tb, tt = sum(ba[i] for i in sel), sum(ta[i] for i in sel)
        g = tb / (tt * 1e-3) / 1e9 if tt > 0 else 0.0
        out[name]["largest_level"] = {"pixels": int(top), "launches": len(sel), "avg_us": tt * 1e3 / len(sel),
                                      "bytes_per_launch": tb / len(sel), "achieved_gbs": g,
                                      "hbm_frac": g / HBM_PEAK_GBS}
    return out


def run_still256(ctx, inputs, cpu_runs, cpu_threads):
    """The 256^2 still (exemplar pair and target 256^2, 5x5 patches, 3 levels,
    4 iterations, orientation and distance transform off) on the GPU and,
    unless cpu_runs is 0, on the oracle with `cpu_threads` host threads
    (median of cpu_runs) and with one thread (one run).  The offsets of the
    two are compared byte for byte."""
    import torch
    from painttween_amd import _ffi
    from painttween_amd.forms import DeviceImages, DeviceOutputs
    style, frame = inputs
    kw = dict(residual_window_size=2, levels=3, orientation_enabled=0, dist_trans_weight=0.0)
    iterations = 4
    img = DeviceImages(0)
    fg = (_ffi.FrameGuides * 1)()
    fg[0] = img.frame(frame)
    sg = img.style(style)
    out = DeviceOutputs(256, 256, 1)
    params = ctx.default_params(**kw)
    o, c, r = out.pointers()

    def run():
        ctx._check(ctx.lib.pt_synthesize_device(ctx.handle, C.byref(params), _ffi.PT_S_FMBM_RM, iterations, 1, 256,
                                                256, 1, fg, 256, 256, C.byref(sg), o, c, r))

    run()
    torch.cuda.synchronize()
    ts = []
    p0 = ctx.pde_count()
    for _ in range(5):
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    pdes = (ctx.pde_count() - p0) // 5
    gpu_ms = float(np.median(ts)) * 1e3
    res = {"image": [256, 256], "exemplar": [256, 256], "patch": 25, "levels": 3, "iterations": iterations,
           "pde_per_frame": int(pdes), "gpu_ms": gpu_ms, "gpu_frames_per_s": 1e3 / gpu_ms}
    if cpu_runs > 0:
        from tests.oracle_ctx import OracleContext
        gpu_off = out.numpy()[0]
        cpu = {}
        for th, runs in ((cpu_threads, cpu_runs), (1, 1)):
            with OracleContext(th) as oc:
                tt = []
                for _ in range(runs):
                    t0 = time.perf_counter()
                    ref = oc.synthesize(oc.default_params(**kw), [frame], style, iterations=iterations)
                    tt.append(time.perf_counter() - t0)
            cpu[th] = float(np.median(tt))
        res.update({"cpu_threads": cpu_threads, "host": host_cpu(), "cpu_s": cpu[cpu_threads], "cpu_1thread_s": cpu[1],
                    "cpu_frames_per_s": 1.0 / cpu[cpu_threads], "cpu_1thread_frames_per_s": 1.0 / cpu[1],
                    "cpu_runs": cpu_runs, "gpu_vs_cpu": cpu[cpu_threads] / (gpu_ms * 1e-3),
                    "offsets_identical": bool(ref[0].tobytes() == gpu_off.tobytes())})
    return res


def run_ext_terms(ctx, steps=5):
    """A 512^2 still (exemplar 512^2, 7x7 patches, 4 levels, 4 iterations,
    orientation and distance-transform guides) with the default patch
    distance, with each extended term alone (object IDs 0.6, coherence 0.3,
    the LAB colour space) and with all three; GPU time per frame of each
    (median of `steps`), inputs resident on the device.  Then a temporal
    sequence (6 frames 256x192, exemplar 256^2, 3 levels, 3 iterations,
    temporal coherence 0.5) with and without LAB and coherence, timed through
    Context.synthesize (host arrays in and out, uploads included)."""
    import torch
    from painttween_amd import _ffi
    from painttween_amd import workloads as W
    from painttween_amd.forms import DeviceImages, DeviceOutputs
    style = dict(W.style_pair(512, 512, seed=6, guides=True), object_ids=W.object_ids(6, 30, 512, 512))
    frame = dict(W.still_frame(512, 512, seed=7, guides=True), object_id=W.object_ids(7, 31, 512, 512))
    base = dict(residual_window_size=3, levels=4, dist_trans_weight=0.5, orientation_enabled=1)
    cases = (("default", base), ("object_ids", dict(base, style_obj_id_weight=0.6)),
             ("coherence", dict(base, coherence_weight=0.3)), ("lab", dict(base, colorspace=1)),
             ("extended", dict(base, coherence_weight=0.3, colorspace=1, style_obj_id_weight=0.6)))
    img = DeviceImages(0)
    fg = (_ffi.FrameGuides * 1)()
    fg[0] = img.frame(frame)
    sg = img.style(style)
    out = DeviceOutputs(512, 512, 1)
    o, c, r = out.pointers()
    res = {"image": [512, 512], "exemplar": [512, 512], "patch": 49, "levels": 4, "iterations": 4,
           "terms": "extended: coherence 0.3, LAB colour space, object IDs 0.6"}
    for name, kw in cases:
        params = ctx.default_params(**kw)

        def run():
            ctx._check(ctx.lib.pt_synthesize_device(ctx.handle, C.byref(params), _ffi.PT_S_FMBM_RM, 4, 1, 512, 512, 1,
                                                    fg, 512, 512, C.byref(sg), o, c, r))

        run()
        torch.cuda.synchronize()
        p0 = ctx.pde_count()
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts)) * 1e3
        pdes = (ctx.pde_count() - p0) // steps
        res[name] = {"ms_per_frame": ms, "pde_per_frame": int(pdes), "pde_per_s": pdes / (ms * 1e-3)}
    res["extended_vs_default"] = res["extended"]["ms_per_frame"] / res["default"]["ms_per_frame"]
    frames = W.animation(256, 192, 6, seed=3)
    st = W.style_pair(256, 256, seed=4, guides=True)
    tkw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5)
    tres = {"frames": 6, "image": [256, 192], "levels": 3, "iterations": 3, "timing": "Context.synthesize, host I/O"}
    for name, kw in (("default", tkw), ("lab_coherence", dict(tkw, coherence_weight=0.3, colorspace=1))):
        params = ctx.default_params(**kw)
        ctx.synthesize(params, frames, st, iterations=3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.synthesize(params, frames, st, iterations=3)
        torch.cuda.synchronize()
        tres[name + "_ms"] = (time.perf_counter() - t0) * 1e3
    res["temporal"] = tres
    return res


def run_batched(ctx, jobs_in, steps, cpu_threads=0):
    """Plain path (one context per job, jobs one after another) against the
    batched form on the same device inputs."""
    import torch
    from painttween_amd import _ffi
    from painttween_amd.forms import Batch, DeviceOutputs
    lib = ctx.lib
    jobs = []
    for spec, (frame, style) in jobs_in:
        kw = dict(residual_window_size=spec["r"], levels=spec["levels"], dist_trans_weight=0.5,
                  input_analogy_weight=1.0, orientation_enabled=1, canvas_output_weight=1.0)
        jobs.append(dict(params=ctx.default_params(**kw), frames=[frame], style=style,
                         iterations=spec["iterations"], fixup_passes=0))
    b = Batch(jobs)
    # plain path on the same device inputs
    from painttween_amd import Context
    ctxs = [Context.gpu(0) for _ in jobs]
    plain_out = [DeviceOutputs(o.w, o.h, 1) for o in b.outputs]

    def plain_all():
        for i, j in enumerate(jobs):
            out = plain_out[i]
            o, c, r = out.pointers()
            ctxs[i]._check(lib.pt_synthesize_device(
                ctxs[i].handle, C.byref(j["params"]), _ffi.PT_S_FMBM_RM, j["iterations"], 0, out.w, out.h, 1,
                b._frames[i], jobs_in[i][0]["sw"], jobs_in[i][0]["sh"], C.pointer(b._styles[i]), o, c, r))

    plain_all()                       # warm: allocations
    torch.cuda.synchronize()
    pde0 = [c.pde_count() for c in ctxs]
    plain_times = []
    for k in range(steps):
        t0 = time.perf_counter()
        plain_all()
        torch.cuda.synchronize()
        plain_times.append(time.perf_counter() - t0)
    job_pdes = [(c.pde_count() - p0) // steps for c, p0 in zip(ctxs, pde0)]
    plain_s = float(np.median(plain_times))
    print("batched: plain path %.1f ms (median of %d) for %d jobs" % (plain_s * 1e3, steps, len(jobs)),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b.tune(1)
    tune_s = time.perf_counter() - t0
    b.run()                           # captures the graph
    times = []
    for k in range(steps):
        t0 = time.perf_counter()
        b.run()
        times.append(time.perf_counter() - t0)
        print("batched run %d: %.1f ms" % (k, times[-1] * 1e3), file=sys.stderr, flush=True)
    exact = all(torch.equal(b.outputs[i].off, plain_out[i].off) and
                torch.equal(b.outputs[i].canvas, plain_out[i].canvas) and
                torch.equal(b.outputs[i].res, plain_out[i].res) for i in range(len(jobs)))
    st = b.stats()
    px = sum(o.w * o.h for o in b.outputs)
    batch_s = float(np.median(times))
    out = {"jobs": len(jobs), "pixels": px, "jobs_per_s": len(jobs) / batch_s, "ms_per_batch": batch_s * 1e3,
           "runs": steps, "plain_jobs_per_s": len(jobs) / plain_s, "plain_ms": plain_s * 1e3,
           "plain_runs": steps, "speedup_vs_plain": plain_s / batch_s, "pde_per_batch": int(sum(job_pdes)),
           "bit_exact_vs_plain": bool(exact), "launches": st["launches"], "multi_launches": st["multi_launches"],
           "ticks": st["ticks"], "recorded_ops": st["ops"], "tuned_shapes": st["shapes"], "tune_s": tune_s,
           "hip_graph": True}
    # the persistent form: the whole batch in one kernel launch
    for o in b.outputs:
        o.off.zero_()
        o.canvas.zero_()
        o.res.zero_()
    b.set(_ffi.PT_BATCH_PERSISTENT, 1)
    t0 = time.perf_counter()
    b.tune(1)
    ptune_s = time.perf_counter() - t0
    b.run()
    ptimes = []
    for k in range(steps):
        t0 = time.perf_counter()
        b.run()
        ptimes.append(time.perf_counter() - t0)
        print("persistent run %d: %.1f ms" % (k, ptimes[-1] * 1e3), file=sys.stderr, flush=True)
    pexact = all(torch.equal(b.outputs[i].off, plain_out[i].off) and
                 torch.equal(b.outputs[i].canvas, plain_out[i].canvas) and
                 torch.equal(b.outputs[i].res, plain_out[i].res) for i in range(len(jobs)))
    pst = b.stats()
    pers_s = float(np.median(ptimes))
    out["persistent"] = {"ms_per_batch": pers_s * 1e3, "jobs_per_s": len(jobs) / pers_s, "runs": steps,
                         "speedup_vs_plain": plain_s / pers_s, "vs_graph": batch_s / pers_s,
                         "bit_exact_vs_plain": bool(pexact), "launches": 1, "grid": pst["persistent_grid"],
                         "tasks": pst["persistent_tasks"], "tiles": pst["persistent_tiles"],
                         "tiles_per_claim": pst["persistent_max_chunk"], "tune_s": ptune_s}
    b.close()
    for c in ctxs:
        c.close()
    if cpu_threads:
        out["cpu_baseline"] = batch_cpu_baseline(jobs_in, job_pdes, cpu_threads)
    return out


def batch_cpu_baseline(jobs_in, job_pdes, threads, budget_s=10.0):
    """The oracle on the host cores over whole jobs of the batch, smallest
    first, until about `budget_s` of CPU work; its rate in patch taps per
    second (patch distances times (2r+1)^2) extrapolates to the whole batch
    with the GPU's per-job patch-distance counts."""
    from tests.oracle_ctx import OracleContext
    order = sorted(range(len(jobs_in)), key=lambda i: job_pdes[i])
    taps = lambda i: (2 * jobs_in[i][0]["r"] + 1) ** 2
    done, spent, tap_sum = [], 0.0, 0
    with OracleContext(threads) as o:
        for i in order:
            spec, (frame, style) = jobs_in[i]
            kw = dict(residual_window_size=spec["r"], levels=spec["levels"], dist_trans_weight=0.5,
                      input_analogy_weight=1.0, orientation_enabled=1, canvas_output_weight=1.0)
            p0 = o.pde_count()
            t0 = time.perf_counter()
            o.synthesize(o.default_params(**kw), [frame], style, iterations=spec["iterations"], fixup_passes=0)
            spent += time.perf_counter() - t0
            tap_sum += (o.pde_count() - p0) * taps(i)
            done.append(i)
            if spent >= budget_s:
                break
    rate = tap_sum / spent
    total_taps = sum(job_pdes[i] * taps(i) for i in range(len(jobs_in)))
    batch_s = total_taps / rate
    return {"value": len(jobs_in) / batch_s, "unit": "jobs/s", "cores": threads, "kind": "port",
            "host": host_cpu(), "taps_per_s": rate,
            "sample": "oracle on the %d smallest jobs of the batch in full (%.1f s), jobs/s extrapolated to the "
                      "%d jobs by patch taps (the GPU's per-job patch-distance counts times (2r+1)^2)"
                      % (len(done), spent, len(jobs_in))}


def _gen_anim(arg):
    from painttween_amd import workloads as W
    w, h, frames, sw = arg
    return W.animation(w, h, frames, seed=3), W.style_pair(sw, sw, seed=3, guides=True)


def run_animation(ctx, anim_in, levels, iterations, cpu_threads=0):
    """The animation workload through pt_synthesize_device: FMBM_RM, two
    passes (forward then backward advection), temporal coherence 0.5."""
    import torch
    from painttween_amd import _ffi
    from painttween_amd.forms import DeviceImages, DeviceOutputs
    frames, style = anim_in
    h, w = frames[0]["color"].shape[:2]
    sh, sw = style["base"].shape[:2]
    F = len(frames)
    img = DeviceImages(0)
    fg = (_ffi.FrameGuides * F)()
    for i, f in enumerate(frames):
        fg[i] = img.frame(f)
    sg = img.style(style)
    out = DeviceOutputs(w, h, F)
    kw = dict(residual_window_size=3, levels=levels, temporal_coherence_weight=0.5, dist_trans_weight=0.5,
              input_analogy_weight=1.0, orientation_enabled=1, canvas_output_weight=1.0)
    params = ctx.default_params(**kw)
    o, c, r = out.pointers()

    def run():
        ctx._check(ctx.lib.pt_synthesize_device(ctx.handle, C.byref(params), _ffi.PT_S_FMBM_RM, iterations, 1, w, h,
                                                F, fg, sw, sh, C.byref(sg), o, c, r))

    # the warm run is recorded with HIP events (bandwidth-bound kernels:
    # advection and merges run only here); the timed run is not
    ctx._check(ctx.lib.pt_profile_enable(ctx.handle, 1))
    run()
    torch.cuda.synchronize()
    kinds = profile_kinds(ctx, BANDWIDTH_KINDS)
    ctx._check(ctx.lib.pt_profile_enable(ctx.handle, 0))
    pde0 = ctx.pde_count()
    t0 = time.perf_counter()
    run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    pdes = ctx.pde_count() - pde0
    res = out.res.cpu().numpy()
    assert np.isfinite(res).all() and (res >= 0).all(), "invalid residual"
    cpu = None
    if cpu_threads:
        cpu = anim_cpu_baseline(frames, style, kw, levels, iterations, cpu_threads, pdes / F)
    return {"bandwidth_kernels": kinds, "frames": F, "cpu_baseline": cpu, "image": [w, h], "exemplar": [sw, sh], "levels": levels, "iterations": iterations,
            "passes": 2, "seconds": dt, "frames_per_s": F / dt, "frame_passes_per_s": 2 * F / dt,
            "pde_per_s": pdes / dt, "schedule": "FMBM_RM, fixup passes 1 (forward then backward), 7x7 patches, "
                                                "temporal coherence 0.5, time derivatives 1"}


def anim_cpu_baseline(frames, style, kw, levels, iterations, threads, gpu_pde_per_frame):
    """The oracle on the host cores over a bounded sample of the animation:
    its first three frames cropped to their top-left 240x136 (velocities,
    guides and the 512^2 exemplar unchanged), the same schedule and
    parameters; its patch-distance rate (temporal terms included)
    extrapolates to frames/s of the full sequence with the GPU's count of
    patch distances per frame."""
    from tests.oracle_ctx import OracleContext
    cw, ch = 240, 136
    crop = [{k: (None if v is None else np.ascontiguousarray(v[:ch, :cw])) for k, v in f.items()} for f in frames[:3]]
    with OracleContext(threads) as o:
        p0 = o.pde_count()
        t0 = time.perf_counter()
        o.synthesize(o.default_params(**kw), crop, style, iterations=iterations, fixup_passes=1)
        dt = time.perf_counter() - t0
        pdes = o.pde_count() - p0
    rate = pdes / dt
    return {"value": rate / gpu_pde_per_frame, "unit": "frames/s", "cores": threads, "kind": "port",
            "host": host_cpu(), "pde_per_s": rate,
            "sample": "oracle on the first 3 frames cropped to 240x136 (%d patch distances with temporal terms in "
                      "%.1f s), frames/s extrapolated by the GPU's %.0f patch distances per 960x540 frame"
                      % (pdes, dt, gpu_pde_per_frame)}


def served_alone_1024(params, iterations, img, frames=4):
    """One 1024x1024 sequence (exemplar 1024^2, the served motion model)
    on an otherwise idle server: per-frame latency from submit to completion,
    beside the plain per-frame path's time for the same frames."""
    import torch
    from painttween_amd import Context, workloads as W
    from painttween_amd.forms import DeviceOutputs, Server, next_frame
    fr, st = W.serving_sequence(dict(w=1024, h=1024, s=1024, seed=99), frames)
    sg = img.style(st)
    fg = [img.frame(f) for f in fr]
    plain = []
    with Context.gpu(0) as c:
        # round 1 continues the sequence of round 0, as the server's resident
        # sequence does: both time the same four temporal frames
        for rnd in range(2):
            for f in fg:
                out = DeviceOutputs(1024, 1024, 1)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                next_frame(c, params, iterations, f, sg, 1024, 1024, out)
                if rnd == 1:
                    plain.append((time.perf_counter() - t0) * 1e3)
    srv = Server(200 << 30, 4)
    lat = []
    try:
        q = srv.open_device(params, iterations, 1024, 1024, sg, 1024, 1024)
        for rnd in range(2):  # round 0 makes the sequence resident
            for f in fg:
                t = srv.submit(q, f)
                srv.drain()
                for tk, ms in srv.poll(max_ticks=0, max_done=16):
                    if rnd == 1 and tk == t:
                        lat.append(ms)
            srv.outputs.clear()
    finally:
        srv.close()
    return {"frames": frames, "served_ms": lat, "plain_ms": plain,
            "served_mean_ms": float(np.mean(lat)) if lat else None, "plain_mean_ms": float(np.mean(plain))}


def run_served(ctx, seqs_in, frames, rates, budget_gb=200.0, cpu_pde_rate=None, in_flight=2):
    """The serving loop under Poisson arrivals against the plain per-frame
    path (one context per sequence, frames one at a time in arrival order).
    The plain path is a FIFO model over each frame's measured service time:
    the mean of a timed round before the rates and one after them, so that a
    drift of the GPU's speed over the leg weighs on both sides alike.  The
    server keeps `in_flight` frames in flight (measured on this mix: 1 and 2
    give the FIFO's latencies, 32 a higher p50; frames of 256^2 and more fill
    most of the GPU alone)."""
    import torch
    from painttween_amd import Context
    from painttween_amd.forms import DeviceImages, DeviceOutputs, Server, next_frame
    iterations = 3
    kw = dict(residual_window_size=3, levels=4, temporal_coherence_weight=0.5, dist_trans_weight=0.5,
              input_analogy_weight=1.0, orientation_enabled=1, canvas_output_weight=1.0)
    params = ctx.default_params(**kw)
    img = DeviceImages(0)
    dev = []
    for spec, (fr, st) in seqs_in:
        dev.append(dict(w=spec["w"], h=spec["h"], sw=spec["s"], sh=spec["s"], style=img.style(st),
                        frames=[img.frame(f) for f in fr]))
    N = len(dev)
    order = [(k, f) for f in range(frames) for k in range(N)]       # arrival order
    torch.cuda.synchronize()
    # plain per-frame path
    pctx = [Context.gpu(0) for _ in range(N)]
    first_round = {}

    def plain_round(keep=None):
        times = []
        for (k, f) in order:
            d = dev[k]
            out = DeviceOutputs(d["w"], d["h"], 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            next_frame(pctx[k], params, iterations, d["frames"][f], d["style"], d["sw"], d["sh"], out)
            times.append(time.perf_counter() - t0)
            if keep is not None:
                keep[(k, f)] = out
        return times

    plain_round(first_round)      # round 0 warms allocations (and is the bit-exactness reference)
    pde0 = sum(c.pde_count() for c in pctx)
    service_before = plain_round()
    served_pdes = sum(c.pde_count() for c in pctx) - pde0
    print("served: plain per-frame path %.1f ms for %d frames" % (sum(service_before) * 1e3, len(service_before)),
          file=sys.stderr, flush=True)
    # server: one warm round (compared with the plain round 0), then the rates
    srv = Server(int(budget_gb * (1 << 30)), min(N, in_flight))
    qs = [srv.open_device(params, iterations, d["w"], d["h"], d["style"], d["sw"], d["sh"]) for d in dev]
    tickets = {}
    for (k, f) in order:
        tickets[srv.submit(qs[k], dev[k]["frames"][f])] = (k, f)
    srv.drain()
    exact = True
    for t, (k, f) in tickets.items():
        o = srv.outputs[t]
        p = first_round[(k, f)]
        exact = exact and torch.equal(o.off, p.off) and torch.equal(o.canvas, p.canvas) and torch.equal(o.res, p.res)
    srv.outputs.clear()
    while srv.poll(max_ticks=0, max_done=1024):   # completions of the warm round
        pass
    rng = np.random.default_rng(1234)
    per_rate = []
    for lam in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(order)))
        arrivals -= arrivals[0]
        outs = {}
        lat = []
        t0 = time.perf_counter()
        i = 0
        done = 0
        last = t0
        while done < len(order):
            now = time.perf_counter() - t0
            while i < len(order) and arrivals[i] <= now:
                k, f = order[i]
                t = srv.submit(qs[k], dev[k]["frames"][f])
                outs[t] = (k, f)
                i += 1
            got = srv.poll(max_ticks=4)
            for t, ms in got:
                if t in outs:
                    lat.append(ms)
                    done += 1
                    last = time.perf_counter()
            if not got and i < len(order) and srv.stats()["in_flight"] == 0:
                time.sleep(max(0.0, min(0.002, arrivals[i] - (time.perf_counter() - t0))))
        span = last - t0
        srv.outputs.clear()
        per_rate.append((lam, arrivals, span, lat))
        print("served at %g fps offered: %.2f fps, p50 %.0f ms, p99 %.0f ms" % (
            lam, len(order) / span, np.percentile(lat, 50), np.percentile(lat, 99)), file=sys.stderr, flush=True)
    st = srv.stats()
    srv.close()
    service_after = plain_round()
    for c in pctx:
        c.close()
    service = [(a + b) / 2 for a, b in zip(service_before, service_after)]
    print("served: plain per-frame path %.1f ms for %d frames after the rates (model: mean %.1f ms)" % (
        sum(service_after) * 1e3, len(service_after), sum(service) * 1e3), file=sys.stderr, flush=True)
    rows = []
    for lam, arrivals, span, lat in per_rate:
        # plain FIFO single server with the measured service times
        c_prev, plat = 0.0, []
        for a, s_ in zip(arrivals, service):
            c_prev = max(a, c_prev) + s_
            plat.append((c_prev - a) * 1e3)
        pspan = c_prev
        r = {"rate_fps": lam, "frames": len(order), "served_fps": len(order) / span,
             "served_p50_ms": float(np.percentile(lat, 50)), "served_p99_ms": float(np.percentile(lat, 99)),
             "plain_fps": len(order) / pspan, "plain_p50_ms": float(np.percentile(plat, 50)),
             "plain_p99_ms": float(np.percentile(plat, 99))}
        r["served_vs_plain_fps"] = r["served_fps"] / r["plain_fps"]
        rows.append(r)
        print("served at %g fps offered: %.2f fps, p50 %.0f ms, p99 %.0f ms (plain %.2f fps, p50 %.0f ms, "
              "p99 %.0f ms)" % (lam, r["served_fps"], r["served_p50_ms"], r["served_p99_ms"], r["plain_fps"],
                                r["plain_p50_ms"], r["plain_p99_ms"]), file=sys.stderr, flush=True)
    per_rate = rows
    alone = served_alone_1024(params, iterations, img)
    px = sum(d["w"] * d["h"] for d in dev)
    return {"sequences": N, "frames_per_sequence_per_round": frames, "pixels_per_round": px * frames,
            "plain_fps": len(service) / sum(service), "plain_ms_before": sum(service_before) * 1e3,
            "plain_ms_after": sum(service_after) * 1e3, "in_flight": min(N, in_flight),
            "bit_exact_vs_plain": bool(exact), "rates": per_rate,
            "budget_gb": budget_gb, "resident_bytes": st["resident_bytes"], "launches": st["launches"],
            "multi_launches": st["multi_launches"], "ticks": st["ticks"],
            "schedule": "per-frame FMBM_RM first pass, 7x7 patches, 4 levels, 3 iterations, temporal coherence 0.5",
            "alone_1024": alone, "fill_tiles": 2048,
            "pde_per_frame": served_pdes / len(service),
            "cpu_baseline": None if not cpu_pde_rate else {
                "value": cpu_pde_rate["pde_per_s"] / (served_pdes / len(service)), "unit": "frames/s",
                "cores": cpu_pde_rate["cores"], "kind": "port",
                "sample": "the animation leg's oracle sample (7x7 patches with temporal terms, %.3g patch "
                          "distances/s), extrapolated by the served frames' mean patch-distance count"
                          % cpu_pde_rate["pde_per_s"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-levels", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-anim", action="store_true")
    ap.add_argument("--no-still256", action="store_true")
    ap.add_argument("--no-ext", action="store_true", help="skip the extended patch-distance terms leg")
    ap.add_argument("--anim-levels", type=int, default=4)
    ap.add_argument("--anim-iterations", type=int, default=3)
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve-seqs", type=int, default=32)
    ap.add_argument("--serve-frames", type=int, default=3)
    ap.add_argument("--serve-rates", default="4,20,100,400")
    ap.add_argument("--serve-in-flight", type=int, default=2, help="frames the server keeps in flight")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 256 MiB five times (calibrates rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE against known bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 4, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the synthesis path runs on a single MI355X "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from painttween_amd import _ffi, Context
    from painttween_amd._ffi import FrameGuides, StyleGuides, PATCHXF_DTYPE

    style, frame, pk, iterations = workload()
    # inputs of the batched and served forms, generated before any GPU use
    from painttween_amd import workloads as W
    batch_in = serve_in = anim_in = None
    still256_in = None if args.no_still256 else (W.style_pair(256, 256, seed=1), W.still_frame(256, 256, seed=1))
    if not args.no_anim:
        anim_in = _pool_map(_gen_anim, [(960, 540, 24, 512)])[0]
        print("animation inputs generated", file=sys.stderr, flush=True)
    if not args.no_batch:
        specs = W.batch_specs(args.batch_jobs)
        batch_in = list(zip(specs, _pool_map(_gen_batch, specs)))
        print("batched inputs generated", file=sys.stderr, flush=True)
    if not args.no_serve:
        sspecs = W.serving_specs(args.serve_seqs)
        serve_in = list(zip(sspecs, _pool_map(_gen_seq, [(sp, args.serve_frames) for sp in sspecs])))
        print("served inputs generated", file=sys.stderr, flush=True)
    h, w = frame["color"].shape[:2]
    sh, sw = style["base"].shape[:2]
    dev = torch.device("cuda:0")

    def dptr(a):
        if a is None:
            return None, _ffi.FP()
        t = torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
        return t, C.cast(C.c_void_p(t.data_ptr()), _ffi.FP)

    keep = []

    def dp(a):
        t, p = dptr(a)
        keep.append(t)
        return p

    fg = (FrameGuides * 1)()
    fg[0] = FrameGuides(dp(frame["color"]), dp(frame.get("orientation")), dp(frame.get("dist_trans")),
                        dp(None), dp(None), dp(None))
    sg = StyleGuides(dp(style["base"]), dp(style["output"]), dp(style.get("orientation")),
                     dp(style.get("dist_trans")))
    out_off = torch.empty(h * w * PATCHXF_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    out_canvas = torch.empty(h * w * 4, dtype=torch.float32, device=dev)
    out_res = torch.empty(h * w, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    ctx = Context.gpu(0)
    lib = ctx.lib
    params = ctx.default_params(**pk)

    def step():
        rc = lib.pt_synthesize_device(ctx.handle, C.byref(params), _ffi.PT_S_FMBM_RM, iterations, 1, w, h, 1, fg,
                                      sw, sh, C.byref(sg), C.c_void_p(out_off.data_ptr()),
                                      C.c_void_p(out_canvas.data_ptr()), C.c_void_p(out_res.data_ptr()))
        ctx._check(rc)

    for i in range(args.warmup):
        step()
        print("warmup step %d done" % i, file=sys.stderr, flush=True)
    pde0 = ctx.pde_count()
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        step()
        lat.append(time.perf_counter() - ts)
        print("timed step %d: %.1f ms" % (i, lat[-1] * 1e3), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    pdes = ctx.pde_count() - pde0

    # per-kernel HIP-event timings from one more, untimed frame (the events
    # around every launch would otherwise add to the timed frames)
    ctx._check(lib.pt_profile_enable(ctx.handle, 1))
    step()
    kstats = {}
    for name in ("propagate", "scatter", "update_residual", "canvas", "upsample", "random"):
        n, ms, px = C.c_uint64(), C.c_double(), C.c_uint64()
        ctx._check(lib.pt_profile_get(ctx.handle, name.encode(), C.byref(n), C.byref(ms), C.byref(px)))
        kstats[name] = dict(launches=n.value, total_ms=ms.value, pixels=px.value)
    still_bw = profile_kinds(ctx, BANDWIDTH_KINDS)
    ctx._check(lib.pt_profile_enable(ctx.handle, 0))

    # sanity of the produced frame (not timed)
    res = out_res.cpu().numpy()
    assert np.isfinite(res).all() and (res >= 0).all(), "invalid residual"

    fps = args.steps / elapsed
    pde_s = pdes / elapsed
    prop = kstats["propagate"]
    prop_avg_ms = prop["total_ms"] / max(prop["launches"], 1)
    prop_px_per_launch = prop["pixels"] / max(prop["launches"], 1)
    prop_bytes = NNF_BYTES_PER_PX * prop_px_per_launch
    achieved_gbs = prop_bytes / (prop_avg_ms * 1e-3) / 1e9
    taps = (2 * pk["residual_window_size"] + 1) ** 2
    valu_tflops = pde_s * taps * FLOP_PER_TAP / 1e12
    lat_ms = np.array(lat) * 1e3

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        dt, cpdes = cpu_baseline(style, frame, pk, iterations, args.cpu_levels, threads)
        pde_per_frame = pdes / args.steps
        cpu = {"value": (cpdes / dt) / pde_per_frame, "unit": "frames/s", "cores": threads, "kind": "port",
               "host": host_cpu(),
               "pde_per_s": cpdes / dt,
               "sample": "oracle (CPU restatement), coarsest %d levels of the same 1024^2 schedule: %d PDEs in "
                         "%.1f s, frames/s extrapolated by the %d PDEs of a full frame"
                         % (args.cpu_levels, cpdes, dt, int(pde_per_frame))}

    line = {
        "metric": METRIC, "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "still frame 1024x1024, exemplar pair 1024x1024, 7x7 patches, 5 levels, "
                               "8 iterations/level, guides colour 1.0 / orientation on / distance transform 0.5, "
                               "FMBM_RM schedule, inputs resident in HBM",
                   "image": [w, h], "exemplar": [sw, sh], "patch": taps, "levels": pk["levels"],
                   "iterations": iterations},
        "pde_per_s": pde_s,
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "roofline": {"bound": "hbm", "kernel": "propagate", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "traffic_source": "profiles/traffic_still1024.md (rocprofv3 FETCH_SIZE+WRITE_SIZE, "
                                       "calibrated by a plain copy)",
                     "avg_launch_ms": prop_avg_ms, "algorithmic_bytes_per_launch": prop_bytes,
                     "binding_unit": "not HBM: the windowed short-jump launches are VALU-issue bound (83% at "
                                     "level 0), the long jumps bound by the gathers' address path (TA busy 69%); "
                                     "profiles/counters_propagate_still1024.md"},
        "valu": {"achieved": valu_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": valu_tflops / FP32_PEAK_TFLOPS, "flop_per_tap": FLOP_PER_TAP},
        "kernels": kstats,
        "bandwidth_kernels": still_bw,
        "cpu_baseline": cpu,
    }
    if still256_in is not None:
        line["still256"] = run_still256(ctx, still256_in, 0 if args.no_cpu else 5, min(os.cpu_count() or 1, 16))
        print("still256: %.2f ms on the GPU" % line["still256"]["gpu_ms"], file=sys.stderr, flush=True)
    if not args.no_ext:
        line["extended_terms"] = run_ext_terms(ctx)
        et = line["extended_terms"]
        print("extended terms (ms per 512^2 frame): default %.1f, object IDs %.1f, coherence %.1f, LAB %.1f, all %.1f;"
              " temporal 6 frames %.0f ms, with LAB and coherence %.0f ms" % (
                  et["default"]["ms_per_frame"], et["object_ids"]["ms_per_frame"], et["coherence"]["ms_per_frame"],
                  et["lab"]["ms_per_frame"], et["extended"]["ms_per_frame"], et["temporal"]["default_ms"],
                  et["temporal"]["lab_coherence_ms"]), file=sys.stderr, flush=True)
    if anim_in is not None:
        line["animation"] = run_animation(ctx, anim_in, args.anim_levels, args.anim_iterations,
                                          0 if args.no_cpu else min(os.cpu_count() or 1, 16))
        # advection and merges run only in the animation: their HBM rooflines
        # come from its (untimed, event-recorded) warm run
        for k in ("advect", "merge"):
            if k in line["animation"]["bandwidth_kernels"]:
                line["bandwidth_kernels"][k] = line["animation"]["bandwidth_kernels"][k]
        print("animation: %.2f frames/s" % line["animation"]["frames_per_s"], file=sys.stderr, flush=True)
    if batch_in is not None:
        line["batched"] = run_batched(ctx, batch_in, args.batch_steps,
                                      0 if args.no_cpu else min(os.cpu_count() or 1, 16))
    if serve_in is not None:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        anim_cpu = (line.get("animation") or {}).get("cpu_baseline")
        line["served"] = run_served(ctx, serve_in, args.serve_frames, rates, cpu_pde_rate=anim_cpu,
                                    in_flight=args.serve_in_flight)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
