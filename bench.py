"""Benchmark of the per-frame synthesis hot path on one MI355X.

One step = one stylized 1024x1024 still frame through the full reference
schedule (device box pyramids of every guide, random init, 5 levels of
jump-flood propagation + random search with 7x7 guided patches, 8 iterations
per level, upsampling between levels), inputs already resident in HBM.
Prints one JSON line.  See DESIGN.md for the roofline figures used.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-levels L]
"""
import argparse
import ctypes as C
import json
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-levels", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the synthesis path runs on a single MI355X "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from painttween_amd import _ffi, Context
    from painttween_amd._ffi import FrameGuides, StyleGuides, PATCHXF_DTYPE

    style, frame, pk, iterations = workload()
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
    ctx._check(lib.pt_profile_enable(ctx.handle, 1))
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

    kstats = {}
    for name in ("propagate", "scatter", "update_residual", "canvas", "upsample", "random"):
        n, ms, px = C.c_uint64(), C.c_double(), C.c_uint64()
        ctx._check(lib.pt_profile_get(ctx.handle, name.encode(), C.byref(n), C.byref(ms), C.byref(px)))
        kstats[name] = dict(launches=n.value, total_ms=ms.value, pixels=px.value)
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
                     "unit": "GB/s", "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_ms": prop_avg_ms, "algorithmic_bytes_per_launch": prop_bytes},
        "valu": {"achieved": valu_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": valu_tflops / FP32_PEAK_TFLOPS, "flop_per_tap": FLOP_PER_TAP},
        "kernels": kstats,
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
