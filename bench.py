#!/usr/bin/env python3
"""Benchmark: APR conversion (get_apr + sample_image) of a 2048x2048x512 uint16
volume on one MI355X, through libaprh's C-ABI.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

One step = APRConverter<uint16_t>::get_apr on the device-resident volume
followed by ParticleData<uint16_t>::sample_image of the same volume.  The input
(background 100, 200k Gaussian blobs with sigma U(1.5,8) px and amplitude
U(300,4000), shot noise, seed 1) is synthesised on the device before timing.
APR parameters: rel_error 0.1, lambda 3, sigma_th 50, Ip_th 120, grad_th 10.

Prints one JSON line (stdout).  Progress goes to stderr.
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

METRIC = "Gvoxels/s get_apr (2048x2048x512 u16) + RL time-to-tol, warm vs plain"
HBM_PEAK_GBS = 8000.0
SHAPE_ZXY = (512, 2048, 2048)
N_BLOBS = 200_000
SEED = 1
SIGMA = (1.5, 8.0)
AMP = (300.0, 4000.0)
BACKGROUND = 100.0
APR_PARAMS = dict(rel_error=0.1, lambda_=3.0, sigma_th=50.0, Ip_th=120.0, grad_th=10.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(stage, N, Nd, P, Npad):
    """Algorithmic HBM bytes of one launch group, from the survey's work model
    (bytes a minimal implementation must move)."""
    return {
        "offset": 2 * N,                      # read the input once (min/max)
        "smooth_bspline_y": 4 * N,            # read input, write smoothed u16
        "smooth_bspline_x": 4 * N,            # read + write u16 in place
        "smooth_bspline_z": 4 * N,
        "gradient_ds": 2 * N + 6 * Nd,        # read u16, write u16 grad + f32 mean
        "inv_bspline": 3 * 8 * Nd,            # three f32 passes
        "local_intensity_scale": 6 * 8 * Npad,
        "sample_image": 2 * N + N // 4 + 4 * P,
    }.get(stage)


def run_gpu(args):
    import libapr_amd as A
    ctx = A.Context(0, timing=True)
    L = A.lib()
    z, x, y = SHAPE_ZXY
    N = x * y * z
    d_img = C.c_void_p()
    A._check(L.aprh_malloc(ctx.h, N * 2, C.byref(d_img)))
    log(f"[bench] synthesising {y}x{x}x{z} u16 volume on the device ({N_BLOBS} blobs)")
    t = time.time()
    A.synth_blobs(SHAPE_ZXY, SEED, N_BLOBS, sigma=SIGMA, amp=AMP, background=BACKGROUND, noise=True,
                  out=d_img.value, ctx=ctx)
    log(f"[bench] synthesised in {time.time() - t:.2f} s")
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(**APR_PARAMS)
    pd = A.ParticleData()
    d_parts = None
    P = None

    def step():
        nonlocal d_parts, P
        apr = conv.get_apr(d_img.value, dims=(y, x, z))
        st = ctx.stage_times()
        if d_parts is None:
            P = apr.total
            d_parts = C.c_void_p()
            A._check(L.aprh_malloc(ctx.h, max(P, 1) * 2, C.byref(d_parts)))
        pd.sample_image(apr, d_img.value, out=d_parts.value)
        st2 = ctx.stage_times()
        st["sample_image"] = st2["sample_image"]
        apr.close()
        return st

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{args.warmup} done (P={P})")
    stage_acc = {}
    A._check(L.aprh_memcpy(ctx.h, None, None, 0))
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = step()
        for k, v in st.items():
            stage_acc[k] = stage_acc.get(k, 0.0) + v
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    log(f"[bench] {args.steps} steps: {ms_per_step:.2f} ms/step, P={P}")
    stage_avg = {k: v / args.steps for k, v in stage_acc.items()}
    for k, v in sorted(stage_avg.items(), key=lambda kv: -kv[1]):
        log(f"[bench]   {k:24s} {v:8.3f} ms")
    yd, xd, zd = (y + 1) // 2, (x + 1) // 2, (z + 1) // 2
    Nd = yd * xd * zd
    Npad = (yd + 6) * (xd + 6) * (zd + 6)
    # dominant launch group among those with a fixed algorithmic-byte model
    cands = {k: v for k, v in stage_avg.items() if algorithmic_bytes(k, N, Nd, P, Npad) and v > 0}
    dom = max(cands, key=cands.get)
    ab = algorithmic_bytes(dom, N, Nd, P, Npad)
    achieved = ab / (cands[dom] / 1000.0) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": ab,
                "avg_ms": round(cands[dom], 4)}
    total_ab = 23 * N + 6 * P
    ctx.close()
    return ms_per_step, P, stage_avg, roofline, total_ab


def run_cpu_baseline():
    """Reference CPU path (oracle/_ref build with the reference's Release flags;
    else the oracle restatement) on a bounded sample of the same workload:
    a 1024x1024x256 volume from the same generator at the same blob density."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    import libapr_amd as A
    kind = "reference"
    lib = oracle_lib.load("ref_rel")
    if lib is None:
        kind = "port"
        lib = oracle_lib.load("oracle")
    if lib is None:
        return None
    shape = (256, 1024, 1024)
    nb = N_BLOBS * (shape[0] * shape[1] * shape[2]) // (SHAPE_ZXY[0] * SHAPE_ZXY[1] * SHAPE_ZXY[2])
    ctx = A.Context(0)
    L = A.lib()
    n = shape[0] * shape[1] * shape[2]
    d = C.c_void_p()
    A._check(L.aprh_malloc(ctx.h, n * 2, C.byref(d)))
    A.synth_blobs(shape, SEED, nb, sigma=SIGMA, amp=AMP, background=BACKGROUND, noise=True, out=d.value, ctx=ctx)
    img = np.zeros(shape, np.uint16)
    A._check(L.aprh_memcpy(ctx.h, C.c_void_p(img.ctypes.data), d, n * 2))
    A._check(L.aprh_free(ctx.h, d))
    ctx.close()
    par = oracle_lib.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    threads = lib.orc_num_threads()
    log(f"[bench] cpu baseline ({kind}, {threads} threads) on {shape[2]}x{shape[1]}x{shape[0]}")
    t = time.perf_counter()
    apr = oracle_lib.get_apr(img, par, "ref_rel" if kind == "reference" else "oracle", keep=False)
    apr.sample(img)
    dt = time.perf_counter() - t
    log(f"[bench] cpu baseline: {dt:.2f} s, P={apr.total}")
    return {"value": round(n / dt / 1e9, 4), "unit": "Gvoxels/s", "cores": threads, "kind": kind,
            "sample": f"get_apr+sample_image on a {shape[2]}x{shape[1]}x{shape[0]} u16 volume of the same generator "
                      f"({nb} blobs, seed {SEED}), {dt:.2f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    ms, P, stages, roofline, total_ab = run_gpu(args)
    z, x, y = SHAPE_ZXY
    N = x * y * z
    cpu = None if args.no_cpu else run_cpu_baseline()
    out = {
        "metric": METRIC,
        "value": round(N / (ms / 1000.0) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {"workload": "get_apr + sample_image, 2048x2048x512 u16, 200k blobs sigma 1.5-8 amp 300-4000, "
                               "bg 100 + shot noise, seed 1; rel_error 0.1 lambda 3 sigma_th 50 Ip_th 120 grad_th 10",
                   "particles": P, "pipeline_algorithmic_bytes": total_ab,
                   "pipeline_GBps": round(total_ab / (ms / 1000.0) / 1e9, 1),
                   "stage_ms": {k: round(v, 3) for k, v in stages.items()},
                   "rl_time_to_tol": "not measured yet"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
