#!/usr/bin/env python3
"""Benchmark of the APR hot path on one MI355X, through libaprh's C-ABI.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-rl] [--no-batch]

Headline (value): APR conversion of a 2048x2048x512 uint16 volume.  One step =
APRConverter<uint16_t>::get_apr on the device-resident volume followed by
ParticleData<uint16_t>::sample_image of the same volume.  The input
(background 100, 200k Gaussian blobs with sigma U(1.5,8) px and amplitude
U(300,4000), shot noise, seed 1) is synthesised on the device before timing.
APR parameters: rel_error 0.1, lambda 3, sigma_th 50, Ip_th 120, grad_th 10.

Also measured (reported under "config"):
  * APR convolution: 3x3x3 mean stencil on the 256x256x128 volume (500 blobs,
    sigma 2-6, amplitude 500-3000, seed 0) and the 5x5x5 Gaussian (sigma 1)
    on the particles of the 2048x2048x512 volume;
  * batched conversion of 64 uneven volumes in one HIP graph (aggregate
    voxels/s, per-volume bit-exactness against single-volume get_apr);
  * Richardson-Lucy time to tolerance: ground truth 1024x1024x256 (background
    100 + 50k blobs, sigma 1-4, amplitude 200-2000), blurred by the 9x9x9
    Gaussian PSF (sigma 1, 1, 2), shot noise, seed 3, converted to an APR; plain
    RL from u = 1 until the relative update is below 1e-3 (cap 200).

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
    """Algorithmic HBM bytes of one launch group, from the survey's work model (bytes a minimal
    implementation must move; its groups split by stage, the gradient counting its read of the
    smoothed volume).  They sum to about 23 N + 6 P, the pipeline figure."""
    return {
        "offset": 2 * N,                      # read the input once (min/max)
        "smooth_bspline_y": 4 * N,            # read input, write smoothed u16
        "smooth_bspline_x": 4 * N,            # read + write u16 in place
        "smooth_bspline_z": 4 * N,
        "gradient_ds": 2 * N + 6 * Nd,        # read u16, write u16 grad + f32 mean
        "inv_bspline": 8 * Nd,                # read and write the f32 mean once (all axes in one sweep)
        "local_intensity_scale": 20 * Nd,     # mean pass 1 (1 B/voxel) + abs-diff and pass 2 (1.5 B/voxel)
        "levels": 11 * Nd,                    # read scale f32, gradient u16, intensity f32; write level u8
        "sample_image": 2 * N + N // 4 + 4 * P,
    }.get(stage)


def oracle_particle_count():
    """the oracle's particle count for this workload, from the committed digests of the CPU
    oracle's results (tests/golden/large_workloads.json; the structure and intensities are
    compared bit for bit in the GPU tests), or None if that file describes another workload"""
    try:
        with open(os.path.join(ROOT, "tests", "golden", "large_workloads.json")) as f:
            d = json.load(f)
        prm = d["params"]["large"]
        if (tuple(prm["shape_zxy"]) == SHAPE_ZXY and prm["seed"] == SEED and prm["n_blobs"] == N_BLOBS
                and tuple(prm["sigma"]) == SIGMA and tuple(prm["amp"]) == AMP):
            return int(d["large"]["total"])
    except (OSError, KeyError, ValueError):
        pass
    return None


def run_copy(args):
    """device-to-device copy of 4 GiB (steps times): the known-size kernel that calibrates FETCH_SIZE /
    WRITE_SIZE (each copy reads and writes 4 GiB)"""
    import libapr_amd as A
    ctx = A.Context(0)
    nbytes = 4 << 30
    a, b = _dev(A, ctx, nbytes), _dev(A, ctx, nbytes)
    A._check(A.lib().aprh_memcpy(ctx.h, b, a, nbytes))
    t = time.perf_counter()
    for _ in range(max(args.steps, 1)):
        A._check(A.lib().aprh_memcpy(ctx.h, b, a, nbytes))
    ms = (time.perf_counter() - t) * 1000.0 / max(args.steps, 1)
    for p in (a, b):
        A._check(A.lib().aprh_free(ctx.h, p))
    ctx.close()
    log(f"[bench] copy of 4 GiB: {ms:.3f} ms, {2 * nbytes / (ms / 1000.0) / 1e9:.0f} GB/s read+write")
    return {"bytes": nbytes, "ms": round(ms, 3), "GBps_read_plus_write": round(2 * nbytes / (ms / 1000.0) / 1e9, 1)}


def run_gpu(args, with_conv=True):
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

    last = {}
    # setup (untimed): one conversion to learn the particle count and size the particle buffer
    apr0 = conv.get_apr(d_img.value, dims=(y, x, z))
    P = apr0.total
    apr0.close()
    P_oracle = oracle_particle_count()
    if P_oracle is not None and P != P_oracle:
        raise RuntimeError(f"particle count {P} differs from the oracle's {P_oracle} for this volume")
    d_parts = C.c_void_p()
    A._check(L.aprh_malloc(ctx.h, max(P, 1) * 2, C.byref(d_parts)))

    def step():
        """get_apr + sample_image of the same image in one call (the row fill gathers the
        particle intensities; the mean pyramid is built inside the structure stage)"""
        if "apr" in last:
            last.pop("apr").close()
        apr, sampled = conv.get_apr_and_sample(d_img.value, d_parts.value, P, dims=(y, x, z))
        if not sampled or apr.total != P:
            raise RuntimeError("particle count changed between conversions of the same image")
        st = ctx.stage_times()
        last["apr"] = apr
        return st

    def step_separate():
        """the same work as two calls: get_apr, then sample_image (pyramid + gather over y_vec)"""
        if "apr" in last:
            last.pop("apr").close()
        apr = conv.get_apr(d_img.value, dims=(y, x, z))
        st = ctx.stage_times()
        pd.sample_image(apr, d_img.value, out=d_parts.value)
        st["sample_image"] = ctx.stage_times()["sample_image"]
        last["apr"] = apr
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
    stage_extra = {"separate_calls_ms": None, "separate_sample_image_ms": None, "oracle_particles": P_oracle}
    if not args.no_separate:
        # reference point (untimed for the metric): the same work as separate get_apr + sample_image calls
        step_separate()
        A._check(L.aprh_memcpy(ctx.h, None, None, 0))
        ts = time.perf_counter()
        sep = step_separate()
        A._check(L.aprh_memcpy(ctx.h, None, None, 0))
        separate_ms = (time.perf_counter() - ts) * 1000.0
        log(f"[bench] as separate get_apr + sample_image calls: {separate_ms:.2f} ms (sample_image {sep['sample_image']:.3f} ms)")
        stage_extra.update(separate_calls_ms=round(separate_ms, 3), separate_sample_image_ms=round(sep["sample_image"], 3))
    stage_avg = {k: v / args.steps for k, v in stage_acc.items()}
    for k, v in sorted(stage_avg.items(), key=lambda kv: -kv[1]):
        log(f"[bench]   {k:24s} {v:8.3f} ms")
    yd, xd, zd = (y + 1) // 2, (x + 1) // 2, (z + 1) // 2
    Nd = yd * xd * zd
    Npad = (yd + 6) * (xd + 6) * (zd + 6)
    # single-kernel stages: (kernel, stages it serves = launches per step); the dominant kernel is
    # the one with the most device time per step, its roofline per launch
    kernels = {"k_tiles<IirOuterCk<unsigned short, 16>>": ["smooth_bspline_x", "smooth_bspline_z"],
               "k_tiles<IirYLdsCk>": ["smooth_bspline_y"], "k_tiles<GradDsVec<true>>": ["gradient_ds"],
               "k_tiles<Levels<unsigned short>>": ["levels"]}
    kt = {k: sum(stage_avg.get(s, 0.0) for s in v) for k, v in kernels.items()}
    dom = max(kt, key=kt.get)
    launch_ms = kt[dom] / len(kernels[dom])
    ab = algorithmic_bytes(kernels[dom][0], N, Nd, P, Npad)
    achieved = ab / (launch_ms / 1000.0) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("per_launch", {}).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "stages": kernels[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": ab,
                "avg_ms": round(launch_ms, 4)}
    # every stage with a byte model: algorithmic GB/s
    # (stages folded into others, such as the offset now found by the y pass and the sampling done by
    # the structure stage, take no time of their own and are left out)
    stage_gbps = {k: round(algorithmic_bytes(k, N, Nd, P, Npad) / (v / 1000.0) / 1e9, 1) for k, v in stage_avg.items()
                  if algorithmic_bytes(k, N, Nd, P, Npad) and v > 0.05}
    roofline["stage_algorithmic_GBps"] = stage_gbps
    total_ab = 23 * N + 6 * P
    if not with_conv:
        last.pop("apr").close()
        A._check(L.aprh_free(ctx.h, d_img))
        A._check(L.aprh_free(ctx.h, d_parts))
        ctx.close()
        return ms_per_step, P, stage_avg, roofline, total_ab, None, stage_extra
    # 5x5x5 Gaussian (sigma 1) APR convolution on these particles (u16 in, f32 out)
    apr = last.pop("apr")
    A._check(L.aprh_free(ctx.h, d_img))
    d_out = C.c_void_p()
    A._check(L.aprh_malloc(ctx.h, max(P, 1) * 4, C.byref(d_out)))
    g5 = A.APRStencil.create_gaussian_filter((1.0, 1.0, 1.0), (5, 5, 5), True)
    c2 = conv_timing(A, apr, g5, d_parts.value, d_out.value, "uint16", 2, args.conv_reps)
    log(f"[bench] 5x5x5 convolution on the 2048x2048x512 APR (reference order): {c2['ms']:.2f} ms")
    ctx.set_conv_mode(A.CONV_SEPARABLE)
    c2s = conv_timing(A, apr, g5, d_parts.value, d_out.value, "uint16", 2, args.conv_reps)
    ctx.set_conv_mode(A.CONV_EXACT)
    log(f"[bench] 5x5x5 convolution on the 2048x2048x512 APR (separable): {c2s['ms']:.2f} ms")
    c2 = {"exact": c2, "separable": c2s}
    apr.close()
    A._check(L.aprh_free(ctx.h, d_out))
    A._check(L.aprh_free(ctx.h, d_parts))
    ctx.close()
    return ms_per_step, P, stage_avg, roofline, total_ab, c2, stage_extra


def conv_algorithmic_bytes(P, T, R, s_in):
    """APR convolution work model: read the input particles, write f32 outputs, read the
    y_vec (2 B), tree means written and read (8 B), row offsets (16 B per row)"""
    return P * (s_in + 6) + 8 * T + 16 * R


def hbm_roofline(ab, ms):
    """roofline entry of a call moving ab algorithmic bytes in ms"""
    gbps = ab / (ms / 1000.0) / 1e9
    return {"bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbps / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(ab)}


def conv_timing(A, apr, stencil, d_in, d_out, dtype, s_in, reps):
    apr.tree_info()   # build the tree once (cached per APR, like APR::initialize_tree_dense)
    A.APRFilter.convolve_pencil(apr, stencil, d_in, d_out, dtype=dtype)
    t = time.perf_counter()
    for _ in range(reps):
        A.APRFilter.convolve_pencil(apr, stencil, d_in, d_out, dtype=dtype)
    ms = (time.perf_counter() - t) * 1000.0 / reps
    P, T = apr.total, apr.total_number_tree_particles()
    R = int(apr.xz_end_size) + int(apr.tree_info().xz_end_size)
    ab = conv_algorithmic_bytes(P, T, R, s_in)
    return {"ms": round(ms, 3), "particles": P, "tree_particles": T, "algorithmic_bytes": ab,
            "algorithmic_GBps": round(ab / (ms / 1000.0) / 1e9, 1), "stencil": list(stencil.shape[::-1]),
            "roofline": hbm_roofline(ab, ms)}


def run_conv_small(args):
    """3x3x3 mean convolution of the 256x256x128 volume's particles"""
    import libapr_amd as A
    ctx = A.Context(0)
    L = A.lib()
    shape = (128, 256, 256)
    z, x, y = shape
    n = x * y * z
    d = C.c_void_p()
    A._check(L.aprh_malloc(ctx.h, n * 2, C.byref(d)))
    A.synth_blobs(shape, 0, 500, sigma=(2.0, 6.0), amp=(500.0, 3000.0), background=100.0, noise=True, out=d.value, ctx=ctx)
    conv = A.APRConverter(ctx)
    conv.par = A.APRParameters(**APR_PARAMS)
    apr = conv.get_apr(d.value, dims=(y, x, z))
    P = apr.total
    dp, do = C.c_void_p(), C.c_void_p()
    A._check(L.aprh_malloc(ctx.h, P * 2, C.byref(dp)))
    A._check(L.aprh_malloc(ctx.h, P * 4, C.byref(do)))
    A.ParticleData().sample_image(apr, d.value, out=dp.value)
    st = A.APRStencil.create_mean_filter((3, 3, 3))
    res = conv_timing(A, apr, st, dp.value, do.value, "uint16", 2, max(args.conv_reps, 5))
    log(f"[bench] 3x3x3 convolution on the 256x256x128 APR: {res['ms']:.3f} ms (P={P})")
    apr.close()
    for p in (d, dp, do):
        A._check(L.aprh_free(ctx.h, p))
    ctx.close()
    return res


BATCH_N = 64
BATCH_SIDES = (128, 192, 256, 384, 512)
BATCH_Z_CAP = 256
BATCH_SEED = 2


def batch_dims():
    """64 (y, x, z) drawn uniformly per axis from {128,192,256,384,512}, z capped at 256 (seed 2)"""
    rng = np.random.default_rng(BATCH_SEED)
    d = rng.choice(BATCH_SIDES, size=(BATCH_N, 3))
    d[:, 2] = np.minimum(d[:, 2], BATCH_Z_CAP)
    return [tuple(int(v) for v in r) for r in d]


def run_batch(args, check=True):
    """Batched uneven volumes: every volume's get_apr recorded into one HIP graph (libapr_amd.APRBatch),
    timed over replays; each volume's structure is compared with a single-volume get_apr of the same
    device image (bit-exact), and the single-volume loop is timed for comparison."""
    import libapr_amd as A
    ctx = A.Context(0)
    dims = batch_dims()
    par = A.APRParameters(**APR_PARAMS)
    t = time.time()
    batch = A.APRBatch(dims, par=par, ctx=ctx)
    nvox = batch.voxels
    for i, (y, x, z) in enumerate(dims):
        nb = max(1, (y * x * z) // 32 ** 3)
        A.synth_blobs((z, x, y), (BATCH_SEED << 16) + i, nb, sigma=SIGMA, amp=AMP, background=BACKGROUND, noise=True,
                      out=batch.input_address(i), ctx=ctx)
    log(f"[bench] batch of {len(dims)} volumes ({nvox / 1e9:.3f} Gvoxels) planned and synthesised in "
        f"{time.time() - t:.1f} s")
    for _ in range(max(args.warmup, 1)):
        batch.run()
    reps = max(args.steps, 1)
    t = time.perf_counter()
    for _ in range(reps):
        batch.run()
    ms = (time.perf_counter() - t) * 1000.0 / reps
    log(f"[bench] batch graph: {ms:.2f} ms per replay, {nvox / (ms / 1000.0) / 1e9:.2f} Gvoxels/s")
    stages = batch.stages()
    log(f"[bench] batch graph holds {len(stages)} stage launches over {len(dims)} volumes")
    for st in sorted(stages, key=lambda r: -r["tuned_ms"])[:8]:
        log(f"[bench]   {st['name']:22s} {str(st['shape'] or ''):18s} vols={st['volumes']:3d} tiles={st['tiles']:9d} "
            f"grid={st['grid']:8d} {st['tuned_ms']:.3f} ms")
    if not check:
        batch.close()
        ctx.close()
        return {"voxels": nvox, "ms_per_batch": round(ms, 3), "Gvoxels_per_s": round(nvox / (ms / 1000.0) / 1e9, 3),
                "stage_launches": len(stages)}
    conv = A.APRConverter(ctx)
    conv.par = par
    exact = 0
    total_p = 0
    t = time.perf_counter()
    for i, (y, x, z) in enumerate(dims):
        single = conv.get_apr(batch.input_address(i), dims=(y, x, z))
        got = batch.get(i)
        total_p += got.total
        s1, s2 = single.structure(), got.structure()
        if got.total == single.total and all(np.array_equal(a, b) for a, b in zip(s1, s2)):
            exact += 1
        single.close()
    loop_ms = (time.perf_counter() - t) * 1000.0
    log(f"[bench] batch: {exact}/{len(dims)} volumes bit-exact vs single-volume get_apr; single-volume loop "
        f"incl. structure copies {loop_ms:.0f} ms")
    # single-volume loop without the copies
    t = time.perf_counter()
    for i, (y, x, z) in enumerate(dims):
        conv.get_apr(batch.input_address(i), dims=(y, x, z)).close()
    seq_ms = (time.perf_counter() - t) * 1000.0
    log(f"[bench] single-volume get_apr loop: {seq_ms:.2f} ms")
    batch.close()
    ctx.close()
    return {"workload": "64 u16 volumes, sides drawn from {128,192,256,384,512} per axis (z capped at 256), "
                        "1 blob per 32^3 voxels (sigma 1.5-8, amp 300-4000), bg 100 + shot noise, seed 2",
            "voxels": nvox, "particles": total_p, "ms_per_batch": round(ms, 3),
            "Gvoxels_per_s": round(nvox / (ms / 1000.0) / 1e9, 3),
            "single_volume_loop_ms": round(seq_ms, 3), "bit_exact_volumes": exact, "volumes": len(dims),
            "stage_launches": len(stages),
            "autotuned_stages": [{k: st[k] for k in ("name", "shape", "shapes_tried", "volumes", "tiles", "grid", "tuned_ms")}
                                 for st in stages]}


RL_SHAPE_ZXY = (256, 1024, 1024)
RL_GEN = dict(n_blobs=50_000, sigma=(1.0, 4.0), amp=(200.0, 2000.0), background=100.0, psf_sigma=(1.0, 1.0, 2.0),
              psf_taps=9, seed=3)
RL_TOL = 1e-3
RL_CAP = 200


def _dev(A, ctx, nbytes):
    p = C.c_void_p()
    A._check(A.lib().aprh_malloc(ctx.h, max(nbytes, 1), C.byref(p)))
    return p


def _to_host(A, ctx, d, n, dtype):
    h = np.empty(n, dtype)
    A._check(A.lib().aprh_memcpy(ctx.h, C.c_void_p(h.ctypes.data), d, h.nbytes))
    return h


def _rel_l2(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run_rl(args):
    """plain and learned-warm-start APR Richardson-Lucy to tolerance on the 1024x1024x256
    deconvolution workload (libapr_amd.warm_start), plus the reference-order iteration time"""
    import libapr_amd as A
    from libapr_amd import warm_start as W
    ctx = A.Context(0)
    z, x, y = RL_SHAPE_ZXY
    g = RL_GEN
    w = W.Workload(ctx, RL_SHAPE_ZXY, g["seed"], g["n_blobs"])
    psf = W.psf_stencil()
    u = A.DeviceArray(w.P * 4, ctx)
    w.apr.tree_info()
    # per-iteration time with every tap in the reference's order, and separable
    per_it = {}
    for name, mode in (("exact", A.CONV_EXACT), ("separable", A.CONV_SEPARABLE)):
        ctx.set_conv_mode(mode)
        w.rl(u, iters=1, tol=0.0)
        t = time.perf_counter()
        w.rl(u, iters=4, tol=0.0)
        per_it[name] = (time.perf_counter() - t) * 250.0
    exact_ms = per_it["exact"]
    log(f"[bench] RL iteration: reference tap order {per_it['exact']:.2f} ms, separable {per_it['separable']:.2f} ms")
    # total-variation RL (APRNumerics::richardson_lucy_tv, reg 0.002): 1 untimed + 3 timed iterations per mode
    tv_it = {}
    for name, mode in (("exact", A.CONV_EXACT), ("separable", A.CONV_SEPARABLE)):
        ctx.set_conv_mode(mode)
        A.APRNumerics.richardson_lucy_tv(w.apr, w.parts.ptr, u.ptr, psf, 1, 0.002, normalize=True, dtype="uint16")
        t = time.perf_counter()
        A.APRNumerics.richardson_lucy_tv(w.apr, w.parts.ptr, u.ptr, psf, 3, 0.002, normalize=True, resume=True, dtype="uint16")
        tv_it[name] = (time.perf_counter() - t) * 1000.0 / 3
    log(f"[bench] RL-TV iteration: reference tap order {tv_it['exact']:.2f} ms, separable {tv_it['separable']:.2f} ms")
    Tn = w.apr.total_number_tree_particles()
    R = int(w.apr.xz_end_size) + int(w.apr.tree_info().xz_end_size)
    rl_ab = 2 * conv_algorithmic_bytes(w.P, Tn, R, 4) + w.P * (2 + 12)
    pf = None
    if os.path.exists(W.DEFAULT_WEIGHTS):
        net = A.WarmStartCNN.load(W.DEFAULT_WEIGHTS, ctx)
        pf = net.post_filter
        net.set_post_filter(None)
        trained = "libapr_amd/data/warm_start_cnn.npy"
    else:
        net = A.WarmStartCNN(ctx, params=None)
        trained = "untrained (weights file missing)"
    ev = W.evaluate(ctx, net, workload=w, log=log)
    evf = None
    if pf:
        net.set_post_filter(pf["sigma"], pf["size"])
        evf = W.evaluate(ctx, net, workload=w, log=log)
    # data residual ||psf * u - input|| / ||input|| of the plain result
    ctx.set_conv_mode(A.CONV_SEPARABLE)
    w.rl(u)
    blur = A.DeviceArray(w.P * 4, ctx)
    A.APRFilter.convolve_pencil(w.apr, psf, u.ptr, blur.ptr, normalize=True, dtype="float32")
    resid = _rel_l2(blur.to_host(np.float32), w.parts.to_host(np.uint16))
    res = {"workload": "1024x1024x256 GT (bg 100 + 50k blobs sigma 1-4 amp 200-2000) * 9^3 Gaussian PSF "
                       "(sigma 1,1,2) + shot noise, seed 3; RL from 1 (plain) or from the network (warm) "
                       "to rel update < 1e-3 (cap 200)",
           "particles": w.P, "iterations": ev["plain_iterations"], "time_to_tol_s": ev["plain_time_s"],
           "ms_per_iteration": round(ev["plain_time_s"] * 1000.0 / max(ev["plain_iterations"], 1), 2),
           "final_rel_update": ev["plain_final_rel_update"], "data_residual": round(resid, 5),
           "rel_l2_error_vs_gt": ev["plain_rel_l2_error_vs_gt"], "conv_mode": "separable",
           "exact_order_ms_per_iteration": round(exact_ms, 2),
           "separable_ms_per_iteration": round(per_it["separable"], 3),
           "iteration_algorithmic_bytes": int(rl_ab),
           "roofline_separable_iteration": hbm_roofline(rl_ab, per_it["separable"]),
           "roofline_exact_iteration": hbm_roofline(rl_ab, per_it["exact"]),
           "tv_ms_per_iteration": {"reg_factor": 0.002, "exact": round(tv_it["exact"], 2), "separable": round(tv_it["separable"], 2)},
           "warm_start": {"weights": trained, "iterations": ev["warm_iterations"], "time_to_tol_s": ev["warm_time_s"],
                          "network_s": ev["warm_network_time_s"], "final_rel_update": ev["warm_final_rel_update"],
                          "rel_l2_error_vs_gt": ev["warm_rel_l2_error_vs_gt"],
                          "time_ratio_warm_over_plain": ev["time_ratio_warm_over_plain"],
                          "error_ratio_warm_over_plain": ev["final_error_ratio_warm_over_plain"],
                          "rel_l2_diff_to_plain_result": ev["warm_vs_plain_result_rel_l2"]}}
    if evf:
        res["warm_start_post_filter"] = {"post_filter": pf, "iterations": evf["warm_iterations"],
                                         "time_to_tol_s": evf["warm_time_s"],
                                         "rel_l2_error_vs_gt": evf["warm_rel_l2_error_vs_gt"],
                                         "time_ratio_warm_over_plain": evf["time_ratio_warm_over_plain"],
                                         "error_ratio_warm_over_plain": evf["final_error_ratio_warm_over_plain"]}
    net.close()
    ctx.close()
    return res


def run_rl_iterations(args):
    """--steps Richardson-Lucy iterations on the deconvolution workload in each convolution mode
    (after one untimed iteration), for profiling"""
    import libapr_amd as A
    from libapr_amd import warm_start as W
    ctx = A.Context(0)
    w = W.Workload(ctx, RL_SHAPE_ZXY, RL_GEN["seed"], RL_GEN["n_blobs"])
    u = A.DeviceArray(w.P * 4, ctx)
    w.apr.tree_info()
    res = {"particles": w.P}
    for name, mode in (("separable", A.CONV_SEPARABLE), ("exact", A.CONV_EXACT)):
        ctx.set_conv_mode(mode)
        w.rl(u, iters=1, tol=0.0)
        t = time.perf_counter()
        w.rl(u, iters=max(args.steps, 1), tol=0.0)
        res[name + "_ms_per_iteration"] = round((time.perf_counter() - t) * 1000.0 / max(args.steps, 1), 3)
        log(f"[bench] RL ({name}): {res[name + '_ms_per_iteration']:.2f} ms per iteration")
    ctx.close()
    return res


def run_tv_iterations(args):
    """--steps total-variation Richardson-Lucy iterations (reg 0.002) on the deconvolution workload
    in each convolution mode (after one untimed iteration), for profiling"""
    import libapr_amd as A
    from libapr_amd import warm_start as W
    ctx = A.Context(0)
    w = W.Workload(ctx, RL_SHAPE_ZXY, RL_GEN["seed"], RL_GEN["n_blobs"])
    psf = W.psf_stencil()
    u = A.DeviceArray(w.P * 4, ctx)
    w.apr.tree_info()
    res = {"particles": w.P}
    for name, mode in (("separable", A.CONV_SEPARABLE), ("exact", A.CONV_EXACT)):
        ctx.set_conv_mode(mode)
        A.APRNumerics.richardson_lucy_tv(w.apr, w.parts.ptr, u.ptr, psf, 1, 0.002, normalize=True, dtype="uint16")
        t = time.perf_counter()
        A.APRNumerics.richardson_lucy_tv(w.apr, w.parts.ptr, u.ptr, psf, max(args.steps, 1), 0.002, normalize=True, resume=True,
                                         dtype="uint16")
        res[name + "_ms_per_iteration"] = round((time.perf_counter() - t) * 1000.0 / max(args.steps, 1), 3)
        log(f"[bench] RL-TV ({name}): {res[name + '_ms_per_iteration']:.2f} ms per iteration")
    ctx.close()
    return res


def run_warm_network(args):
    """--steps passes of the warm-start network (tile extraction, four layers, sampling back to
    particles, no post filter) over the deconvolution workload's particles, after one untimed pass,
    for profiling"""
    import libapr_amd as A
    from libapr_amd import warm_start as W
    ctx = A.Context(0)
    w = W.Workload(ctx, RL_SHAPE_ZXY, RL_GEN["seed"], RL_GEN["n_blobs"])
    net = (A.WarmStartCNN.load(W.DEFAULT_WEIGHTS, ctx, post_filter=None) if os.path.exists(W.DEFAULT_WEIGHTS)
           else A.WarmStartCNN(ctx, params=None))
    u = A.DeviceArray(w.P * 4, ctx)
    net.warm_start(w.apr, w.parts.ptr, out=u.ptr, dtype="uint16")
    t = time.perf_counter()
    for _ in range(max(args.steps, 1)):
        net.warm_start(w.apr, w.parts.ptr, out=u.ptr, dtype="uint16")
    ms = (time.perf_counter() - t) * 1000.0 / max(args.steps, 1)
    log(f"[bench] warm-start network: {ms:.2f} ms per pass")
    ctx.close()
    return {"particles": w.P, "ms_per_pass": round(ms, 3)}


def run_cpu_rl():
    """Reference CPU Richardson-Lucy (oracle/_ref Release build, else the oracle) timed for 3
    iterations on a 256x256x128 volume of the deconvolution generator (same blob density)"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    import libapr_amd as A
    kind, which = "reference", "ref_rel"
    if oracle_lib.load(which) is None:
        kind, which = "port", "oracle"
    if oracle_lib.load(which) is None:
        return None
    shape = (128, 256, 256)
    n = shape[0] * shape[1] * shape[2]
    g = RL_GEN
    nb = g["n_blobs"] * n // (RL_SHAPE_ZXY[0] * RL_SHAPE_ZXY[1] * RL_SHAPE_ZXY[2])
    ctx = A.Context(0)
    gt, img = _dev(A, ctx, n * 4), _dev(A, ctx, n * 2)
    A.synth_deconv(shape, g["seed"], nb, gt.value, img.value, sigma=g["sigma"], amp=g["amp"],
                   background=g["background"], psf_sigma=g["psf_sigma"], psf_taps=g["psf_taps"], ctx=ctx)
    host = _to_host(A, ctx, img, n, np.uint16).reshape(shape)
    for p in (gt, img):
        A._check(A.lib().aprh_free(ctx.h, p))
    ctx.close()
    par = oracle_lib.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)
    apr = oracle_lib.get_apr(host, par, which, keep=False)
    parts = apr.sample(host)
    psf = A.APRStencil.create_gaussian_filter(g["psf_sigma"], (g["psf_taps"],) * 3, True)
    threads = oracle_lib.load(which).orc_num_threads()
    t = time.perf_counter()
    apr.richardson_lucy(parts, psf, 3)
    dt = (time.perf_counter() - t) / 3
    log(f"[bench] cpu RL ({kind}, {threads} threads): {dt * 1000:.1f} ms/iteration at P={apr.total}")
    return {"ms_per_iteration": round(dt * 1000.0, 2), "particles": apr.total, "cores": threads, "kind": kind, **host_cpu(),
            "ns_per_particle_iteration": round(dt * 1e9 / apr.total, 2),
            "sample": f"3 RL iterations (9^3 PSF) on the APR of a 256x256x128 volume of the same generator ({nb} blobs)"}


def host_cpu():
    """CPU model and the host's logical CPUs (all of them, and those this process may use)"""
    model = None
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
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "cpu_count": os.cpu_count(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def fp64_errors():
    """error of the float stages against float64 restatements, as recorded by
    tests/test_fp64_error.py (APRH_FP64_ERRORS=tests/golden/fp64_errors.json)"""
    path = os.path.join(ROOT, "tests", "golden", "fp64_errors.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return {"source": "tests/test_fp64_error.py (CPU; the GPU results are bit-equal to the oracle)", "max_abs_over_max_and_rel_l2": json.load(f)}


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
    return {"value": round(n / dt / 1e9, 4), "unit": "Gvoxels/s", "cores": threads, "kind": kind, **host_cpu(),
            "sample": f"get_apr+sample_image on a {shape[2]}x{shape[1]}x{shape[0]} u16 volume of the same generator "
                      f"({nb} blobs, seed {SEED}), {dt:.2f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rl", action="store_true")
    ap.add_argument("--conv-reps", type=int, default=3)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-separate", action="store_true",
                    help="skip the reference timing of get_apr and sample_image as separate calls (profiling runs)")
    ap.add_argument("--only", choices=["get_apr", "batch", "copy", "conv", "rl", "tv", "warm"],
                    help="run one part (for profiling): the headline conversion steps, the batched volumes, "
                         "the device copy of 4 GiB that calibrates the HBM byte counters, the 5x5x5 convolution "
                         "of the headline volume's particles (both modes), --steps Richardson-Lucy iterations "
                         "on the deconvolution workload (both modes), --steps total-variation RL iterations "
                         "there (both modes), or --steps passes of the warm-start network over that workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if args.only:
        part = {"get_apr": lambda: (lambda r: list(r[:4]) + [r[6]])(run_gpu(args, with_conv=False)), "batch": lambda: run_batch(args, check=False),
                "copy": lambda: run_copy(args), "conv": lambda: run_gpu(args)[5], "rl": lambda: run_rl_iterations(args),
                "tv": lambda: run_tv_iterations(args), "warm": lambda: run_warm_network(args)}[args.only]()
        print(json.dumps({"only": args.only, "result": part}), flush=True)
        return
    ms, P, stages, roofline, total_ab, conv5, extra = run_gpu(args)
    z, x, y = SHAPE_ZXY
    N = x * y * z
    conv3 = run_conv_small(args)
    batch = None if args.no_batch else run_batch(args)
    rl = None if args.no_rl else run_rl(args)
    cpu = None if args.no_cpu else run_cpu_baseline()
    if rl is not None and not args.no_cpu:
        rl["cpu_baseline"] = run_cpu_rl()
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
                   "call": "aprh_get_apr_sample_u16 (conversion and sampling in one call)",
                   "as_separate_calls_ms": extra["separate_calls_ms"],
                   "particles": P, "particles_equal_oracle": extra["oracle_particles"] == P if extra["oracle_particles"] else None,
                   "pipeline_algorithmic_bytes": total_ab,
                   "pipeline_GBps": round(total_ab / (ms / 1000.0) / 1e9, 1),
                   "stage_ms": {k: round(v, 3) for k, v in stages.items()},
                   "conv_3x3x3_256x256x128": conv3, "conv_5x5x5_2048x2048x512": conv5,
                   "batched_uneven_volumes": batch,
                   "rl_time_to_tol": rl,
                   "fp64_error": fp64_errors()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
