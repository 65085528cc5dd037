"""Digests of the benchmark workloads at their full sizes, computed by the CPU oracle.

    python tests/golden/make_large_fixtures.py [--only large,batch,deconv]

Writes tests/golden/large_workloads.json, which tests/test_gpu_large.py compares the
GPU results with.  The inputs come from the host generators of libaprh
(aprh_synth_blobs_host, aprh_synth_deconv_host), whose bytes equal the device
generators' (the digest of each input is stored too, and the GPU test checks its own
device-generated input against it first).  Every array is recorded as the xxh3-64
digest of its bytes plus a few values at fixed strides, so a mismatch can be located.

Workloads (sizes and seeds of the benchmark):
  large   2048x2048x512 u16, 200k blobs (sigma 1.5-8, amplitude 300-4000), seed 1:
          get_apr + sample_image, then the 5x5x5 Gaussian (sigma 1) convolve_pencil of the
          u16 particles (reflective boundary, downsampled stencils)
  batch   the 64 uneven volumes (sides from {128,...,512}, z <= 256, 1 blob per 32^3
          voxels, seeds (2 << 16) + i): get_apr of each
  deconv  1024x1024x256 deconvolution volume (50k blobs sigma 1-4 amplitude 200-2000,
          9^3 Gaussian PSF sigma (1, 1, 2), shot noise, seed 3): get_apr + sample_image,
          the tree means and the 9^3 convolution of the particles times 1.37 (float),
          then 3 Richardson-Lucy iterations from u = 1 (normalised, downsampled stencils)
Test infrastructure only.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import xxhash

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

OUT = os.path.join(HERE, "large_workloads.json")
APR_PARAMS = dict(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10)

LARGE = dict(shape_zxy=(512, 2048, 2048), seed=1, n_blobs=200_000, sigma=(1.5, 8.0), amp=(300.0, 4000.0))
BATCH = dict(n=64, sides=(128, 192, 256, 384, 512), z_cap=256, seed=2, sigma=(1.5, 8.0), amp=(300.0, 4000.0))
DECONV = dict(shape_zxy=(256, 1024, 1024), seed=3, n_blobs=50_000, sigma=(1.0, 4.0), amp=(200.0, 2000.0),
              psf_sigma=(1.0, 1.0, 2.0), psf_taps=9, rl_iters=3)
N_PROBES = 64


def digest(a):
    a = np.ascontiguousarray(a)
    return xxhash.xxh3_64_hexdigest(memoryview(a.reshape(-1).view(np.uint8)))


def probes(a):
    """N_PROBES values at evenly spaced indices (index, value)"""
    a = np.asarray(a).reshape(-1)
    if a.size == 0:
        return []
    idx = np.unique(np.linspace(0, a.size - 1, N_PROBES).astype(np.int64))
    return [[int(i), a[i].item()] for i in idx]


def record(a):
    return {"digest": digest(a), "size": int(np.asarray(a).size), "probes": probes(a)}


def batch_dims():
    rng = np.random.default_rng(BATCH["seed"])
    d = rng.choice(BATCH["sides"], size=(BATCH["n"], 3))
    d[:, 2] = np.minimum(d[:, 2], BATCH["z_cap"])
    return [tuple(int(v) for v in r) for r in d]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_large(O, A):
    g = LARGE
    t = time.time()
    img = A.synth_blobs(g["shape_zxy"], g["seed"], g["n_blobs"], sigma=g["sigma"], amp=g["amp"])
    log(f"large: synthesised in {time.time() - t:.1f} s")
    res = {"input": digest(img)}
    t = time.time()
    apr = O.get_apr(img, O.params(**APR_PARAMS), "oracle", keep=False)
    log(f"large: oracle get_apr {time.time() - t:.1f} s, P={apr.total}")
    t = time.time()
    parts = apr.sample(img)
    del img
    log(f"large: sample {time.time() - t:.1f} s")
    res.update(total=apr.total, l_min=apr.l_min, l_max=apr.l_max, level_xz_vec=[int(v) for v in apr.level_xz_vec],
               xz_end_vec=record(apr.xz_end_vec), y_vec=record(apr.y_vec), particles=record(parts))
    st = O.gaussian_stencil((1.0, 1.0, 1.0), (5, 5, 5), True)
    stencils = O.downsampled_stencils(st, apr.l_max - apr.l_min, False)
    t = time.time()
    out = apr.convolve_pencil(stencils, parts)
    log(f"large: 5x5x5 convolve_pencil {time.time() - t:.1f} s")
    res["conv_gauss5"] = record(out)
    return res


def run_batch(O, A):
    vols = []
    t = time.time()
    for i, (y, x, z) in enumerate(batch_dims()):
        nb = max(1, (y * x * z) // 32 ** 3)
        img = A.synth_blobs((z, x, y), (BATCH["seed"] << 16) + i, nb, sigma=BATCH["sigma"], amp=BATCH["amp"])
        apr = O.get_apr(img, O.params(**APR_PARAMS), "oracle", keep=False)
        vols.append({"dims": [y, x, z], "seed": (BATCH["seed"] << 16) + i, "n_blobs": nb, "input": digest(img),
                     "total": apr.total, "level_xz_vec": [int(v) for v in apr.level_xz_vec],
                     "xz_end_vec": digest(apr.xz_end_vec), "y_vec": digest(apr.y_vec)})
    log(f"batch: {len(vols)} volumes in {time.time() - t:.1f} s")
    return {"volumes": vols}


def run_deconv(O, A):
    g = DECONV
    t = time.time()
    _, img = A.synth_deconv_host(g["shape_zxy"], g["seed"], g["n_blobs"], sigma=g["sigma"], amp=g["amp"],
                                 psf_sigma=g["psf_sigma"], psf_taps=g["psf_taps"])
    log(f"deconv: synthesised in {time.time() - t:.1f} s")
    res = {"input": digest(img)}
    apr = O.get_apr(img, O.params(**APR_PARAMS), "oracle", keep=False)
    parts = apr.sample(img)
    del img
    res.update(total=apr.total, l_min=apr.l_min, l_max=apr.l_max, level_xz_vec=[int(v) for v in apr.level_xz_vec],
               xz_end_vec=record(apr.xz_end_vec), y_vec=record(apr.y_vec), particles=record(parts))
    psf = O.gaussian_stencil(g["psf_sigma"], (g["psf_taps"],) * 3, True)
    # float particles with fractional values: tree means and the reference-order convolution
    # depend on the summation order there
    pf = (parts.astype(np.float32) * np.float32(1.37)).astype(np.float32)
    res["tree_mean_f32"] = record(apr.fill_tree_mean(pf))
    stack = O.downsampled_stencils(psf, apr.l_max - apr.l_min, True)
    res["conv_psf_f32"] = record(apr.convolve_pencil(stack, pf))
    t = time.time()
    u, it, rel = apr.richardson_lucy(parts, psf, g["rl_iters"], use_ds=True, normalize=True)
    log(f"deconv: {it} RL iterations {time.time() - t:.1f} s, P={apr.total}")
    res["rl"] = {"iterations": it, "last_rel_update": rel, "u": record(u)}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="large,batch,deconv")
    args = ap.parse_args()
    import oracle_lib as O
    import libapr_amd as A
    if O.load("oracle") is None:
        O.build(ref=False)
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    data["about"] = ("digests (xxh3-64 of the array bytes) of the CPU oracle's results on the benchmark "
                     "workloads; written by tests/golden/make_large_fixtures.py")
    data["params"] = {"large": LARGE, "batch": BATCH, "deconv": DECONV, "apr": APR_PARAMS}
    for part in args.only.split(","):
        data[part] = {"large": run_large, "batch": run_batch, "deconv": run_deconv}[part](O, A)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1)
        log(f"wrote {OUT} ({part})")


if __name__ == "__main__":
    main()
