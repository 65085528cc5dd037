"""Training and evaluation of the learned Richardson-Lucy warm start.

    python -m libapr_amd.warm_start train [--volumes 200] [--steps 2000] [--out PATH]
    python -m libapr_amd.warm_start eval  [--weights PATH]

Training pairs come from the plain APR Richardson-Lucy itself: for each seed a
deconvolution workload (ground truth = background 100 + blobs sigma 1-4 px,
amplitude 200-2000; blurred by the 9x9x9 Gaussian PSF with sigma 1, 1, 2;
shot noise) is converted to an APR, its particles are deconvolved by plain RL
until the relative update is below 1e-3 (cap 200), and the network learns the
map from the piecewise-constant reconstruction of the observed particles to
that of the converged estimate, on 64^3 tiles (4-voxel halo), both divided by
the mean of the observed reconstruction.

Evaluation: on a held-out volume, plain RL from u = 1 against RL started from
the network's estimate (aprh_cnn_warm_start, without and with the post filter
that is recorded beside the weights), both to the same tolerance;
reports iterations, wall time (the warm start's includes the network), the
relative L2 error of each reconstruction against the ground truth, and the
difference between the two results.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import libapr_amd as A

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "warm_start_cnn.npy")

APR_PARAMS = dict(rel_error=0.1, lambda_=3.0, sigma_th=50.0, Ip_th=120.0, grad_th=10.0)
GEN = dict(sigma=(1.0, 4.0), amp=(200.0, 2000.0), background=100.0, psf_sigma=(1.0, 1.0, 2.0), psf_taps=9)
BLOB_DENSITY = 50_000 / (1024 * 1024 * 256)   # blobs per voxel of the 1024x1024x256 workload
TOL, CAP = 1e-3, 200


def psf_stencil():
    return A.APRStencil.create_gaussian_filter(GEN["psf_sigma"], (GEN["psf_taps"],) * 3, True)


class Workload:
    """one deconvolution volume on the device: ground truth, observed APR and particles"""

    def __init__(self, ctx, shape_zxy, seed, n_blobs=None):
        z, x, y = shape_zxy
        self.dims = (y, x, z)
        self.n = x * y * z
        nb = n_blobs if n_blobs is not None else max(1, int(round(BLOB_DENSITY * self.n)))
        self.gt = A.DeviceArray(self.n * 4, ctx)
        img = A.DeviceArray(self.n * 2, ctx)
        A.synth_deconv(shape_zxy, seed, nb, self.gt.ptr, img.ptr, ctx=ctx, **GEN)
        conv = A.APRConverter(ctx)
        conv.par = A.APRParameters(**APR_PARAMS)
        self.apr = conv.get_apr(img.ptr, dims=self.dims)
        self.P = self.apr.total
        self.parts = A.DeviceArray(self.P * 2, ctx)
        A.ParticleData().sample_image(self.apr, img.ptr, out=self.parts.ptr)
        img.free()
        self.ctx = ctx

    def rl(self, u, iters=CAP, tol=TOL, resume=False):
        _, st = A.APRNumerics.richardson_lucy(self.apr, self.parts.ptr, u.ptr, psf_stencil(), iters, normalize=True,
                                              resume=resume, tol=tol, dtype="uint16")
        return st

    def recon(self, parts_ptr, dtype, out):
        A.APRReconstruction.reconstruct_constant(self.apr, parts_ptr, 0, np.float32, out=out.ptr, dtype=dtype)
        return out

    def error_vs_gt(self, u):
        rec = self.recon(u.ptr, "float32", A.DeviceArray(self.n * 4, self.ctx))
        r = rec.to_host(np.float32).astype(np.float64)
        g = self.gt.to_host(np.float32).astype(np.float64)
        return float(np.linalg.norm(r - g) / np.linalg.norm(g))


def make_pairs(ctx, seeds, shape_zxy, target_iters=0, log=print):
    """(input reconstruction, target reconstruction, scale) per seed, kept on the device.
    The target is plain RL run to the tolerance, or for exactly target_iters iterations."""
    pairs = []
    t = time.time()
    for k, seed in enumerate(seeds):
        w = Workload(ctx, shape_zxy, seed)
        u = A.DeviceArray(w.P * 4, ctx)
        st = w.rl(u, iters=target_iters, tol=0.0) if target_iters > 0 else w.rl(u)
        xin = w.recon(w.parts.ptr, "uint16", A.DeviceArray(w.n * 4, ctx))
        tgt = w.recon(u.ptr, "float32", A.DeviceArray(w.n * 4, ctx))
        scale = A.image_mean(xin.ptr, w.n, "float32", ctx)
        pairs.append((xin, tgt, scale, w.dims))
        if k % 20 == 0:
            log(f"[warm_start] pair {k + 1}/{len(seeds)}: P={w.P} RL {st['iterations']} iterations ({time.time() - t:.1f} s)")
    return pairs


def train(ctx, pairs, steps=2000, batch=8, lr=1e-3, seed=0, log=print):
    net = A.WarmStartCNN(ctx, seed=seed)
    rng = np.random.default_rng(seed)
    S, H = A.CNN_TILE, A.CNN_HALO
    s_in = S + 2 * H
    xb = A.DeviceArray(batch * s_in ** 3 * 4, ctx)
    tb = A.DeviceArray(batch * S ** 3 * 4, ctx)
    losses = []
    t0 = time.time()
    for it in range(steps):
        xin, tgt, scale, (y, x, z) = pairs[rng.integers(len(pairs))]
        org = np.stack([rng.integers(0, max(y - S, 0) + 1, batch), rng.integers(0, max(x - S, 0) + 1, batch),
                        rng.integers(0, max(z - S, 0) + 1, batch)], axis=1).astype(np.int32)
        A.extract_tiles(xin.ptr, (y, x, z), org, S, H, scale, out=xb.ptr, dtype="float32", ctx=ctx)
        A.extract_tiles(tgt.ptr, (y, x, z), org, S, 0, scale, out=tb.ptr, dtype="float32", ctx=ctx)
        cur = lr * (0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * it / steps)))
        losses.append(net.train_step(xb.ptr, tb.ptr, S, cur, T=batch))
        if it % 200 == 0 or it == steps - 1:
            log(f"[warm_start] step {it}/{steps} loss {np.mean(losses[-50:]):.4e} ({time.time() - t0:.1f} s)")
    return net, losses


def evaluate(ctx, net, shape_zxy=(256, 1024, 1024), seed=3, n_blobs=50_000, log=print, workload=None):
    w = workload or Workload(ctx, shape_zxy, seed, n_blobs)
    ctx.set_conv_mode(A.CONV_SEPARABLE)
    u = A.DeviceArray(w.P * 4, ctx)
    w.rl(u, iters=1)                      # warm the pools and kernels
    t = time.perf_counter()
    st_p = w.rl(u)
    t_plain = time.perf_counter() - t
    err_plain = w.error_vs_gt(u)
    uw = A.DeviceArray(w.P * 4, ctx)
    net.warm_start(w.apr, w.parts.ptr, out=uw.ptr, dtype="uint16")   # warm the network path
    t = time.perf_counter()
    net.warm_start(w.apr, w.parts.ptr, out=uw.ptr, dtype="uint16")
    t_cnn = time.perf_counter() - t
    st_w = w.rl(uw, resume=True)
    t_warm = time.perf_counter() - t
    err_warm = w.error_vs_gt(uw)
    up = u.to_host(np.float32).astype(np.float64)
    uwh = uw.to_host(np.float32).astype(np.float64)
    diff = float(np.linalg.norm(uwh - up) / np.linalg.norm(up))
    res = {"particles": w.P, "plain_iterations": st_p["iterations"], "plain_time_s": round(t_plain, 4),
           "plain_final_rel_update": st_p["last_rel_update"], "warm_final_rel_update": st_w["last_rel_update"],
           "plain_rel_l2_error_vs_gt": round(err_plain, 5), "warm_iterations": st_w["iterations"],
           "warm_time_s": round(t_warm, 4), "warm_network_time_s": round(t_cnn, 4),
           "warm_rel_l2_error_vs_gt": round(err_warm, 5), "time_ratio_warm_over_plain": round(t_warm / t_plain, 4),
           "final_error_ratio_warm_over_plain": round(err_warm / err_plain, 5),
           "warm_vs_plain_result_rel_l2": round(diff, 5), "tolerance": TOL, "conv_mode": "separable"}
    log(f"[warm_start] plain {st_p['iterations']} it {t_plain:.3f} s err {err_plain:.4f} | warm {st_w['iterations']} it "
        f"{t_warm:.3f} s (network {t_cnn:.3f} s) err {err_warm:.4f}")
    ctx.set_conv_mode(A.CONV_EXACT)
    return res


HELD_OUT_SEEDS = (400, 401, 402, 403)   # used by neither training (100-299) nor evaluation (3)
POST_FILTER_CANDIDATES = [((sy, sy, 0.3), (3, 3, 1)) for sy in (0.4, 0.45, 0.5, 0.55, 0.6, 0.65)] + \
    [((0.5, 0.5, 0.5), (3, 3, 3)), ((0.6, 0.6, 0.5), (3, 3, 3))]


def select_post_filter(ctx, net, seeds=HELD_OUT_SEEDS, shape_zxy=(128, 512, 512), max_error_ratio=1.01, log=print):
    """The post filter chosen on held-out volumes (seeds outside training and evaluation): for
    each candidate, RL from the filtered network estimate to the tolerance on every volume;
    the candidate with the fewest mean iterations whose mean final-error ratio against plain RL
    is at most max_error_ratio (else the lowest error ratio).  Returns (choice, table)."""
    ctx.set_conv_mode(A.CONV_SEPARABLE)
    vols = []
    for seed in seeds:
        w = Workload(ctx, shape_zxy, seed)
        u = A.DeviceArray(w.P * 4, ctx)
        st = w.rl(u)
        vols.append((w, st["iterations"], w.error_vs_gt(u)))
        u.free()
    table = []
    for sigma, size in POST_FILTER_CANDIDATES:
        net.set_post_filter(sigma, size)
        its, ratios = [], []
        for w, it_plain, err_plain in vols:
            uw = A.DeviceArray(w.P * 4, ctx)
            net.warm_start(w.apr, w.parts.ptr, out=uw.ptr, dtype="uint16")
            st = w.rl(uw, resume=True)
            its.append(st["iterations"] / it_plain)
            ratios.append(w.error_vs_gt(uw) / err_plain)
            uw.free()
        row = {"sigma": list(sigma), "size": list(size), "iterations_over_plain": round(float(np.mean(its)), 4),
               "error_ratio": round(float(np.mean(ratios)), 5)}
        log(f"[warm_start] held-out post filter {sigma} {size}: iterations/plain {row['iterations_over_plain']:.3f} "
            f"error ratio {row['error_ratio']:.4f}")
        table.append(row)
    ok = [r for r in table if r["error_ratio"] <= max_error_ratio]
    best = min(ok, key=lambda r: r["iterations_over_plain"]) if ok else min(table, key=lambda r: r["error_ratio"])
    ctx.set_conv_mode(A.CONV_EXACT)
    return {"sigma": best["sigma"], "size": best["size"]}, table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval", "select-post-filter"])
    ap.add_argument("--volumes", type=int, default=200)
    ap.add_argument("--first-seed", type=int, default=100)
    ap.add_argument("--shape", default="128,256,256", help="z,x,y of the training volumes")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval-shape", default="256,1024,1024")
    ap.add_argument("--report", default=None)
    ap.add_argument("--post-sigma", default="0.55,0.55,0.3", help="y,x,z sigma of the post filter evaluated")
    ap.add_argument("--post-size", default="3,3,1", help="y,x,z size of the post filter evaluated")
    ap.add_argument("--target-iters", type=int, default=0,
                    help="train on plain RL after this many iterations (0: run to the tolerance)")
    args = ap.parse_args()
    ctx = A.Context(0)
    ctx.set_conv_mode(A.CONV_SEPARABLE)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    report = {}
    if args.cmd == "train":
        shape = tuple(int(v) for v in args.shape.split(","))
        pairs = make_pairs(ctx, range(args.first_seed, args.first_seed + args.volumes), shape, args.target_iters, log)
        net, losses = train(ctx, pairs, args.steps, args.batch, args.lr, log=log)
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        net.save(args.out)
        report["train"] = {"volumes": args.volumes, "shape_zxy": shape, "steps": args.steps, "batch": args.batch,
                           "target": f"plain RL, {args.target_iters} iterations" if args.target_iters else "plain RL to 1e-3",
                           "first_loss": float(np.mean(losses[:20])), "last_loss": float(np.mean(losses[-50:]))}
        del pairs
    else:
        net = A.WarmStartCNN.load(args.weights, ctx, post_filter=None)
    if args.cmd == "select-post-filter":
        choice, table = select_post_filter(ctx, net, log=log)
        rec_path = os.path.splitext(args.weights)[0] + ".json"
        rec = json.load(open(rec_path)) if os.path.exists(rec_path) else {}
        rec["post_filter"] = choice
        rec["post_filter_selection"] = {"held_out_seeds": list(HELD_OUT_SEEDS), "shape_zxy": [128, 512, 512],
                                        "rule": "fewest iterations with mean final-error ratio <= 1.01 vs plain RL",
                                        "candidates": table}
        with open(args.report or rec_path, "w") as f:
            json.dump(rec, f, indent=1)
        print(json.dumps({"post_filter": choice}))
        return
    es = tuple(int(v) for v in args.eval_shape.split(","))
    w = Workload(ctx, es, 3, 50_000)
    net.set_post_filter(None)
    report["eval"] = evaluate(ctx, net, es, log=log, workload=w)
    pf = {"sigma": [float(v) for v in args.post_sigma.split(",")], "size": [int(v) for v in args.post_size.split(",")]}
    net.set_post_filter(pf["sigma"], pf["size"])
    report["eval_with_post_filter"] = dict(evaluate(ctx, net, es, log=log, workload=w), post_filter=pf)
    report["post_filter"] = pf
    print(json.dumps(report))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
