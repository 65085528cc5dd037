"""GPU parity of the learned warm start's network and of the float sampling it
uses, against the CPU oracle (oracle/cnn_oracle.hpp, pinned by finite
differences in test_oracle_cnn.py).

The oracle emulates the GPU's bf16 roundings (weights of layers 1-3, layer-1
input taps, activations, back-propagated gradients), so the remaining
difference is FP32 accumulation order and the rare bf16 rounding flip it
causes: forward outputs agree to 1e-3 relative (L2), gradients to 1e-2 per
parameter group.
"""
import numpy as np
import pytest

import libapr_amd as A

pytestmark = pytest.mark.gpu

GROUPS = {"W1": (0, 432), "b1": (432, 448), "W2": (448, 7360), "b2": (7360, 7376), "W3": (7376, 14288),
          "b3": (14288, 14304), "W4": (14304, 14736), "b4": (14736, 14737)}


def rand_params(seed, last=0.05):
    rng = np.random.default_rng(seed)
    p = np.zeros(A.CNN_PARAMS, np.float32)
    p[0:432] = rng.normal(0, np.sqrt(2 / 27), 432)
    p[432:448] = rng.normal(0, 0.05, 16)
    p[448:7360] = rng.normal(0, np.sqrt(2 / 432), 6912)
    p[7360:7376] = rng.normal(0, 0.05, 16)
    p[7376:14288] = rng.normal(0, np.sqrt(2 / 432), 6912)
    p[14288:14304] = rng.normal(0, 0.05, 16)
    p[14304:14736] = rng.normal(0, last, 432)
    p[14736] = 0.02
    return p


def tiles(seed, T, S):
    """smooth positive tiles (sums of Gaussians) so activations are not all kinks"""
    rng = np.random.default_rng(seed)
    s = S + 8
    z, x, y = np.meshgrid(np.arange(s), np.arange(s), np.arange(s), indexing="ij")
    out = np.zeros((T, s, s, s), np.float32)
    for t in range(T):
        v = np.full((s, s, s), 0.3)
        for _ in range(6):
            c = rng.random(3) * s
            v += rng.random() * np.exp(-((z - c[0]) ** 2 + (x - c[1]) ** 2 + (y - c[2]) ** 2) / (2 * (1 + 3 * rng.random()) ** 2))
        out[t] = v + 0.05 * rng.random((s, s, s))
    return out


def test_cnn_param_count():
    assert A.lib().aprh_cnn_param_count() == A.CNN_PARAMS == 14737


@pytest.mark.parametrize("S,T", [(8, 3), (13, 2), (24, 1)])
def test_forward_matches_oracle(oracle, ctx, S, T):
    L = oracle.load("oracle")
    p = rand_params(S)
    net = A.WarmStartCNN(ctx, params=p)
    x = tiles(S + 1, T, S)
    print(f"forward S={S} T={T}", flush=True)
    got = net.forward_tiles(x, S)
    exp = np.zeros(T * S ** 3, np.float32)
    L.orc_cnn_forward(p, x.reshape(-1), T, S, 1, exp)
    exp = exp.reshape(got.shape)
    rel = np.linalg.norm(got - exp) / np.linalg.norm(exp - x[:, 4:4 + S, 4:4 + S, 4:4 + S])
    print(f"relative L2 error of the network branch: {rel:.2e}, max abs {np.abs(got - exp).max():.2e}")
    assert rel < 1e-3


@pytest.mark.parametrize("S,T", [(8, 2), (11, 1)])
def test_gradients_match_oracle(oracle, ctx, S, T):
    L = oracle.load("oracle")
    p = rand_params(100 + S, last=0.3)
    net = A.WarmStartCNN(ctx, params=p)
    x = tiles(200 + S, T, S)
    tgt = (x[:, 4:4 + S, 4:4 + S, 4:4 + S] * 1.5 + 0.1).astype(np.float32)
    print(f"gradients S={S} T={T}", flush=True)
    loss, g = net.train_step(x, tgt, S, 0.0, grads=True)
    ge = np.zeros(A.CNN_PARAMS, np.float64)
    le = L.orc_cnn_grad(p, x.reshape(-1), tgt.reshape(-1), T, S, 1, ge)
    assert abs(loss - le) <= 1e-4 * abs(le)
    for name, (a, b) in GROUPS.items():
        e = np.linalg.norm(g[a:b] - ge[a:b]) / np.linalg.norm(ge[a:b])
        print(f"  {name}: relative L2 {e:.2e}")
        assert e < 1e-2, name
    # no update with lr = 0
    assert np.array_equal(net.get_params(), p)


def test_training_reduces_loss(ctx):
    """a learnable target (a smoothed, scaled copy of the input) is fitted by Adam"""
    S, T = 16, 4
    net = A.WarmStartCNN(ctx, seed=3)
    rng_seeds = range(6)
    losses = []
    for it in range(120):
        x = tiles(rng_seeds[it % 6], T, S)
        c = x[:, 3:5 + S, 3:5 + S, 3:5 + S]
        tgt = 0.5 * (c[:, 1:-1, 1:-1, 1:-1] * 2 + c[:, :-2, 1:-1, 1:-1] + c[:, 2:, 1:-1, 1:-1]) * 0.8
        losses.append(net.train_step(x, np.ascontiguousarray(tgt, np.float32), S, 2e-3))
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert np.mean(losses[-6:]) < 0.2 * np.mean(losses[:6])


def test_apply_equals_tiles(ctx):
    """whole-image application = forward of reflect-padded 64^3 tiles"""
    net = A.WarmStartCNN(ctx, params=rand_params(7))
    shape = (70, 64, 130)   # z, x, y: partial tiles along z and y
    rng = np.random.default_rng(1)
    img = (rng.random(shape) * 500 + 100).astype(np.float32)
    scale = float(img.mean())
    out = net.apply(img, scale=scale)
    z, x, y = shape
    org = [(0, 0, 0), (64, 0, 64), (128, 0, 0)]
    xt = A.extract_tiles(img, (y, x, z), org, 64, 4, scale)
    pad = np.pad(img, 4, mode="reflect")
    for k, (oy, ox, oz) in enumerate(org):
        ref = pad[oz:oz + 72, ox:ox + 72, oy:oy + 72]
        ref = np.pad(ref, [(0, 72 - ref.shape[0]), (0, 72 - ref.shape[1]), (0, 72 - ref.shape[2])])
        n = [min(72, s + 8 - o) for s, o in zip((z, x, y), (oz, ox, oy))]
        np.testing.assert_allclose(xt[k][:n[0], :n[1], :n[2]], ref[:n[0], :n[1], :n[2]] / scale, rtol=1e-6)
    yt = net.forward_tiles(xt, 64) * scale
    for k, (oy, ox, oz) in enumerate(org):
        blk = out[oz:oz + 64, ox:ox + 64, oy:oy + 64]
        np.testing.assert_allclose(blk, yt[k][:blk.shape[0], :blk.shape[1], :blk.shape[2]], rtol=1e-6, atol=1e-3)


def test_sample_image_f32_matches_oracle(oracle, ctx):
    img16 = A.synth_blobs((29, 45, 61), 3, 12)
    o = oracle.get_apr(img16, oracle.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10), "oracle")
    g = A.APR.from_structure(o.dims, o.xz_end_vec, o.y_vec, ctx)
    rng = np.random.default_rng(0)
    img = (img16.astype(np.float32) * np.float32(0.37) + rng.random(img16.shape, dtype=np.float32)).astype(np.float32)
    got = A.ParticleData().sample_image(g, img)
    assert got.dtype == np.float32
    assert np.array_equal(got.view(np.uint32), o.sample_f32(img).view(np.uint32))


def test_warm_start_shape_and_positivity(oracle, ctx):
    img16 = A.synth_blobs((32, 64, 64), 0, 20)
    o = oracle.get_apr(img16, oracle.params(rel_error=0.1, sigma_th=50, Ip_th=120, grad_th=10), "oracle")
    g = A.APR.from_structure(o.dims, o.xz_end_vec, o.y_vec, ctx)
    parts = o.sample(img16)
    net = A.WarmStartCNN(ctx, params=rand_params(9, last=0.0))   # W4 = 0: the network is the identity + b4
    u0 = net.warm_start(g, parts)
    assert u0.shape == (o.total,) and (u0 > 0).all()
    # identity network: u0 = sample(reconstruct(parts) + b4 * mean) = parts + b4 * mean
    rec = o.reconstruct(parts, 0, np.float32)
    mean = float(rec.astype(np.float64).mean())
    np.testing.assert_allclose(u0, parts.astype(np.float64) + 0.02 * mean, rtol=1e-5)


def test_trained_warm_start_gain(ctx):
    """What the trained network buys on a volume it never saw (training used seeds 100-299):
    RL from its estimate reaches the plain method's stopping tolerance in far fewer
    iterations, its estimate is nearer the converged plain result than the observation is,
    and the final error against the ground truth stays close to plain RL's."""
    from libapr_amd import warm_start as W
    w = W.Workload(ctx, (128, 256, 256), seed=7)
    net = A.WarmStartCNN.load(W.DEFAULT_WEIGHTS, ctx, post_filter=None)
    ctx.set_conv_mode(A.CONV_SEPARABLE)
    try:
        u = A.DeviceArray(w.P * 4, ctx)
        st_p = w.rl(u)
        err_p = w.error_vs_gt(u)
        up = u.to_host(np.float32).astype(np.float64)
        uw = A.DeviceArray(w.P * 4, ctx)
        net.warm_start(w.apr, w.parts.ptr, out=uw.ptr, dtype="uint16")
        u0 = uw.to_host(np.float32).astype(np.float64)
        obs = w.parts.to_host(np.uint16).astype(np.float64)
        d_net = np.linalg.norm(u0 - up) / np.linalg.norm(up)
        d_obs = np.linalg.norm(obs - up) / np.linalg.norm(up)
        st_w = w.rl(uw, resume=True)
        err_w = w.error_vs_gt(uw)
        print(f"plain: {st_p['iterations']} iterations, error {err_p:.4f}; warm: {st_w['iterations']} iterations, "
              f"error {err_w:.4f}; distance to the plain result: network {d_net:.4f}, observation {d_obs:.4f}")
        assert st_w["last_rel_update"] < W.TOL and st_p["last_rel_update"] < W.TOL
        assert st_w["iterations"] <= 0.6 * st_p["iterations"]
        assert d_net < 0.5 * d_obs
        assert err_w < 1.1 * err_p
        # with the post filter recorded beside the weights: more iterations, error nearer plain RL's
        netf = A.WarmStartCNN.load(W.DEFAULT_WEIGHTS, ctx)
        assert netf.post_filter is not None
        netf.warm_start(w.apr, w.parts.ptr, out=uw.ptr, dtype="uint16")
        st_f = w.rl(uw, resume=True)
        err_f = w.error_vs_gt(uw)
        print(f"with the post filter {netf.post_filter}: {st_f['iterations']} iterations, error {err_f:.4f}")
        assert st_f["iterations"] <= 0.8 * st_p["iterations"]
        assert err_f < err_w and err_f < 1.03 * err_p
    finally:
        ctx.set_conv_mode(A.CONV_EXACT)


def test_warm_start_post_filter_is_a_convolution(ctx):
    """the post filter is convolve_pencil of the network's estimate with the normalised
    Gaussian (bit for bit), and switching it off restores the plain estimate"""
    from libapr_amd import warm_start as W
    w = W.Workload(ctx, (64, 96, 128), seed=11)
    net = A.WarmStartCNN.load(W.DEFAULT_WEIGHTS, ctx, post_filter=None)
    plain = net.warm_start(w.apr, w.parts.ptr, dtype="uint16").copy()
    sigma, size = (0.55, 0.55, 0.3), (3, 3, 1)
    net.set_post_filter(sigma, size)
    got = net.warm_start(w.apr, w.parts.ptr, dtype="uint16").copy()
    st = A.APRStencil.create_gaussian_filter(sigma, size, True)
    exp = A.APRFilter.convolve_pencil(w.apr, st, plain, normalize=True)
    assert np.array_equal(got.view(np.uint32), exp[:w.P].view(np.uint32))
    assert not np.array_equal(got, plain)
    net.set_post_filter(None)
    assert np.array_equal(net.warm_start(w.apr, w.parts.ptr, dtype="uint16"), plain)
    with pytest.raises(A.AprhError):
        net.set_post_filter((0.5, 0.5, 0.5), (2, 3, 3))
