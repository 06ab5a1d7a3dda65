"""Denoiser: the FP64 checker pinned by finite differences (CPU), and the BF16
MFMA kernels (GPU) against the checker."""
import numpy as np
import pytest

from oracle import denoise_ref as R


def random_params(seed, scale=0.05, head="residual", cin=13):
    rng = np.random.default_rng(seed)
    used = 3 if head == "residual" else 27
    parts = []
    for l, (co, ci) in enumerate(R.shapes(head)):
        w = rng.normal(scale=scale, size=(co, 27, ci))
        if l == 0:
            w[:, :, cin:] = 0
        if l == R.LAYERS - 1:
            w[used:] = 0
        b = rng.normal(scale=0.01, size=co)
        if l == R.LAYERS - 1:
            b[used:] = 0
        parts += [w.ravel(), b]
    return np.concatenate(parts)


def test_checker_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 3, 3, 4, 13))
    t = rng.normal(size=(1, 3, 3, 4, 3))
    p = random_params(2, scale=0.2)
    loss, g = R.loss_grad(p, x, t, 3)
    offs = np.cumsum([0] + [co * 27 * ci + co for co, ci in R.shapes()])
    idx = []
    for l, (co, ci) in enumerate(R.shapes()):
        cr = 3 if l == R.LAYERS - 1 else co
        cir = 13 if l == 0 else ci
        for _ in range(3):
            o, tp, i = rng.integers(cr), rng.integers(27), rng.integers(cir)
            idx.append(offs[l] + (o * 27 + tp) * ci + i)
        idx.append(offs[l] + co * 27 * ci + rng.integers(cr))  # a bias
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_grad(pp, x, t, 3)[0] - R.loss_grad(pm, x, t, 3)[0]) / (2 * h)
        # central differences: truncation O(h^2 f''') and cancellation O(eps f / h)
        assert abs(fd - g[k]) <= 2e-5 * max(1e-3, abs(g[k])) + 1e-8, (k, fd, g[k])


def test_checker_kernel_head_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(1, 3, 4, 3, 16))
    t = rng.normal(size=(1, 3, 4, 3, 3))
    p = random_params(5, scale=0.05, head="kernel", cin=16)
    loss, g = R.loss_grad(p, x, t, 3, head="kernel")
    offs = np.cumsum([0] + [co * 27 * ci + co for co, ci in R.shapes("kernel")])
    idx = []
    for l, (co, ci) in enumerate(R.shapes("kernel")):
        cr = 27 if l == R.LAYERS - 1 else co
        for _ in range(3):
            o, tp, i = rng.integers(cr), rng.integers(27), rng.integers(ci)
            idx.append(offs[l] + (o * 27 + tp) * ci + i)
        idx.append(offs[l] + co * 27 * ci + rng.integers(cr))  # a bias
    h = 1e-6
    nonzero = 0
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_grad(pp, x, t, 3, head="kernel")[0] - R.loss_grad(pm, x, t, 3, head="kernel")[0]) / (2 * h)
        nonzero += abs(g[k]) > 1e-9
        assert abs(fd - g[k]) <= 2e-5 * max(1e-3, abs(g[k])) + 1e-8, (k, fd, g[k])
    assert nonzero >= len(idx) // 2


def test_checker_kernel_head_closed_forms():
    # zero logits: a box filter over the in-volume taps; a dominant centre logit: the identity
    rng = np.random.default_rng(6)
    rad = rng.normal(size=(1, 3, 4, 5, 3))
    views, valid = R._tap_views(rad)
    logits = np.zeros((1, 3, 4, 5, 32))
    w = R.kernel_weights(logits, valid)
    box = (views * valid[..., None]).sum(axis=0) / valid.sum(axis=0)[..., None]
    assert np.allclose(np.einsum("t...,t...c->...c", w, views), box)
    logits[..., 13] = 200.0
    w = R.kernel_weights(logits, valid)
    assert np.allclose(np.einsum("t...,t...c->...c", w, views), rad)


def test_checker_conv_closed_forms():
    # a delta kernel at tap t shifts the volume by -off(t) with zero fill
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 4, 5, 6, 16))
    for t in (0, 13, 26, 5):
        w = np.zeros((16, 27, 16))
        w[:, t, :] = np.eye(16)
        y = R.conv3d(x, w)
        dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
        assert np.array_equal(y, xp[:, 1 + dz:5 + dz, 1 + dy:6 + dy, 1 + dx:7 + dx, :])


@pytest.mark.gpu
@pytest.mark.parametrize("cin,cout,shape", [(16, 32, (2, 5, 6, 7)), (32, 32, (2, 5, 6, 7)), (32, 16, (2, 5, 6, 7)),
                                            (16, 16, (2, 5, 6, 7)), (32, 32, (1, 9, 19, 35)),
                                            (16, 32, (2, 4, 8, 16))])
def test_conv3d_bf16_matches_checker(ctx, cin, cout, shape):
    # shapes: inside one 16 x 8 x 4 output tile; spread over 3 x 3 x 3 tiles with
    # partial tiles on every axis; exactly one full tile per volume
    from math_amd import native
    rng = np.random.default_rng(cin + cout)
    x = R.bf16(rng.normal(size=shape + (cin,)))
    w = R.bf16(rng.normal(scale=0.1, size=(cout, 27, cin)))
    b = rng.normal(size=cout).astype(np.float32)
    y = native.conv3d_bf16(ctx, x, w, b, relu=False)
    ref = R.conv3d(x, w, b.astype(np.float64))
    err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    print(f"conv {cin}->{cout}: max rel err {err:.2e}")
    assert err < 1e-5  # inputs are exact in BF16; only FP32 accumulation order differs


@pytest.mark.gpu
def test_denoiser_forward_and_gradient_match_checker(ctx):
    from math_amd import native
    rng = np.random.default_rng(7)
    d = native.Denoiser(ctx, 13, residual_channel=3, seed=11)
    p = d.params()
    x = rng.normal(size=(2, 6, 7, 8, 13)).astype(np.float32)
    t = rng.normal(size=(2, 6, 7, 8, 3)).astype(np.float32)
    y = d.forward(x)
    ref, _ = R.forward(p, x, 3, emulate_bf16=True)
    ferr = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    loss, g = d.loss_grad(x, t)
    rl, rg = R.loss_grad(p, x, t, 3, emulate_bf16=True)
    offs = np.cumsum([0] + [co * 27 * ci + co for co, ci in R.shapes()])
    errs = []
    for l in range(R.LAYERS):
        a, b = g[offs[l]:offs[l + 1]], rg[offs[l]:offs[l + 1]]
        errs.append(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    print(f"forward rel err {ferr:.2e}, loss {loss:.6f} vs {rl:.6f}, grad rel err per layer {np.round(errs, 4)}")
    assert ferr < 2e-2
    assert abs(loss - rl) <= 2e-2 * rl
    assert max(errs) < 5e-2


@pytest.mark.gpu
def test_denoiser_gradient_over_many_tiles_matches_checker(ctx):
    # 3 x 16 x 64 x 64 voxels: 768 weight-gradient tiles of 16 x 8 x 2, more than
    # the 512 blocks of the gradient kernel, so blocks sum several tiles; the
    # gradients are also bit-reproducible from run to run
    from math_amd import native
    rng = np.random.default_rng(17)
    d = native.Denoiser(ctx, 13, residual_channel=3, seed=5)
    p = d.params()
    x = rng.normal(size=(3, 16, 64, 64, 13)).astype(np.float32)
    t = rng.normal(size=(3, 16, 64, 64, 3)).astype(np.float32)
    loss, g = d.loss_grad(x, t)
    loss2, g2 = d.loss_grad(x, t)
    assert loss == loss2 and np.array_equal(g, g2)
    rl, rg = R.loss_grad(p, x, t, 3, emulate_bf16=True)
    offs = np.cumsum([0] + [co * 27 * ci + co for co, ci in R.shapes()])
    errs = [np.linalg.norm(g[offs[l]:offs[l + 1]] - rg[offs[l]:offs[l + 1]]) /
            max(np.linalg.norm(rg[offs[l]:offs[l + 1]]), 1e-30) for l in range(R.LAYERS)]
    print(f"many tiles: loss {loss:.6f} vs {rl:.6f}, grad rel err per layer {np.round(errs, 4)}")
    assert abs(loss - rl) <= 2e-2 * rl
    assert max(errs) < 5e-2


@pytest.mark.gpu
def test_denoiser_kernel_head_matches_checker(ctx):
    from math_amd import native
    rng = np.random.default_rng(9)
    d = native.Denoiser(ctx, 16, residual_channel=3, seed=12, head=native.HEAD_KERNEL)
    p = random_params(13, scale=0.05, head="kernel", cin=16).astype(np.float32)
    d.set_params(p)
    x = rng.normal(size=(2, 5, 6, 7, 16)).astype(np.float32)
    t = rng.normal(size=(2, 5, 6, 7, 3)).astype(np.float32)
    y = d.forward(x)
    ref, _ = R.forward(p, x, 3, emulate_bf16=True, head="kernel")
    ferr = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
    loss, g = d.loss_grad(x, t)
    rl, rg = R.loss_grad(p, x, t, 3, emulate_bf16=True, head="kernel")
    offs = np.cumsum([0] + [co * 27 * ci + co for co, ci in R.shapes("kernel")])
    errs = []
    for l in range(R.LAYERS):
        a, b = g[offs[l]:offs[l + 1]], rg[offs[l]:offs[l + 1]]
        errs.append(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    print(f"kernel head: forward rel err {ferr:.2e}, loss {loss:.6f} vs {rl:.6f}, grad rel err {np.round(errs, 4)}")
    assert ferr < 2e-2
    assert abs(loss - rl) <= 2e-2 * rl
    assert max(errs) < 5e-2


@pytest.mark.gpu
def test_denoiser_training_reduces_loss(ctx):
    from math_amd import native
    rng = np.random.default_rng(5)
    clean = np.cumsum(rng.normal(size=(4, 24, 24, 24, 3)), axis=1).astype(np.float32) * 0.05
    clean -= clean.min()
    noisy = clean + rng.normal(scale=0.1, size=clean.shape).astype(np.float32)
    x = np.concatenate([noisy, noisy, np.ones_like(clean[..., :1])], axis=-1)
    d = native.Denoiser(ctx, 7, residual_channel=0, seed=3)
    losses = d.train(x, clean, patch=16, batch=4, steps=60, lr=2e-3, seed=1)
    print("losses", losses[:3], losses[-3:])
    assert np.all(np.isfinite(losses))
    assert losses[-10:].mean() < 0.5 * losses[:5].mean()


@pytest.mark.gpu
def test_forward_scene_equals_host_input(ctx):
    # the device-built input of a painted scene (nsp_denoiser_forward_scene) gives
    # the image of the host-built volume through nsp_denoiser_forward, blended
    from math_amd import denoise_eval as DE
    from math_amd import native
    v, f, cfg = DE.variant(205, 128, ([40, 50, 60], 24), g=8)
    mesh = ctx.mesh(v, f)
    p = DE.ProgressiveInputs(ctx, mesh, cfg, seed=3)
    p.advance()
    p.advance()
    d = native.Denoiser(ctx, DE.IN_CHANNELS, DE.RESIDUAL, seed=4)
    x = p.volume()
    a = 0.7
    y_host = a * d.forward(x[None])[0] + (1 - a) * x[..., DE.RESIDUAL:DE.RESIDUAL + 3]
    y_dev = DE.to_volume(d.forward_scene(p.s4, 8, blend=a), cfg["screen"])
    err = np.max(np.abs(y_dev - y_host)) / np.max(np.abs(y_host))
    print(f"device vs host input: max rel err {err:.2e}")
    assert err < 1e-2
    ms, flops = d.last_forward()
    assert ms > 0 and flops == pytest.approx(2 * 27 * (13 * 32 + 6 * 32 * 32 + 32 * 3) * 24 ** 3)
    # left on the device, then copied
    d.forward_scene(p.s4, 8, blend=a, copy=False)
    assert np.array_equal(DE.to_volume(d.output(p.s4.pixels), cfg["screen"]), y_dev)


@pytest.mark.gpu
def test_training_resumes_from_checkpoint(ctx):
    # 40 steps in one call equal 25 steps, a checkpoint (parameters, Adam moments,
    # step count) restored into a new denoiser, and 15 more steps: bit for bit
    from math_amd import native
    rng = np.random.default_rng(8)
    x = rng.normal(size=(3, 20, 20, 20, 7)).astype(np.float32)
    y = (x[..., :3] * 0.5).astype(np.float32)
    a = native.Denoiser(ctx, 7, residual_channel=0, seed=4)
    la = a.train(x, y, patch=16, batch=2, steps=40, lr=1e-3, seed=9)
    b = native.Denoiser(ctx, 7, residual_channel=0, seed=4)
    lb1 = b.train(x, y, patch=16, batch=2, steps=25, lr=1e-3, seed=9)
    st = b.state()
    assert st[3] == 25
    b.close()
    c = native.Denoiser(ctx, 7, residual_channel=0, seed=99)
    c.set_state(*st)
    lb2 = c.train(x, y, patch=16, batch=2, steps=15, lr=1e-3, seed=9)
    assert np.array_equal(np.concatenate([lb1, lb2]), la)
    pa, ma, va, ta = a.state()
    pc, mc, vc, tc = c.state()
    assert ta == tc == 40
    assert np.array_equal(pa, pc) and np.array_equal(ma, mc) and np.array_equal(va, vc)


@pytest.mark.gpu
def test_trained_denoiser_beats_its_noisy_input_on_heldout_render(ctx):
    # what the learned part gains: trained on the tracer's own low/high-spp
    # renders, the network's output on a held-out scene is closer to a
    # high-spp render than the 4-spp image it starts from
    from math_amd import denoise_eval as DE
    from math_amd import native
    X, Y = DE.training_set(ctx, list(range(100, 108)), window=32, target_spp=256, levels=(1, 2), log=print)
    d = native.Denoiser(ctx, DE.IN_CHANNELS, DE.RESIDUAL, seed=1)
    losses = d.train(X, Y, patch=16, batch=16, steps=3000, lr=2e-4, seed=3)
    print("loss", losses[:5].mean(), "->", losses[-25:].mean(), flush=True)
    v, f, cfg = DE.variant(200, 128, ([48, 48, 48], 32))
    mesh = ctx.mesh(v, f)
    ref, _ = DE.render_mean(ctx, mesh, cfg, 1024, seed=2001)
    x, _ = DE.render_inputs(ctx, mesh, cfg, seed=2002)
    y = d.forward(x[None])[0]
    e_in, e_out = DE.rmse(x[..., DE.RESIDUAL:DE.RESIDUAL + 3], ref), DE.rmse(y, ref)
    print(f"held-out rmse: 4-spp input {e_in:.4f}, denoised {e_out:.4f}", flush=True)
    assert np.all(np.isfinite(y))
    assert e_out < 0.9 * e_in


def test_blend_weights_closed_form():
    # fit_blend's least-squares weight: a network output equal to the target gives 1, one that
    # overshoots the target twice as far from the 4-spp input gives 0.5, the input itself 0
    from math_amd import denoise_eval as DE
    rng = np.random.default_rng(3)
    levels = (1, 2, 4)
    V = 2
    X = rng.random((V * len(levels), 4, 4, 4, DE.IN_CHANNELS)).astype(np.float32)
    Y = rng.random((V * len(levels), 4, 4, 4, 3)).astype(np.float32)
    r4 = X[..., DE.RESIDUAL:DE.RESIDUAL + 3]

    class Fake:
        def __init__(self, out):
            self.out, self.k = out, 0

        def forward(self, x):
            y = self.out[self.k:self.k + x.shape[0]]
            self.k += x.shape[0]
            return y
    for out, expect in ((Y, 1.0), (r4 + 2 * (Y - r4), 0.5), (r4 - (Y - r4), 0.0)):
        a = DE.fit_blend(Fake(out), X, Y, levels=levels, batch=4)
        assert set(a) == set(levels)
        assert all(abs(v - expect) < 1e-5 for v in a.values()), (a, expect)
    a = {1: 1.0, 4: 0.0}
    assert DE.blend_weight(a, 2) == 0.5 and DE.blend_weight(a, 1000) == 0.0 and DE.blend_weight(a, 1) == 1.0
    # above the last fitted level the weight falls as 1 / level
    b = {1: 1.0, 8: 0.4}
    assert DE.blend_weight(b, 16) == pytest.approx(0.2) and DE.blend_weight(b, 80) == pytest.approx(0.04)
