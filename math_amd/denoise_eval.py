"""Denoiser workload (C3): training data from the tracer itself, training, and
the time-to-RMSE comparison of denoised against plain FP32 rendering.

  variants   n=4 tesseract-boundary meshes of the C1 generator with radial noise
             sigma ~ U[0, 0.05] and a point light at distance 4 in a direction
             uniform on S^3 (seeded)
  inputs     13 channels per voxel of the 3D screen: the 1-spp radiance (the
             first sample of each pixel), the 4-spp (8-sample) radiance, and the
             first-hit albedo and normal, all from one 4-spp pass with AOVs.
             Rendering is progressive: at level j the slots hold j
             eight-sample passes (the 1-spp slot the j first samples), so
             level 1 is the plain 1-spp + 4-spp input and the network can be
             applied at any sample budget; a level costs what a plain 4-spp
             pass costs, plus the AOV writes
  target     FP32 render of target_spp samples per pixel (1024 spp requests
             11^3 = 1331 stratified samples)
  network    8 layers 3x3x3, 32 channels, BF16/FP32-accumulate, residual on the
             4-spp radiance, trained on 32^3 patches, batch 16, Adam
  variants   64 for training (seeds 100-163), 16 held out (seeds 200-215)
  judge      RMSE over pixels x 3 linear-RGB channels against a 4096-sample FP64
             render of the same scene; the plain render is the fallback
"""
import time

import numpy as np

from . import configs, native

IN_CHANNELS = 13
RESIDUAL = 3  # the 4-spp radiance channels (after the 1-spp radiance)
TRAIN_SEEDS = tuple(range(100, 164))
HELDOUT_SEEDS = tuple(range(200, 216))
LEVELS = (1, 2, 4, 8, 16, 32, 64, 128)  # progressive levels stored for training
THRESHOLDS = (0.03, 0.02, 0.01)


def variant(seed, screen=128, window=None, g=24):
    """Mesh and scene dict of variant `seed`; window=(origin[3], size) renders a
    sub-cube of the screen with the same pixel footprint."""
    rng = np.random.default_rng(1000003 + seed)
    sigma = float(rng.uniform(0.0, 0.05))
    ldir = rng.normal(size=4)
    ldir /= np.linalg.norm(ldir)
    light = tuple((4.0 * ldir).tolist())
    v, f = native.meshgen_cube_sphere(4, g, True, sigma, 0.0, seed)
    cfg = configs.scene_dict(v, f, screen, light)
    if window is not None:
        org, size = window
        cam = np.array(cfg["camera_org"])
        upp = cfg["units_per_pixel"]
        for a in range(3):
            cam += np.array(cfg["axes"][a]) * upp * (org[a] + size / 2.0 - screen / 2.0)
        cfg["camera_org"] = cam.tolist()
        cfg["screen"] = [size, size, size]
    return v, f, cfg


def to_volume(arr, screen):
    """[pixels, c] in GlobalIndex order (x0 fastest) -> [s2, s1, s0, c]."""
    return arr.reshape(screen[2], screen[1], screen[0], -1)


def _paint_timed(scene, spp, depth, seed, **kw):
    """One paint call; returns its wall seconds (nsp_paint synchronises its stream before returning)."""
    t0 = time.perf_counter()
    scene.paint(spp, depth, seed=seed, **kw)
    return time.perf_counter() - t0


class ProgressiveInputs:
    """The denoiser input of one scene, rendered progressively: each advance()
    adds one 4-spp (8-sample) pass with its AOVs."""

    def __init__(self, ctx, mesh, cfg, seed, depth=5):
        self.s4 = native.Scene(ctx, mesh, cfg)
        self.screen = cfg["screen"]
        self.seed, self.depth = seed, depth
        self.level = 0
        self.secs = 0.0  # paint seconds so far

    def advance(self):
        self.secs += _paint_timed(self.s4, 4, self.depth, self.seed + 7919, aov=True)
        self.level += 1

    def volume(self):
        r4, c4 = self.s4.image()
        alb, nrm, _ = self.s4.aov()
        r1 = self.s4.first_sample()
        k4 = np.maximum(c4, 1)[:, None]
        x = np.concatenate([r1 / max(self.level, 1), r4 / k4, alb / k4, nrm / k4], axis=1)
        return to_volume(x.astype(np.float32), self.screen)


def render_inputs(ctx, mesh, cfg, seed, depth=5):
    """Level-1 (4-spp) denoiser input volume and the paint seconds spent."""
    p = ProgressiveInputs(ctx, mesh, cfg, seed, depth)
    p.advance()
    return p.volume(), p.secs


def render_mean(ctx, mesh, cfg, spp, seed, depth=5, precision=native.PRECISION_FP32, passes=1, log=None):
    """Mean image of `passes` progressive passes of `spp` samples each."""
    sc = native.Scene(ctx, mesh, cfg)
    t0 = time.perf_counter()
    for p in range(passes):
        sc.paint(spp, depth, seed=seed, precision=precision)
        if log:
            log(f"  pass {p + 1}/{passes}: {time.perf_counter() - t0:.1f} s")
    secs = time.perf_counter() - t0
    s, c = sc.image()
    return to_volume((s / np.maximum(c, 1)[:, None]).astype(np.float32), cfg["screen"]), secs


def rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))


def training_set(ctx, variants, screen=128, window=64, target_spp=1024, depth=5, levels=LEVELS, log=print):
    """Inputs at each progressive level in `levels` and the target of every
    variant: X [len(variants) * len(levels), w, w, w, IN_CHANNELS], Y [.., 3]."""
    xs, ys = [], []
    for i, seed in enumerate(variants):
        rng = np.random.default_rng(seed)
        org = [int(rng.integers(0, screen - window + 1)) for _ in range(3)]
        v, f, cfg = variant(seed, screen, (org, window))
        mesh = ctx.mesh(v, f)
        y, secs = render_mean(ctx, mesh, cfg, target_spp, seed=10 * seed + 2, depth=depth)
        p = ProgressiveInputs(ctx, mesh, cfg, seed=10 * seed + 1, depth=depth)
        for lv in range(1, max(levels) + 1):
            p.advance()
            if lv in levels:
                xs.append(p.volume())
                ys.append(y)
        log(f"training volume {i}: seed {seed}, window {org}, target {secs:.2f} s, inputs {p.secs:.2f} s")
        mesh.close()
    return np.stack(xs), np.stack(ys)


def fit_blend(denoiser, X, Y, levels=LEVELS, batch=4):
    """Per progressive level, the least-squares weight alpha in [0, 1] of the
    network output against the 4-spp input channels: minimises
    |alpha * net + (1 - alpha) * r4 - target|^2 over the training volumes of
    that level (X, Y as training_set returns them, variant-major, level-minor).
    Returns {level: alpha}."""
    L = len(levels)
    num = np.zeros(L)
    den = np.zeros(L)
    for i0 in range(0, X.shape[0], batch):
        xs = X[i0:i0 + batch]
        ys = denoiser.forward(xs)
        for j in range(xs.shape[0]):
            li = (i0 + j) % L
            r4 = xs[j, ..., RESIDUAL:RESIDUAL + 3].astype(np.float64)
            dv = ys[j].astype(np.float64) - r4
            num[li] += float(np.sum(dv * (Y[i0 + j].astype(np.float64) - r4)))
            den[li] += float(np.sum(dv * dv))
    return {lv: float(np.clip(num[k] / den[k], 0.0, 1.0)) if den[k] > 0 else 0.0 for k, lv in enumerate(levels)}


def blend_weight(alphas, level):
    """alpha at a progressive level: linear in log2(level) between the fitted
    levels, the first value below them, and above the last fitted level L the
    last value times L / level (the input's variance, which the network
    removes, falls as 1 / level, while its own error does not)."""
    lv = sorted(alphas)
    if level > lv[-1]:
        return float(alphas[lv[-1]] * lv[-1] / level)
    x = np.log2(np.asarray(lv, np.float64))
    return float(np.interp(np.log2(max(level, 1)), x, [alphas[k] for k in lv]))


def _mean_volume(scene, screen):
    s, c = scene.image()
    return to_volume(s / np.maximum(c, 1)[:, None], screen), c


def first_time(curve, threshold):
    """First time of a (samples, seconds, rmse) curve at which rmse <= threshold, or None."""
    for _, t, e in curve:
        if e <= threshold:
            return t
    return None


def time_to_rmse(ctx, mesh, cfg, ref, denoiser, depth=5, thresholds=THRESHOLDS, max_passes=256, max_levels=256,
                 seed=77, alphas=None):
    """Error-versus-time curves of both methods against `ref`, and the time at
    which each first reaches each threshold.  Only GPU work is timed (paint
    calls and the network forward); image read-back and the RMSE evaluation
    between passes are not.  Each curve stops at the smallest threshold (or
    after max_passes plain passes, max_levels denoised levels).

      plain     progressive 4-spp (8-sample) FP32 passes
      denoised  progressive levels of ProgressiveInputs (8 samples per
                level); after each level the network maps the input, built on
                the device from the scene's sums, to an image (level 1 being
                the standard 1-spp + 4-spp input), blended with the 4-spp input
                by the level's fitted weight when alphas (fit_blend) are
                given.  The time at a level is its passes plus one inference.
    """
    screen = cfg["screen"]
    goal = min(thresholds)
    out = {}
    sc = native.Scene(ctx, mesh, cfg)
    t, curve = 0.0, []
    for p in range(max_passes):
        t += _paint_timed(sc, 4, depth, seed)
        img, c = _mean_volume(sc, screen)
        e = rmse(img, ref)
        curve.append((int(c[0]), t, e))
        if e <= goal:
            break
    out["plain_curve"] = curve
    out["plain_rays"] = sc.statistics()["rays"]
    out["plain_paint_s"] = t
    p = ProgressiveInputs(ctx, mesh, cfg, seed + 1, depth)
    dcurve = []
    conv_ms, conv_flops = [], 0.0
    P = int(np.prod(screen))
    for k in range(max_levels):
        p.advance()
        # the network reads the scene's sums on the device and leaves its
        # (blended) image there; reading it back to score it is not timed
        a = blend_weight(alphas, p.level) if alphas is not None else 1.0
        t1 = time.perf_counter()
        denoiser.forward_scene(p.s4, 8, blend=a, copy=False)
        dt = time.perf_counter() - t1
        y = to_volume(denoiser.output(P), screen)
        ms, conv_flops = denoiser.last_forward()
        conv_ms.append(ms)
        e = rmse(y, ref)
        if k == 0:
            x = p.volume()
            out["denoised_rmse"] = e
            out["input4_rmse"] = rmse(x[..., RESIDUAL:RESIDUAL + 3], ref)
            out["denoiser_inference_s"] = dt
        # a render stopped at this level costs its passes and one inference
        dcurve.append((8 * p.level, p.secs + dt, e))
        if e <= goal:
            break
    out["denoised_curve"] = dcurve
    out["denoised_rays"] = p.s4.statistics()["rays"]
    out["denoised_paint_s"] = p.secs
    out["conv_forward_ms"] = float(np.median(conv_ms))
    out["conv_forward_flops"] = conv_flops
    out["thresholds"] = {}
    for thr in thresholds:
        tp, td = first_time(curve, thr), first_time(dcurve, thr)
        out["thresholds"][str(thr)] = dict(plain_s=tp, denoised_s=td,
                                           speedup=(tp / td) if (tp is not None and td is not None) else None)
    return out
