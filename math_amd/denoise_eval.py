"""Denoiser workload (C3): training data from the tracer itself, training, and
the time-to-RMSE comparison of denoised against plain FP32 rendering.

  variants   n=4 tesseract-boundary meshes of the C1 generator with radial noise
             sigma ~ U[0, 0.05] and a point light at distance 4 in a direction
             uniform on S^3 (seeded)
  inputs     1-spp and 4-spp (8 samples) radiance, first-hit albedo and normal
             (13 channels per voxel of the 3D screen)
  target     1024-spp (1331 samples) FP32 render
  network    8 layers 3x3x3, 32 channels, BF16/FP32-accumulate, residual on the
             4-spp radiance, trained on 32^3 patches, batch 16, Adam
  judge      RMSE over pixels x 3 linear-RGB channels against a 4096-spp FP64
             render of the same scene; the plain render is the fallback
"""
import time

import numpy as np

from . import configs, native

IN_CHANNELS = 13
RESIDUAL = 3  # the 4-spp radiance channels


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


def render_inputs(ctx, mesh, cfg, seed, depth=5):
    """13-channel denoiser input volume and the render seconds spent."""
    t0 = time.perf_counter()
    s1 = native.Scene(ctx, mesh, cfg)
    s1.paint(1, depth, seed=seed, aov=True)
    r1, c1 = s1.image()
    alb, nrm = s1.aov()
    s4 = native.Scene(ctx, mesh, cfg)
    s4.paint(4, depth, seed=seed + 7919)
    r4, c4 = s4.image()
    secs = time.perf_counter() - t0
    k1 = np.maximum(c1, 1)[:, None]
    x = np.concatenate([r1 / k1, r4 / np.maximum(c4, 1)[:, None], alb / k1, nrm / k1], axis=1)
    return to_volume(x.astype(np.float32), cfg["screen"]), secs


def render_mean(ctx, mesh, cfg, spp, seed, depth=5, precision=native.PRECISION_FP32, passes=1):
    sc = native.Scene(ctx, mesh, cfg)
    t0 = time.perf_counter()
    sc.paint(spp, depth, seed=seed, precision=precision, passes=passes)
    secs = time.perf_counter() - t0
    s, c = sc.image()
    return to_volume((s / np.maximum(c, 1)[:, None]).astype(np.float32), cfg["screen"]), secs


def rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))


def training_set(ctx, variants, screen=128, window=64, target_spp=1024, depth=5, log=print):
    xs, ys = [], []
    for i, seed in enumerate(variants):
        rng = np.random.default_rng(seed)
        org = [int(rng.integers(0, screen - window + 1)) for _ in range(3)]
        v, f, cfg = variant(seed, screen, (org, window))
        mesh = ctx.mesh(v, f)
        x, _ = render_inputs(ctx, mesh, cfg, seed=10 * seed + 1, depth=depth)
        y, secs = render_mean(ctx, mesh, cfg, target_spp, seed=10 * seed + 2, depth=depth)
        xs.append(x)
        ys.append(y)
        log(f"training volume {i}: seed {seed}, window {org}, target {secs:.2f} s")
        mesh.close()
    return np.stack(xs), np.stack(ys)


def time_to_rmse(ctx, mesh, cfg, ref, denoiser, depth=5, threshold=0.01, max_passes=64, seed=77):
    """Plain: progressive 4-spp (8-sample) passes until RMSE <= threshold.
    Denoised: 1-spp + 4-spp inputs -> network.  Returns a dict of times and errors."""
    out = {}
    sc = native.Scene(ctx, mesh, cfg)
    t_plain = None
    curve = []
    t0 = time.perf_counter()
    for p in range(max_passes):
        sc.paint(4, depth, seed=seed)
        s, c = sc.image()
        img = to_volume(s / np.maximum(c, 1)[:, None], cfg["screen"])
        e = rmse(img, ref)
        el = time.perf_counter() - t0
        curve.append((int(c[0]), el, e))
        if e <= threshold and t_plain is None:
            t_plain = el
            break
    out["plain_curve"] = curve
    out["plain_time_s"] = t_plain
    x, t_in = render_inputs(ctx, mesh, cfg, seed=seed + 1, depth=depth)
    t1 = time.perf_counter()
    y = denoiser.forward(x[None])[0]
    t_net = time.perf_counter() - t1
    out["denoised_rmse"] = rmse(y, ref)
    out["input4_rmse"] = rmse(x[..., 3:6], ref)
    out["denoised_time_s"] = t_in + t_net
    out["denoiser_inference_s"] = t_net
    return out
