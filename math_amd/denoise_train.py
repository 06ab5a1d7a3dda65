"""Training recipe of the shipped denoiser weights (math_amd/denoiser_c3.npz).

  python -m math_amd.denoise_train [--steps N] [--out PATH]

Renders the training set of the C3 workload on the GPU (64 n=4 variants,
seeds 100-163: inputs at the progressive levels of denoise_eval.LEVELS from
4-spp passes with AOVs, 1024-spp targets, 64^3 windows of the 128^3 screen),
trains the 8-layer BF16 network with the library's own MFMA forward and
backward kernels and Adam (32^3 patches, batch 16), fits the per-level blend
weights, and writes the parameters, the blend weights and the recipe into one
.npz file (plain arrays and a JSON string: no pickled objects).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "denoiser_c3.npz")


def train(ctx, steps=16000, variants=None, window=64, target_spp=1024, log=print):
    """Trains a denoiser; returns (denoiser, blend weights, recipe dict)."""
    from . import denoise_eval as DE
    from . import native

    variants = list(DE.TRAIN_SEEDS if variants is None else variants)
    t0 = time.perf_counter()
    X, Y = DE.training_set(ctx, variants, window=window, target_spp=target_spp, log=log)
    data_s = time.perf_counter() - t0
    log(f"training set: {X.shape[0]} volumes of {X.shape[1]}^3 x {X.shape[-1]} channels in {data_s:.1f} s")
    dn = native.Denoiser(ctx, DE.IN_CHANNELS, DE.RESIDUAL, seed=1)
    t0 = time.perf_counter()
    chunk = 1000
    losses = []
    for c in range(0, steps, chunk):
        k = min(chunk, steps - c)
        # step decay: 2e-4, then 5e-5 for the last fifth
        lr = 2e-4 if c < 0.8 * steps else 5e-5
        losses.append(dn.train(X, Y, patch=32, batch=16, steps=k, lr=lr, seed=2))
        log(f"steps {c + k}: lr {lr:g}, loss {losses[-1][-100:].mean():.6f} ({time.perf_counter() - t0:.1f} s)")
    losses = np.concatenate(losses)
    train_s = time.perf_counter() - t0
    alphas = DE.fit_blend(dn, X, Y)
    log(f"blend weights per level {alphas}")
    recipe = {"inputs": "1-spp radiance, 4-spp radiance, albedo, normal (13 channels)",
              "variants": variants, "window": window, "target_spp": target_spp, "levels": list(DE.LEVELS),
              "steps": steps, "patch": 32, "batch": 16, "lr": "2e-4, 5e-5 for the last fifth", "seed": 2,
              "data_seconds": round(data_s, 1), "train_seconds": round(train_s, 1),
              "loss_first": float(losses[:10].mean()), "loss_last": float(losses[-200:].mean())}
    return dn, alphas, recipe


def save(path, dn, alphas, recipe):
    levels = np.array(sorted(alphas), np.int64)
    np.savez(path, params=dn.params(), blend_levels=levels,
             blend_weights=np.array([alphas[int(k)] for k in levels], np.float64),
             recipe=np.array(json.dumps(recipe)))


def load(path=WEIGHTS):
    """(parameters float32, blend weights {level: weight}, recipe dict) of a weights file."""
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        alphas = {int(k): float(w) for k, w in zip(z["blend_levels"], z["blend_weights"])}
        recipe = json.loads(str(z["recipe"]))
    return params, alphas, recipe


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=16000)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--variants", type=int, default=64, help="training variants (seeds from 100)")
    args = ap.parse_args()
    sys.path.insert(0, os.path.dirname(HERE))
    from math_amd import denoise_eval as DE
    from math_amd import native

    ctx = native.Context(0)
    dn, alphas, recipe = train(ctx, steps=args.steps, variants=list(DE.TRAIN_SEEDS[:args.variants]),
                               log=lambda m: print(m, flush=True))
    save(args.out, dn, alphas, recipe)
    print(json.dumps(recipe), flush=True)


if __name__ == "__main__":
    main()
