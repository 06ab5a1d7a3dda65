"""Benchmark: Mrays/s of the n-space path tracer at n=4 on one MI355X.

Workload (C1): the tesseract boundary split into 8 x 24^3 cubes x 6 Kuhn
tetrahedra (663,552 3-simplices) on S^3 with radial N(0, 0.02) noise (seed 1),
point light (2,2,2,2), parallel projection onto a 128^3 screen, 16 spp
requested (27 samples per pixel under the reference's k^(n-1) rounding), max
depth 5, FP32 tracing with FP64 near-boundary re-test.  One step = one full
render pass of the screen.  Rays are counted by the reference's rules
(Paintbrush statistics): every trace_path call and every shadow ray.

The JSON line carries the closest-hit kernel's roofline (algorithmic bytes per
launch from in-kernel traversal counters / its average launch time from HIP
events on the render stream) and a CPU baseline (the FP64 CPU restatement on
a bounded crop of the same workload, all host threads).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s at n=4; time to RMSE≤0.01 vs 4096-spp ref, denoised vs plain"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def closest_bytes(st, n):
    """Algorithmic bytes of all closest-hit launches of one pass: ray read
    4(2N+2) + hit write 8 per ray, 8 per node visit, 4 per leaf candidate and
    4(N^2+N) per simplex test (the FP32 SoA record)."""
    rays = st["closest_rays"]
    return (rays * (4 * (2 * n + 2) + 8) + 8 * st["node_visits"] + 4 * st["simplex_tests"]
            + 4 * (n * n + n) * st["simplex_tests"])


def cpu_baseline(v, f, cfg, spp, depth, seconds_target=15.0):
    """FP64 CPU restatement on a crop of the C1 screen, all host threads."""
    import numpy as np  # noqa: F401

    from oracle import oracle as O

    threads = min(os.cpu_count() or 1, 16)
    log(f"cpu baseline: building the checker's tree ({len(f)} simplices)")
    m = O.Mesh(v, f)
    m.build_tree()
    sc = O.make_scene(cfg, spp, depth, 12345)
    screen = cfg["screen"]
    # crop: rows of pixels through the middle of the screen, grown until ~seconds_target
    mid = (screen[2] // 2) * screen[0] * screen[1] + (screen[1] // 2) * screen[0]
    count = 64 * threads
    while True:
        t0 = time.perf_counter()
        _, rays = O.render(m, sc, mid, count, threads)
        dt = time.perf_counter() - t0
        if dt > seconds_target / 4 or count >= screen[0] * screen[1]:
            break
        count *= 4
    if dt < seconds_target:
        count2 = min(int(count * seconds_target / max(dt, 1e-3)), screen[0] * screen[1] * 4)
        t0 = time.perf_counter()
        _, rays = O.render(m, sc, mid - count2 // 2, count2, threads)
        dt = time.perf_counter() - t0
        count = count2
    return dict(value=rays / dt / 1e6, unit="Mrays/s", cores=threads, kind="port",
                sample=f"{count} pixels x 27 samples from the middle of the 128^3 screen, FP64, "
                       f"{rays} rays in {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C1")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-paths", type=int, default=0)
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (to scale the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 26, dtype=torch.float32, device="cuda")  # 256 MiB
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the path tracer shards by screen tiles and "
              "needs no collective; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    from math_amd import configs, native

    v, f, cfg, spp, depth = configs.config(args.config)
    n = cfg["n"]
    log(f"config {args.config}: n={n} F={len(f)} screen={cfg['screen']} spp={spp} depth={depth}")
    ctx = native.Context(0)
    t0 = time.perf_counter()
    mesh = ctx.mesh(v, f)
    build_s = time.perf_counter() - t0
    info = mesh.info()
    log(f"mesh build {build_s:.2f} s (basis {info.basis_ms:.1f} ms, tree {info.tree_ms:.1f} ms): "
        f"{info.tree_nodes} nodes, {info.tree_leaves} leaves, {info.tree_entries} entries, {info.tree_levels} levels")
    scene = native.Scene(ctx, mesh, cfg)
    for w in range(args.warmup):
        scene.paint(spp, depth, seed=1000 + w, batch_paths=args.batch_paths, use_graph=args.graph)
        log(f"warmup {w}: {scene.stage_stats()['ms_total']:.1f} ms")
    r0 = scene.statistics()["rays"]
    closest_ms = 0.0
    closest_launches = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        scene.paint(spp, depth, seed=k, batch_paths=args.batch_paths, use_graph=args.graph)
        st = scene.stage_stats()
        closest_ms += st["ms_closest"]
        closest_launches += st["launches_closest"]
        log(f"step {k}: {st['ms_total']:.1f} ms, closest {st['ms_closest']:.1f} ms, shade {st['ms_shade']:.1f} ms, "
            f"shadow {st['ms_shadow']:.1f} ms")
    elapsed = time.perf_counter() - t0
    rays = scene.statistics()["rays"] - r0
    mrays = rays / elapsed / 1e6
    # traversal counters in a separate pass (the atomics would perturb the timing)
    scene.set_counters(True)
    scene.paint(spp, depth, seed=0, batch_paths=args.batch_paths)
    cst = scene.stage_stats()
    scene.set_counters(False)
    bytes_pass = closest_bytes(cst, n)
    achieved = bytes_pass / (closest_ms / args.steps * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="k_closest", launches_per_step=closest_launches // max(args.steps, 1),
                    avg_launch_ms=round(closest_ms / max(closest_launches, 1), 4),
                    bytes_per_step=int(bytes_pass),
                    per_ray=dict(nodes=cst["node_visits"] / max(cst["closest_rays"], 1),
                                 leaves=cst["leaf_visits"] / max(cst["closest_rays"], 1),
                                 tests=cst["simplex_tests"] / max(cst["closest_rays"], 1),
                                 retest_frac=cst["retests"] / max(cst["simplex_tests"], 1)))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(v, f, cfg, spp, depth)
    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: n=4 plain path trace, 663,552-tetrahedron noisy S^3 mesh, "
                               "128^3 parallel-projected screen, 16 spp requested (27 samples/pixel), depth 5, "
                               "point light (2,2,2,2), FP32 trace + FP64 re-test",
                   "rays_per_step": rays // args.steps, "mesh_build_s": round(build_s, 3),
                   "tree": dict(nodes=info.tree_nodes, leaves=info.tree_leaves, entries=info.tree_entries)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
