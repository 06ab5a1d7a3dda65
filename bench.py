"""Benchmark: Mrays/s of the n-space path tracer at n=4 on one MI355X.

Workload (C1): the tesseract boundary split into 8 x 24^3 cubes x 6 Kuhn
tetrahedra (663,552 3-simplices) on S^3 with radial N(0, 0.02) noise (seed 1),
point light (2,2,2,2), parallel projection onto a 128^3 screen, 16 spp
requested (27 samples per pixel under the reference's k^(n-1) rounding), max
depth 5, FP32 tracing with FP64 near-boundary re-test.  One step = one full
render pass of the screen.  Rays are counted by the reference's rules
(Paintbrush statistics): every trace_path call and every shadow ray.

The JSON line carries the roofline of the traversal kernel with the most time
per step (algorithmic bytes per launch from in-kernel traversal counters / its
average launch time from HIP events on the stream it runs on), a CPU baseline
(the FP64 CPU restatement on a seeded sample of pixels spread over the whole
screen, on every CPU the process may run on) and a parity object (one GPU pass
against the CPU restatement on those pixels and sample streams).
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


def walk_bytes(st, n):
    """Algorithmic bytes of the closest-hit tree walks of one pass (the bounce
    rays; rasterised primary rays are not walked): ray read 4(2N+2) + hit write
    8 per ray, 8 per node visit, 4 per leaf candidate (its simplex index) and
    4(N^2+N) per simplex test (the FP32 record)."""
    rays = st["closest_rays"] - st["primary_rays"]
    return (rays * (4 * (2 * n + 2) + 8) + 8 * st["node_visits"]
            + (4 + 4 * (n * n + n)) * st["simplex_tests"])


def shadow_walks(st):
    """Shadow rays walked through the tree (the shadow rays of the primary hits
    are rasterised from the light when the scene has one point light)."""
    return st["shadow_rays"] - st.get("shadow_rays_rasterised", 0)


def shadow_bytes(st, n):
    """Algorithmic bytes of the shadow (any-hit) walks of one pass: ray read
    4(2N+1) + path 4 + contribution 12 per walked ray, 8 per node visit,
    4 + 4(N^2+N) per simplex test."""
    return (shadow_walks(st) * (4 * (2 * n + 1) + 16) + 8 * st["shadow_node_visits"]
            + (4 + 4 * (n * n + n)) * st["shadow_simplex_tests"])


def profiled_traffic(config, kernel, layout):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE
    / WRITE_SIZE summary of the same workload and record layout (scaled by
    the copy-kernel calibration stated there) and that source, or (None, None)."""
    name = f"{kernel}_{config.lower()}.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    if f"Record layout: {layout}" not in text:
        return None, None
    for line in text.splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1]), (f"profiles/{name}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of the "
                                             f"same workload, copy-calibrated (not measured in this run)")
    return None, None


def cpu_baseline(ctx, mesh, v, f, cfg, spp, depth, seconds_target=15.0):
    """The FP64 CPU restatement on every CPU this process may run on, timed on a
    seeded sample of pixels spread over the whole screen, with all their
    samples; the sample grows until its render takes about seconds_target.
    The same pixels of one GPU pass with the same sample streams (outside the
    timed region) are compared with the checker's: the `parity` object."""
    import numpy as np

    from math_amd import native
    from oracle import oracle as O

    threads = len(os.sched_getaffinity(0))
    log(f"cpu baseline: building the checker's tree ({len(f)} simplices, {threads} threads)")
    m = O.Mesh(v, f)
    m.build_tree(threads=threads)
    seed = 12345
    sc = O.make_scene(cfg, spp, depth, seed)
    total = int(np.prod(cfg["screen"]))
    order = np.random.default_rng(seed).permutation(total)
    count = 256 * threads
    while True:
        px = np.sort(order[:count])
        t0 = time.perf_counter()
        ref, rays = O.render_pixels(m, sc, px, threads)
        dt = time.perf_counter() - t0
        if dt > seconds_target / 3 or count >= total:
            break
        count = min(total, int(count * min(8.0, max(1.5, seconds_target / 2.5 / max(dt, 1e-3)))))
    cpu = dict(value=rays / dt / 1e6, unit="Mrays/s", cores=threads, kind="port",
               sample=f"{count} of {total} pixels ({cfg['screen']}) drawn uniformly over the screen (seed {seed}), "
                      f"all samples, FP64, {rays} rays in {dt:.1f} s on {threads} threads "
                      f"(the host's CPUs available to this process: {threads}; os.cpu_count() {os.cpu_count()})")
    # parity: one GPU pass of the benchmark's defaults on the same sample streams
    gs = native.Scene(ctx, mesh, cfg)
    gs.paint(spp, depth, seed=seed)
    img, cnt = gs.image()
    gs.close()
    g = img[px] / cnt[px][:, None]
    o = ref / cnt[px][:, None]
    diff = g.astype(np.float64) - o
    parity = dict(checker="FP64 CPU restatement, same pixels and sample streams", pixels=int(count),
                  rmse=float(np.sqrt(np.mean(diff ** 2))), mean_diff=float(diff.mean()),
                  max_abs=float(np.abs(diff).max()), bar_rmse=1e-2)
    return cpu, parity


def denoise_eval(ctx, scenes, args):
    """C3: time to RMSE <= 0.03, 0.02 and 0.01 against a 4096-sample FP64 render,
    FP32 + denoiser vs plain FP32, on held-out 128^3 scenes (seeds 200, 201, ...).
    The network is the shipped one (math_amd/denoiser_c3.npz, trained by
    math_amd/denoise_train.py with the library's kernels) unless --denoise-train
    trains it here first."""
    import numpy as np

    from math_amd import denoise_eval as DE
    from math_amd import denoise_train as DT
    from math_amd import native

    if args.denoise_train:
        dn, alphas, recipe = DT.train(ctx, steps=args.train_steps,
                                      variants=list(DE.TRAIN_SEEDS[:args.train_variants]), log=log)
    else:
        params, alphas, recipe = DT.load()
        dn = native.Denoiser(ctx, DE.IN_CHANNELS, DE.RESIDUAL, seed=1)
        dn.set_params(params)
    results = []
    for seed in DE.HELDOUT_SEEDS[:scenes]:
        v, f, cfg = DE.variant(seed, 128)
        mesh = ctx.mesh(v, f)
        # 8 passes of 8^3 = 512 stratified samples: 4096 samples per pixel
        ref, ref_s = DE.render_mean(ctx, mesh, cfg, 512, seed=seed * 7, precision=native.PRECISION_FP64, passes=8)
        log(f"held-out scene {seed}: FP64 4096-sample reference in {ref_s:.1f} s")
        r = DE.time_to_rmse(ctx, mesh, cfg, ref, dn, seed=seed * 11, alphas=alphas)
        r["seed"] = seed
        log(f"held-out scene {seed}: {r['thresholds']}")
        results.append(r)
        mesh.close()
    thr = [str(t) for t in DE.THRESHOLDS]
    per_scene = []
    for r in results:
        per_scene.append({"seed": r["seed"], **{t: r["thresholds"][t] for t in thr},
                          "level1_rmse": {"input_4spp": round(r["input4_rmse"], 5),
                                          "denoised": round(r["denoised_rmse"], 5)}})
    geo, reached = {}, {}
    for t in thr:
        sp = [r["thresholds"][t]["speedup"] for r in results if r["thresholds"][t]["speedup"]]
        geo[t] = float(np.exp(np.mean(np.log(sp)))) if sp else None
        # scenes where both curves reached the threshold within 256 levels of 8 samples
        reached[t] = f"{len(sp)} of {len(results)}"
    plain_rays = sum(r["plain_rays"] for r in results)
    plain_s = sum(r["plain_paint_s"] for r in results)
    dn_rays = sum(r["denoised_rays"] for r in results)
    dn_s = sum(r["denoised_paint_s"] for r in results)
    conv_ms = float(np.median([r["conv_forward_ms"] for r in results]))
    conv_flops = results[0]["conv_forward_flops"]
    tflops = conv_flops / (conv_ms * 1e-3) / 1e12
    return {
        "metric": "time to RMSE <= thr against a 4096-sample FP64 render at n=4, FP32 + denoiser vs plain FP32",
        "scenes": f"{len(results)} of the 16 held-out 128^3 scenes (seeds {DE.HELDOUT_SEEDS[0]}-"
                  f"{DE.HELDOUT_SEEDS[len(results) - 1]})",
        "per_scene": per_scene,
        "geomean_speedup": geo,
        "geomean_over": reached,
        "plain": {"Mrays_s": round(plain_rays / plain_s / 1e6, 1),
                  "roofline": "the traversal kernels of the Mrays/s line (same kernels, 4-spp passes)"},
        "denoised": {"Mrays_s": round(dn_rays / dn_s / 1e6, 1),
                     "denoiser_roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": 2500.0,
                                           "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 4),
                                           "kernel": "k_conv (8 layers, 128^3, BF16 MFMA, FP32 accumulate)",
                                           "forward_ms": round(conv_ms, 3),
                                           "flops": conv_flops}},
        "blend_weights": alphas,
        "network": "shipped weights (math_amd/denoiser_c3.npz)" if not args.denoise_train else "trained in this run",
        "recipe": recipe,
    }


def denoiser_forward(args):
    """Five forwards of the shipped network on one 128^3 volume of 13 random input
    channels (the inference shape of the time-to-RMSE evaluation), with the
    device time of the eight convolutions: the workload of the denoiser's
    profiles."""
    import numpy as np

    from math_amd import denoise_eval as DE
    from math_amd import denoise_train as DT
    from math_amd import native

    ctx = native.Context(0)
    dn = native.Denoiser(ctx, DE.IN_CHANNELS, DE.RESIDUAL, seed=1)
    dn.set_params(DT.load()[0])
    x = np.random.default_rng(0).random(size=(1, 128, 128, 128, DE.IN_CHANNELS)).astype(np.float32)
    ms = []
    for k in range(5):
        dn.forward(x)
        ms.append(dn.last_forward()[0])
        log(f"forward {k}: {ms[-1]:.3f} ms on the device")
    flops = dn.last_forward()[1]
    print(json.dumps({"workload": "denoiser forward, 128^3 voxels, 13 -> 32 x 7 -> 3 channels, 3x3x3",
                      "device_ms": ms, "flops": flops,
                      "tflops_best": round(flops / (min(ms) * 1e-3) / 1e12, 1)}), flush=True)


def denoise_bench(args):
    from math_amd import native

    ctx = native.Context(0)
    print(json.dumps(denoise_eval(ctx, args.eval_scenes, args)), flush=True)


def obj_bench(args):
    """.objN ingest: the C1 mesh written as an OBJ4 file, read back with the
    library's chunked reader (all host threads, and one), and a mesh built on
    the GPU from the read file; the oracle's line-by-line reader timed on a
    bounded prefix of the same file."""
    import os
    import tempfile

    import numpy as np

    from math_amd import configs, native
    from oracle import obj_ref

    v, f, _, _, _ = configs.config("C1")
    out = {"metric": ".obj4 ingest of the C1 mesh", "unit": "MB/s"}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c1.obj4")
        t0 = time.perf_counter()
        native.write_obj(path, v, f)
        out["write_s"] = time.perf_counter() - t0
        size = os.path.getsize(path)
        out["file_mb"] = size / 1e6
        for threads in (0, 1):
            best = None
            for _ in range(3):
                t0 = time.perf_counter()
                obj = native.read_obj(path, threads)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            key = "all_threads" if threads == 0 else "one_thread"
            out[key] = {"seconds": best, "mb_per_s": size / 1e6 / best}
            log(f"read ({key}): {best * 1e3:.1f} ms, {size / 1e6 / best:.0f} MB/s")
        assert obj.facets.shape == f.shape
        ctx = native.Context(0) if not args.no_gpu_mesh else None
        if ctx is not None:
            t0 = time.perf_counter()
            m = native.Mesh.from_obj(ctx, obj)
            out["mesh_from_obj_s"] = time.perf_counter() - t0
            info = m.info()
            out["mesh_build_ms"] = {"basis": info.basis_ms, "tree": info.tree_ms}
            m.close()
        # the oracle on a bounded sample: the first facets of the mesh with the
        # vertices they use
        used, inv = np.unique(f[:args.obj_sample_facets], return_inverse=True)
        sample = os.path.join(d, "sample.obj4")
        native.write_obj(sample, v[used], inv.reshape(-1, f.shape[1]).astype(np.int32))
        ssize = os.path.getsize(sample)
        t0 = time.perf_counter()
        obj_ref.read(sample)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ssize / 1e6 / dt, "unit": "MB/s", "cores": 1, "kind": "port",
                               "sample": f"oracle line-by-line reader on the first {args.obj_sample_facets} facets "
                                         f"of the mesh and their vertices ({ssize / 1e6:.1f} MB)"}
    print(json.dumps(out), flush=True)


def hull_bench(args):
    """Convex hull in 4 dimensions of 100,000 points in the unit ball, of the
    same count in a hyperplane with one apex, and of 20,000 points on the
    sphere (test_convex_hull_speed and sphere_mesh's workloads): exact
    predicates, the first 4096 points on the host, the conflict graph of the rest
    in one GPU pass, then the host again; beside it the same library with every
    point on the host, and the oracle (exact Python) on a bounded sample."""
    import numpy as np

    from math_amd import configs, native
    from oracle import hull_ref

    ctx = native.Context(0)
    native.convex_hull(ctx, configs.hull_points(4, 2000, 1))  # warm-up (kernel load)
    out = {"metric": "convex hull in 4 dimensions", "unit": "ms", "workloads": []}
    for name, count, sphere, zero in (("ball", 100000, False, False), ("hyperplane+apex", 100000, False, True),
                                      ("sphere", 20000, True, False)):
        p = configs.hull_points(4, count, 80, on_sphere=sphere, zero=zero)
        t0 = time.perf_counter()
        a = native.convex_hull(ctx, p)
        ta = time.perf_counter() - t0
        t0 = time.perf_counter()
        b = native.convex_hull(ctx, p, count)
        tb = time.perf_counter() - t0
        same = set(map(tuple, a.vertices.tolist())) == set(map(tuple, b.vertices.tolist()))
        w = {"workload": f"{name}, {count} points", "facets": len(a.vertices), "ms": ta * 1e3,
             "ms_split": {"host_first_4096": a.ms[0], "gpu_conflict_pass": a.ms[1], "host_rest": a.ms[2]},
             "dropped_by_gpu_pass": a.dropped, "host_only_ms": tb * 1e3, "same_facets_as_host_only": same}
        log(f"hull {name}: {ta * 1e3:.0f} ms ({a.ms[0]:.0f} + GPU {a.ms[1]:.1f} + {a.ms[2]:.0f}), dropped "
            f"{a.dropped}, {len(a.vertices)} facets; host only {tb * 1e3:.0f} ms; same {same}")
        out["workloads"].append(w)
    # the oracle on a bounded sample of the ball workload
    p = configs.hull_points(4, args.hull_sample, 80)
    t0 = time.perf_counter()
    hull_ref.convex_hull(p, np.arange(len(p)))
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": dt * 1e3, "unit": "ms", "cores": 1, "kind": "port",
                           "sample": f"oracle (exact Python integers) on {args.hull_sample} points of the ball workload"}
    print(json.dumps(out), flush=True)


def cocone_bench(args):
    """Cocone and BoundCocone (rho 0.3, alpha 0.14) of the reconstruction
    test's notched spheres at its largest sizes: 3000 points in 3 dimensions,
    the same with its 8 clones (27,000 points), 20,000 points in 4 dimensions.
    Per workload: the constructor (exact Delaunay, facets, per-point GPU pass),
    then each reconstruction; beside it the oracle (C, one core) on the same
    triangulation, i.e. every step after the Delaunay."""
    from math_amd import configs, native
    from oracle import oracle as O

    ctx = native.Context(0)
    native.ManifoldConstructor(ctx, configs.sphere_with_notch(3, 200)).cocone()  # warm-up (kernel load)
    out = {"metric": "Cocone and BoundCocone reconstruction", "unit": "ms", "workloads": []}
    oracle_ms = lib_ms = 0.0
    for name, n, count, clones in (("notched 2-sphere", 3, 3000, False), ("notched 2-sphere, 9 objects", 3, 3000, True),
                                   ("notched 3-sphere", 4, 20000, False)):
        p = configs.sphere_with_notch(n, count)
        if clones:
            p = configs.clone_points(p)
        t0 = time.perf_counter()
        m = native.ManifoldConstructor(ctx, p)
        t_create = time.perf_counter() - t0
        create_ms = m.stats()
        t0 = time.perf_counter()
        _, fc = m.cocone()
        t_cocone = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, fb = m.bound_cocone(0.3, 0.14)
        t_bound = time.perf_counter() - t0
        bound_ms = m.stats()
        w = {"workload": f"{name}, {len(p)} points", "delaunay_objects": m.objects, "cocone_facets": len(fc),
             "bound_cocone_facets": len(fb), "constructor_ms": t_create * 1e3, "cocone_ms": t_cocone * 1e3,
             "bound_cocone_ms": t_bound * 1e3,
             "constructor_split_ms": {"delaunay": create_ms[0], "facets_host": create_ms[1], "gpu": create_ms[2]},
             "bound_cocone_split_ms": {"gpu": bound_ms[3], "prune_host": bound_ms[4], "extract_host": bound_ms[5]}}
        m.close()
        if not args.no_cpu:
            d = native.delaunay(ctx, p)
            t0 = time.perf_counter()
            O.cocone(d.points, d.vertices, d.orthos)
            O.cocone(d.points, d.vertices, d.orthos, bound=True, rho=0.3, alpha=0.14)
            o_ms = (time.perf_counter() - t0) * 1e3
            l_ms = (t_create * 1e3 - create_ms[0]) + (t_cocone + t_bound) * 1e3
            w["after_delaunay_ms"] = l_ms
            w["oracle_after_delaunay_ms"] = o_ms
            oracle_ms += o_ms
            lib_ms += l_ms
        log(f"cocone {w['workload']}: {m.objects} simplices, Cocone {len(fc)}, BoundCocone {len(fb)}; constructor "
            f"{t_create * 1e3:.0f} ms (Delaunay {create_ms[0]:.0f}, facets {create_ms[1]:.0f}, GPU {create_ms[2]:.1f}), "
            f"cocone {t_cocone * 1e3:.1f} ms, bound {t_bound * 1e3:.1f} ms"
            + (f"; after Delaunay {w['after_delaunay_ms']:.0f} ms vs oracle {w['oracle_after_delaunay_ms']:.0f} ms"
               if not args.no_cpu else ""))
        out["workloads"].append(w)
    if not args.no_cpu:
        out["cpu_baseline"] = {"value": oracle_ms, "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "oracle (C) Cocone + BoundCocone on the library's triangulations of the "
                                         f"three workloads; the library's same steps take {lib_ms:.0f} ms"}
    print(json.dumps(out), flush=True)


def batch_bench(args):
    """C4: many scenes of uneven size and dimension, 4 spp, FP32: the batched
    launch (scenes on concurrent lanes) vs one nsp_paint call per scene."""
    import numpy as np

    from math_amd import configs, native

    ctx = native.Context(0)
    meshes = []
    t0 = time.perf_counter()
    for n, g, edge, seed in configs.c4_scenes(256)[:args.scenes]:
        v, f, cfg = configs.c4_scene(n, g, edge, seed)
        meshes.append((ctx.mesh(v, f), cfg, len(f)))
    log(f"{args.scenes} meshes built in {time.perf_counter() - t0:.1f} s "
        f"({min(m[2] for m in meshes)}..{max(m[2] for m in meshes)} simplices)")
    per = [native.Scene(ctx, m, c) for m, c, _ in meshes]
    # warm-up: per-scene set-up (raster boxes, workspaces) is not timed, as for the batch
    for s in per:
        s.paint(4, 5, seed=1)
        s.set_image(*[np.zeros_like(a) for a in s.image()])
    r0 = sum(s.statistics()["rays"] for s in per)
    t0 = time.perf_counter()
    for i, s in enumerate(per):
        s.paint(4, 5, seed=1)
        log(f"per-scene {i}: n={s.n} {s.pixels} pixels, {time.perf_counter() - t0:.1f} s cumulative")
    per_s = time.perf_counter() - t0
    per_rays = sum(s.statistics()["rays"] for s in per) - r0
    bat = [native.Scene(ctx, m, c) for m, c, _ in meshes]
    b = native.Batch(ctx, bat)
    if args.lanes > 0:
        b.set_lanes(args.lanes)
    if args.batch_autotune:
        b.autotune(4, 5, seed=1)  # not timed; images are restored
        log(f"autotuned batch: {b.lanes} lanes")
    b.paint(4, 5, seed=1)  # warm-up: workspaces, raster set-up
    for s in bat:
        s.set_image(*[np.zeros_like(a) for a in s.image()])
    t0 = time.perf_counter()
    b.paint(4, 5, seed=1)
    bat_s = time.perf_counter() - t0
    bat_rays = b.statistics()["rays"]
    log(f"batched pass {bat_s:.2f} s on {b.lanes} lanes")
    graph = None
    if args.batch_graph:
        # the same pass replayed from a captured hipGraph (captured and
        # instantiated by the untimed first call)
        t0 = time.perf_counter()
        b.paint(4, 5, seed=1, use_graph=True)
        log(f"graph capture and first replay {time.perf_counter() - t0:.1f} s")
        t0 = time.perf_counter()
        b.paint(4, 5, seed=1, use_graph=True)
        gs = time.perf_counter() - t0
        graph = {"seconds": round(gs, 3), "Mrays_s": round(b.statistics()["rays"] / gs / 1e6, 2)}
    out = {"metric": "C4 batched launch (concurrent lanes), Mrays/s", "scenes": args.scenes,
           "lanes": b.lanes, "autotuned": bool(args.batch_autotune),
           "per_scene": {"seconds": round(per_s, 3), "Mrays_s": round(per_rays / per_s / 1e6, 2)},
           "batched": {"seconds": round(bat_s, 3), "Mrays_s": round(bat_rays / bat_s / 1e6, 2)},
           "speedup": round(per_s / bat_s, 3), "graph_replay": graph}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C1")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-paths", type=int, default=0)
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--primary", choices=["auto", "walk"], default="auto",
                    help="primary visibility: rasterised (auto) or the tree walk")
    ap.add_argument("--layout", type=int, default=-1,
                    help="FP32 record layout: 0 leaf-ordered SoA, 1 simplex-ordered AoS (default: the library's)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (to scale the FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--denoise", action="store_true",
                    help="C3: train the denoiser on renders of n=4 variants, then time to RMSE<=0.01, "
                         "denoised vs plain")
    ap.add_argument("--denoise-train", action="store_true",
                    help="with --denoise: train the network here (recipe of math_amd/denoise_train.py) instead of "
                         "loading the shipped weights")
    ap.add_argument("--train-variants", type=int, default=64)
    ap.add_argument("--train-steps", type=int, default=16000)
    ap.add_argument("--eval-scenes", type=int, default=16, help="with --denoise: held-out scenes to evaluate")
    ap.add_argument("--denoise-scenes", type=int, default=16,
                    help="held-out scenes of the denoise object in the default run (0: none)")
    ap.add_argument("--obj", action="store_true", help=".obj4 ingest of the C1 mesh (reader MB/s)")
    ap.add_argument("--hull", action="store_true", help="convex hulls of 4-dimensional point sets (ms)")
    ap.add_argument("--hull-sample", type=int, default=3000, help="with --hull: points the oracle takes")
    ap.add_argument("--cocone", action="store_true", help="Cocone and BoundCocone reconstructions (ms)")
    ap.add_argument("--no-gpu-mesh", action="store_true", help="with --obj: skip the GPU mesh build")
    ap.add_argument("--obj-sample-facets", type=int, default=60000, help="with --obj: facets read by the oracle")
    ap.add_argument("--batch-scenes", action="store_true",
                    help="C4: batched persistent launch over many scenes vs one launch sequence per scene")
    ap.add_argument("--scenes", type=int, default=16,
                    help="C4: the first scenes of the 256 seeded scenes (all 256 meshes do not fit beside the workspaces)")
    ap.add_argument("--lanes", type=int, default=0, help="batch lanes (0: the library's default)")
    ap.add_argument("--batch-graph", action="store_true",
                    help="also time the batch replayed from its captured hipGraph")
    ap.add_argument("--batch-autotune", action="store_true",
                    help="time lane counts first (two whole batched passes per candidate)")
    ap.add_argument("--denoiser-forward", action="store_true",
                    help="five 128^3 forwards of the shipped denoiser (the workload of its profiles)")
    args = ap.parse_args()
    if args.denoiser_forward:
        return denoiser_forward(args)
    if args.denoise:
        return denoise_bench(args)
    if args.batch_scenes:
        return batch_bench(args)
    if args.hull:
        return hull_bench(args)
    if args.cocone:
        return cocone_bench(args)
    if args.obj:
        return obj_bench(args)
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
    if args.layout >= 0:
        mesh.record_layout = args.layout
    log(f"record layout {mesh.record_layout}")
    scene = native.Scene(ctx, mesh, cfg)
    primary = native.PRIMARY_AUTO if args.primary == "auto" else native.PRIMARY_WALK
    for w in range(args.warmup):
        scene.paint(spp, depth, seed=1000 + w, batch_paths=args.batch_paths, use_graph=args.graph, primary=primary)
        log(f"warmup {w}: {scene.stage_stats()['ms_total']:.1f} ms")
    r0 = scene.statistics()["rays"]
    stage_ms = dict(raster=0.0, closest=0.0, shade=0.0, shadow=0.0, walk=0.0)
    launches = dict(closest=0, shadow=0, walk=0)
    t0 = time.perf_counter()
    for k in range(args.steps):
        scene.paint(spp, depth, seed=k, batch_paths=args.batch_paths, use_graph=args.graph, primary=primary)
        st = scene.stage_stats()
        for key in stage_ms:
            stage_ms[key] += st["ms_" + key]
        for key in launches:
            launches[key] += st["launches_" + key]
        log(f"step {k}: {st['ms_total']:.1f} ms, raster {st['ms_raster']:.1f} ms, closest {st['ms_closest']:.1f} ms, "
            f"shade {st['ms_shade']:.1f} ms, shadow {st['ms_shadow']:.1f} ms, walk {st['ms_walk']:.1f} ms")
    elapsed = time.perf_counter() - t0
    rays = scene.statistics()["rays"] - r0
    mrays = rays / elapsed / 1e6
    # traversal counters in a separate pass (the atomics would perturb the timing)
    scene.set_counters(True)
    scene.paint(spp, depth, seed=0, batch_paths=args.batch_paths, primary=primary)
    cst = scene.stage_stats()
    scene.set_counters(False)
    # roofline of the dominant traversal kernel (the closest-hit walks of the bounce rays or the
    # shadow walks, whichever took longer; both are HBM/gather kernels with no MFMA work); the
    # other one is reported beside it
    layout_name = ["leaf_soa", "simplex_aos"][mesh.record_layout]
    walk_rays = max(cst["closest_rays"] - cst["primary_rays"], 1)
    sr = max(shadow_walks(cst), 1)
    per_closest = dict(rays=cst["closest_rays"] - cst["primary_rays"], nodes=cst["node_visits"] / walk_rays,
                       leaves=cst["leaf_visits"] / walk_rays, tests=cst["simplex_tests"] / walk_rays,
                       fp64_frac=cst["retests"] / max(cst["simplex_tests"], 1))
    per_shadow = dict(rays=shadow_walks(cst), nodes=cst["shadow_node_visits"] / sr,
                      leaves=cst["shadow_leaf_visits"] / sr, tests=cst["shadow_simplex_tests"] / sr,
                      fp64_frac=cst["shadow_retests"] / max(cst["shadow_simplex_tests"], 1))
    kernels = {
        "k_closest": dict(bytes=walk_bytes(cst, n), ms=stage_ms["closest"], launches=launches["closest"],
                          per_ray=per_closest),
        "k_shadow": dict(bytes=shadow_bytes(cst, n), ms=stage_ms["shadow"], launches=launches["shadow"],
                         per_ray=per_shadow),
        # one launch per depth d holding the closest-hit walks of depth d + 1 and the shadow walks of depth d
        "k_walk": dict(bytes=walk_bytes(cst, n) + shadow_bytes(cst, n), ms=stage_ms["walk"],
                       launches=launches["walk"], per_ray=dict(closest=per_closest, shadow=per_shadow)),
    }

    def line(name):
        k = kernels[name]
        achieved = k["bytes"] / (k["ms"] / args.steps * 1e-3) / 1e9
        traffic, source = profiled_traffic(args.config, name, layout_name)
        return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=source,
                    kernel=name, launches_per_step=k["launches"] // max(args.steps, 1),
                    avg_launch_ms=round(k["ms"] / max(k["launches"], 1), 4),
                    bytes_per_launch=int(k["bytes"] / max(k["launches"] // max(args.steps, 1), 1)),
                    per_ray=k["per_ray"])

    if kernels["k_walk"]["ms"] > 0:
        # the merged walks: one kernel holds all traversal of the bounce and shadow rays
        roofline = line("k_walk")
        roofline["stage_ms_per_step"] = {k: round(v / args.steps, 2) for k, v in stage_ms.items()}
        if cst["shadow_rays_rasterised"]:
            # the shadow rays of the primary hits, rasterised from the light (stage "shadow")
            # (on the side stream, beside the depth-0 walk: the stage time includes sharing the GPU)
            roofline["shadow_raster"] = dict(rays=cst["shadow_rays_rasterised"],
                                             ms_per_step_beside_walk=round(stage_ms["shadow"] / args.steps, 3))
    elif min(kernels[k]["ms"] for k in ("k_closest", "k_shadow")) > 0:
        dominant = max(("k_closest", "k_shadow"), key=lambda name: kernels[name]["ms"])
        roofline = line(dominant)
        roofline["other"] = line("k_shadow" if dominant == "k_closest" else "k_closest")
        roofline["stage_ms_per_step"] = {k: round(v / args.steps, 2) for k, v in stage_ms.items()}
        # the shadow walks run on a second stream beside the next depth's closest-hit walks, so
        # each kernel's launch time includes sharing the GPU with the other; both walks' bytes
        # over the whole step bound their joint rate from below
        both = (kernels["k_closest"]["bytes"] + kernels["k_shadow"]["bytes"]) / (elapsed / args.steps) / 1e9
        roofline["both_walks_over_step"] = dict(achieved=round(both, 1), unit="GB/s",
                                                frac=round(both / HBM_PEAK_GBS, 4))
    else:
        roofline = None  # hipGraph replay: no per-stage events inside the graph
        log("stage times are not measured inside a captured graph: no roofline line")
    cpu = parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(ctx, mesh, v, f, cfg, spp, depth)
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
        "config": {"workload": configs.WORKLOADS[args.config] + ", FP32 trace + FP64 re-test",
                   "rays_per_step": rays // args.steps, "mesh_build_s": round(build_s, 3),
                   "record_layout": layout_name,
                   "tree": dict(nodes=info.tree_nodes, leaves=info.tree_leaves, entries=info.tree_entries)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    if args.denoise_scenes > 0:
        out["denoise"] = denoise_eval(ctx, args.denoise_scenes, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
