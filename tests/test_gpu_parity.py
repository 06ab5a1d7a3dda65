"""GPU parity tests: the HIP path (through the C ABI) against the FP64 CPU checker.

Bars: simplex indices bit-exact except on rays the checker or the GPU flags as
within the near-boundary band; hit distance within 1e-5 relative; FP64 basis
records bit-equal; tree leaf lists equal to the checker's; renders within the
stated RMSE of the checker's FP64 render on the same counter-based sample streams.
"""
import numpy as np
import pytest

from math_amd import configs, native
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def sphere_mesh(n, g, sigma=0.02, cap=0.0, seed=1):
    return native.meshgen_cube_sphere(n, g, True, sigma, cap, seed)


def rays_at(n, K, seed, radius=3.0, spread=0.5):
    rng = np.random.default_rng(seed)
    org = rng.normal(size=(K, n))
    org = org / np.linalg.norm(org, axis=1, keepdims=True) * radius
    d = rng.normal(size=(K, n)) * spread - org
    return org.astype(np.float32), d.astype(np.float32)


def parse_dump(dump, n):
    out, k = {}, 0
    while k < len(dump):
        key = (int(dump[k]),) + tuple(int(x) for x in dump[k + 1:k + 1 + n])
        c = int(dump[k + 1 + n])
        out[key] = dump[k + 2 + n:k + 2 + n + c].tolist()
        k += 2 + n + c
    return out


@pytest.mark.parametrize("n,g", [(3, 8), (4, 3), (5, 2), (6, 1)])
def test_basis_records_bit_equal(ctx, n, g):
    v, f = sphere_mesh(n, g)
    print(f"basis n={n} g={g} F={len(f)}", flush=True)
    gm = ctx.mesh(v, f)
    om = O.Mesh(v, f)
    gn, gpn, gpd = gm.records()
    on, opn, opd, _ = om.records()
    assert np.array_equal(gn, on)
    assert np.array_equal(gpn, opn)
    assert np.array_equal(gpd, opd)


@pytest.mark.parametrize("n,g", [(3, 8), (4, 3), (5, 2)])
def test_tree_equals_checker(ctx, n, g):
    v, f = sphere_mesh(n, g)
    print(f"tree n={n} g={g}", flush=True)
    gm = ctx.mesh(v, f)
    om = O.Mesh(v, f)
    om.build_tree()
    a = parse_dump(gm.tree_dump(), n)
    b = parse_dump(om.tree_dump(), n)
    assert a.keys() == b.keys()
    for k in a:
        assert a[k] == b[k], k
    info = gm.info()
    oi = om.tree_info()
    assert info.tree_nodes == oi["nodes"] and info.tree_entries == oi["entries"]


@pytest.mark.parametrize("n,g", [(3, 10), (4, 4), (5, 2), (6, 1)])
def test_first_hit_parity_small(ctx, n, g):
    v, f = sphere_mesh(n, g)
    print(f"first-hit n={n} g={g}", flush=True)
    gm = ctx.mesh(v, f)
    om = O.Mesh(v, f)
    org, d = rays_at(n, 4000, 100 + n)
    gi, gt, gnear = gm.first_hit(org, d)
    oi, ot, onear = om.trace_brute(org, d, eps_bary=1e-9, rel_t=1e-9)
    ok = (gnear == 0) & (onear == 0)
    assert ok.mean() > 0.95
    assert np.array_equal(gi[ok], oi[ok])
    # with the FP64 re-test the flagged rays agree as well, except true near-ties
    assert np.mean(gi == oi) > 0.999
    h = (gi == oi) & (oi >= 0)
    assert h.sum() > 400
    assert np.max(np.abs(gt[h] - ot[h]) / ot[h]) <= 1e-5


def test_first_hit_parity_c1_full_size(ctx):
    v, f, cfg, _, _ = configs.config("C1")
    print(f"first-hit C1 F={len(f)}", flush=True)
    gm = ctx.mesh(v, f)
    om = O.Mesh(v, f)
    org, d = rays_at(4, 300, 7, radius=3.0, spread=0.6)
    gi, gt, gnear = gm.first_hit(org, d)
    oi, ot, onear = om.trace_brute(org, d, eps_bary=1e-9, rel_t=1e-9)
    ok = (gnear == 0) & (onear == 0)
    assert np.array_equal(gi[ok], oi[ok])
    h = (gi == oi) & (oi >= 0)
    assert h.sum() > 100
    assert np.max(np.abs(gt[h] - ot[h]) / ot[h]) <= 1e-5


def test_closed_mesh_invariant_c1(ctx):
    # test_mesh.cpp:106-223 at full size: enter, exit, then nothing; <= 0.05 % failures
    v, f, _, _, _ = configs.config("C1", sigma=0.0)
    gm = ctx.mesh(v, f)
    rng = np.random.default_rng(11)
    K = 200000
    org = rng.normal(size=(K, 4))
    org = org / np.linalg.norm(org, axis=1, keepdims=True) * 4
    tgt = rng.normal(size=(K, 4))
    tgt *= (0.6 * rng.random(K) / np.linalg.norm(tgt, axis=1))[:, None]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    i1, t1, _ = gm.first_hit(org.astype(np.float32), d.astype(np.float32))
    o2 = org + d * (t1[:, None] + 1e-4)
    i2, t2, _ = gm.first_hit(o2.astype(np.float32), d.astype(np.float32))
    o3 = o2 + d * (t2[:, None] + 1e-4)
    i3, _, _ = gm.first_hit(o3.astype(np.float32), d.astype(np.float32))
    hit = i1 >= 0
    errors = np.sum(~hit) + np.sum(hit & (i2 < 0)) + np.sum(hit & (i2 >= 0) & (i3 >= 0))
    print(f"closed-mesh errors {errors} / {K}", flush=True)
    assert errors <= 0.0005 * K


def small_scene(n=4, g=4, edge=10, spp=4, depth=3):
    v, f = sphere_mesh(n, g)
    light = tuple([2.0] * n)
    cfg = configs.scene_dict(v, f, edge, light)
    return v, f, cfg, spp, depth


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_render_fp64_matches_checker(ctx, n):
    # paint<N,T> is instantiated for N = 3..6 (painter.cpp:627-643)
    g = {3: 8, 4: 4, 5: 2, 6: 1}[n]
    edge = {3: 24, 4: 10, 5: 5, 6: 4}[n]
    v, f, cfg, spp, depth = small_scene(n, g, edge)
    print(f"render fp64 n={n}", flush=True)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(spp, depth, precision=native.PRECISION_FP64, seed=5)
    img, cnt = sc.image()
    st = sc.statistics()
    om = O.Mesh(v, f)
    om.build_tree()
    ref, rays = O.render(om, O.make_scene(cfg, spp, depth, 5), threads=8)
    mean_g = img / cnt[:, None]
    mean_o = ref / cnt[:, None]
    diff = np.abs(mean_g - mean_o)
    rmse = float(np.sqrt(np.mean(diff ** 2)))
    print(f"fp64 rmse {rmse:.3e} max {diff.max():.3e} rays gpu {st['rays']} cpu {rays}", flush=True)
    assert rmse < 1e-4
    assert abs(st["rays"] - rays) <= 0.001 * rays


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_render_fp32_matches_checker(ctx, n):
    g = {3: 8, 4: 4, 5: 2, 6: 1}[n]
    edge = {3: 32, 4: 12, 5: 6, 6: 4}[n]
    v, f, cfg, spp, depth = small_scene(n, g, edge, spp=16)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(spp, depth, precision=native.PRECISION_FP32, seed=9)
    img, cnt = sc.image()
    om = O.Mesh(v, f)
    om.build_tree()
    ref, rays = O.render(om, O.make_scene(cfg, spp, depth, 9), threads=8)
    rmse = float(np.sqrt(np.mean((img / cnt[:, None] - ref / cnt[:, None]) ** 2)))
    print(f"fp32 n={n} rmse vs FP64 checker {rmse:.3e}", flush=True)
    assert rmse < 5e-3


@pytest.mark.parametrize("n,spp,edge", [(3, 36, 12), (3, 64, 12), (4, 125, 4)])
def test_render_many_samples_per_pixel(ctx, n, spp, edge):
    # per-pixel sums over 36 samples (LDS tile of fewer than 256 pixels), 64 and 125
    # samples (one wave per pixel)
    g = {3: 8, 4: 4}[n]
    v, f, cfg, _, depth = small_scene(n, g, edge, spp=spp, depth=2)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(spp, depth, precision=native.PRECISION_FP64, seed=12)
    img, cnt = sc.image()
    om = O.Mesh(v, f)
    om.build_tree()
    ref, rays = O.render(om, O.make_scene(cfg, spp, depth, 12), threads=8)
    assert np.all(cnt == spp)
    rmse = float(np.sqrt(np.mean((img / cnt[:, None] - ref / cnt[:, None]) ** 2)))
    print(f"n={n} spp={spp} rmse {rmse:.3e}", flush=True)
    assert rmse < 1e-4


@pytest.mark.parametrize("n,precision", [(3, native.PRECISION_FP64), (4, native.PRECISION_FP64),
                                         (4, native.PRECISION_FP32)])
def test_render_smooth_normals_matches_checker(ctx, n, precision):
    # vertex normals (radial, so they disagree with the noisy facets' normals) with smooth_normal:
    # MeshHyperplaneSimplex normal classification, shading_normal and the back-face probes of
    # direct_diffuse_lighting (painter.cpp:238-284), whose rays are counted as the reference counts them
    g = {3: 8, 4: 4}[n]
    edge = {3: 24, 4: 10}[n]
    v, f = native.meshgen_cube_sphere(n, g, True, 0.05, 0.0, 3)
    vn = v / np.linalg.norm(v, axis=1, keepdims=True)
    cfg = configs.scene_dict(v, f, edge, tuple([2.0] * n))
    gm = ctx.mesh(v, f, normals=vn, facet_normals=f)
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(4, 3, precision=precision, seed=8, smooth=True)
    img, cnt = sc.image()
    st = sc.statistics()
    om = O.Mesh(v, f, normals=vn, facet_normals=f)
    om.build_tree()
    ref, rays = O.render(om, O.make_scene(cfg, 4, 3, 8, smooth=True), threads=8)
    rmse = float(np.sqrt(np.mean((img / cnt[:, None] - ref / cnt[:, None]) ** 2)))
    print(f"smooth n={n} precision={precision} rmse {rmse:.3e} rays gpu {st['rays']} cpu {rays}", flush=True)
    assert rmse < (1e-4 if precision == native.PRECISION_FP64 else 5e-3)
    assert abs(st["rays"] - rays) <= 0.002 * rays


@pytest.mark.parametrize("n,kind,precision", [(3, native.PROJECTOR_PERSPECTIVE, native.PRECISION_FP64),
                                              (4, native.PROJECTOR_PERSPECTIVE, native.PRECISION_FP64),
                                              (4, native.PROJECTOR_SPHERICAL, native.PRECISION_FP64),
                                              (4, native.PROJECTOR_PERSPECTIVE, native.PRECISION_FP32),
                                              (3, native.PROJECTOR_SPHERICAL, native.PRECISION_FP32)])
def test_render_other_projectors_match_checker(ctx, n, kind, precision):
    # PerspectiveProjector and SphericalProjector (projector.h:102-151, 204-275): primary rays from
    # the camera point; FP32 renders walk the tree for them (rasterisation needs one direction)
    g = {3: 8, 4: 4}[n]
    edge = {3: 24, 4: 10}[n]
    v, f = sphere_mesh(n, g)
    cfg = configs.projector_dict(v, f, edge, tuple([2.0] * n), kind, 50.0)
    print(f"projector {kind} n={n} precision={precision}", flush=True)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(4, 3, precision=precision, seed=12)
    img, cnt = sc.image()
    st = sc.statistics()
    om = O.Mesh(v, f)
    om.build_tree()
    ref, rays = O.render(om, O.make_scene(cfg, 4, 3, 12), threads=8)
    rmse = float(np.sqrt(np.mean((img / cnt[:, None] - ref / cnt[:, None]) ** 2)))
    lit = float(np.mean(np.abs(ref / cnt[:, None] - np.array(configs.BACKGROUND_RGB)) > 1e-3))
    print(f"rmse {rmse:.3e} rays gpu {st['rays']} cpu {rays}, object pixels {lit:.2f}", flush=True)
    assert 0.05 < lit < 0.95  # the object fills part of the view
    assert rmse < (1e-4 if precision == native.PRECISION_FP64 else 5e-3)
    if precision == native.PRECISION_FP64:
        assert abs(st["rays"] - rays) <= 0.001 * rays
    else:
        walk = native.Scene(ctx, gm, cfg)
        walk.paint(4, 3, precision=precision, seed=12, primary=native.PRIMARY_WALK)
        assert np.array_equal(walk.image()[0], img)


def test_record_layouts_identical(ctx):
    # both FP32 record copies hold the same values: hits, distances and images are bit-identical
    v, f = sphere_mesh(4, 4)
    gm = ctx.mesh(v, f)
    org, d = rays_at(4, 4000, 21)
    out = {}
    for layout in (native.LAYOUT_LEAF_SOA, native.LAYOUT_SIMPLEX_AOS):
        print(f"record layout {layout}", flush=True)
        gm.record_layout = layout
        assert gm.record_layout == layout
        hits = gm.first_hit(org, d)
        cfg = configs.scene_dict(v, f, 10, (2.0,) * 4)
        sc = native.Scene(ctx, gm, cfg)
        sc.paint(8, 3, precision=native.PRECISION_FP32, seed=4)
        out[layout] = hits + sc.image()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    with pytest.raises(native.NspError):
        gm.record_layout = 7


def test_direct_light_aov(ctx):
    # without bounces the whole radiance of a path is direct: the AOV equals the image bit for bit;
    # with bounces the direct part is what a depth-0 render of the same samples collects
    v, f, cfg, _, _ = small_scene(4, 4, 10)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(8, 0, seed=6, aov=True)
    img, cnt = sc.image()
    alb, nrm, dl = sc.aov()
    assert np.array_equal(img, dl)
    hit = alb[:, 0] > 0.5 * cnt  # white object, diffuse 0.8, against the darker background
    assert hit.any() and np.all(np.linalg.norm(nrm[hit], axis=1) <= cnt[hit] * (1 + 1e-5))
    sc2 = native.Scene(ctx, gm, cfg)
    sc2.paint(8, 4, seed=6, aov=True)
    img2, _ = sc2.image()
    _, _, dl2 = sc2.aov()
    assert np.array_equal(dl2, dl)
    assert np.all(img2 >= dl2 - 1e-5 * np.abs(dl2))
    assert img2.sum() > dl2.sum()


def test_render_deterministic_and_batch_invariant(ctx):
    v, f, cfg, spp, depth = small_scene(4, 4, 12, spp=8)
    gm = ctx.mesh(v, f)
    imgs = []
    for batch in (0, 1000, 27 * 50):
        sc = native.Scene(ctx, gm, cfg)
        sc.paint(spp, depth, seed=3, batch_paths=batch)
        imgs.append(sc.image())
    for s, c in imgs[1:]:
        assert np.array_equal(s, imgs[0][0]) and np.array_equal(c, imgs[0][1])


def test_progressive_passes_and_checkpoint(ctx):
    v, f, cfg, spp, depth = small_scene(4, 4, 8, spp=4)
    gm = ctx.mesh(v, f)
    a = native.Scene(ctx, gm, cfg)
    a.paint(spp, depth, seed=4, passes=2)
    s2, c2 = a.image()
    b = native.Scene(ctx, gm, cfg)
    b.paint(spp, depth, seed=4, passes=1)
    s1, c1 = b.image()
    c = native.Scene(ctx, gm, cfg)
    c.set_image(s1, c1)
    # resume: the pass index continues from the restored scene's own counter
    c.paint(spp, depth, seed=4, passes=1)
    assert np.array_equal(c2, 2 * c1)
    st = a.statistics()
    assert st["passes"] == 2 and st["pixels"] == 2 * a.pixels


def test_pixels_after_callback(ctx):
    v, f, cfg, spp, depth = small_scene(4, 3, 6, spp=1)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    got = {}

    def cb(first, count, rgb):
        got[first] = rgb

    sc.paint(spp, depth, seed=1, on_pixels=cb)
    means = np.concatenate([got[k] for k in sorted(got)])
    s, c = sc.image()
    assert np.allclose(means, s / c[:, None])


def test_error_paths(ctx):
    v, f = sphere_mesh(4, 2)
    bad = f.copy()
    bad[0] = [0, 0, 1, 2]  # degenerate simplex: normal not finite
    with pytest.raises(native.NspError) as e:
        ctx.mesh(v, bad)
    assert e.value.code == native.NSP_E_INVALID
    gm = ctx.mesh(v, f)
    native.lib().nsp_debug_fail_next_launch(ctx.h)
    with pytest.raises(native.NspError) as e:
        gm.first_hit(np.zeros((4, 4), np.float32), np.ones((4, 4), np.float32))
    assert e.value.code == native.NSP_E_HIP
    # the context stays usable after the injected failure
    idx, _, _ = gm.first_hit(np.full((1, 4), 3.0, np.float32), np.full((1, 4), -1.0, np.float32))
    assert idx[0] >= 0


def test_cpp_painter_api_drop_in():
    """The reference's headless painter test written against the C++17 drop-in
    (paint<4,double>, single_object_scene, VisibleBarPaintbrush) on the GPU."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_painter_api")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe, "4", "12", "4"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout
    _, pixels, rays, lum, passes = r.stdout.split()
    assert int(pixels) > 0 and int(rays) > int(pixels) and float(lum) > 0 and int(passes) == 1


def test_graph_replay_equals_eager(ctx):
    """Each batch's launch sequence captured once in a hipGraph and replayed per
    batch (several replays per pass, eager work and syncs between passes) gives
    the eager image bit for bit, pass after pass."""
    v, f, cfg, spp, depth = small_scene(4, 4, 12, spp=8)
    gm = ctx.mesh(v, f)
    eager = native.Scene(ctx, gm, cfg)
    graph = native.Scene(ctx, gm, cfg)
    for p in range(3):
        eager.paint(spp, depth, seed=6, batch_paths=27 * 200)
        graph.paint(spp, depth, seed=6, batch_paths=27 * 200, use_graph=True)
        gm.first_hit(np.full((8, 4), 3.0, np.float32), np.full((8, 4), -1.0, np.float32))  # eager work between
        a, ca = eager.image()
        b, cb = graph.image()
        assert np.array_equal(a, b) and np.array_equal(ca, cb), p
    assert eager.statistics()["rays"] == graph.statistics()["rays"]


def test_passes_in_one_call_equal_separate_calls(ctx):
    """k passes in one call (eager, and with one graph replayed across every
    batch of every pass) equal k calls of one pass, bit for bit: the kernels
    read the pass index from device memory, written by the pass reset."""
    v, f, cfg, spp, depth = small_scene(4, 4, 10, spp=8)
    gm = ctx.mesh(v, f)
    one = native.Scene(ctx, gm, cfg)
    for _ in range(3):
        one.paint(spp, depth, seed=8, batch_paths=27 * 300)
    ref = one.image()
    for use_graph in (False, True):
        print(f"three passes in one call, graph={use_graph}", flush=True)
        sc = native.Scene(ctx, gm, cfg)
        sc.paint(spp, depth, seed=8, batch_paths=27 * 300, passes=3, use_graph=use_graph)
        got = sc.image()
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        assert sc.statistics()["rays"] == one.statistics()["rays"]


def test_unbounded_passes_until_stop(ctx):
    # passes=-1 (BarPaintbrush max_pass_count -1) runs until the stop flag, here
    # set by the pass callback after the third pass
    import ctypes
    v, f, cfg, spp, depth = small_scene(4, 3, 8, spp=4)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    stop = ctypes.c_int32(0)
    seen = []

    def on_pass(k, secs):
        seen.append(k)
        if k == 3:
            stop.value = 1
    with pytest.raises(native.NspError) as e:
        sc.paint(spp, depth, seed=2, passes=-1, stop=stop, on_pass=on_pass, use_graph=True)
    assert e.value.code == native.NSP_E_STOPPED
    assert seen == [1, 2, 3] and sc.statistics()["passes"] == 3
    ref = native.Scene(ctx, gm, cfg)
    ref.paint(spp, depth, seed=2, passes=3)
    assert np.array_equal(sc.image()[0], ref.image()[0])
    # an unbounded paint without a stop flag is refused before any launch
    params = native.PaintParams(spp, depth, 0, native.PRECISION_FP32, -1, 2, 0, 0, 0, 0)
    assert native.lib().nsp_paint(sc.h, ctypes.byref(params), None, None) == native.NSP_E_INVALID


def test_pixels_delivered_per_batch(ctx):
    # each wavefront batch's pixels arrive once per pass, in order, with the
    # running mean; after the last pass they equal the image
    v, f, cfg, spp, depth = small_scene(4, 3, 9, spp=8)
    gm = ctx.mesh(v, f)
    sc = native.Scene(ctx, gm, cfg)
    P = sc.pixels
    runs = []
    # 8 requested samples in n = 4 are 2^3 = 8 per pixel: batches of 100 pixels
    sc.paint(spp, depth, seed=5, batch_paths=8 * 100, passes=2,
             on_pixels=lambda first, count, rgb: runs.append((first, count, rgb)))
    starts = [r[0] for r in runs]
    assert len(runs) == 2 * ((P + 99) // 100)
    half = len(runs) // 2
    assert starts[:half] == starts[half:] == list(range(0, P, 100))
    s, c = sc.image()
    last = np.concatenate([r[2] for r in runs[half:]])
    assert np.allclose(last, s / c[:, None], rtol=1e-6, atol=0)


@pytest.mark.parametrize("n", [3, 4, 5])
def test_rasterised_primary_visibility_equals_tree_walk(ctx, n):
    # FP32 renders find the primary hits by rasterising the simplices into the
    # parallel projection; the tree walk gives the same nearest hits, so the
    # images are bit-identical
    g = {3: 10, 4: 4, 5: 2}[n]
    edge = {3: 40, 4: 14, 5: 6}[n]
    v, f = sphere_mesh(n, g, sigma=0.05)
    cfg = configs.scene_dict(v, f, edge, tuple([2.0] * n))
    gm = ctx.mesh(v, f)
    imgs = []
    for primary in (native.PRIMARY_AUTO, native.PRIMARY_WALK):
        print(f"primary mode {primary} n={n}", flush=True)
        sc = native.Scene(ctx, gm, cfg)
        sc.paint(8, 3, seed=12, primary=primary, batch_paths=3000)
        imgs.append(sc.image() + (sc.statistics()["rays"],))
    for a, b in zip(imgs[0], imgs[1]):
        assert np.array_equal(a, b)


def test_cpp_single_object_scene_conventions():
    """single_object_scene(background, colour, diffuse, min, max, mesh) of the C++
    drop-in against the formulas of src/painter/scenes/single_object.h:102-179:
    camera at the mesh centre moved to max[N-1] + |size| on the last axis,
    looking along -e_{N-1}, axes e_0..e_{N-2}, units per pixel = largest
    projected extent / max screen, screen edges lround(size_i / largest *
    max) clamped to [min, max], a constant light of colour 1 at max + 100 (max -
    centre); the mesh min/max over the facets' vertices (mesh.cpp:85-98)."""
    import math
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_conventions")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe, "scene", "3", "5", "23"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for ln in r.stdout.splitlines():
        k, *vals = ln.split()
        out.setdefault(k, []).append([float(x) for x in vals])
    v, f = native.meshgen_cube_sphere(4, 3, True, 0.03, 0.0, 7)
    used = v[np.unique(f)].astype(np.float64)
    mn, mx = used.min(axis=0), used.max(axis=0)
    assert np.array_equal(out["min"][0], mn) and np.array_equal(out["max"][0], mx)
    size = mx - mn
    centre = mn + size / 2
    maxp = size[:3].max()
    cam = centre.copy()
    cam[3] = mx[3] + math.sqrt(float(size @ size))
    assert out["projector"][0] == [native.PROJECTOR_PARALLEL]
    assert np.allclose(out["camera_org"][0], cam, rtol=0, atol=1e-12)
    assert out["camera_dir"][0] == [0.0, 0.0, 0.0, -1.0]
    assert np.array_equal(np.array(out["axis"]), np.eye(4)[:3])
    assert out["units_per_pixel"][0][0] == pytest.approx(maxp / 23, rel=1e-14)
    assert out["screen"][0] == [min(max(math.floor(size[i] / maxp * 23 + 0.5), 5), 23) for i in range(3)]
    assert len(out["light"]) == 1
    light = out["light"][0]
    assert light[0] == native.LIGHT_CONSTANT
    assert np.allclose(light[1:5], mx + 100 * (mx - centre), rtol=1e-14, atol=1e-12)
    assert light[5:] == [1.0, 1.0, 1.0]


def test_first_sample_aov(ctx):
    """nsp_image_first_sample: per pass, the radiance of each pixel's first sample
    (the denoiser's 1-spp input).  With one sample per pixel it is the image
    itself; with 8 it is the checker's path colour of sample 0, summed over the
    passes."""
    v, f, cfg, _, _ = small_scene(4, 4, 8)
    gm = ctx.mesh(v, f)
    one = native.Scene(ctx, gm, cfg)
    one.paint(1, 3, seed=6, aov=True)
    assert np.array_equal(one.first_sample(), one.image()[0])
    sc = native.Scene(ctx, gm, cfg)
    sc.paint(4, 3, precision=native.PRECISION_FP64, seed=6, aov=True, passes=2)
    fs = sc.first_sample()
    om = O.Mesh(v, f)
    om.build_tree()
    rng = np.random.default_rng(2)
    for pixel in rng.choice(sc.pixels, size=60, replace=False):
        expect = np.zeros(3, np.float32)
        for k in range(2):
            rgb, _ = O.path_sample(om, O.make_scene(cfg, 4, 3, 6, pass_index=k), int(pixel), 0)
            expect += rgb
        assert np.allclose(fs[pixel], expect, rtol=1e-5, atol=1e-6), (pixel, fs[pixel], expect)
