"""Extended mode on the GPU: convergence to the FP64 oracle (oracle/ptext.c)
within the Monte Carlo bound, closed-form radiometry, and determinism under
re-batching.

Statistical bar: on blocks of pixels, |GPU mean - oracle mean| <= 4 sigma for
>= 99 % of blocks (sigma from the oracle's per-sample variance scaled to both
sample counts), and the image means agree within 3 sigma."""
import json

import numpy as np
import pytest

import pathtracer_amd as pt
from pathtracer_amd import scenes
from oracle import ptext

pytestmark = pytest.mark.gpu


def gpu_render(ctx, text, W, H, spp, depth, spf=None, frames=None):
    scene = ctx.scene_from_json(text)
    spf = spf or spp
    r = pt.Renderer(ctx, scene, W, H, spf=spf, path_length=depth, mode=pt.MODE_EXT)
    for _ in range(frames or spp // spf):
        r.render_frame()
    return r.acc.read()[..., :3].astype(np.float64)


def block(a, b):
    h, w = a.shape[0] // b * b, a.shape[1] // b * b
    return a[:h, :w].reshape(h // b, b, w // b, b, -1).mean(axis=(1, 3))


def converge_check(gpu, gspp, ora, blk=4, z_max=4.0, frac=0.99):
    mean_o, var_o = ora["mean"], ora["var_of_mean"]
    ospp_ratio = 1.0  # var_of_mean is already per the oracle's sample count
    # GPU variance of its mean, from the oracle's per-sample variance
    var_g = var_o * (ora["spp"] / gspp)
    gb, ob = block(gpu, blk), block(mean_o, blk)
    vb = block(var_o + var_g, blk) / (blk * blk) * ospp_ratio
    sig = np.sqrt(np.maximum(vb, 1e-30))
    z = (gb - ob) / sig
    ok = np.abs(z) <= z_max
    live = vb > 1e-20
    share = ok[live].mean() if live.any() else 1.0
    # whole-image mean
    n = gpu.size / 3
    sig_img = np.sqrt((var_o + var_g).sum(axis=(0, 1))) / n
    dimg = gpu.mean(axis=(0, 1)) - mean_o.mean(axis=(0, 1))
    return share, dimg / np.maximum(sig_img, 1e-30), z


def no_caustic_cornell(wavelengths):
    """The Cornell box with the glass sphere made diffuse: no path reaches the
    light only through specular bounces, so per-pixel sample variances are
    finite and well estimated."""
    s = scenes.cornell_dict(wavelengths, 0)
    s["material"][4] = {"type": "lambert", "gaussians": [[0.6, 600.0, 80.0]]}
    return json.dumps(s)


@pytest.mark.parametrize("wavelengths", [1, 4])
def test_cornell_converges_to_fp64_oracle(gpu_ctx, wavelengths):
    text = no_caustic_cornell(wavelengths)
    W = H = 48
    gspp, ospp = 4096, 1024
    gpu = gpu_render(gpu_ctx, text, W, H, gspp, 8, spf=512)
    ora = ptext.render(text, W, H, ospp, 8, seed=12345)
    ora["spp"] = ospp
    share, zimg, z = converge_check(gpu, gspp, ora)
    print(f"\ncornell {wavelengths} wavelength(s): {share * 100:.2f}% of 4x4 blocks within 4 sigma, "
          f"image-mean z = {np.round(zimg, 2)}")
    assert share >= 0.99
    assert np.all(np.abs(zimg) < 3.0)


def same_seed_agreement(ctx, text, W, H, spp, depth):
    """GPU (FP32) and oracle (FP64) with the same seed draw the same numbers in
    the same frames, so they trace the same paths; returns the share of pixels
    within 1e-3 and 1e-2 relative, and the number of diverged pixels."""
    g = gpu_render(ctx, text, W, H, spp, depth)
    o = ptext.render(text, W, H, spp, depth)["mean"]
    den = np.abs(o).sum(-1)
    rel = np.where(den > 0, np.abs(g - o).sum(-1) / np.maximum(den, 1e-30), np.abs(g).sum(-1))
    return float((rel <= 1e-3).mean()), float((rel <= 1e-2).mean()), int((rel > 1e-3).sum())


@pytest.mark.parametrize("name,min3,min2", [("cornell", 0.999, 0.999), ("spheres", 0.95, 0.97), ("sdf", 0.90, 0.95)])
def test_same_seed_paths_match_fp64_oracle(gpu_ctx, name, min3, min2):
    text, depth, (W, H) = {
        "cornell": (scenes.cornell(4), 8, (48, 48)),
        "spheres": (scenes.random_spheres(4096, 1, 4), 12, (64, 36)),
        "sdf": (scenes.sdf_scene(), 8, (64, 36)),
    }[name]
    f3, f2, diverged = same_seed_agreement(gpu_ctx, text, W, H, 16, depth)
    print(f"\n{name}: {f3 * 100:.2f}% of pixels within 1e-3, {f2 * 100:.2f}% within 1e-2, "
          f"{diverged} pixels with diverged paths")
    assert f3 >= min3 and f2 >= min2


@pytest.mark.parametrize("name", ["cornell_glass", "spheres", "sdf"])
def test_image_mean_converges_to_fp64_oracle(gpu_ctx, name):
    """Scenes with caustics (light reached only through specular bounces) have
    heavy-tailed pixel estimates; their image means must still agree within 3 sigma."""
    text, depth, (W, H), gspp, ospp = {
        "cornell_glass": (scenes.cornell(4), 8, (48, 48), 4096, 1024),
        "spheres": (scenes.random_spheres(4096, 1, 4), 12, (48, 32), 4096, 1024),
        "sdf": (scenes.sdf_scene(), 8, (48, 27), 1024, 256),
    }[name]
    gpu = gpu_render(gpu_ctx, text, W, H, gspp, depth, spf=min(gspp, 512))
    ora = ptext.render(text, W, H, ospp, depth, seed=4242)
    ora["spp"] = ospp
    _, zimg, _ = converge_check(gpu, gspp, ora)
    print(f"\n{name}: image-mean z = {np.round(zimg, 2)}")
    assert np.all(np.abs(zimg) < 3.0)


def _furnace(material, env=1.0, wavelengths=4):
    return json.dumps({
        "camera": {"numShots": 1, "position": [[0.0, 0.0, -3.0]], "angle": [[0.0, 0.0]], "ISO": 1, "size": 0.057,
                   "apertureSize": 1.0, "apertureDistance": 0.1, "lensRadius": 0.01, "lensFocalLength": 0.03,
                   "lensThickness": 0.0, "lensDistance": 0.05, "model": "pinhole"},
        "sphere": [{"position": [0.0, 0.0, 0.0], "radius": 1.0, "materialID": 1, "lightID": 0}],
        "material": [material],
        "light": [],
        "environment": {"radiance": env},
        "render": {"wavelengths": wavelengths, "russianRouletteDepth": 64, "seed": 3},
    })


def flat_xyz(ctx, env=1.0):
    """XYZ of the bare uniform environment (no object in view)."""
    t = json.loads(_furnace({"type": "lambert", "gaussians": [[0.5, 550.0, 1e7]]}, env))
    t["sphere"] = []
    return gpu_render(ctx, json.dumps(t), 16, 16, 256, 4).mean(axis=(0, 1))


def test_white_furnace_lambert(gpu_ctx):
    """Convex Lambert sphere of albedo 0.5 under a uniform sky: every pixel on it
    sees 0.5 x the sky (one bounce, the sphere cannot see itself)."""
    sky = flat_xyz(gpu_ctx)
    img = gpu_render(gpu_ctx, _furnace({"type": "lambert", "gaussians": [[0.5, 550.0, 1e7]]}), 32, 32, 256, 8)
    centre = img[12:20, 12:20].mean(axis=(0, 1))
    np.testing.assert_allclose(centre / sky, 0.5, rtol=0.02)


def test_lossless_dielectric_keeps_energy(gpu_ctx):
    """A smooth non-dispersive dielectric absorbs nothing: under a uniform sky
    the sphere disappears (R + T = 1 along every path)."""
    sky = flat_xyz(gpu_ctx)
    img = gpu_render(gpu_ctx, _furnace({"type": "dielectric", "cauchy": [1.5, 0.0]}), 32, 32, 256, 64)
    np.testing.assert_allclose(img.mean(axis=(0, 1)) / sky, 1.0, rtol=0.02)


def test_conductor_normal_incidence(gpu_ctx):
    """Mirror sphere: the centre pixel sees the sky scaled by the normal-incidence
    reflectance ((n-1)^2 + k^2) / ((n+1)^2 + k^2)."""
    n, k = 1.2, 4.0
    f0 = ((n - 1) ** 2 + k ** 2) / ((n + 1) ** 2 + k ** 2)
    # A sky-only render of the same pixels draws the same wavelengths (the RNG is
    # keyed on pixel and sample), so the per-pixel ratio is the reflectance.
    t = json.loads(_furnace({"type": "conductor", "n": n, "k": k}))
    t["sphere"] = []
    sky = gpu_render(gpu_ctx, json.dumps(t), 33, 33, 64, 4)
    img = gpu_render(gpu_ctx, _furnace({"type": "conductor", "n": n, "k": k}), 33, 33, 64, 4)
    # the centre 3 x 3 pixels see the sphere within ~6 degrees of normal incidence
    np.testing.assert_allclose(img[15:18, 15:18] / sky[15:18, 15:18], f0, rtol=3e-3)


def test_rebatching_is_invisible(scene_texts):
    """Counter-based RNG: the same pixels and samples give the same bits whatever
    the wave capacity (paths per batch) and frame split."""
    text = scenes.random_spheres(n=512, scene_seed=2, wavelengths=4, seed=5)
    with pt.Context(0, wave_paths=1500) as small, pt.Context(0) as big:
        a = gpu_render(small, text, 40, 30, 8, 12, spf=8)
        b = gpu_render(big, text, 40, 30, 8, 12, spf=8)
    assert np.array_equal(a, b)


def test_bvh_shapes_agree(scene_texts):
    """BVH width and leaf size change the traversal order only: identical closest
    hits, identical images."""
    text = scenes.random_spheres(n=1024, scene_seed=4, wavelengths=4, seed=1)
    imgs = []
    for w, leaf in ((2, 4), (2, 1), (4, 4), (8, 8), (2, 16)):
        with pt.Context(0, bvh_width=w, bvh_leaf=leaf) as c:
            imgs.append(gpu_render(c, text, 32, 24, 4, 12))
    for im in imgs[1:]:
        assert np.array_equal(imgs[0], im)


def test_counters_and_no_nans(gpu_ctx):
    text = scenes.random_spheres(n=4096, scene_seed=1, wavelengths=4, seed=0)
    gpu_ctx.reset_counters()
    img = gpu_render(gpu_ctx, text, 64, 36, 4, 12)
    c = gpu_ctx.counters()
    assert not np.isnan(img).any() and c["nan_samples"] == 0
    assert c["paths"] == 64 * 36 * 4
    # no hit record names a primitive outside the scene (such a record would be
    # classified as a miss and hide in the class sum below)
    assert c["bad_refs"] == 0
    # every segment ends in exactly one class queue
    assert c["segments"] == c["escaped"] + c["emitter_hits"] + c["lambert_hits"] + c["dielectric_hits"] + c["conductor_hits"]
    assert c["dielectric_hits"] > 0 and c["conductor_hits"] > 0 and c["shadow_segments"] > 0
    # every ray into the sphere BVH visits its root at least
    assert c["bvh_nodes"] >= c["segments"] and c["shadow_bvh_nodes"] > 0


def lit_sdf_scene(kind: str) -> str:
    """The mandelbulb + menger scene with one of the two removed and a 6500 K
    quad light above, so that shadow rays march the SDF too."""
    d = scenes.sdf_scene_dict(seed=3)
    d["sdf"] = [s for s in d["sdf"] if s["kind"] == kind]
    d["quad"] = [{"position": [-0.5, 4.0, -0.5], "edge1": [1.0, 0.0, 0.0], "edge2": [0.0, 0.0, 1.0],
                  "materialID": 1, "lightID": 1}]
    d["light"] = [{"emission": {"temperature": 6500.0, "luminosity": 12.0}}]
    return json.dumps(d)


@pytest.mark.parametrize("kind", ["menger", "mandelbulb"])
def test_march_flop_counters(gpu_ctx, kind):
    """The sphere-tracing FP32 operation counters against the per-step cost:
    every step of a menger of 4 levels costs 10 + 16 + 33 * 4; a mandelbulb
    step costs 10 + 10 + 65 per iteration executed, between 1 and 8."""
    gpu_ctx.reset_counters()
    gpu_render(gpu_ctx, lit_sdf_scene(kind), 64, 48, 4, 6)
    c = gpu_ctx.counters()
    steps = c["sdf_steps"]
    assert steps > 0 and c["shadow_segments"] > 0 and c["shadow_march_flops"] > 0
    own = "march_flops_" + kind
    for k in ("mandelbulb", "menger", "blob", "terrain"):
        if "march_flops_" + k != own:
            assert c["march_flops_" + k] == 0, k
    total = c[own] + c["shadow_march_flops"]
    if kind == "menger":
        assert total == steps * (10 + 16 + 33 * 4)
    else:
        iters, rem = divmod(total - steps * 20, 65)
        assert rem == 0 and steps <= iters <= 8 * steps
        assert iters > 2 * steps  # most evaluations near the surface run past the first iterations


def four_kind_sdf_scene() -> str:
    """Every SDF kind, two mengers (so a ray marches several SDFs of one kind),
    and a quad light so that shadow rays march them too."""
    d = json.loads(lit_sdf_scene("mandelbulb"))
    d["sdf"] += [
        {"kind": "menger", "iterations": 4, "position": [1.4, 1.0, 0.0], "boundingSize": [2.02, 2.02, 2.02],
         "material": 2.0},
        {"kind": "menger", "iterations": 3, "position": [0.2, 0.5, 2.5], "boundingSize": [2.02, 2.02, 2.02],
         "material": 1.0},
        {"kind": "blob", "position": [-0.2, 0.9, -1.4], "boundingSize": [2.2, 3.2, 2.2], "material": 2.0},
        {"kind": "terrain", "position": [0.0, 1.5, 7.0], "boundingSize": [10.0, 3.0, 2.0], "material": 1.0},
    ]
    return json.dumps(d)


@pytest.mark.parametrize("name", ["sdf", "four_kinds"])
def test_persistent_march_is_invisible(name):
    """Sphere tracing with lanes refilled from the queue (the default) and one
    ray per lane give identical images, step counts, exhausted marches and
    operation counts."""
    text, depth = {"sdf": (scenes.sdf_scene(seed=5), 8), "four_kinds": (four_kind_sdf_scene(), 6)}[name]
    imgs, cnts = [], []
    for march in (0, 1):
        with pt.Context(0, march=march, sdf_max_steps=96) as c:
            imgs.append(gpu_render(c, text, 64, 48, 4, depth))
            cnts.append(c.counters())
    assert np.array_equal(imgs[0], imgs[1])
    assert cnts[0]["sdf_steps"] > 0 and cnts[0]["sdf_exhausted"] > 0
    for k in ("segments", "shadow_segments", "sdf_steps", "sdf_exhausted", "lambert_hits", "march_flops_mandelbulb",
              "march_flops_menger", "march_flops_blob", "march_flops_terrain", "shadow_march_flops"):
        assert cnts[0][k] == cnts[1][k], k
    if name == "four_kinds":
        assert all(cnts[0]["march_flops_" + k] > 0 for k in ("mandelbulb", "menger", "blob", "terrain"))


@pytest.mark.parametrize("name", ["spheres", "sdf", "cornell"])
def test_persistent_traversal_is_invisible(scene_texts, name):
    """Persistent lanes refilled from the queue visit the same nodes per ray as
    one ray per lane: identical hits, identical images and counters."""
    text, depth = {
        "spheres": (scenes.random_spheres(n=4096, scene_seed=1, wavelengths=4, seed=7), 12),
        "sdf": (scenes.sdf_scene(seed=7), 8),
        "cornell": (scenes.cornell(4, seed=7), 8),
    }[name]
    imgs, cnts = [], []
    for tr in (2, 1):  # one ray per lane, persistent
        with pt.Context(0, traversal=tr) as c:
            imgs.append(gpu_render(c, text, 64, 36, 4, depth))
            cnts.append(c.counters())
    assert np.array_equal(imgs[0], imgs[1])
    assert cnts[0]["bad_refs"] == 0 and cnts[1]["bad_refs"] == 0
    for k in ("segments", "shadow_segments", "sdf_steps", "lambert_hits", "dielectric_hits", "conductor_hits",
              "bvh_nodes", "shadow_bvh_nodes", "march_flops_mandelbulb", "march_flops_menger", "shadow_march_flops"):
        assert cnts[0][k] == cnts[1][k], k
