"""The reference's lenses (src/shader.comp:366-448) and Dupin cyclides
(src/shader.comp:633-679) in the extended mode, against the FP64 oracle
(oracle/ptext.c hit_lens, hit_cyclide): same-seed per-pixel agreement,
block-wise convergence with independent seeds, emitting lenses and cyclides
(found by BSDF sampling only), a glass lens, and every scene the reference
ships rendering in the extended mode."""
import json

import numpy as np
import pytest

import pathtracer_amd as pt
from pathtracer_amd import scenes
from tests.test_gpu_ext_paths import converges, gpu_render, same_seed

pytestmark = pytest.mark.gpu

GOLDEN_CYC = dict(a=3.36, b=-3.17, c=-1.06, d=-1.5)


def shapes_cornell(emissive=False, glass_lens=False):
    """The Cornell box with its glass sphere made diffuse, a converging and a
    diverging lens and a cyclide (the reference scenes' shape parameters,
    scaled into the box)."""
    d = scenes.cornell_dict(4, 3)
    d["material"][4] = {"type": "lambert", "gaussians": [[0.6, 600.0, 80.0]]}
    d["sphere"] = d["sphere"][:1]
    if glass_lens:
        d["material"].append({"type": "dielectric", "sellmeier": [1.03961212, 0.231792344, 1.01046945,
                                                                  6.00069867e-3, 2.00179144e-2, 103.560653]})
    lens_mat = len(d["material"]) if glass_lens else 4
    d["lens"] = [
        {"position": [0.7, 0.45, 0.45], "rotation": [0.0, 30.0, 10.0], "radius": 0.15, "focalLength": 0.12,
         "thickness": 0.02, "isConverging": True, "materialID": lens_mat, "lightID": 0},
        {"position": [0.45, 0.7, 0.7], "rotation": [0.0, 60.0, 0.0], "radius": 0.1, "focalLength": 0.1,
         "thickness": 0.05, "isConverging": False, "materialID": 1, "lightID": 0},
    ]
    d["cyclide"] = [{"position": [0.3, 0.35, 0.45], "rotation": [20.0, 0.0, 0.0], "scale": [0.04, 0.04, 0.04],
                     **GOLDEN_CYC, "boundingRadius": 0.16, "materialID": 4, "lightID": 0}]
    if emissive:
        d["light"].append({"emission": {"temperature": 3500.0, "luminosity": 6.0}})
        d["light"].append({"emission": {"temperature": 5000.0, "luminosity": 4.0}})
        d["lens"][0]["lightID"] = len(d["light"]) - 1
        d["cyclide"][0]["lightID"] = len(d["light"])
    return json.dumps(d)


@pytest.mark.parametrize("emissive", [False, True], ids=["diffuse", "emitting"])
def test_same_seed_shapes_match_fp64_oracle(gpu_ctx, emissive):
    text = shapes_cornell(emissive)
    f3, f2, _ = same_seed(gpu_ctx, text, 48, 48, 16, 8)
    print(f"\nlenses + cyclide ({'emitting' if emissive else 'diffuse'}): {f3 * 100:.2f}% of pixels within "
          f"1e-3, {f2 * 100:.2f}% within 1e-2")
    assert f3 >= 0.99 and f2 >= 0.995


@pytest.mark.parametrize("emissive", [False, True], ids=["diffuse", "emitting"])
def test_shapes_converge_to_fp64_oracle(gpu_ctx, emissive):
    share, zimg = converges(gpu_ctx, shapes_cornell(emissive), 32, 32, 2048, 512, 8)
    print(f"\nlenses + cyclide: {share * 100:.1f}% of blocks within 4 sigma (largest |z| {converges.zmax:.2f}), "
          f"image-mean z = {np.round(zimg, 2)}")
    assert np.all(np.abs(zimg) < 3.0)
    assert share >= 0.99


def test_glass_lens_converges_to_fp64_oracle(gpu_ctx):
    """A Sellmeier glass lens refracts through both caps (its normals must be
    the solid's outward ones); caustics: the image mean."""
    share, zimg = converges(gpu_ctx, shapes_cornell(glass_lens=True), 32, 32, 2048, 512, 8)
    print(f"\nglass lens: {share * 100:.1f}% of blocks within 4 sigma, image-mean z = {np.round(zimg, 2)}")
    assert np.all(np.abs(zimg) < 3.0)


def test_every_shipped_scene_renders_in_extended_mode(scene_texts):
    """The reference's 11 scenes (lenses, cyclides, an emitting lens, SDFs) in
    the extended mode: accepted, finite, lit, and no hit record outside the
    scene's tables."""
    with pt.Context(0) as c:
        for name, text in scene_texts.items():
            img = gpu_render(c, text, 64, 48, 4, 6)
            assert np.isfinite(img).all() and img.max() > 0.0, name
        assert c.counters()["bad_refs"] == 0
