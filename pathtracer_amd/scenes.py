"""Seeded generators of the benchmark scenes, in the reference's scene-JSON
format extended with the keys the extended mode reads:

  "quad":        [{"position": corner, "edge1": e1, "edge2": e2, "materialID", "lightID"}]
  "material":    {"type": "lambert", "gaussians": [[w, mean_nm, sigma_nm], ...]}
                 {"type": "dielectric", "cauchy": [A, B_um2]} | {"sellmeier": [B1, B2, B3, C1, C2, C3]}
                 {"type": "conductor", "n": n, "k": k}
                 (the reference's {"reflection": {...}} lobe stays valid)
  "sdf":         {"kind": "mandelbulb"|"menger"|..., "iterations", "position", "boundingSize", "material"}
  "environment": {"radiance": L}           uniform spectral radiance of the sky
  "render":      {"wavelengths": 1|4, "lambdaMin", "lambdaMax", "russianRouletteDepth", "seed"}
  camera "model": "pinhole" | "lens"

Workloads:
  cornell()        Cornell box of 5 diffuse quads with smooth 3-Gaussian
                   spectra, one Lambert and one dielectric sphere (Cauchy
                   A = 1.5, B = 0.01 um^2) and a 6500 K blackbody quad light.
  random_spheres() cornell() plus N spheres with centres uniform in the unit
                   cube, radius log-uniform in [0.002, 0.05], 60 % Lambert /
                   25 % dielectric (A in [1.4, 1.8], B in [0.003, 0.02]) /
                   15 % conductor (n in [0.2, 2.5], k in [2, 8]).
  sdf_scene()      power-8 mandelbulb and menger sponge over a ground plane
                   under a uniform white environment.
  serving_request(i)  request i of the serving workload, seeded i.
"""
from __future__ import annotations

import json
import math

import numpy as np

# Smooth reflectance spectra as three Gaussians (weight, mean nm, sigma nm).
WHITE = [[0.25, 450.0, 60.0], [0.25, 560.0, 60.0], [0.25, 670.0, 60.0]]
RED = [[0.05, 450.0, 40.0], [0.08, 530.0, 40.0], [0.65, 660.0, 50.0]]
GREEN = [[0.08, 450.0, 30.0], [0.55, 540.0, 35.0], [0.06, 650.0, 40.0]]


def _camera(pos, angle, model="pinhole", size=0.057, ap=0.0025, apdist=0.049, iso=1600):
    return {
        "numShots": 1,
        "position": [list(map(float, pos))],
        "angle": [list(map(float, angle))],
        "ISO": iso,
        "size": size,
        "apertureSize": ap,
        "apertureDistance": apdist,
        "lensRadius": 0.01,
        "lensFocalLength": 0.03,
        "lensThickness": 0.0,
        "lensDistance": 0.05,
        "model": model,
    }


def cornell_dict(wavelengths: int = 1, seed: int = 0) -> dict:
    """The unit-cube Cornell box [0,1]^3, open toward the camera at z = 0.

    The camera sits at z < 0 looking along +z (yaw 0 faces +z in the
    reference's camera, src/shader.comp:1456-1465). That camera shows +x on the
    right of the image, so the red wall is placed at x = 1 to appear on the left.
    """
    q = lambda c, e1, e2, m, l=0: {"position": c, "edge1": e1, "edge2": e2, "materialID": m, "lightID": l}
    quads = [
        q([0, 0, 0], [0, 0, 1], [1, 0, 0], 1),          # floor (normal +y)
        q([0, 1, 0], [1, 0, 0], [0, 0, 1], 1),          # ceiling (normal -y)
        q([0, 0, 1], [0, 1, 0], [1, 0, 0], 1),          # back wall at z = 1 (normal -z)
        q([1, 0, 0], [0, 0, 1], [0, 1, 0], 2),          # red wall at x = 1 (normal -x)
        q([0, 0, 0], [0, 1, 0], [0, 0, 1], 3),          # green wall at x = 0 (normal +x)
        q([0.375, 0.999, 0.375], [0.25, 0, 0], [0, 0, 0.25], 6, 1),  # 6500 K light (normal -y)
    ]
    spheres = [
        {"position": [0.68, 0.18, 0.62], "radius": 0.18, "materialID": 4, "lightID": 0},
        {"position": [0.3, 0.18, 0.4], "radius": 0.18, "materialID": 5, "lightID": 0},
    ]
    materials = [
        {"type": "lambert", "gaussians": WHITE},
        {"type": "lambert", "gaussians": RED},
        {"type": "lambert", "gaussians": GREEN},
        {"type": "lambert", "gaussians": [[0.3, 500.0, 50.0], [0.4, 600.0, 50.0], [0.1, 420.0, 30.0]]},
        {"type": "dielectric", "cauchy": [1.5, 0.01]},
        {"type": "lambert", "gaussians": [[0.0, 550.0, 10.0]]},  # the light's surface: black
    ]
    # Camera outside the open face looking along +z; exposure apertureSize^2 * ISO = 0.01.
    cam = _camera([0.5, 0.5, -1.6], [0.0, 0.0], iso=1, ap=0.1, apdist=0.049 * 1.6)
    return {
        "camera": cam,
        "sphere": spheres,
        "quad": quads,
        "material": materials,
        "light": [{"emission": {"temperature": 6500.0, "luminosity": 12.0}}],
        "render": {"wavelengths": wavelengths, "lambdaMin": 380.0, "lambdaMax": 780.0,
                   "russianRouletteDepth": 3, "seed": seed},
    }


def cornell(wavelengths: int = 1, seed: int = 0) -> str:
    return json.dumps(cornell_dict(wavelengths, seed))


def random_spheres_dict(n: int = 4096, scene_seed: int = 1, wavelengths: int = 4, seed: int = 0) -> dict:
    s = cornell_dict(wavelengths, seed)
    rng = np.random.default_rng(scene_seed)
    mats = s["material"]
    centres = rng.uniform(0.0, 1.0, (n, 3))
    radii = np.exp(rng.uniform(math.log(0.002), math.log(0.05), n))
    kind = rng.uniform(0.0, 1.0, n)
    for i in range(n):
        if kind[i] < 0.60:
            mu = rng.uniform(400.0, 700.0, 3)
            sig = rng.uniform(20.0, 80.0, 3)
            w = rng.dirichlet([1.0, 1.0, 1.0]) * rng.uniform(0.3, 0.9)
            m = {"type": "lambert", "gaussians": [[float(w[j]), float(mu[j]), float(sig[j])] for j in range(3)]}
        elif kind[i] < 0.85:
            m = {"type": "dielectric", "cauchy": [float(rng.uniform(1.4, 1.8)), float(rng.uniform(0.003, 0.02))]}
        else:
            m = {"type": "conductor", "n": float(rng.uniform(0.2, 2.5)), "k": float(rng.uniform(2.0, 8.0))}
        mats.append(m)
        s["sphere"].append({"position": [float(v) for v in centres[i]], "radius": float(radii[i]),
                            "materialID": len(mats), "lightID": 0})
    return s


def random_spheres(n: int = 4096, scene_seed: int = 1, wavelengths: int = 4, seed: int = 0) -> str:
    return json.dumps(random_spheres_dict(n, scene_seed, wavelengths, seed))


def sdf_scene_dict(jitter: float = 0.0, scene_seed: int = 0, wavelengths: int = 4, seed: int = 0,
                   mandel_iters: int = 8, menger_levels: int = 4) -> dict:
    rng = np.random.default_rng(scene_seed)
    j = (lambda: 1.0 + rng.uniform(-jitter, jitter)) if jitter > 0 else (lambda: 1.0)
    mb_pos = [-1.4 * j(), 1.25 * j(), 0.0]
    mg_pos = [1.4 * j(), 1.0 * j(), 0.0]
    return {
        "camera": _camera([0.0, 2.2, -6.0], [0.0, -12.0], iso=1, ap=0.1, apdist=0.049 * 1.2),
        "plane": [{"position": [0.0, 0.0, 0.0], "materialID": 1, "lightID": 0}],
        "sdf": [
            {"kind": "mandelbulb", "iterations": mandel_iters, "bailout": 16.0, "position": mb_pos,
             "boundingSize": [2.5 * j(), 2.5 * j(), 2.5 * j()], "material": 1.0},
            {"kind": "menger", "iterations": menger_levels, "position": mg_pos,
             "boundingSize": [2.02, 2.02, 2.02], "material": 2.0},
        ],
        "material": [
            {"type": "lambert", "gaussians": WHITE},
            {"type": "lambert", "gaussians": [[0.7, 600.0, 60.0], [0.2, 480.0, 40.0], [0.0, 550.0, 10.0]]},
            {"type": "lambert", "gaussians": [[0.6, 470.0, 50.0], [0.2, 560.0, 50.0], [0.0, 550.0, 10.0]]},
        ],
        "light": [],
        "environment": {"radiance": 1.0},
        "render": {"wavelengths": wavelengths, "lambdaMin": 380.0, "lambdaMax": 780.0,
                   "russianRouletteDepth": 3, "seed": seed},
    }


def sdf_scene(**kw) -> str:
    return json.dumps(sdf_scene_dict(**kw))


SERVING_RESOLUTIONS = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]


def serving_request(i: int) -> dict:
    """Request i of the serving workload, from a generator seeded i: resolution
    uniform over SERVING_RESOLUTIONS; sample budget round(exp(U[ln 16, ln 1024]));
    with probability 1/2 the Cornell box with N = round(exp(U[0, ln 4096]))
    random spheres (scene seed i, depth 12), else the mandelbulb + menger scene
    with its parameters jittered by U[0.9, 1.1] (scene seed i, depth 8); 4 hero
    wavelengths; render seed i."""
    rng = np.random.default_rng(i)
    w, h = SERVING_RESOLUTIONS[int(rng.integers(len(SERVING_RESOLUTIONS)))]
    spp = int(round(math.exp(rng.uniform(math.log(16.0), math.log(1024.0)))))
    if rng.uniform() < 0.5:
        n = int(round(math.exp(rng.uniform(0.0, math.log(4096.0)))))
        return {"id": i, "kind": f"spheres{n}", "text": random_spheres(n=n, scene_seed=i, wavelengths=4, seed=i),
                "width": w, "height": h, "spp": spp, "depth": 12}
    return {"id": i, "kind": "sdf", "text": sdf_scene(jitter=0.1, scene_seed=i, wavelengths=4, seed=i),
            "width": w, "height": h, "spp": spp, "depth": 8}
