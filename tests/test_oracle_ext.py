"""Pins the FP64 extended-mode oracle (oracle/ptext.c) to closed forms."""
import math

import numpy as np
import pytest

from oracle import ptext
from pathtracer_amd import scenes


def test_fresnel_dielectric_closed_forms():
    f = ptext.lib.ptext_fresnel_dielectric
    # normal incidence ((n1 - n2) / (n1 + n2))^2
    assert abs(f(1.0, 1.0, 1.5) - 0.04) < 1e-12
    # Brewster angle: the p component vanishes, R = rs^2 / 2
    tb = math.atan(1.5)
    ci, ct = math.cos(tb), math.cos(math.asin(math.sin(tb) / 1.5))
    rs = (ci - 1.5 * ct) / (ci + 1.5 * ct)
    assert abs(f(ci, 1.0, 1.5) - 0.5 * rs * rs) < 1e-12
    # total internal reflection beyond the critical angle
    assert f(math.cos(math.radians(60)), 1.5, 1.0) == 1.0
    # reciprocity R(theta_i, n1 -> n2) = R(theta_t, n2 -> n1)
    ci = math.cos(0.7)
    ct = math.sqrt(1 - (math.sin(0.7) / 1.5) ** 2)
    assert abs(f(ci, 1.0, 1.5) - f(ct, 1.5, 1.0)) < 1e-12


def test_fresnel_conductor_closed_forms():
    f = ptext.lib.ptext_fresnel_conductor
    for n, k in [(0.2, 3.0), (1.2, 4.0), (2.5, 8.0)]:
        assert abs(f(1.0, n, k) - ((n - 1) ** 2 + k ** 2) / ((n + 1) ** 2 + k ** 2)) < 1e-12
        assert abs(f(1e-9, n, k) - 1.0) < 1e-6  # grazing incidence reflects everything
    # k = 0 reduces to the dielectric formula
    for ci in [1.0, 0.8, 0.3]:
        assert abs(f(ci, 1.5, 0.0) - ptext.lib.ptext_fresnel_dielectric(ci, 1.0, 1.5)) < 1e-12


def test_planck_normalisation():
    for T in [2200.0, 6500.0]:
        lmax = 2.897771955e-3 / T * 1e9
        assert abs(ptext.lib.ptext_planck(lmax, T, 1.0) - 1.0) < 2e-5
        assert abs(ptext.lib.ptext_planck(lmax, T, 3.0) - 3.0) < 6e-5


def test_oracle_furnace_and_energy():
    """Convex Lambert sphere (albedo 0.5) under a uniform sky: 0.5 x sky on the
    sphere; a lossless dielectric sphere is invisible."""
    import json

    def scene(mat):
        return json.dumps({
            "camera": {"numShots": 1, "position": [[0, 0, -3.0]], "angle": [[0, 0]], "ISO": 1, "size": 0.057,
                       "apertureSize": 1.0, "apertureDistance": 0.1, "lensRadius": 0.01, "lensFocalLength": 0.03,
                       "lensThickness": 0.0, "lensDistance": 0.05, "model": "pinhole"},
            "sphere": [{"position": [0, 0, 0], "radius": 1.0, "materialID": 1, "lightID": 0}],
            "material": [mat], "light": [], "environment": {"radiance": 1.0},
            "render": {"wavelengths": 4, "russianRouletteDepth": 64}})
    sky = json.loads(scene({"type": "lambert", "gaussians": [[0.5, 550, 1e7]]}))
    sky["sphere"] = []
    s = ptext.render(json.dumps(sky), 8, 8, 64, 4)["mean"].mean(axis=(0, 1))
    lam = ptext.render(scene({"type": "lambert", "gaussians": [[0.5, 550, 1e7]]}), 8, 8, 64, 8)["mean"]
    # the image corners look past the sphere's limb: use the centre pixels
    np.testing.assert_allclose(lam[2:6, 2:6].mean(axis=(0, 1)) / s, 0.5, rtol=0.02)
    gl = ptext.render(scene({"type": "dielectric", "cauchy": [1.5, 0.0]}), 8, 8, 64, 64)["mean"]
    np.testing.assert_allclose(gl.mean(axis=(0, 1)) / s, 1.0, rtol=0.02)


def test_oracle_equal_energy_sky_is_white():
    """An equal-energy sky integrated against the CIE table gives x = y = 1/3."""
    import json
    t = json.loads(scenes.sdf_scene())
    t["sdf"], t["plane"] = [], []
    r = ptext.render(json.dumps(t), 8, 8, 256, 2)["mean"].mean(axis=(0, 1))
    np.testing.assert_allclose(r / r.sum(), [1 / 3, 1 / 3, 1 / 3], atol=3e-3)
