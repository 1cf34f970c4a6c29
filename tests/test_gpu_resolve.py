"""Display resolve (pt_resolve_srgb8) against a numpy restatement of the
reference's SaveRender / shader.frag math (src/pathtracer.cpp:3491-3518,
1040-1071; src/shader.frag:30-99): Bradford E -> D65 and XYZ -> linear sRGB
(Lindbloom), clamp at 0, tonemap (none / Reinhard / ACES / DEUCES), the sRGB
OETF, truncation to 8 bits."""
import numpy as np
import pytest

import pathtracer_amd as pt

# GLSL `v * mat3(...)` with the constructors' argument order: out[i] = dot(v, row i as written)
BRADFORD_E_D65 = np.array([[0.9531874, -0.0265906, 0.0238731],
                           [-0.0382467, 1.0288406, 0.0094060],
                           [0.0026068, -0.0030332, 1.0892565]], np.float32)
XYZ_TO_SRGB = np.array([[3.2404542, -1.5371385, -0.4985314],
                        [-0.9692660, 1.8760108, 0.0415560],
                        [0.0556434, -0.2040259, 1.0572252]], np.float32)


def mat_apply(m, c):
    """(m[i,0] x + m[i,1] y) + m[i,2] z in float32, the order the kernel computes."""
    x, y, z = c[..., 0], c[..., 1], c[..., 2]
    return np.stack([(m[i, 0] * x + m[i, 1] * y) + m[i, 2] * z for i in range(3)], -1).astype(np.float32)


def resolve_numpy(xyz: np.ndarray, tonemap: int) -> np.ndarray:
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        r = np.maximum(mat_apply(XYZ_TO_SRGB, mat_apply(BRADFORD_E_D65, xyz.astype(np.float32))), np.float32(0))
        if tonemap == 1:
            r = r / (np.float32(1) + r)
        elif tonemap == 2:
            r = r * (np.float32(2.51) * r + np.float32(0.03)) / (r * (np.float32(2.43) * r + np.float32(0.59)) + np.float32(0.14))
        elif tonemap == 3:
            r = np.exp(np.float32(-0.25) / r).astype(np.float32)
        r = np.clip(r, np.float32(0), np.float32(1))
        lo = np.float32(12.92) * r
        hi = np.float32(1.055) * np.power(r, np.float32(1.0 / 2.4)) - np.float32(0.055)
        v = np.where(r <= np.float32(0.0031308), lo, hi).astype(np.float32)
    return (v.astype(np.float64) * 255.0).astype(np.int64).astype(np.uint8)


def test_equal_energy_white_maps_to_srgb_white():
    """Pins the restatement: E-white (1, 1, 1) -> D65 white -> linear sRGB (1, 1, 1);
    a transposed matrix fails this."""
    lin = XYZ_TO_SRGB.astype(np.float64) @ (BRADFORD_E_D65.astype(np.float64) @ np.ones(3))
    np.testing.assert_allclose(lin, 1.0, atol=2e-4)
    bad = XYZ_TO_SRGB.T.astype(np.float64) @ (BRADFORD_E_D65.astype(np.float64) @ np.ones(3))
    assert np.abs(bad - 1.0).max() > 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("tonemap", [0, 1, 2, 3])
def test_resolve_matches_the_restatement(gpu_ctx, tonemap):
    rng = np.random.default_rng(tonemap)
    W, H = 97, 61
    xyz = np.exp(rng.uniform(np.log(1e-4), np.log(20.0), (H, W, 3))).astype(np.float32)
    xyz[0, :8] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1e-9, 1e-9, 1e-9], [100, 100, 100],
                  [0.2, 0.1, 0.9], [0.9, 0.1, 0.02], [-0.1, 0.2, 0.3]]
    tex = np.concatenate([xyz, np.ones((H, W, 1), np.float32)], -1)
    acc = gpu_ctx.accum(W, H)
    acc.write(tex)
    got = acc.resolve_srgb8(tonemap).reshape(H, W, 3).astype(np.int64)
    want = resolve_numpy(xyz, tonemap).astype(np.int64)
    diff = np.abs(got - want)
    # the kernel's pow/exp are this library's FP32 built-ins: a value within an
    # ulp of a 1/255 step may truncate to the neighbouring code
    assert diff.max() <= 1 and (diff == 0).mean() >= 0.999
    assert (got[0, 0] == 0).all()
    if tonemap == 0:
        # saturated: 1.055f * 1 - 0.055f = 0.99999994f in FP32, truncated to 254 as
        # the reference's SaveRender does (src/pathtracer.cpp:1061-1070, 3508-3510)
        assert (got[0, 4] == 254).all() and (want[0, 4] == 254).all()
