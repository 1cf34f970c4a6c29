"""The device build of ptmath.h returns the same bits as the host build, so
the CPU oracle and the gfx950 kernels evaluate GLSL built-ins identically."""
import numpy as np
import pytest

import pathtracer_amd as pt

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(77)


@pytest.mark.parametrize("fn,lo,hi", [(0, -4000, 4000), (1, -4000, 4000), (2, -110, 90), (3, 0, 1e6), (5, -1.1, 1.1),
                                      (6, 0, 1e30), (7, 1e-30, 1e30)])
def test_device_equals_host(gpu_ctx, fn, lo, hi):
    x = rng.uniform(lo, hi, 1 << 20).astype(np.float32)
    if fn == 3:
        x = np.concatenate([x, np.ldexp(rng.uniform(1, 2, 1 << 18), rng.integers(-149, 128, 1 << 18)).astype(np.float32)])
    x = np.concatenate([x, np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 1.0, -1.0], np.float32)])
    a = pt.math_eval_host(fn, x)
    b = gpu_ctx.math_eval(fn, x)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_device_pow_equals_host(gpu_ctx):
    x = np.ldexp(rng.uniform(0.5, 1, 1 << 20), rng.integers(-40, 40, 1 << 20)).astype(np.float32)
    y = rng.uniform(-9, 9, 1 << 20).astype(np.float32)
    a = pt.math_eval_host(4, x, y)
    b = gpu_ctx.math_eval(4, x, y)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
