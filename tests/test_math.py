"""ptmath.h (the tracer's FP32 GLSL built-ins) pinned against numpy binary64.

The reference leaves sin/cos/exp/log/pow/acos to the Vulkan driver; the
tracer fixes one definition for the CPU oracle and the GPU. Here each is
checked against the binary64 value: a dropped series term, a wrong reduction
constant or a wrong quadrant shows up as an error of many ulp.
"""
import numpy as np
import pytest

import pathtracer_amd as pt

SIN, COS, EXP, LOG, POW, ACOS, SQRT, RSQRT = range(8)


def ulp_err(got32, ref64):
    ref32 = ref64.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    spacing = np.maximum(spacing, np.float64(np.finfo(np.float32).smallest_subnormal))
    return np.abs(got32.astype(np.float64) - ref64) / spacing


rng = np.random.default_rng(1234)


@pytest.mark.parametrize("fn,npf,lo,hi,tol", [
    (SIN, np.sin, -3000.0, 3000.0, 2.0),
    (COS, np.cos, -3000.0, 3000.0, 2.0),
    (SIN, np.sin, -4.0, 4.0, 2.0),
    (COS, np.cos, -4.0, 4.0, 2.0),
    (EXP, np.exp, -87.0, 88.0, 1.5),
    (ACOS, np.arccos, -1.0, 1.0, 1.0),
])
def test_unary_ulp(fn, npf, lo, hi, tol):
    x = rng.uniform(lo, hi, 400_000).astype(np.float32)
    got = pt.math_eval_host(fn, x)
    err = ulp_err(got, npf(x.astype(np.float64)))
    assert err.max() <= tol, (fn, err.max(), x[err.argmax()])


def test_log_ulp_over_exponents():
    m = rng.uniform(1.0, 2.0, 400_000)
    e = rng.integers(-140, 127, 400_000)
    x = np.ldexp(m, e).astype(np.float32)
    x = x[x > 0]
    got = pt.math_eval_host(LOG, x)
    err = ulp_err(got, np.log(x.astype(np.float64)))
    assert err.max() <= 1.5


def test_pow_ulp():
    x = np.ldexp(rng.uniform(0.5, 1.0, 300_000), rng.integers(-30, 30, 300_000)).astype(np.float32)
    y = rng.uniform(-8.0, 8.0, 300_000).astype(np.float32)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    ok = (ref > 1e-37) & (ref < 3e38)
    got = pt.math_eval_host(POW, x, y)
    assert ulp_err(got[ok], ref[ok]).max() <= 1.0


def test_exp_subnormal_and_limits():
    x = np.array([-87.5, -95.0, -103.0, -104.5, 88.7, 89.0, 0.0, -0.0], np.float32)
    got = pt.math_eval_host(EXP, x)
    ref = np.exp(x.astype(np.float64))
    assert got[3] == 0.0 and np.isinf(got[5])
    assert ulp_err(got[:3], ref[:3]).max() <= 1.5
    assert got[6] == 1.0 and got[7] == 1.0
    assert ulp_err(got[4:5], ref[4:5]).max() <= 1.5


def test_special_values():
    nan = np.float32(np.nan)
    assert np.isnan(pt.math_eval_host(SIN, np.array([np.inf, nan], np.float32))).all()
    assert np.isnan(pt.math_eval_host(LOG, np.array([-1.0], np.float32)))[0]
    assert pt.math_eval_host(LOG, np.array([0.0], np.float32))[0] == -np.inf
    assert pt.math_eval_host(LOG, np.array([1.0], np.float32))[0] == 0.0
    p = pt.math_eval_host(POW, np.array([0.0, 0.0, 2.0, 1.0], np.float32), np.array([2.0, -1.0, 0.0, 7.0], np.float32))
    assert list(p) == [0.0, np.inf, 1.0, 1.0]
    assert np.isnan(pt.math_eval_host(ACOS, np.array([1.5], np.float32)))[0]
    # sin(k*pi) tiny, cos(0) exact, quadrant signs
    assert pt.math_eval_host(COS, np.array([0.0], np.float32))[0] == 1.0
    s = pt.math_eval_host(SIN, np.array([np.pi / 2, np.pi, 3 * np.pi / 2], np.float32))
    assert abs(s[0] - 1) < 1e-7 and abs(s[1]) < 1e-6 and abs(s[2] + 1) < 1e-7


def test_sqrt_rsqrt_correctly_rounded():
    x = np.abs(rng.standard_normal(200_000)).astype(np.float32) * 100
    assert np.array_equal(pt.math_eval_host(SQRT, x), np.sqrt(x))
    r = pt.math_eval_host(RSQRT, x)
    assert np.array_equal(r, (np.float32(1.0) / np.sqrt(x)).astype(np.float32))
