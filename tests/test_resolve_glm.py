"""The resolve's colour matrices against the glm library the reference vendors.

oracle/_ref/glm_probe (built by oracle/Makefile where the reference tree is
present) evaluates SaveRender's `XYZ * glm::mat3(...)` Bradford and sRGB stages
and the clamp at zero (src/pathtracer.cpp:970-986, 3501-3503) with glm itself.
The numpy restatement that tests/test_gpu_resolve.py holds the GPU kernel to,
bit for bit, must give the same binary32 values: this pins the operand
orientation (glm's column-major constructor with the vector on the left) and
the order of the products and sums."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.test_gpu_resolve import BRADFORD_E_D65, XYZ_TO_SRGB, mat_apply

ROOT = Path(__file__).resolve().parent.parent
PROBE = ROOT / "oracle" / "_ref" / "glm_probe"

pytestmark = pytest.mark.skipif(not PROBE.exists(), reason="oracle/_ref/glm_probe not built (needs the reference tree)")


def glm_stage(xyz: np.ndarray) -> np.ndarray:
    r = subprocess.run([str(PROBE)], input=np.ascontiguousarray(xyz, np.float32).tobytes(), capture_output=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return np.frombuffer(r.stdout, np.float32).reshape(-1, 3)


def restated(xyz: np.ndarray, srgb=XYZ_TO_SRGB) -> np.ndarray:
    return np.maximum(mat_apply(srgb, mat_apply(BRADFORD_E_D65, xyz.astype(np.float32))), np.float32(0))


def test_matrix_stage_is_bit_identical_to_glm():
    rng = np.random.default_rng(11)
    xyz = np.exp(rng.uniform(np.log(1e-6), np.log(1e3), (100000, 3))).astype(np.float32)
    xyz[:500] *= np.float32(-1)  # negative texels (a negative-lobe sample) reach the clamp
    xyz[500:508] = [[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [1e-30, 1e-30, 1e-30], [100, 100, 100],
                    [0.2, 0.1, 0.9], [0.9, 0.1, 0.02], [-0.1, 0.2, 0.3]]
    got = glm_stage(xyz)
    assert np.array_equal(got.view(np.uint32), restated(xyz).view(np.uint32))


def test_a_transposed_matrix_is_caught():
    rng = np.random.default_rng(12)
    xyz = rng.uniform(0.0, 2.0, (1000, 3)).astype(np.float32)
    wrong = restated(xyz, np.ascontiguousarray(XYZ_TO_SRGB.T))
    assert not np.array_equal(glm_stage(xyz), wrong)
