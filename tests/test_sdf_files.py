"""The SDF recogniser on the four distance-estimator files the reference ships
(sdfs/*.glsl), read where the reference tree is present: each is recognised as
its kind with the parameters its text sets (loop counts, bail-out, and the
material ids and radial mix of its sdfmaterial)."""
from pathlib import Path

import numpy as np
import pytest

import pathtracer_amd as pt

SDFS = Path("/root/reference/sdfs")

pytestmark = pytest.mark.skipif(not SDFS.is_dir(), reason="reference tree not present")

# file -> (kind, material kind, iterations, material parameters)
EXPECTED = {
    # `i < 4`, bail-out `r > 16.0`; sdfmaterial mix(4.0, 3.0, factor / (0.8 + factor))
    "mandelbulb.glsl": (pt.SDF_MANDELBULB, pt.SDFMAT_RADIAL, 4, [4.0, 3.0, 0.8]),
    # `m < 5` levels; constant material 0
    "menger.glsl": (pt.SDF_MENGER, pt.SDFMAT_CONSTANT, 5, [0.0]),
    # four smooth-min spheres with material ids 4, 5, 6, 6
    "blob.glsl": (pt.SDF_BLOB, pt.SDFMAT_BLOB, None, [4.0, 5.0, 6.0, 6.0]),
    "terrain.glsl": (pt.SDF_TERRAIN, pt.SDFMAT_CONSTANT, None, [0.0]),
}


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_reference_sdf_files_are_recognised(name):
    kind, mat_kind, iters, mat = EXPECTED[name]
    d = pt.recognize_sdf((SDFS / name).read_text())
    assert (d.kind, d.material_kind) == (kind, mat_kind)
    if iters is not None:
        assert d.iterations == iters
    if kind == pt.SDF_MANDELBULB:
        assert d.bailout == 16.0
    np.testing.assert_array_equal(np.float32(list(d.material)[:len(mat)]), np.float32(mat))


def test_every_shipped_file_is_covered():
    assert sorted(p.name for p in SDFS.glob("*.glsl")) == sorted(EXPECTED)
