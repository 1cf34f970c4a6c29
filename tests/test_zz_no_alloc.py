"""pt_dispatch_noalloc (reference mode through a static device wave, no
allocation) against pt_dispatch(PT_MODE_REF): texels bit for bit, on two calls
in a row, the second with fewer pixels (a smaller accumulator) so the static
wave holds the first call's records when the second starts. Called through the
C ABI."""
import ctypes as C

import numpy as np
import pytest

import pathtracer_amd as pt
from tests.conftest import SCENES

pytestmark = pytest.mark.gpu


def run(ctx, scene, W, H, spf, frames, noalloc):
    r = pt.Renderer(ctx, scene, W, H, spf=spf, path_length=6, mode=pt.MODE_REF)
    for f in range(1, frames + 1):
        r.pc.frame = r.pc.currentSamples = f * spf
        if noalloc:
            st = pt.lib.pt_dispatch_noalloc(ctx.h, scene.h, r.acc.h, C.byref(r.pc))
            assert st == pt.PT_OK, pt.lib.pt_last_error(ctx.h)
        else:
            st = pt.lib.pt_dispatch(ctx.h, scene.h, r.acc.h, C.byref(r.pc), pt.MODE_REF)
            assert st == pt.PT_OK
    assert pt.lib.pt_sync(ctx.h) == pt.PT_OK
    return r.acc.read()


@pytest.mark.parametrize("scene_idx", [1, 4])
def test_noalloc_dispatch_equals_dispatch(scene_idx):
    text = SCENES[scene_idx].read_text()
    with pt.Context(0) as ctx:
        scene = ctx.scene_from_json(text)
        # 1st: 320 x 240 x 2 spp = 153,600 paths, more than the static wave's
        # 65,536, so pixels run in several chunks; 2nd: smaller, after it
        for (W, H, spf, frames) in ((320, 240, 2, 2), (72, 40, 3, 2)):
            c0 = ctx.counters()
            a = run(ctx, scene, W, H, spf, frames, noalloc=True)
            c1 = ctx.counters()
            b = run(ctx, scene, W, H, spf, frames, noalloc=False)
            c2 = ctx.counters()
            assert np.array_equal(a, b), (W, H)
            assert np.isfinite(a).all() and a[..., 1].mean() > 0
            for k in ("segments", "shadow_segments", "sdf_steps", "paths"):
                assert c1[k] - c0[k] == c2[k] - c1[k], k


def test_noalloc_dispatch_rejects_bad_arguments(gpu_ctx):
    scene = gpu_ctx.scene_from_json(SCENES[1].read_text())
    acc = gpu_ctx.accum(16, 8)
    pc = pt.push_constants(16, 9, spf=1, path_length=4)
    scene.camera(1, pc)
    pc.frame = pc.currentSamples = 1
    assert pt.lib.pt_dispatch_noalloc(gpu_ctx.h, scene.h, acc.h, C.byref(pc)) == pt.PT_ERR_INVALID_ARG
    assert pt.lib.pt_dispatch_noalloc(None, None, None, None) == pt.PT_ERR_INVALID_ARG
