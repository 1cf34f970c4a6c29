"""The denoiser's gain on renders it was not trained on: relative RMSE against
a 8192-sample render, noisy input vs denoised output. Training runs on the GPU
through the library's own forward/backward kernels (pt_dn_train_step)."""
import os

import numpy as np
import pytest

import pathtracer_amd as pt
from pathtracer_amd import denoise_train as T

pytestmark = pytest.mark.gpu

# training_scene(i): 100 and 106 are random-sphere Cornell boxes, 101 and 104
# the mandelbulb + menger scene; none is in the training sets below
HELD = {"spheres": 100, "sdf": 101}


def held_out(ctx, i, spp, size=128):
    text, depth = T.training_scene(i)
    sc = ctx.scene_from_json(text)
    try:
        feat = T.render_features(ctx, sc, size, size, spp, depth)
        ref = T.render_mean(ctx, sc, size, size, 8192, depth, first=10 ** 5)
    finally:
        sc.release()
    return feat, ref


def gains(ctx, params, spp):
    out = {}
    with pt.Denoiser(ctx, 1, 128, 128) as dn:
        dn.set_params(params)
        for kind, i in HELD.items():
            feat, ref = held_out(ctx, i, spp)
            noisy = T.rel_rmse(feat[..., :3], ref)
            den = T.rel_rmse(T.denoise_image(dn, feat), ref)
            print(f"\n{kind} at {spp} spp: noisy {noisy:.4f} denoised {den:.4f}")
            out[kind] = (noisy, den)
    return out


def test_training_on_the_gpu_gives_a_gain(gpu_ctx):
    pairs = []
    for i in range(8):
        text, depth = T.training_scene(i)
        pairs.append(T.make_pair(gpu_ctx, text, depth, 128, 128, (4, 16)[i % 2], 1024))
    with pt.Denoiser(gpu_ctx, 8, 64, 64, seed=1) as dn:
        losses = T.train_two_phase(dn, pairs, 2400, 8, 64)
        params = dn.params()
    assert np.mean(losses[1100:1200]) < 0.8 * np.mean(losses[:100])
    # Weight gradients are summed with FP32 atomics, so a run is not
    # bit-reproducible; runs of 1600 steps gave SDF ratios on both sides of 0.5
    # (0.55 on one), and the bar leaves room for that spread.
    g = gains(gpu_ctx, params, 16)
    assert g["sdf"][1] < 0.6 * g["sdf"][0]
    # on the glossy, caustic sphere boxes a network trained on 8 scenes gains
    # 4-12 % at 16 spp (run to run); it must not be worse than its input
    assert g["spheres"][1] < g["spheres"][0]


@pytest.mark.skipif(not os.path.exists(T.DEFAULT_WEIGHTS), reason="no trained weights shipped")
def test_shipped_weights_give_a_gain(gpu_ctx):
    g = gains(gpu_ctx, np.load(T.DEFAULT_WEIGHTS, allow_pickle=False), 16)
    assert g["sdf"][1] < 0.35 * g["sdf"][0]
    assert g["spheres"][1] < 0.85 * g["spheres"][0]
