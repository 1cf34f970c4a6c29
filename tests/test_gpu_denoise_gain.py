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


# The shipped network alone, denoised / noisy relative RMSE on the held-out
# scenes (measured 0.24 / 0.33 / 0.51 on the SDF scene and 0.75 / 0.94 / 1.33
# on the sphere box at 16 / 64 / 256 spp): a gain everywhere on the SDF scene;
# on the sphere box (thousands of tiny glossy and glass spheres, caustics) a
# gain up to 64 spp and a bias that outweighs the input's noise from about
# 128 spp, which the served form catches (next test).
SHIPPED_BARS = {16: {"sdf": 0.35, "spheres": 0.85}, 64: {"sdf": 0.45, "spheres": 1.0},
                256: {"sdf": 0.65}}


@pytest.mark.skipif(not os.path.exists(T.DEFAULT_WEIGHTS), reason="no trained weights shipped")
@pytest.mark.parametrize("spp", sorted(SHIPPED_BARS))
def test_shipped_weights_give_a_gain(gpu_ctx, spp):
    g = gains(gpu_ctx, np.load(T.DEFAULT_WEIGHTS, allow_pickle=False), spp)
    for kind, bar in SHIPPED_BARS[spp].items():
        assert g[kind][1] < bar * g[kind][0], (kind, g[kind])


@pytest.mark.skipif(not os.path.exists(T.DEFAULT_WEIGHTS), reason="no trained weights shipped")
@pytest.mark.parametrize("spp", [16, 64, 256])
def test_served_denoising_is_never_worse_than_plain(gpu_ctx, spp):
    """The denoiser as served: a request run to its sample budget returns the
    denoised image only when the dual-buffer estimate rates it better than the
    plain mean. Against an 8192-sample render of other samples, the served image
    is never worse than the plain one (beyond the estimate's noise), on both
    held-out scenes at 16, 64 and 256 spp, and on the SDF scene it gains."""
    from tests.test_gpu_serve import run_until_done
    res = {}
    for kind, i in HELD.items():
        text, depth = T.training_scene(i)
        with pt.Server(gpu_ctx) as srv, pt.Denoiser(gpu_ctx, 2, 128, 128) as dn:
            T.load_weights(dn)
            srv.set_denoiser(dn)
            a = srv.submit(text, 128, 128, spp_max=spp, path_length=depth, target=1e-6, denoise=True)
            b = srv.submit(text, 128, 128, spp_max=spp, path_length=depth, target=1e-6)
            st = run_until_done(srv, [a, b])
            served, plain = srv.read(a)[..., :3], srv.read(b)[..., :3]
        sc = gpu_ctx.scene_from_json(text)
        ref = T.render_mean(gpu_ctx, sc, 128, 128, 8192, depth, first=10 ** 5, chunk=512)
        sc.release()
        e_s, e_p = T.rel_rmse(served, ref), T.rel_rmse(plain, ref)
        res[kind] = (e_p, e_s, bool(st[a].denoised))
        print(f"\n{kind} at {spp} spp: plain {e_p:.4f} served {e_s:.4f} (denoised: {st[a].denoised}, "
              f"estimate {st[a].rel_error:.4f}, plain estimate {st[a].rel_plain:.4f})")
        # a 1e-6 target is out of reach: both run to their budgets (an estimate
        # that came out as zero once finished a sphere request at 8 samples)
        assert st[a].spp == spp and st[b].spp == spp and not st[a].converged
    for kind, (e_p, e_s, _) in res.items():
        assert e_s <= 1.05 * e_p, (kind, res[kind])
    assert res["sdf"][1] < 0.7 * res["sdf"][0] and res["sdf"][2]
