"""Training pairs for the U-Net denoiser (pt_dn_*) rendered by the tracer, and
its training on the GPU by the library's own forward/backward kernels.

A pair is (features of a low-sample render, mean of an independent
high-sample render of the same view): the features come from
pt_render_features over samples [0, spp_in), the target from pt_dispatch over
samples [spp_in, spp_in + spp_ref). Inputs and targets are divided by the
RMS of the noisy XYZ mean (`image_scale`), so the network sees radiance of
order one whatever the scene's exposure; denoised images are scaled back.

Training runs in two phases: the mean absolute log error first (robust to
the fireflies of low-sample inputs), then the squared error of the scaled
radiance, the quantity whose sum the relative RMSE measures.

    python -m pathtracer_amd.denoise_train --out pathtracer_amd/data/denoiser_weights.npy
"""
from __future__ import annotations

import argparse
import math
import os
import time

import numpy as np

import pathtracer_amd as pt
from pathtracer_amd import scenes


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "denoiser_weights.npy")


def load_weights(dn: pt.Denoiser, path: str = DEFAULT_WEIGHTS) -> None:
    """Parameters saved by main() (float32 .npy, no pickles)."""
    dn.set_params(np.load(path, allow_pickle=False))


def image_scale(xyz: np.ndarray) -> float:
    """RMS over pixels and channels of an XYZ image (a positive floor for black ones)."""
    return max(float(np.sqrt(np.mean(np.square(xyz.astype(np.float64))))), 1e-12)


def render_features(ctx: pt.Context, scene: pt.Scene, w: int, h: int, spp: int, depth: int, first: int = 0,
                    shot: int = 1) -> np.ndarray:
    """(h, w, FEATURES) features of samples [first, first + spp)."""
    pc = pt.push_constants(w, h, spp, depth)
    scene.camera(shot, pc)
    pc.frame = first + spp
    pc.currentSamples = spp
    return ctx.features(scene, pc)


def render_mean(ctx: pt.Context, scene: pt.Scene, w: int, h: int, spp: int, depth: int, first: int = 0,
                shot: int = 1, chunk: int = 64) -> np.ndarray:
    """(h, w, 3) XYZ mean over samples [first, first + spp), traced `chunk` samples per dispatch."""
    acc = ctx.accum(w, h)
    try:
        done = 0
        while done < spp:
            m = min(chunk, spp - done)
            pc = pt.push_constants(w, h, m, depth)
            scene.camera(shot, pc)
            pc.frame = first + done + m
            pc.currentSamples = done + m
            ctx.dispatch(scene, acc, pc, pt.MODE_EXT)
            done += m
        return acc.read()[..., :3].copy()
    finally:
        acc.release()


def training_scene(i: int) -> tuple[str, int]:
    """Scene i of the training set: the serving workload's two families with
    seeds disjoint from its requests (10**6 + i)."""
    rng = np.random.default_rng(10 ** 6 + i)
    seed = 10 ** 6 + i
    if rng.uniform() < 0.5:
        n = int(round(math.exp(rng.uniform(0.0, math.log(4096.0)))))
        return scenes.random_spheres(n=n, scene_seed=seed, wavelengths=4, seed=seed), 12
    return scenes.sdf_scene(jitter=0.1, scene_seed=seed, wavelengths=4, seed=seed), 8


def make_pair(ctx: pt.Context, text: str, depth: int, w: int, h: int, spp_in: int, spp_ref: int, chunk: int = 64):
    """-> (features (h, w, FEATURES), target (h, w, 3)) from independent samples."""
    sc = ctx.scene_from_json(text)
    try:
        feat = render_features(ctx, sc, w, h, spp_in, depth)
        target = render_mean(ctx, sc, w, h, spp_ref, depth, first=spp_in, chunk=chunk)
    finally:
        sc.release()
    return feat, target


def normalise(feat: np.ndarray, target: np.ndarray | None = None):
    """Radiance channels (XYZ, standard error of Y) and target divided by image_scale."""
    s = image_scale(feat[..., 0:3])
    f = feat.copy()
    f[..., 0:3] /= s
    f[..., 10] /= s
    return (f, None if target is None else target / s, s)


def denoise_image(dn: pt.Denoiser, feat: np.ndarray) -> np.ndarray:
    """Denoised XYZ (h, w, 3) of one feature image, in the input's units."""
    f, _, s = normalise(feat)
    return dn.denoise(f[None])[0] * s


def rel_rmse(img: np.ndarray, ref: np.ndarray) -> float:
    """sqrt(sum of squared error) / sqrt(sum of ref^2) over pixels and channels."""
    img, ref = img.astype(np.float64), ref.astype(np.float64)
    return float(np.sqrt(np.sum((img - ref) ** 2)) / max(np.sqrt(np.sum(ref ** 2)), 1e-300))


def train(dn: pt.Denoiser, pairs, steps: int, batch: int = 8, crop: int = 64, lr: float = 1e-3,
          lr_end: float = 1e-4, seed: int = 0, log_every: int = 0) -> list[float]:
    """Adam on random crops of the normalised pairs, learning rate decayed
    geometrically from lr to lr_end; returns the per-step losses."""
    rng = np.random.default_rng(seed)
    norm = [normalise(f, t) for f, t in pairs]
    losses = []
    for step in range(steps):
        fb = np.empty((batch, crop, crop, pt.FEATURES), np.float32)
        tb = np.empty((batch, crop, crop, 3), np.float32)
        for b in range(batch):
            f, t, _ = norm[int(rng.integers(len(norm)))]
            y = int(rng.integers(0, f.shape[0] - crop + 1))
            x = int(rng.integers(0, f.shape[1] - crop + 1))
            fb[b] = f[y:y + crop, x:x + crop]
            tb[b] = t[y:y + crop, x:x + crop]
        a = lr * (lr_end / lr) ** (step / max(steps - 1, 1))
        losses.append(dn.train_step(fb, tb, a))
        if log_every and (step % log_every == 0 or step == steps - 1):
            print(f"step {step:5d} lr {a:.2e} loss {np.mean(losses[-log_every:]):.4f}", flush=True)
    return losses


def train_two_phase(dn: pt.Denoiser, pairs, steps: int, batch: int = 8, crop: int = 96, seed: int = 0,
                    log_every: int = 0) -> list[float]:
    """steps / 2 of mean absolute log error (lr 1e-3 -> 2e-4), then steps / 2
    of squared error (lr 2e-4 -> 2e-5)."""
    dn.set_loss(pt.DN_LOSS_L1_LOG)
    losses = train(dn, pairs, steps // 2, batch, crop, 1e-3, 2e-4, seed, log_every)
    dn.set_loss(pt.DN_LOSS_L2)
    losses += train(dn, pairs, steps - steps // 2, batch, crop, 2e-4, 2e-5, seed + 1, log_every)
    return losses


# Held-out scenes (never trained on): a random-sphere box and the SDF scene.
HELD_OUT = {"spheres": 100, "sdf": 101}


def evaluate(ctx: pt.Context, params: np.ndarray, spp_list, size: int = 128, spp_ref: int = 8192) -> dict:
    """Relative RMSE of the noisy input and of the denoised image against a
    spp_ref-sample render, on the held-out scenes at each input sample count."""
    out = {}
    with pt.Denoiser(ctx, 1, size, size) as dn:
        dn.set_params(params)
        for kind, i in HELD_OUT.items():
            text, depth = training_scene(i)
            sc = ctx.scene_from_json(text)
            try:
                ref = render_mean(ctx, sc, size, size, spp_ref, depth, first=10 ** 5, chunk=512)
                for spp in spp_list:
                    feat = render_features(ctx, sc, size, size, spp, depth)
                    out[(kind, spp)] = (rel_rmse(feat[..., :3], ref), rel_rmse(denoise_image(dn, feat), ref))
            finally:
                sc.release()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--eval", nargs="+", metavar="NPY",
                    help="instead of training: held-out relative RMSE, noisy and denoised, of these parameter files")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--first", type=int, default=0,
                    help="index of the first training scene (scenes 100 and 101 are the gain tests' held-out ones)")
    ap.add_argument("--size", type=int, default=192)
    ap.add_argument("--spp-in", type=str, default="4,16,64,256,1024")
    ap.add_argument("--spp-ref", type=int, default=4096, help="samples per pixel of a target (at least)")
    ap.add_argument("--ref-factor", type=int, default=16,
                    help="a target has at least this many times its input's samples (up to --spp-ref-max), so "
                         "that the target's own noise stays well below the input's at every input level")
    ap.add_argument("--spp-ref-max", type=int, default=16384)
    ap.add_argument("--steps", type=int, default=16000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=96)
    ap.add_argument("--chunk", type=int, default=512, help="samples per pixel per dispatch of the targets")
    ap.add_argument("--out", type=str, default="")
    a = ap.parse_args(argv)
    if a.eval:
        with pt.Context(0) as ctx:
            for w in a.eval:
                res = evaluate(ctx, np.load(w, allow_pickle=False), [4, 16, 64, 256, 1024])
                for (kind, spp), (noisy, den) in sorted(res.items()):
                    print(f"{w}: {kind} {spp:5d} spp: noisy {noisy:.4f} denoised {den:.4f} ratio {den / noisy:.3f}",
                          flush=True)
        return 0
    if not a.out:
        ap.error("--out is required for training")
    spp_in = [int(v) for v in a.spp_in.split(",")]
    t0 = time.perf_counter()
    with pt.Context(0) as ctx:
        pairs = []
        for i in range(a.first, a.first + a.scenes):
            if a.first <= 101 < a.first + a.scenes and i in (100, 101):
                continue  # held out for the gain tests
            text, depth = training_scene(i)
            s_in = spp_in[i % len(spp_in)]
            s_ref = min(max(a.spp_ref, a.ref_factor * s_in), max(a.spp_ref, a.spp_ref_max))
            feat, target = make_pair(ctx, text, depth, a.size, a.size, s_in, s_ref, chunk=a.chunk)
            pairs.append((feat, target))
            if len(pairs) % 50 == 1 or a.scenes <= 100:
                print(f"pair {i}: spp {s_in} -> {s_ref} noisy rel {rel_rmse(feat[..., :3], target):.3f} "
                      f"({time.perf_counter() - t0:.0f} s)", flush=True)
        with pt.Denoiser(ctx, a.batch, a.crop, a.crop, seed=1) as dn:
            train_two_phase(dn, pairs, a.steps, a.batch, a.crop, log_every=500)
            np.save(a.out, dn.params())
    print(f"saved {a.out} ({time.perf_counter() - t0:.0f} s)")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
