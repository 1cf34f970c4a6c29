"""The U-Net denoiser (pt_dn_*): forward pass and gradients of the GPU's BF16
MFMA kernels against the FP64 CPU restatement (oracle/denoise_ref.py), which is
itself pinned by finite differences; training reduces the loss."""
import numpy as np
import pytest
import torch

from oracle import denoise_ref as D


def random_features(rng, n, h, w):
    f = np.empty((n, h, w, 12), np.float32)
    f[..., 0:3] = np.exp(rng.normal(-1.0, 1.0, (n, h, w, 3)))          # XYZ, heavy-ish tails
    f[..., 3:6] = rng.uniform(0.0, 1.0, (n, h, w, 3))                   # albedo
    nv = rng.normal(size=(n, h, w, 3))
    f[..., 6:9] = nv / np.linalg.norm(nv, axis=-1, keepdims=True)       # normal
    f[..., 9] = rng.uniform(0.5, 5.0, (n, h, w))                        # distance
    f[..., 10] = f[..., 1] * rng.uniform(0.05, 1.0, (n, h, w))           # standard error of Y
    f[..., 11] = rng.integers(1, 256, (n, h, w))                        # samples
    return f


def random_params(rng):
    out = np.zeros(D.param_count(), np.float64)
    for (c1, c2, co, _), (name, sl) in zip([s for s in D.SPEC for _ in (0, 1)], D.layer_slices()):
        if name.startswith("w"):
            out[sl] = rng.normal(0.0, np.sqrt(2.0 / (9 * (c1 + c2))), sl.stop - sl.start)
        else:
            out[sl] = rng.normal(0.0, 0.05, sl.stop - sl.start)
    return out


@pytest.mark.parametrize("kind", ["l1log", "l2"])
def test_oracle_gradient_matches_finite_differences(kind):
    """Central differences of the FP64 loss on a tiny input pin the oracle's
    gradient (and so its layer wiring) independently of autograd."""
    rng = np.random.default_rng(0)
    p = random_params(rng)
    feat = random_features(rng, 1, 8, 8)
    target = np.exp(rng.normal(-1.0, 1.0, (1, 8, 8, 3)))
    loss, g = D.loss_and_grad(p, feat, target, kind)
    h = 1e-6
    checked = 0
    for name, sl in D.layer_slices():
        blk = np.abs(g[sl])
        if blk.max() < 1e-6:
            continue
        i = sl.start + int(np.argmax(blk))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (D.loss_and_grad(pp, feat, target, kind)[0] - D.loss_and_grad(pm, feat, target, kind)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 * abs(g[i]) + 1e-9, (name, fd, g[i])
        checked += 1
    assert checked >= 18


def test_oracle_layout_is_the_library_layout():
    assert D.param_count() == 479888
    # weight [t][co][ci] with tap t at offset (t // 3 - 1, t % 3 - 1): a single
    # unit weight at tap 5 (dy 0, dx +1), co 0, ci 0 shifts channel 0 left by one
    p = np.zeros(D.param_count())
    w0, _ = D.layer_slices()[0]
    p[D.layer_slices()[0][1].start + (5 * 32 + 0) * 16 + 0] = 1.0
    x = torch.zeros(1, 16, 4, 4, dtype=torch.float64)
    x[0, 0] = torch.arange(16, dtype=torch.float64).reshape(4, 4)
    w, b = D.unpack(p)[0]
    y = torch.nn.functional.conv2d(x, w, b, padding=1)[0, 0]
    assert torch.equal(y[:, :3], x[0, 0][:, 1:]) and torch.all(y[:, 3] == 0)


@pytest.mark.gpu
def test_forward_matches_the_oracle(gpu_ctx):
    import pathtracer_amd as pt
    rng = np.random.default_rng(1)
    with pt.Denoiser(gpu_ctx, 2, 32, 32, seed=3) as dn:
        assert dn.n == D.param_count()
        p = random_params(rng).astype(np.float32)
        dn.set_params(p)
        assert np.array_equal(dn.params(), p)
        feat = random_features(rng, 2, 32, 32)
        got = dn.denoise(feat)
    want = D.denoise(p.astype(np.float64), feat.astype(np.float64))
    err = np.abs(np.log1p(got) - np.log1p(want))
    print(f"\nforward: log-space error max {err.max():.4f} mean {err.mean():.5f}")
    # BF16 activations through 11 layers: a few parts in a thousand
    assert err.mean() < 0.01 and np.quantile(err, 0.999) < 0.05


@pytest.mark.gpu
def test_block_tiling_is_invisible(gpu_ctx):
    """A conv block takes several consecutive pixel tiles when the image is large
    (about two rounds of the blocks the CUs hold at once) and one tile when it
    is small. The interior of a 1024x1024 image (3 tiles per block in the
    full-resolution layers, a last block with fewer) and the interior of a
    256x256 crop of it, farther from the crop's edge than the network's
    receptive field (about 30 pixels), are bit-identical."""
    import pathtracer_amd as pt
    rng = np.random.default_rng(7)
    p = random_params(rng).astype(np.float32)
    feat = random_features(rng, 1, 1024, 1024)
    y0, x0, m = 384, 256, 64  # crop origin (multiples of 4: the same pooling grid) and margin
    with pt.Denoiser(gpu_ctx, 1, 1024, 1024, seed=3) as dn:
        dn.set_params(p)
        full = dn.denoise(feat)
        crop = dn.denoise(feat[:, y0:y0 + 256, x0:x0 + 256])
    assert np.isfinite(full).all()
    assert np.array_equal(crop[:, m:-m, m:-m], full[:, y0 + m:y0 + 256 - m, x0 + m:x0 + 256 - m])


def grad_agreement(g_gpu, g_ref):
    rows = []
    for name, sl in D.layer_slices():
        a, b = g_gpu[sl].astype(np.float64), g_ref[sl]
        nb = np.linalg.norm(b)
        cos = float(a @ b / (np.linalg.norm(a) * nb + 1e-300))
        rel = float(np.linalg.norm(a - b) / (nb + 1e-300))
        rows.append((name, nb, cos, rel))
    return rows


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["l1log", "l2"])
def test_gradients_match_the_oracle(gpu_ctx, kind):
    import pathtracer_amd as pt
    rng = np.random.default_rng(2)
    with pt.Denoiser(gpu_ctx, 2, 32, 32, seed=4) as dn:
        dn.set_loss(pt.DN_LOSS_L2 if kind == "l2" else pt.DN_LOSS_L1_LOG)
        p = random_params(rng).astype(np.float32)
        dn.set_params(p)
        feat = random_features(rng, 2, 32, 32)
        target = np.exp(rng.normal(-1.0, 1.0, (2, 32, 32, 3))).astype(np.float32)
        loss, g = dn.gradients(feat, target)
    ref_loss, ref_g = D.loss_and_grad(p.astype(np.float64), feat.astype(np.float64), target.astype(np.float64), kind)
    rows = grad_agreement(g, ref_g)
    for name, nb, cos, rel in rows:
        print(f"{name:4s} |g| {nb:.3e} cos {cos:.5f} rel {rel:.4f}")
    assert abs(loss - ref_loss) <= 0.01 * ref_loss
    whole = float(g.astype(np.float64) @ ref_g / (np.linalg.norm(g) * np.linalg.norm(ref_g)))
    assert whole > 0.995
    # every block whose gradient is not negligible points the same way
    big = max(r[1] for r in rows)
    for name, nb, cos, rel in rows:
        if nb > 1e-3 * big:
            assert cos > 0.98 and rel < 0.25, name


@pytest.mark.gpu
def test_a_planted_mistake_in_the_oracle_is_caught(gpu_ctx):
    """The parity check has power: an oracle whose skip concatenation is in the
    other order disagrees with the GPU by far more than the tolerance."""
    import pathtracer_amd as pt
    rng = np.random.default_rng(5)
    with pt.Denoiser(gpu_ctx, 1, 16, 16, seed=5) as dn:
        p = random_params(rng).astype(np.float32)
        dn.set_params(p)
        feat = random_features(rng, 1, 16, 16)
        got = dn.denoise(feat)
    orig = D.forward

    def swapped(params, f):
        L = D.unpack(params if isinstance(params, torch.Tensor) else torch.tensor(np.asarray(params, np.float64)))
        Fn = torch.nn.functional

        def conv(i, x):
            w, b = L[i]
            y = Fn.conv2d(x, w, b, padding=1)
            return torch.relu(y) if D.SPEC[i][3] else y
        x = D.to_input(torch.tensor(np.asarray(f, np.float64)))
        e0 = conv(1, conv(0, x))
        e1 = conv(3, conv(2, Fn.avg_pool2d(e0, 2)))
        bb = conv(5, conv(4, Fn.avg_pool2d(e1, 2)))
        up1 = Fn.interpolate(bb, scale_factor=2, mode="nearest")
        d1 = conv(7, conv(6, torch.cat([up1[:, :64], e1, up1[:, 64:]], 1)))  # wrong order
        d0 = conv(9, conv(8, torch.cat([Fn.interpolate(d1, scale_factor=2, mode="nearest"), e0], 1)))
        return x[:, 0:3] + conv(10, d0)[:, 0:3]
    try:
        D.forward = swapped
        wrong = D.denoise(p.astype(np.float64), feat.astype(np.float64))
    finally:
        D.forward = orig
    right = D.denoise(p.astype(np.float64), feat.astype(np.float64))
    e_right = np.abs(np.log1p(got) - np.log1p(right)).mean()
    e_wrong = np.abs(np.log1p(got) - np.log1p(wrong)).mean()
    assert e_wrong > 10 * e_right


@pytest.mark.gpu
def test_training_reduces_the_loss(gpu_ctx):
    import pathtracer_amd as pt
    rng = np.random.default_rng(6)
    feat = random_features(rng, 4, 32, 32)
    clean = feat[..., 0:3] * np.exp(rng.normal(0.0, 0.5, (4, 32, 32, 3))).astype(np.float32)
    with pt.Denoiser(gpu_ctx, 4, 32, 32, seed=7) as dn:
        first = dn.train_step(feat, clean, 1e-3)
        for _ in range(150):
            last = dn.train_step(feat, clean, 1e-3)
    print(f"\nloss {first:.4f} -> {last:.4f}")
    assert last < 0.7 * first
