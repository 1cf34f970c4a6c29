"""FP64 CPU restatement of the denoiser network of pathtracer_amd/csrc/denoise.hip.

Test infrastructure only: tests/ use it to check the GPU forward pass and
gradients; nothing in the product path imports it. The parameter vector has
the library's layout (pt_dn_get_params): per layer, weights [9][Cout][Cin]
with tap t at offset (t // 3 - 1, t % 3 - 1), then biases [Cout].

Pinned by finite differences on a tiny input (tests/test_denoise.py) besides
autograd's own correctness.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as Fn

# (input channels, concatenated skip channels, output channels, ReLU)
SPEC = [(16, 0, 32, True), (32, 0, 32, True), (32, 0, 64, True), (64, 0, 64, True), (64, 0, 128, True),
        (128, 0, 128, True), (128, 64, 64, True), (64, 0, 64, True), (64, 32, 32, True), (32, 0, 32, True),
        (32, 0, 16, False)]


def param_count() -> int:
    return sum(9 * co * (c1 + c2) + co for c1, c2, co, _ in SPEC)


def unpack(params: np.ndarray | torch.Tensor):
    """-> [(weight [Cout, Cin, 3, 3], bias [Cout])] as float64 tensors (views of one leaf)."""
    p = params if isinstance(params, torch.Tensor) else torch.tensor(np.asarray(params, np.float64))
    out, off = [], 0
    for c1, c2, co, _ in SPEC:
        ci = c1 + c2
        w = p[off:off + 9 * co * ci].reshape(9, co, ci).permute(1, 2, 0).reshape(co, ci, 3, 3)
        off += 9 * co * ci
        b = p[off:off + co]
        off += co
        out.append((w, b))
    assert off == p.numel()
    return out


def to_input(feat: torch.Tensor) -> torch.Tensor:
    """features (N, H, W, 12) -> network input (N, 16, H, W): log1p(XYZ),
    albedo, normal, 1 / (1 + distance), log1p(standard error of Y),
    1 / sqrt(samples), zeros."""
    n, h, w, _ = feat.shape
    x = torch.zeros(n, 16, h, w, dtype=torch.float64)
    f = feat.permute(0, 3, 1, 2)
    x[:, 0:3] = torch.log1p(torch.clamp(f[:, 0:3], min=0.0))
    x[:, 3:9] = f[:, 3:9]
    x[:, 9] = 1.0 / (1.0 + torch.clamp(f[:, 9], min=0.0))
    x[:, 10] = torch.log1p(torch.clamp(f[:, 10], min=0.0))
    x[:, 11] = 1.0 / torch.sqrt(torch.clamp(f[:, 11], min=1.0))
    return x


def forward(params, feat: np.ndarray | torch.Tensor) -> torch.Tensor:
    """-> z = log1p(input XYZ) + correction, (N, 3, H, W)."""
    f = feat if isinstance(feat, torch.Tensor) else torch.tensor(np.asarray(feat, np.float64))
    L = unpack(params)

    def conv(i, x):
        w, b = L[i]
        y = Fn.conv2d(x, w, b, padding=1)
        return torch.relu(y) if SPEC[i][3] else y

    x = to_input(f)
    e0 = conv(1, conv(0, x))
    e1 = conv(3, conv(2, Fn.avg_pool2d(e0, 2)))
    bb = conv(5, conv(4, Fn.avg_pool2d(e1, 2)))
    d1 = conv(7, conv(6, torch.cat([Fn.interpolate(bb, scale_factor=2, mode="nearest"), e1], 1)))
    d0 = conv(9, conv(8, torch.cat([Fn.interpolate(d1, scale_factor=2, mode="nearest"), e0], 1)))
    out = conv(10, d0)
    return x[:, 0:3] + out[:, 0:3]


def denoise(params, feat) -> np.ndarray:
    z = forward(params, feat)
    return torch.clamp(torch.expm1(z), min=0.0).permute(0, 2, 3, 1).numpy()


def loss_and_grad(params: np.ndarray, feat: np.ndarray, target: np.ndarray, kind: str = "l1log"):
    """The loss and its gradient w.r.t. the parameter vector: mean |z - log1p(target)|
    ("l1log") or mean (expm1(z) - target)^2 ("l2")."""
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    z = forward(p, feat)
    t = torch.clamp(torch.tensor(np.asarray(target, np.float64)), min=0.0).permute(0, 3, 1, 2)
    if kind == "l2":
        loss = ((torch.expm1(z) - t) ** 2).mean()
    else:
        loss = (z - torch.log1p(t)).abs().mean()
    loss.backward()
    return loss.item(), p.grad.numpy()


def layer_slices():
    """(name, slice) of every weight and bias block in the parameter vector."""
    out, off = [], 0
    for i, (c1, c2, co, _) in enumerate(SPEC):
        n = 9 * co * (c1 + c2)
        out.append((f"w{i}", slice(off, off + n)))
        off += n
        out.append((f"b{i}", slice(off, off + co)))
        off += co
    return out
