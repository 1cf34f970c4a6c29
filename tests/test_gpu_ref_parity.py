"""Reference-mode parity on the GPU: the gfx950 wavefront tracer against the
oracle (oracle/ptref.c, FP32 restatement of src/shader.comp) on the
reference's 11 scenes, through the C ABI.

Bar (floating point): per-pixel |dXYZ| / |XYZ| <= 1e-3 on >= 99.9 % of the
pixels; pixels beyond it are reported as diverged paths. Both sides use the
same definitions of the GLSL built-ins (ptmath.h), so most pixels are
expected to be bit-identical."""
import json

import numpy as np
import pytest

import pathtracer_amd as pt
from oracle import ptref
from tests.conftest import SCENES

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3
MIN_FRACTION = 0.999


def oracle_render(text, W, H, spf, path_length, frames, shot=1):
    pc = pt.push_constants(W, H, spf=spf, path_length=path_length)
    ubo, pc, sdfs = pt.pack_json(text, shot, pc)
    sd = b"".join(bytes(s) for s in sdfs)
    tex = np.zeros((H, W, 4), np.float32)
    cnt = None
    for f in range(1, frames + 1):
        pc.frame = f * spf
        pc.currentSamples = f * spf
        c = ptref.dispatch(bytes(ubo), bytes(pc), sd, len(sdfs), tex)
        cnt = c if cnt is None else {k: cnt[k] + c[k] for k in c}
    return tex, cnt


def gpu_render(ctx, text, W, H, spf, path_length, frames, shot=1):
    scene = ctx.scene_from_json(text)
    r = pt.Renderer(ctx, scene, W, H, spf=spf, path_length=path_length, shot=shot, mode=pt.MODE_REF)
    ctx.reset_counters()
    for _ in range(frames):
        r.render_frame()
    out = r.acc.read()
    return out, ctx.counters()


def compare(gpu, ref):
    g, r = gpu[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    exact = np.all(gpu[..., :3].view(np.uint32) == ref[..., :3].view(np.uint32), axis=-1)
    denom = np.maximum(np.linalg.norm(r, axis=-1), 1e-12)
    rel = np.linalg.norm(g - r, axis=-1) / denom
    rel = np.where(np.linalg.norm(r, axis=-1) == 0, np.linalg.norm(g, axis=-1), rel)
    within = rel <= REL_TOL
    return exact.mean(), within.mean(), int((~within).sum())


@pytest.mark.parametrize("path", SCENES, ids=lambda p: p.stem)
def test_ref_parity_scene(gpu_ctx, path):
    text = path.read_text()
    W, H, spf, L, frames = 96, 64, 2, 6, 2
    ref, rc = oracle_render(text, W, H, spf, L, frames)
    got, gc = gpu_render(gpu_ctx, text, W, H, spf, L, frames)
    exact, frac, diverged = compare(got, ref)
    print(f"\n{path.stem}: bit-exact {exact:.5f}, within {REL_TOL:g}: {frac:.5f}, diverged pixels {diverged}; "
          f"segments gpu {gc['segments']} oracle {rc['segments']}, shadow {gc['shadow_segments']}/{rc['shadow_segments']}, "
          f"sdf steps {gc['sdf_steps']}/{rc['sdf_steps']}")
    assert frac >= MIN_FRACTION
    assert np.all(got[..., 3] == 1.0)
    assert gc["paths"] == rc["paths"] == W * H * spf * frames
    assert gc["dispatches"] == frames
    for k in ("segments", "shadow_segments", "sdf_steps"):
        assert abs(gc[k] - rc[k]) <= 1e-3 * max(rc[k], 1), k


def test_ref_wave_chunking_is_invisible(scene_texts):
    """Tiny wave capacity: pixels and samples are split over many waves; the
    per-pixel sample order is kept, so the result is bit-identical."""
    text = scene_texts["scene2"]
    W, H, spf = 40, 30, 3
    with pt.Context(0, wave_paths=1024) as small, pt.Context(0) as big:
        a, _ = gpu_render(small, text, W, H, spf, 5, 2)
        b, _ = gpu_render(big, text, W, H, spf, 5, 2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_ref_persistence_blend(gpu_ctx, scene_texts):
    """currentSamples == spf with frame > spf selects the exponential blend
    (src/shader.comp:1500-1502); the oracle restates it."""
    text = scene_texts["scene1"]
    W, H, spf = 48, 32, 2
    pc = pt.push_constants(W, H, spf=spf, path_length=4, persistence=0.25, fps=30.0)
    ubo, pc, sdfs = pt.pack_json(text, 1, pc)
    tex = np.full((H, W, 4), 0.5, np.float32)
    pc.frame, pc.currentSamples = 3 * spf, spf
    ptref.dispatch(bytes(ubo), bytes(pc), b"", 0, tex)
    scene = gpu_ctx.scene_from_json(text)
    acc = gpu_ctx.accum(W, H)
    acc.write(np.full((H, W, 4), 0.5, np.float32))
    gpu_ctx.dispatch(scene, acc, pc, pt.MODE_REF)
    got = acc.read()
    _, frac, _ = compare(got, tex)
    assert frac >= MIN_FRACTION


def test_ref_from_packed_equals_from_json(gpu_ctx, scene_texts):
    text = scene_texts["scene6"]
    W, H = 32, 24
    pc = pt.push_constants(W, H, spf=1, path_length=4)
    ubo, pc, sdfs = pt.pack_json(text, 1, pc)
    pc.frame = pc.currentSamples = 1
    s1 = gpu_ctx.scene_from_json(text)
    s2 = gpu_ctx.scene_from_packed(ubo, sdfs)
    a1, a2 = gpu_ctx.accum(W, H), gpu_ctx.accum(W, H)
    gpu_ctx.dispatch(s1, a1, pc)
    gpu_ctx.dispatch(s2, a2, pc)
    assert np.array_equal(a1.read().view(np.uint32), a2.read().view(np.uint32))


def test_ref_rejects_bad_dispatch(gpu_ctx, scene_texts):
    scene = gpu_ctx.scene_from_json(scene_texts["scene0"])
    acc = gpu_ctx.accum(16, 16)
    pc = pt.push_constants(16, 8, spf=1)
    pc.frame = pc.currentSamples = 1
    with pytest.raises(pt.PTError) as e:
        gpu_ctx.dispatch(scene, acc, pc)
    assert e.value.status == pt.PT_ERR_INVALID_ARG
    s = json.loads(scene_texts["scene0"])
    s["sphere"] = [dict(s["sphere"][0]) for _ in range(300)]
    big = gpu_ctx.scene_from_json(json.dumps(s))  # loads (extended mode can hold it)
    pc = pt.push_constants(16, 16, spf=1)
    pc.frame = pc.currentSamples = 1
    with pytest.raises(pt.PTError) as e:
        gpu_ctx.dispatch(big, acc, pc, pt.MODE_REF)
    assert e.value.status == pt.PT_ERR_CAPACITY


def test_scene_to_json_matches_host_rewrite(gpu_ctx, scene_texts):
    """pt_scene_write_json on a loaded scene writes what pt_json_rewrite writes
    for the same text; a scene built from a packed buffer has no JSON model."""
    for text in list(scene_texts.values())[:4]:
        sc = gpu_ctx.scene_from_json(text)
        assert sc.to_json() == pt.json_rewrite(text)
    ubo, _, sdfs = pt.pack_json(list(scene_texts.values())[0])
    packed = gpu_ctx.scene_from_packed(ubo, sdfs)
    with pytest.raises(pt.PTError):
        packed.to_json()
