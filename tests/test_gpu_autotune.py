"""The autotuner on the GPU: the configuration it picks renders the same image
as the default one (BVH shape, traversal, wave and grid sizes change the order
of work only), and on an SDF scene it picks the smallest step budget whose
exhausted-march fraction is within tolerance."""
import numpy as np
import pytest

import pathtracer_amd as pt
from pathtracer_amd import autotune as A
from pathtracer_amd import scenes

pytestmark = pytest.mark.gpu


def render(choice, text, w, h, spp, depth):
    with choice.context() as ctx:
        sc = ctx.scene_from_json(text)
        r = pt.Renderer(ctx, sc, w, h, spf=spp, path_length=depth, mode=pt.MODE_EXT)
        r.render_frame()
        img = r.acc.read().copy()
        r.acc.release()
        sc.release()
    return img


def test_tuned_configuration_renders_the_same_image():
    text = scenes.random_spheres(n=512, scene_seed=7, wavelengths=4, seed=2)
    res = A.autotune(text, 96, 64, 8, spp=64, repeats=1)
    rates = [t.segments_per_s for t in res.trials]
    assert max(rates) > 0
    # every configuration timed once, and the choice is the fastest of them
    assert len({tuple(sorted(t.params.items())) for t in res.trials}) == len(res.trials)
    assert max(res.trials, key=lambda t: t.segments_per_s).params == A.asdict(res.best)
    print(f"\nbest {A.asdict(res.best)}: {max(rates) / 1e9:.2f} G segments/s over {len(res.trials)} trials")
    ref = render(A.Choice(), text, 96, 64, 16, 8)
    for c in (res.best, A.Choice(blocks_per_cu=4, wave_paths=1 << 20), A.Choice(blocks_per_cu=16)):
        assert np.array_equal(render(c, text, 96, 64, 16, 8), ref)


def test_sdf_budget_choice():
    text = scenes.sdf_scene(seed=1)
    res = A.autotune(text, 64, 48, 6, spp=16, repeats=1, sdf_tolerance=1e-4)
    fr = {t.params["sdf_max_steps"]: t.sdf_exhausted_frac for t in res.trials[:len(A.SDF_BUDGETS)]
          if t.params["bvh_width"] == 2 and t.params["bvh_leaf"] == 1 and t.params["traversal"] == 1}
    print(f"\nexhausted fraction by budget: {fr}; chosen {res.best.sdf_max_steps}")
    b = res.best.sdf_max_steps
    assert fr[b] <= 1e-4 or b == A.SDF_BUDGETS[-1]
    smaller = [x for x in A.SDF_BUDGETS if x < b]
    assert all(fr[x] > 1e-4 for x in smaller)
    # exhausted fractions fall as the budget grows
    vals = [fr[x] for x in sorted(fr)]
    assert all(a >= b2 for a, b2 in zip(vals, vals[1:]))
