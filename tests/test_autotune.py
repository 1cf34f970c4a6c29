"""The autotuner's search (pathtracer_amd/autotune.py) on a synthetic cost
model: coordinate descent finds the fastest configuration, and the SDF step
budget is the smallest within the exhausted-march tolerance."""
import pathtracer_amd.autotune as A


def fake_measure(optimum, exhausted_at):
    calls = []

    def measure(choice, text, w, h, depth, spp, repeats):
        p = A.asdict(choice)
        calls.append(p)
        # separable cost: the rate falls off with each parameter's distance from its optimum
        rate = 1e9
        for k, v in optimum.items():
            rate /= 1.0 + 0.1 * abs(v - p[k]) / max(v, 1)
        return A.Trial(p, rate, exhausted_at.get(p["sdf_max_steps"], 0.0))
    return measure, calls


def test_descent_finds_the_separable_optimum(monkeypatch):
    opt = {"bvh_width": 4, "bvh_leaf": 2, "traversal": 2, "wave_paths": 1 << 25, "blocks_per_cu": 16}
    m, calls = fake_measure(opt, {})
    monkeypatch.setattr(A, "measure", m)
    r = A.autotune('{"sphere": []}', 64, 64, 8)
    for k, v in opt.items():
        assert getattr(r.best, k) == v, k
    assert r.best.sdf_max_steps == 256  # no SDF in the scene: budget untouched
    # each configuration is measured once: the incumbent of every coordinate
    # after the first is not timed again
    assert len({tuple(sorted(c.items())) for c in calls}) == len(calls)
    assert len(calls) == len(A.WIDTHS) + (len(A.LEAVES) - 1) + 1 + (len(A.WAVES) - 1) + (len(A.BLOCKS_PER_CU) - 1)
    assert max(r.trials, key=lambda t: t.segments_per_s).params == A.asdict(r.best)


def test_sdf_budget_is_the_smallest_within_tolerance(monkeypatch):
    exhausted = {32: 0.2, 48: 0.05, 64: 3e-3, 96: 5e-5, 128: 1e-5, 192: 0.0, 256: 0.0}
    m, calls = fake_measure({"bvh_width": 2}, exhausted)
    monkeypatch.setattr(A, "measure", m)
    r = A.autotune('{"sdf": [{"kind": "menger"}]}', 64, 64, 8, sdf_tolerance=1e-4)
    assert r.best.sdf_max_steps == 96
    assert [c["sdf_max_steps"] for c in calls[:4]] == [32, 48, 64, 96]
    # the rest of the descent keeps the chosen budget
    assert all(c["sdf_max_steps"] == 96 for c in calls[4:])
