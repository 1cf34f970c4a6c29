"""Autotuner for the extended tracer's launch and data-structure parameters.

For one workload (scene JSON, resolution, path length) it measures traced
segments per second of short renders under candidate pt_config values and
returns the fastest configuration:

  bvh_width x bvh_leaf   BVH arity (2, 4, 8) and primitives per leaf (1, 2, 4)
  traversal              persistent lanes refilled from the queue (1), or one ray per lane (2)
  wave_paths             paths in flight per wave (queue capacity); each
                         candidate is measured on dispatches that fill its wave
  blocks_per_cu          grid shape of the queue kernels (blocks of 256 lanes per CU)
  march_blocks_per_cu    grid shape of the persistent SDF march kernels (SDF scenes)
  sdf_max_steps          the smallest sphere-tracing budget at which at most
                         `sdf_tolerance` of the traced rays (extension and
                         shadow) have a march run out of steps (an exhausted
                         march is a miss, so the budget is a quality bound
                         first, a speed knob second)

The search is coordinate descent in that order, each parameter tried with the
others at their best so far. A configuration is measured once: the incumbent
is not re-timed for every coordinate, so timing noise cannot make the same
configuration both win and lose, and the choice is always the fastest
configuration measured. Every candidate renders the same samples, so
the image does not depend on the choice except through sdf_max_steps.
Measurement is wall time around pt_dispatch + pt_sync after one warm-up
dispatch, median of `repeats`.

    python -m pathtracer_amd.autotune --spheres 4096 --width 1920 --height 1080
"""
from __future__ import annotations

import argparse
import json
import statistics
import time
from dataclasses import asdict, dataclass, field

import pathtracer_amd as pt
from pathtracer_amd import scenes

WIDTHS = (2, 4, 8)
LEAVES = (1, 2, 4)
WAVES = (1 << 24, 1 << 25, 1 << 26, 1 << 27)
BLOCKS_PER_CU = (4, 8, 16)
MARCH_BLOCKS_PER_CU = (2, 4, 8, 16)
SDF_BUDGETS = (32, 48, 64, 96, 128, 192, 256)


@dataclass
class Choice:
    bvh_width: int = 2
    bvh_leaf: int = 1
    traversal: int = 1
    wave_paths: int = 1 << 27
    blocks_per_cu: int = 8
    march_blocks_per_cu: int = 8
    sdf_max_steps: int = 256

    def context(self, **over) -> pt.Context:
        kw = asdict(self)
        kw.update(over)
        return pt.Context(0, **kw)


@dataclass
class Trial:
    params: dict
    segments_per_s: float
    sdf_exhausted_frac: float = 0.0


@dataclass
class Result:
    best: Choice
    trials: list[Trial] = field(default_factory=list)

    def as_dict(self) -> dict:
        return {"best": asdict(self.best), "trials": [asdict(t) for t in self.trials]}


def measure(choice: Choice, text: str, w: int, h: int, depth: int, spp: int, repeats: int) -> Trial:
    """Segments per second of `repeats` dispatches of spp samples per pixel (one warm-up)."""
    with choice.context() as ctx:
        sc = ctx.scene_from_json(text)
        acc = ctx.accum(w, h)
        pc = pt.push_constants(w, h, spp, depth)
        sc.camera(1, pc)
        times, segs, marches, exhausted = [], 0, 0, 0
        for k in range(repeats + 1):
            pc.frame = (k + 1) * spp
            pc.currentSamples = (k + 1) * spp
            ctx.sync()
            c0 = ctx.counters()
            t0 = time.perf_counter()
            ctx.dispatch(sc, acc, pc, pt.MODE_EXT)
            ctx.sync()
            dt = time.perf_counter() - t0
            c1 = ctx.counters()
            if k == 0:
                continue  # warm-up: BVH build, wave allocation, code loading
            times.append(dt)
            segs = c1["segments"] - c0["segments"]
            marches += c1["segments"] - c0["segments"] + c1["shadow_segments"] - c0["shadow_segments"]
            exhausted += c1["sdf_exhausted"] - c0["sdf_exhausted"]
        acc.release()
        sc.release()
    return Trial(asdict(choice), segs / statistics.median(times), exhausted / max(marches, 1))


def autotune(text: str, width: int, height: int, depth: int, spp: int = 0, repeats: int = 3,
             sdf_tolerance: float = 1e-4, log=None) -> Result:
    """Coordinate descent over the parameters above for this workload; spp
    defaults to the samples per pixel that fill each candidate's wave."""
    has_sdf = '"sdf"' in text and '"kind"' in text
    res = Result(Choice())
    seen: dict[str, Trial] = {}

    def run(c: Choice) -> Trial:
        key = json.dumps(asdict(c), sort_keys=True)
        if key in seen:
            return seen[key]
        t = measure(c, text, width, height, depth, spp if spp > 0 else max(1, c.wave_paths // (width * height)),
                    repeats)
        seen[key] = t
        res.trials.append(t)
        if log:
            log(f"{json.dumps(t.params)} -> {t.segments_per_s / 1e9:.3f} G segments/s"
                + (f", {t.sdf_exhausted_frac:.2e} marches exhausted" if has_sdf else ""))
        return t

    # SDF budget first: the smallest that keeps exhausted marches within tolerance
    if has_sdf:
        for b in SDF_BUDGETS:
            c = Choice(**{**asdict(res.best), "sdf_max_steps": b})
            t = run(c)
            if t.sdf_exhausted_frac <= sdf_tolerance or b == SDF_BUDGETS[-1]:
                res.best = c
                break
    best_rate = 0.0

    def descend(name: str, values) -> None:
        nonlocal best_rate
        for v in values:
            c = Choice(**{**asdict(res.best), name: v})
            t = run(c)
            if t.segments_per_s > best_rate:
                best_rate, res.best = t.segments_per_s, c

    descend("bvh_width", WIDTHS)
    descend("bvh_leaf", LEAVES)
    descend("traversal", (1, 2))
    descend("wave_paths", WAVES)
    descend("blocks_per_cu", BLOCKS_PER_CU)
    if has_sdf:
        descend("march_blocks_per_cu", MARCH_BLOCKS_PER_CU)
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Autotune the extended tracer for one workload.")
    ap.add_argument("--scene", type=str, default="", help="scene JSON file (default: the random-sphere box)")
    ap.add_argument("--spheres", type=int, default=4096)
    ap.add_argument("--sdf", action="store_true", help="the mandelbulb + menger scene")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--depth", type=int, default=12)
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--out", type=str, default="")
    a = ap.parse_args(argv)
    if a.scene:
        with open(a.scene) as f:
            text = f.read()
    elif a.sdf:
        text = scenes.sdf_scene()
    else:
        text = scenes.random_spheres(n=a.spheres)
    r = autotune(text, a.width, a.height, a.depth, a.spp, a.repeats, log=lambda s: print(s, flush=True))
    out = r.as_dict()
    chosen = [t.segments_per_s for t in r.trials if t.params == out["best"]]
    print(json.dumps({"best": out["best"], "segments_per_s": chosen[0] if chosen else None}))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
