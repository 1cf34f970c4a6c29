"""Benchmark of the extended spectral wavefront tracer on one MI355X.

Workloads (--workload):
  spheres (default)  Cornell box plus 4096 random spheres (60 % Lambert, 25 %
                     dispersive dielectric, 15 % conductor), 1920x1080, path
                     depth 12, 4 hero wavelengths, 64 spp per step (one wave
                     of 133 M paths; 16 steps are the 1024-spp render).
  ref                the reference's own scene file scene1.json (3 spheres, a
                     plane, 2 rotated boxes, the BK7 camera lens) in reference
                     mode: the shader's primitives, sequential PCG chain, 4
                     rotated wavelengths and quirks, bit-identical to the FP32
                     CPU restatement of the shader, which is timed beside it;
                     512x512, path length 8, 8 spp per step.
  cornell            the Cornell box alone (5 quads, a Lambert and a Cauchy
                     dielectric sphere, 6500 K quad light), 512x512, depth 8,
                     one wavelength per path, 64 spp per step: a step is the
                     whole 64-spp render.
  sdf                power-8 mandelbulb (8 iterations) and level-4 menger
                     sponge over a ground plane under a uniform white sky,
                     sphere tracing of at most 256 steps to 1e-4, 1920x1080,
                     depth 8, 4 hero wavelengths, 64 spp per step.
A step is one dispatch tracing `--spp` samples of every pixel (one wave of
paths); the timed region is exactly `--steps` dispatches bracketed by stream
synchronisation, with the scene, BVH and path buffers resident in HBM.

Prints one JSON line: metric, value (spectral path-segments/s), roofline of
the dominant kernel (algorithmic bytes, or for a sphere-tracing kernel FP32
operations, over its average HIP-event duration), and the CPU baseline (the
oracle timed on this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "spectral path-segments/s; p50/p99 time-to-RMSE≤1% per render request"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

# Algorithmic HBM bytes per unit of work of each kernel: the SoA records the
# kernel must read and write once (path state, hit, queue entries), excluding
# the BVH and scene tables that stay in L2. A kernel's bytes per launch are the
# sum over its terms of (bytes per unit x units traced in the timed region)
# divided by its launches.
#   extension ray (segment), in k_ext_trace (which also classifies):
#     queue index 4 + ray origin/direction 24 + hit (t, primitive) 8
#     + primitive material/light lookup 8 + material-queue entry (path, t, primitive) 12 = 56 B
#   in k_ext_intersect (hit only; k_ext_classify appends the material queues):
#     extension ray (path index 4, origin/direction 24) + hit (t, primitive) 8 = 36 B
#     (the ray is read as two 16-byte records kept in queue order, 32 B, so the
#     kernel needs no path index, and the hit is written in queue order)
#   in k_ext_classify: queue index 4 + hit 8 + material/light lookup 8
#     + material-queue entry (path, t, primitive) 12                    = 32 B
#   shadow ray, in k_ext_trace / k_ext_shadow:
#     record (origin 12, direction 12, tmax 4, XYZ 12, path 4) 44 + path XYZ read+write 24 = 68 B
#   Lambert vertex, in k_ext_shade_lambert:
#     queue entry (path, hit) 12 + ray 24 + throughput 16 + wavelength 4 + key 4 + flags 4
#     + writes: ray 24, throughput 16, pdf 4, flags 4, shadow record 44, next-queue 4 = 160 B
#     (a continuing ray is also written into its next-queue slot, 32 B more, so the
#     intersect kernel reads it in queue order)
BYTES_PER_UNIT = {
    "k_ext_trace": [(56, "segments"), (68, "shadow_segments")],
    "k_ext_intersect": [(36, "segments")],
    "k_ext_classify": [(32, "segments")],
    "k_ext_shadow": [(68, "shadow_segments")],
    "k_ext_shade_lambert": [(160, "lambert_hits")],
    # reference mode (SoA scalar records):
    #   k_ref_intersect: queue index 4 + ray 24 + hit (t, normal, material, light) 24 = 52 B per segment
    #   k_ref_shade: reads queue 4, hit 24, ray 24, throughput 16, MIS weight 4, seed 4, hero
    #     wavelength 4; writes ray 24, throughput 16, weight 4, seed 4, next-queue 4 = 132 B per
    #     segment; + shadow record (path 4, ray 24, light object 4, contribution 16) 48 B per shadow ray
    #   k_ref_shadow: reads ray 24 + light object 4, writes hit distance 4 + object 4 = 36 B per shadow ray
    "k_ref_intersect": [(52, "segments")],
    "k_ref_shade": [(132, "segments"), (48, "shadow_segments")],
    "k_ref_shadow": [(36, "shadow_segments")],
}

# Sphere-tracing kernels are bound by FP32 VALU work, not HBM. Their algorithmic
# FP32 operations are counted on the device (pt_counters.march_flops_*,
# shadow_march_flops): per distance-estimator evaluation, the operations the
# estimator's source performs (one per add, multiply, divide, min, max, floor,
# sqrt, rsqrt, log or sin, two per fused multiply-add): mandelbulb 10 + 65 per
# iteration executed, menger 16 + 33 per level, blob 65, terrain 93; plus 10 per
# march step for the point, its offset and the step update. The shadow kernel
# of an SDF scene also traverses the analytic scene; only its marching counts.
FLOPS_COUNTER = {
    "k_ext_march_mandelbulb": "march_flops_mandelbulb",
    "k_ext_march_menger": "march_flops_menger",
    "k_ext_march_blob": "march_flops_blob",
    "k_ext_march_terrain": "march_flops_terrain",
    "k_ext_shadow": "shadow_march_flops",
}


def workload(args):
    """(scene JSON, width, height, depth, spp per step, description) of --workload."""
    from pathtracer_amd import scenes
    if args.workload == "ref":
        text = (ROOT / "tests" / "golden" / "scenes" / "scene1.json").read_text()
        dflt = (512, 512, 8, 8)
        what = ("the reference's scene1.json (3 spheres, plane, 2 rotated boxes, BK7 lens) in reference mode "
                "(the shader's method, 4 rotated wavelengths), {w}x{h}, path length {d}")
    elif args.workload == "cornell":
        text = scenes.cornell(wavelengths=1, seed=0)
        dflt = (512, 512, 8, 64)
        what = ("Cornell box (5 quads with 3-Gaussian spectra, Lambert sphere, Cauchy dielectric sphere "
                "A=1.5 B=0.01, 6500 K quad light), {w}x{h}, depth {d}, 1 wavelength per path in 380-780 nm")
    elif args.workload == "sdf":
        text = scenes.sdf_scene(seed=0)
        dflt = (1920, 1080, 8, 64)
        what = ("mandelbulb (power 8, 8 iterations, bailout 16) + level-4 menger sponge over a ground plane, "
                "uniform white environment, sphere tracing <= 256 steps to 1e-4, {w}x{h}, depth {d}, "
                "4 hero wavelengths")
    else:
        text = scenes.random_spheres(n=args.spheres, scene_seed=1, wavelengths=4, seed=0)
        dflt = (1920, 1080, 12, 64)
        what = ("Cornell box + " + str(args.spheres) + " random spheres (60% Lambert / 25% dielectric / "
                "15% conductor), {w}x{h}, depth {d}, 4 hero wavelengths")
    w = args.width or dflt[0]
    h = args.height or dflt[1]
    d = args.depth or dflt[2]
    spp = args.spp or dflt[3]
    return text, w, h, d, spp, what.format(w=w, h=h, d=d)


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16,
                    help="timed steps (16 steps of the default workload are its whole 1024-spp render)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("spheres", "cornell", "sdf", "ref"), default="spheres",
                    help="scene of the tracer bench (module docstring)")
    ap.add_argument("--spp", type=int, default=0, help="samples per pixel per step (default: the workload's)")
    ap.add_argument("--width", type=int, default=0)
    ap.add_argument("--height", type=int, default=0)
    ap.add_argument("--depth", type=int, default=0)
    ap.add_argument("--spheres", type=int, default=4096)
    ap.add_argument("--bvh-width", type=int, default=0)
    ap.add_argument("--bvh-leaf", type=int, default=0)
    ap.add_argument("--wave-paths", type=int, default=0)
    ap.add_argument("--traversal", type=int, default=0, help="0 default (persistent), 1 persistent lanes refilled from the queue, 2 one ray per lane")
    ap.add_argument("--march", type=int, default=0, help="SDF march: 0 persistent lanes refilled from the queue, 1 one ray per lane")
    ap.add_argument("--march-blocks", type=int, default=0, help="blocks per CU of the persistent march (0 = default)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-kernel event timing")
    ap.add_argument("--serve", type=int, default=0, metavar="N",
                    help="serving workload instead: requests 0..N-1 of the serving stream (256 = the full workload)")
    ap.add_argument("--serve-target", type=float, default=0.01, help="relative RMSE at which a request is done")
    ap.add_argument("--serve-load", type=float, default=0.85,
                    help="utilisation of the one-request-at-a-time tracer that sets the Poisson arrival rate")
    ap.add_argument("--serve-check", type=int, default=0, metavar="N",
                    help="accuracy instead: the requests among 0..N-1 of at most 512x512, served with the "
                         "denoiser, against independent plain renders of 4096-8192 samples per pixel")
    ap.add_argument("--serve-graphs", type=int, default=1,
                    help="1: batched waves replayed from captured hipGraphs; 0: kernels launched one by one")
    ap.add_argument("--serve-wave", type=int, default=0,
                    help="paths per batched wave of the server's context, both systems (0 = the library's 2^22)")
    ap.add_argument("--serve-max-active", type=int, default=0, help="requests the batching server traces together (0 = 16)")
    ap.add_argument("--serve-profile", action="store_true",
                    help="GPU time by kernel family in both systems (HIP events per launch, graphs off)")
    ap.add_argument("--serve-denoise", action="store_true",
                    help="the batching server may finish requests with the shipped denoiser's image")
    ap.add_argument("--serve-no-observe", action="store_true",
                    help="one-at-a-time p99 from the service times only (no real-time one-at-a-time run)")
    ap.add_argument("--serve-slice", type=int, default=24, metavar="N",
                    help="the default (spheres) bench line also carries p50/p99 time-to-RMSE of requests 0..N-1 "
                         "of the serving stream, one at a time and the batching server + denoiser at 1x and 3x arrivals (0 = off)")
    ap.add_argument("--serve-policy", choices=("srpt", "hrrn"), default="hrrn",
                    help="the served runs' ordering: highest response ratio (time in system + estimated "
                         "remaining GPU time) / remaining first, or shortest estimated remaining GPU time first")
    ap.add_argument("--serve-speedup", type=str, default="1",
                    help="comma-separated factors: the batching server's arrivals come this many times faster than the "
                         "one-at-a-time ones (one served run per factor)")
    ap.add_argument("--denoiser", action="store_true",
                    help="denoiser workload instead: forward passes at --width x --height and training steps")
    ap.add_argument("--denoiser-eval", action="store_true",
                    help="the shipped denoiser on held-out 1080p scenes against a plain --eval-ref-spp render")
    ap.add_argument("--eval-ref", type=int, default=16384, help="samples per pixel of the --denoiser-eval reference")
    ap.add_argument("--weights", type=str, default="", help="denoiser parameters (.npy) instead of the shipped ones")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run --steps device-to-device copies of MIB MiB (calibrates the HBM byte counters)")
    ap.add_argument("--peak-probe", action="store_true",
                    help="only measure the device peaks: streaming copy (GB/s) and unpacked FP32 FMA loop (TFLOP/s)")
    return ap.parse_args()


def copy_probe(mib: int, steps: int):
    """Plain device copies of a known size: FETCH_SIZE / WRITE_SIZE per dispatch
    of this kernel against mib MiB give the counters' scale factor."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    assert torch.equal(dst, src)
    print(json.dumps({"copy_bytes": n * 4, "steps": steps, "GB/s_read+write": 2 * n * 4 * steps / dt / 1e9}))


def peak_probe(steps: int):
    """Measured peaks beside the published ones: a streaming copy of 2 x 4 GiB
    and the FP32 FMA loop (8 independent chains per thread, no packed FP32),
    each launched `steps` times after a warm-up and timed with HIP events."""
    import pathtracer_amd as pt
    with pt.Context(0) as ctx:
        bw, nbytes, copy_ms = ctx.peak_probe(0, 1 << 30, max(1, steps))
        fl, nflop, fma_ms = ctx.peak_probe(1, 1 << 16, max(1, steps))
    print(json.dumps({
        "copy": {"kernel": "k_peak_copy", "bytes_per_launch": nbytes, "ms": round(copy_ms, 4),
                 "GB/s": round(bw / 1e9, 1), "published_GB/s": HBM_PEAK_GBS,
                 "frac_of_published": round(bw / 1e9 / HBM_PEAK_GBS, 4)},
        "fp32_fma": {"kernel": "k_peak_fma", "flops_per_launch": nflop, "ms": round(fma_ms, 4),
                     "TFLOP/s": round(fl / 1e12, 2), "published_TFLOP/s": FP32_PEAK_TFLOPS,
                     "frac_of_published": round(fl / 1e12 / FP32_PEAK_TFLOPS, 4)},
        "steps": steps}))


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from a committed rocprofv3
    --pmc summary (profiles/pmc_<kernel>_<workload>.txt)."""
    for f in sorted((ROOT / "profiles").glob(f"pmc_{kernel}_{workload}.txt")):
        for line in f.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                try:
                    return float(line.split(":", 1)[1].split()[0])
                except ValueError:
                    return None
    return None


def valu_issue_from_profiles(kernel: str, workload: str):
    """The kernel's VALU issue share (wave-wide VALU instructions over what the
    CUs can issue in its launches' time) from the committed rocprofv3 --pmc
    summary of the same workload, or None."""
    for f in sorted((ROOT / "profiles").glob(f"pmc_{kernel}_{workload}.txt")):
        for line in f.read_text().splitlines():
            if line.startswith("valu_issue_share:"):
                try:
                    return float(line.split(":", 1)[1].split()[0])
                except ValueError:
                    return None
    return None


def measured_peaks():
    """Device peaks measured by `bench.py --peak-probe` (streaming copy GB/s,
    unpacked FP32 FMA loop TFLOP/s) from the committed summary
    profiles/peaks_copy_fma_mi355x.txt, or None when it is absent."""
    f = ROOT / "profiles" / "peaks_copy_fma_mi355x.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith('{"copy"'):
            j = json.loads(line)
            return {"GB/s": j["copy"]["GB/s"], "TFLOP/s": j["fp32_fma"]["TFLOP/s"]}
    return None


def roofline_of(prof: dict, pcnt: dict, workload: str, tag: str):
    """The roofline object of the dominant kernel: prof maps kernel names to
    (launches, total ms) from the HIP events, pcnt holds the device counters of
    the same dispatches. SDF marches are priced in FP32 flops against the VALU
    peak, the other kernels in algorithmic bytes against HBM; both carry the
    measured peak beside the published one."""
    roofline = None
    mp = measured_peaks()
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1][1])
        name, (launches, total_ms) = dom
        terms = BYTES_PER_UNIT.get(name)
        fl = FLOPS_COUNTER.get(name) if workload == "sdf" else None
        if fl and launches > 0:
            flops = pcnt.get(fl, 0)
            achieved = flops / (total_ms * 1e-3) / 1e12
            roofline = {
                "bound": "valu",
                "achieved": round(achieved, 3),
                "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                "traffic": traffic_from_profiles(name, tag),
                # beside the published peak: the FMA loop measured on this device
                "peak_measured": mp and mp["TFLOP/s"],
                "frac_measured": mp and round(achieved / mp["TFLOP/s"], 5),
                # what bounds it: VALU issue (rocprofv3 counters of the same workload)
                "valu_issue_share": valu_issue_from_profiles(name, tag),
                "kernel": name,
                "avg_launch_ms": round(total_ms / launches, 4),
                "algorithmic_flops_per_launch": round(flops / launches),
            }
        elif terms and launches > 0:
            total_bytes = sum(bpu * pcnt.get(unit, 0) for bpu, unit in terms)
            achieved = total_bytes / (total_ms * 1e-3) / 1e9
            roofline = {
                "bound": "hbm",
                "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic_from_profiles(name, tag),
                # beside the published peak: the streaming copy measured on this device
                "peak_measured": mp and mp["GB/s"],
                "frac_measured": mp and round(achieved / mp["GB/s"], 5),
                # the divergent traversal kernels are bound by VALU issue and
                # dependent fetches, not bandwidth: their issue share from the
                # rocprofv3 counters of the same workload
                "valu_issue_share": valu_issue_from_profiles(name, tag),
                "kernel": name,
                "avg_launch_ms": round(total_ms / launches, 4),
                "algorithmic_bytes_per_launch": round(total_bytes / launches),
            }
    return roofline


def cpu_baseline(args, scene_text: str, W: int, H: int, depth: int, full_spp: int = 0):
    """The FP64 oracle tracer (CPU restatement with its own BVH) on a bounded
    sample of the same workload: the full frame at a number of samples per
    pixel chosen to take about --cpu-seconds, on this process's host cores
    (at most 16, the share of one GPU). With full_spp, the whole render of
    that many samples per pixel when it fits in about 3 x --cpu-seconds."""
    try:
        from oracle import ptext
    except Exception as e:  # oracle not built
        return {"value": None, "unit": "segments/s", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    # time of one sample per pixel over the frame, from 8 bands spread over its height
    band = max(1, H // 16)
    probe_s = 0.0
    for k in range(8):
        r0 = min(H - band, k * H // 8)
        probe_s += ptext.render(scene_text, W, H, spp=1, depth=depth, threads=threads, rows=(r0, r0 + band))["seconds"]
    frame_s = probe_s * H / (8 * band)
    spp = max(1, int(round(args.cpu_seconds / max(frame_s, 1e-9))))
    if full_spp and frame_s * full_spp <= 3.0 * args.cpu_seconds:
        spp = full_spp
    run = ptext.render(scene_text, W, H, spp=spp, depth=depth, threads=threads)
    return {
        "value": run["segments"] / run["seconds"],
        "unit": "segments/s",
        "cores": threads,
        # host CPUs visible to this process; a GPU box's share for one GPU is 16
        # threads, so the run uses at most 16 of them
        "nproc": len(os.sched_getaffinity(0)),
        "kind": "port",
        "sample": f"FP64 CPU oracle with its own BVH, the full {W}x{H} frame x {spp} spp, "
                  f"depth {depth}: {run['segments']} segments in {run['seconds']:.2f} s",
    }


def ref_cpu_baseline(args, ctx, text: str, W: int, H: int, depth: int):
    """The FP32 CPU restatement of the reference shader (oracle/ptref.c) on the
    same scene in reference mode, on at most 16 host threads: one dispatch of the
    number of samples per pixel that takes about --cpu-seconds. Also the
    per-pixel parity of a 2-spp frame between it and the GPU."""
    import numpy as np
    import pathtracer_amd as pt
    try:
        from oracle import ptref
    except Exception as e:  # oracle not built
        return {"value": None, "unit": "segments/s", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}
    threads = max(1, min(16, len(os.sched_getaffinity(0))))

    def packed(spf):
        pc = pt.push_constants(W, H, spf=spf, path_length=depth)
        ubo, pc, sdfs = pt.pack_json(text, 1, pc)
        pc.frame = pc.currentSamples = spf
        return ubo, pc, sdfs

    # parity: the same 2-spp frame on the GPU and the CPU
    ubo, pc, sdfs = packed(2)
    sc = ctx.scene_from_json(text)
    acc = ctx.accum(W, H)
    ctx.dispatch(sc, acc, pc, pt.MODE_REF)
    gpu = acc.read().copy()
    acc.release()
    sc.release()
    cpu = np.zeros((H, W, 4), np.float32)
    ptref.dispatch(bytes(ubo), bytes(pc), b"", 0, cpu, threads=threads)
    g, c = gpu[..., :3].astype(np.float64), cpu[..., :3].astype(np.float64)
    rel = np.linalg.norm(g - c, axis=-1) / np.maximum(np.linalg.norm(c, axis=-1), 1e-12)
    within = float(np.mean(rel <= 1e-3))
    exact = float(np.mean(np.all(gpu[..., :3] == cpu[..., :3], axis=-1)))
    # timing: one spp, then the sample
    ubo, pc, sdfs = packed(1)
    t0 = time.perf_counter()
    ptref.dispatch(bytes(ubo), bytes(pc), b"", 0, np.zeros((H, W, 4), np.float32), threads=threads)
    t1 = time.perf_counter() - t0
    spf = max(1, int(round(args.cpu_seconds / max(t1, 1e-9))))
    ubo, pc, sdfs = packed(spf)
    t0 = time.perf_counter()
    cnt = ptref.dispatch(bytes(ubo), bytes(pc), b"", 0, np.zeros((H, W, 4), np.float32), threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": cnt["segments"] / dt,
        "unit": "segments/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 CPU restatement of the reference shader (oracle/ptref.c), the full {W}x{H} frame x {spf} spp, "
                  f"path length {depth}: {cnt['segments']} segments in {dt:.2f} s; per-pixel parity with the GPU on a "
                  f"2-spp frame: {within * 100:.3f}% of pixels within 1e-3, {exact * 100:.3f}% bit-identical",
    }


def percentile(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def serving_measure(args, n_requests: int, speedups, denoise: bool, observe_fifo: bool = True, log=None):
    """Requests 0..n-1 of scenes.serving_request, served two ways.

    1. Each request's service time alone: all requests traced one at a time to
       completion (FIFO server, one active request, all queued at once).
    2. Poisson arrivals (seeded) at the rate that loads that one-at-a-time
       tracer to --serve-load.
    3. With observe_fifo, the one-at-a-time server run in real time on those
       arrivals: its observed response times (time from arrival to the
       relative-RMSE target or to the sample budget).
    4. The same arrival sequence, compressed `speedup` times, served in real
       time by the batching server (--serve-policy; continuous batching; with `denoise` requests
       may finish with the shipped denoiser's image).
    Returns the measurement as a dict."""
    import numpy as np
    import pathtracer_amd as pt
    from pathtracer_amd import scenes

    log = log or (lambda msg: print(msg, file=sys.stderr, flush=True))
    reqs = [scenes.serving_request(i) for i in range(n_requests)]
    ctx = pt.Context(0, graph_mode=0 if args.serve_graphs else 1, wave_paths=args.serve_wave)
    t_start = time.perf_counter()

    def kernel_split(prof):
        """GPU ms by kernel family from pt_profile_read totals (--serve-profile)."""
        fam = {"trace": 0.0, "denoise": 0.0, "moments+estimate": 0.0, "other": 0.0}
        for name, (_, ms) in prof.items():
            if name.startswith(("k_ext_moments", "k_ext_rel", "k_ext_features", "k_dn_assemble", "k_dn_cross",
                                "k_dn_scale")):
                fam["moments+estimate"] += ms
            elif name.startswith("k_ext_") or name.startswith("k_bump"):
                fam["trace"] += ms
            elif name.startswith(("k_conv", "k_pool", "k_up", "k_relu", "k_prep", "k_head", "k_dn", "k_wcopy")):
                fam["denoise"] += ms
            else:
                fam["other"] += ms
        out = {k: round(v, 1) for k, v in fam.items()}
        top = sorted(prof.items(), key=lambda kv: -kv[1][1])[:16]
        out["by_kernel"] = {name: [n, round(ms, 1)] for name, (n, ms) in top}
        return out

    # 1. service times alone
    service = {}
    if args.serve_profile:
        ctx.profile(True)
        ctx.profile_read(reset=True)
    with pt.Server(ctx, policy=pt.SERVE_FIFO, max_active=1) as srv:
        ids = {srv.submit(r["text"], r["width"], r["height"], r["spp"], r["depth"], args.serve_target,
                          t_arrive=0.0): r["id"] for r in reqs}
        while len(service) < len(reqs):
            sts = srv.step()
            if sts is None:
                break
            for st in sts:
                if st.state in (pt.REQ_DONE, pt.REQ_FAILED):
                    service[ids[st.id]] = st.as_dict()
                    srv.release(st.id)
                    if len(service) % 32 == 0:
                        log(f"one at a time: {len(service)}/{len(reqs)} done ({time.perf_counter() - t_start:.0f} s)")
    s_alone = np.array([service[r["id"]]["gpu_seconds"] for r in reqs])
    alone_kernels = kernel_split(ctx.profile_read(reset=True)) if args.serve_profile else None
    alone_converged = sum(service[r["id"]]["converged"] for r in reqs)

    # 2. arrivals
    rate = args.serve_load / s_alone.mean()
    gaps = np.random.default_rng(20261202).exponential(1.0 / rate, len(reqs))
    offsets = np.cumsum(gaps) - gaps[0]

    dn = None
    if denoise:
        from pathtracer_amd import denoise_train
        hmax = max(r["height"] for r in reqs)
        wmax = max(r["width"] for r in reqs)
        # room for a batch of 2: the error estimate's two half images in one pass
        dn = pt.Denoiser(ctx, 2, (hmax + 3) // 4 * 4, (wmax + 3) // 4 * 4)
        denoise_train.load_weights(dn)

    def realtime(policy, speedup, max_active, use_dn):
        """One server run in real time on the arrivals compressed `speedup` times."""
        done, segments = {}, 0
        if args.serve_profile:
            ctx.profile_read(reset=True)
        with pt.Server(ctx, policy=policy, max_active=max_active) as srv:
            if use_dn:
                srv.set_denoiser(dn)
            base = srv.now()
            arrive = base + offsets / speedup
            ids, nxt = {}, 0
            while len(done) < len(reqs):
                now = srv.now()
                while nxt < len(reqs) and arrive[nxt] <= now:
                    r = reqs[nxt]
                    ids[srv.submit(r["text"], r["width"], r["height"], r["spp"], r["depth"], args.serve_target,
                                   t_arrive=float(arrive[nxt]), denoise=use_dn)] = r["id"]
                    nxt += 1
                sts = srv.step()
                if sts is None:
                    if nxt < len(reqs):
                        time.sleep(max(0.0, min(0.002, arrive[nxt] - srv.now())))
                    continue
                for st in sts:
                    if st.state in (pt.REQ_DONE, pt.REQ_FAILED):
                        d = st.as_dict()
                        d["response"] = d["t_done"] - d["t_arrive"]
                        done[ids[st.id]] = d
                        segments += d["segments"]
                        srv.release(st.id)
            makespan = srv.now() - base
        conv = [r["id"] for r in reqs if done[r["id"]]["converged"]]
        resp = np.array([done[r["id"]]["response"] for r in reqs])
        resp_conv = np.array([done[i]["response"] for i in conv])
        return {"arrival_speedup": speedup, "arrival_rate_per_s": rate * speedup,
                "p50_time_to_rmse_s": percentile(resp_conv, 50), "p99_time_to_rmse_s": percentile(resp_conv, 99),
                "converged": len(conv), "p50_response_s": percentile(resp, 50),
                "p99_response_s": percentile(resp, 99), "makespan_s": makespan,
                "segments_per_s": segments / makespan,
                "denoised": int(sum(done[r["id"]]["denoised"] for r in reqs)),
                "mean_spp": float(np.mean([done[r["id"]]["spp"] for r in reqs])),
                "gpu_seconds": float(sum(done[r["id"]]["gpu_seconds"] for r in reqs)),
                "kernel_ms": kernel_split(ctx.profile_read(reset=True)) if args.serve_profile else None}

    # 3. one request at a time in real time, uncompressed arrivals
    fifo_obs = None
    if observe_fifo:
        fifo_obs = realtime(pt.SERVE_FIFO, 1.0, 1, False)
        log(f"one at a time, observed: p99 response {fifo_obs['p99_response_s']:.2f} s "
            f"({time.perf_counter() - t_start:.0f} s)")
    # the same from the service times (a FIFO queue is exact given service times)
    fifo_resp, c = [], 0.0
    for k in range(len(reqs)):
        c = max(offsets[k], c) + s_alone[k]
        fifo_resp.append(c - offsets[k])
    fifo_resp = np.array(fifo_resp)
    fifo_conv = np.array([fifo_resp[i] for i in range(len(reqs)) if service[reqs[i]["id"]]["converged"]])
    p99_fifo = fifo_obs["p99_response_s"] if fifo_obs else percentile(fifo_resp, 99)

    # 4. the batching server in real time, once per arrival speedup
    runs = []
    policy = pt.SERVE_HRRN if args.serve_policy == "hrrn" else pt.SERVE_SRPT
    for k in speedups:
        runs.append(realtime(policy, float(k), args.serve_max_active, dn is not None))
        log(f"{args.serve_policy.upper()} at {k}x arrivals: p99 response {runs[-1]['p99_response_s']:.2f} s, "
            f"converged {runs[-1]['converged']} ({time.perf_counter() - t_start:.0f} s)")
    if dn is not None:
        dn.close()
    ctx.close()
    # the largest arrival speedup at which the batching server's p99 response is no
    # worse than the one-at-a-time server's at 1x
    equal_p99 = max([r["arrival_speedup"] for r in runs if r["p99_response_s"] <= p99_fifo], default=None)
    one = {"p50_time_to_rmse_s": percentile(fifo_conv, 50), "p99_time_to_rmse_s": percentile(fifo_conv, 99),
           "converged": int(alone_converged),
           "p50_response_s": percentile(fifo_resp, 50), "p99_response_s": percentile(fifo_resp, 99),
           "mean_service_s": float(s_alone.mean()),
           "mean_spp": float(np.mean([service[r["id"]]["spp"] for r in reqs])),
           "gpu_seconds": float(s_alone.sum()), "kernel_ms": alone_kernels}
    if fifo_obs:
        one["observed"] = fifo_obs
    return {"requests": len(reqs), "target_rel": args.serve_target, "load": args.serve_load,
            "arrival_rate_per_s": rate, "srpt_denoise": dn is not None,
            "wave_paths": args.serve_wave or (1 << 22), "srpt_max_active": args.serve_max_active or 16,
            "served_policy": args.serve_policy,
            "p99_one_at_a_time_s": p99_fifo, "p99_basis": "observed" if fifo_obs else "from service times",
            "srpt": runs[-1], "srpt_runs": runs, "throughput_ratio_at_equal_p99": equal_p99,
            "one_at_a_time": one, "bench_seconds": time.perf_counter() - t_start}


SERVE_WORKLOAD = ("serving stream requests 0..{last}: 50% Cornell+N random spheres (N log-uniform 1..4096, "
                  "depth 12) / 50% mandelbulb+menger (parameters x U[0.9,1.1], depth 8); resolution "
                  "U{{256^2, 512^2, 1024x768, 1920x1080}}; spp budget log-uniform 16..1024; Poisson arrivals")


def serve_bench(args):
    """Serving workload (requests 0..--serve - 1): prints one JSON line whose
    value is the batching server's p50 time to the relative-RMSE target."""
    m = serving_measure(args, args.serve, args.serve_speedup.split(","), args.serve_denoise,
                        observe_fifo=not args.serve_no_observe)
    srpt = m["srpt"]
    line = {
        "metric": METRIC,
        "value": srpt["p50_time_to_rmse_s"],
        "unit": f"s (p50 time-to-RMSE<=1%, {args.serve_policy.upper()} server)",
        "n_gpus": 1,
        "higher_is_better": False,
        "data": "synthetic",
        "dtype": "fp32",
        "config": {"workload": SERVE_WORKLOAD.format(last=m["requests"] - 1),
                   **{k: m[k] for k in ("requests", "target_rel", "load", "arrival_rate_per_s", "srpt_denoise",
                                        "wave_paths", "srpt_max_active", "served_policy")}},
        **{k: m[k] for k in ("srpt", "srpt_runs", "throughput_ratio_at_equal_p99", "p99_one_at_a_time_s",
                             "p99_basis", "one_at_a_time", "bench_seconds")},
    }
    print(json.dumps(line))


# The U-Net's layers as built (input channels padded to 16, output to 16):
# (input channels, output channels, resolution divisor) per 3x3 convolution.
DN_LAYERS = [(16, 32, 1), (32, 32, 1), (32, 64, 4), (64, 64, 4), (64, 128, 16), (128, 128, 16), (192, 64, 4),
             (64, 64, 4), (96, 32, 1), (32, 32, 1), (32, 16, 1)]


def dn_forward_flops(n_pixels: int) -> float:
    """Multiply-add flops (x2) of one forward pass's convolutions over n_pixels full-resolution pixels."""
    return sum(2.0 * 9 * ci * co * n_pixels / div for ci, co, div in DN_LAYERS)


def denoiser_eval(args):
    """The learned denoiser at its benchmark size: held-out scenes of the
    serving families (one mandelbulb + menger scene, one random-sphere box),
    first-hit features of an spp-sample render at --width x --height
    (default 1920 x 1080) denoised by the shipped network, against a plain
    render of --eval-ref samples per pixel (default 16384) from disjoint
    samples. Prints the relative RMSE of the plain mean and of the denoised
    image at each input sample count, with the reference's own relative
    standard error (its noise floor)."""
    import numpy as np
    import pathtracer_amd as pt
    from pathtracer_amd import denoise_train as T, scenes

    W, H = ((args.width or 1920) + 3) // 4 * 4, ((args.height or 1080) + 3) // 4 * 4
    ref_spp = args.eval_ref
    spps = [4, 16, 64]
    held_out = [
        ("sdf", scenes.sdf_scene(jitter=0.1, scene_seed=2_000_001, wavelengths=4, seed=2_000_001), 8),
        ("spheres1000", scenes.random_spheres(n=1000, scene_seed=2_000_002, wavelengths=4, seed=2_000_002), 12),
    ]
    out = []
    t0 = time.perf_counter()
    with pt.Context(0) as ctx, pt.Denoiser(ctx, 1, H, W) as dn:
        if args.weights:
            T.load_weights(dn, args.weights)
        else:
            T.load_weights(dn)
        for name, text, depth in held_out:
            sc = ctx.scene_from_json(text)
            ctx.sync()
            tr = time.perf_counter()
            ref = T.render_mean(ctx, sc, W, H, ref_spp, depth, first=max(spps), chunk=16)
            ref_s = time.perf_counter() - tr
            # noise floor of the reference: its mean's standard error, from the
            # per-pixel spread of 64 samples scaled to ref_spp
            f64 = T.render_features(ctx, sc, W, H, 64, depth, first=max(spps) + ref_spp)
            se_y = f64[..., 10].astype(np.float64) * np.sqrt(64.0 / ref_spp)
            floor = float(np.sqrt(np.sum(se_y ** 2) * 3.0) / np.sqrt(np.sum(ref.astype(np.float64) ** 2)))
            row = {"scene": name, "depth": depth, "ref_spp": ref_spp, "ref_seconds": round(ref_s, 2),
                   "ref_rel_se_estimate": round(floor, 5)}
            for spp in spps:
                feat = T.render_features(ctx, sc, W, H, spp, depth)
                ctx.sync()
                td = time.perf_counter()
                den = T.denoise_image(dn, feat)
                dt = time.perf_counter() - td
                row[f"plain_{spp}spp"] = round(T.rel_rmse(feat[..., :3], ref), 5)
                row[f"denoised_{spp}spp"] = round(T.rel_rmse(den, ref), 5)
                row["denoise_seconds"] = round(dt, 4)
            sc.release()
            out.append(row)
            print(json.dumps(row), flush=True)
    print(json.dumps({
        "metric": "denoiser relative RMSE against a plain render",
        "config": {"workload": f"held-out serving-family scenes, {W}x{H}, inputs of 4/16/64 spp "
                               f"(XYZ, albedo, normal, depth, standard error of Y), reference {ref_spp} spp"},
        "results": out, "seconds": round(time.perf_counter() - t0, 1)}))


def denoiser_bench(args):
    """Denoiser: --steps forward passes of the shipped network over the
    first-hit features of a 4-sample render at --width x --height (batch 1),
    then --steps training steps on batches of 8 crops of 96 x 96. Per-kernel
    GPU time from HIP events on the library's stream; the dominant kernel's
    MFMA roofline is its convolution flops / its time against the dense BF16
    peak."""
    import numpy as np
    import pathtracer_amd as pt
    from pathtracer_amd import denoise_train as T, scenes

    W, H = ((args.width or 1920) + 3) // 4 * 4, ((args.height or 1080) + 3) // 4 * 4
    ctx = pt.Context(0)
    sc = ctx.scene_from_json(scenes.sdf_scene(seed=1))
    feat = T.render_features(ctx, sc, W, H, 4, 8)
    f, _, _ = T.normalise(feat)
    sc.release()
    with pt.Denoiser(ctx, 1, H, W) as dn:
        T.load_weights(dn)
        for _ in range(args.warmup):
            dn.denoise(f[None])
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dn.denoise(f[None])
        dt_host = (time.perf_counter() - t0) / args.steps
        ctx.profile(True)
        ctx.profile_read(reset=True)
        for _ in range(args.steps):
            dn.denoise(f[None])
        prof = ctx.profile_read(reset=True)
        ctx.profile(False)
    fwd_ms = sum(ms for _, ms in prof.values()) / args.steps
    conv_launches, conv_ms = prof["k_dn_conv"]
    conv_flops = dn_forward_flops(W * H)
    achieved = conv_flops * args.steps / (conv_ms * 1e-3) / 1e12
    # training: batch 8 of 96 x 96 crops, forward + backward-data + backward-weights
    rng = np.random.default_rng(0)
    B, C = 8, 96
    fb = np.stack([f[y:y + C, x:x + C] for y, x in zip(rng.integers(0, H - C, B), rng.integers(0, W - C, B))])
    tb = fb[..., 0:3].copy()
    with pt.Denoiser(ctx, B, C, C, seed=1) as dn:
        for _ in range(args.warmup):
            dn.train_step(fb, tb, 1e-4)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            dn.train_step(fb, tb, 1e-4)
        dt_train = (time.perf_counter() - t0) / args.steps
        ctx.profile(True)
        ctx.profile_read(reset=True)
        for _ in range(args.steps):
            dn.train_step(fb, tb, 1e-4)
        tprof = ctx.profile_read(reset=True)
        ctx.profile(False)
    train_ms = sum(ms for _, ms in tprof.values()) / args.steps
    wg_launches, wg_ms = tprof["k_dn_wgrad"]
    wg_flops = dn_forward_flops(B * C * C)  # one weight-gradient GEMM per layer, the forward's shape
    line = {
        "metric": "denoised frames/s (GPU time)",
        "value": 1e3 / fwd_ms,
        "unit": f"frames/s at {W}x{H}",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": fwd_ms,
        "higher_is_better": True, "dtype": "bf16 (fp32 accumulation)", "data": "synthetic",
        "config": {"workload": f"U-Net forward over the first-hit features of a 4-sample render of the "
                               f"mandelbulb + menger scene at {W}x{H}; training on 8 crops of 96x96"},
        "host_inclusive_ms": 1e3 * dt_host,
        "roofline": {"bound": "mfma", "kernel": "k_dn_conv (forward)", "achieved": achieved, "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": achieved / 2500.0,
                     "flops_per_pass": conv_flops, "conv_ms_per_pass": conv_ms / args.steps},
        "kernels_ms_per_pass": {k: ms / args.steps for k, (n, ms) in sorted(prof.items(), key=lambda kv: -kv[1][1])},
        "training": {"ms_per_step_gpu": train_ms, "ms_per_step_host": 1e3 * dt_train,
                     "wgrad_tflops": wg_flops * args.steps / (wg_ms * 1e-3) / 1e12,
                     "kernels_ms_per_step": {k: ms / args.steps
                                             for k, (n, ms) in sorted(tprof.items(), key=lambda kv: -kv[1][1])}},
    }
    print(json.dumps(line))
    ctx.close()


def serve_check(args):
    """Actual relative RMSE of served images. The requests among 0..N-1 of at
    most 512 x 512 are served together (SRPT, shipped denoiser, target
    --serve-target). Each one's reference is a plain render of the same scene
    with another render seed (independent samples), 8192 samples per pixel at
    256 x 256 and 4096 at 512 x 512, served plain to its budget; its own
    variance of the mean is taken off the squared difference:
    actual^2 = (|img - ref|^2 - sum var(ref)) / |ref|^2. Prints one JSON line."""
    import numpy as np
    import pathtracer_amd as pt
    from pathtracer_amd import denoise_train, scenes

    reqs = [r for r in (scenes.serving_request(i) for i in range(args.serve_check))
            if r["width"] * r["height"] <= 512 * 512]
    ctx = pt.Context(0)
    t0 = time.perf_counter()

    def run(srv, items):
        ids = {srv.submit(t, r["width"], r["height"], spp, r["depth"], target, denoise=dn): r["id"]
               for r, t, spp, target, dn in items}
        out = {}
        while len(out) < len(ids):
            sts = srv.step()
            if sts is None:
                break
            for st in sts:
                if st.state in (pt.REQ_DONE, pt.REQ_FAILED):
                    img = srv.read(st.id)[..., :3].astype(np.float64)
                    out[ids[st.id]] = (st.as_dict(), img)
                    srv.release(st.id)
        return out

    with pt.Denoiser(ctx, 1, 512, 512) as dn, pt.Server(ctx) as srv:
        denoise_train.load_weights(dn)
        srv.set_denoiser(dn)
        served = run(srv, [(r, r["text"], r["spp"], args.serve_target, True) for r in reqs])
    print(f"served {len(served)} requests ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    rows = []
    for r in reqs:
        d = json.loads(r["text"])
        d["render"]["seed"] = int(d["render"].get("seed", 0)) + 7919
        ref_spp = 8192 if r["width"] * r["height"] <= 256 * 256 else 4096
        with pt.Server(ctx) as srv:
            (st_ref, ref), = run(srv, [(r, json.dumps(d), ref_spp, 1e-9, False)]).values()
        st, img = served[r["id"]]
        n2 = float(np.sum(ref ** 2))
        var_ref = st_ref["rel_plain"] ** 2 * n2
        actual = float(np.sqrt(max(float(np.sum((img - ref) ** 2)) - var_ref, 0.0) / n2))
        rows.append({"id": r["id"], "kind": r["kind"], "res": f"{r['width']}x{r['height']}", "spp": st["spp"],
                     "spp_max": r["spp"], "denoised": st["denoised"], "converged": st["converged"],
                     "estimate": st["rel_error"], "actual": actual, "ref_rel": st_ref["rel_plain"]})
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    conv = [x for x in rows if x["converged"]]
    line = {
        "metric": "actual relative RMSE of served images against independent converged renders",
        "target_rel": args.serve_target, "requests": len(rows),
        "converged": len(conv),
        "converged_actual_within_target": sum(x["actual"] <= args.serve_target for x in conv),
        "converged_actual_within_1.1x_target": sum(x["actual"] <= 1.1 * args.serve_target for x in conv),
        "converged_actual_max": max((x["actual"] for x in conv), default=None),
        "median_actual_over_estimate": float(np.median([x["actual"] / x["estimate"] for x in rows if x["estimate"] > 0])),
        "rows": rows,
        "bench_seconds": time.perf_counter() - t0,
    }
    print(json.dumps(line))
    ctx.close()


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is sample-split with one accumulator reduce, not built)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe, args.steps)
        return
    if args.peak_probe:
        peak_probe(args.steps)
        return
    if args.serve_check:
        serve_check(args)
        return
    if args.serve:
        serve_bench(args)
        return
    if args.denoiser:
        denoiser_bench(args)
        return
    if args.denoiser_eval:
        denoiser_eval(args)
        return
    trace_bench(args)


def trace_bench(args):
    import pathtracer_amd as pt

    text, W, H, depth, spp, what = workload(args)
    # name of this workload's profile summaries under profiles/
    tag = {"spheres": f"spheres{args.spheres}_{W}x{H}", "sdf": f"sdf_{W}x{H}", "cornell": f"cornell_{W}x{H}",
           "ref": f"ref_scene1_{W}x{H}"}[args.workload]
    mode = pt.MODE_REF if args.workload == "ref" else pt.MODE_EXT
    ctx = pt.Context(0, wave_paths=args.wave_paths, bvh_width=args.bvh_width, traversal=args.traversal,
                     bvh_leaf=args.bvh_leaf, march=args.march, march_blocks_per_cu=args.march_blocks)
    scene = ctx.scene_from_json(text)
    r = pt.Renderer(ctx, scene, W, H, spf=spp, path_length=depth, mode=mode)
    for _ in range(args.warmup):
        r.render_frame()
    ctx.sync()
    # Timed region: exactly --steps dispatches, no instrumentation.
    ctx.reset_counters()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render_frame()
    ctx.sync()
    dt = time.perf_counter() - t0
    cnt = ctx.counters()
    seg_rate = cnt["segments"] / dt

    # The same --steps dispatches again with a HIP event pair around every
    # kernel launch (on the library's stream): per-kernel average durations.
    prof, pcnt, dt_prof = {}, cnt, None
    if not args.no_profile:
        ctx.reset_counters()
        ctx.profile(True)
        ctx.profile_read(reset=True)
        ctx.sync()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            r.render_frame()
        ctx.sync()
        dt_prof = time.perf_counter() - t1
        pcnt = ctx.counters()
        prof = ctx.profile_read(reset=True)
        ctx.profile(False)

    roofline = roofline_of(prof, pcnt, args.workload, tag)
    full = spp if args.workload == "cornell" else 0
    if args.no_cpu_baseline:
        cpu = None
    elif args.workload == "ref":
        cpu = ref_cpu_baseline(args, ctx, text, W, H, depth)
    else:
        cpu = cpu_baseline(args, text, W, H, depth, full_spp=full)
    config = {
        "workload": what,
        "spp_per_step": spp,
        "paths_per_step": W * H * spp,
        "segments": cnt["segments"],
        "shadow_segments": cnt["shadow_segments"],
        "segments_per_path": cnt["segments"] / max(1, cnt["paths"]),
    }
    if args.workload == "ref":
        pass
    elif args.workload == "sdf":
        config["sdf_steps_per_segment"] = cnt["sdf_steps"] / max(1, cnt["segments"] + cnt["shadow_segments"])
        config["sdf_exhausted"] = cnt["sdf_exhausted"]
        config["march_tflops"] = {k: round(pcnt[v] / 1e12, 4) for k, v in FLOPS_COUNTER.items() if pcnt.get(v)}
    else:
        config["bvh_nodes_per_segment"] = cnt["bvh_nodes"] / max(1, cnt["segments"])
        config["bvh_nodes_per_shadow_ray"] = cnt["shadow_bvh_nodes"] / max(1, cnt["shadow_segments"])
    if args.workload != "ref":
        # segments whose ray left the scene (camera rays outside the box's
        # opening among them): the rate over the segments that hit something
        config["escaped"] = cnt["escaped"]
        config["hit_segments_per_s"] = (cnt["segments"] - cnt["escaped"]) / dt
    line = {
        "metric": METRIC,
        "value": seg_rate,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if prof:
        line["profiled_ms_per_step"] = dt_prof * 1e3 / args.steps
        line["kernels_ms"] = {k: round(v[1], 3) for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    ctx.close()
    if args.workload == "spheres" and args.serve_slice > 0:
        line["serving"] = serving_slice(args)
        config["serving_slice"] = SERVE_WORKLOAD.format(last=args.serve_slice - 1) + \
            " (the first requests of the 256-request stream)"
    print(json.dumps(line))


def serving_slice(args):
    """The metric's second half on a bounded slice of the serving stream: p50 /
    p99 time to relative RMSE <= 1 % (over the requests that reach it) and
    response times, one request at a time (observed in real time) and with the
    batching server (--serve-policy) plus denoiser at 1x and 3x the arrival rate."""
    m = serving_measure(args, args.serve_slice, ["1", "3"], denoise=True, observe_fifo=True)
    one = m["one_at_a_time"]["observed"]
    pick = ("p50_time_to_rmse_s", "p99_time_to_rmse_s", "converged", "p50_response_s", "p99_response_s",
            "denoised", "mean_spp", "gpu_seconds")
    return {"requests": m["requests"], "target_rel": m["target_rel"], "load": m["load"],
            "arrival_rate_per_s": round(m["arrival_rate_per_s"], 4),
            "p50_time_to_rmse_s": m["srpt_runs"][0]["p50_time_to_rmse_s"],
            "p99_time_to_rmse_s": m["srpt_runs"][0]["p99_time_to_rmse_s"],
            "one_at_a_time": {k: one[k] for k in pick if k in one},
            "served_policy": args.serve_policy,
            "served_denoise_1x": {k: m["srpt_runs"][0][k] for k in pick},
            "served_denoise_3x": {k: m["srpt_runs"][1][k] for k in pick},
            "throughput_ratio_at_equal_p99": m["throughput_ratio_at_equal_p99"],
            "seconds": round(m["bench_seconds"], 1)}


if __name__ == "__main__":
    main()
