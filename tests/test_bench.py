"""bench.py's contract on the CPU: one GPU only, the workloads' shapes, and the
roofline tables naming kernels and counters that exist, so that a renamed
launch cannot silently drop the roofline of the dominant kernel."""
import json
import re
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402
import pathtracer_amd as pt  # noqa: E402


def args(*argv):
    old = sys.argv
    sys.argv = ["bench.py", *argv]
    try:
        return bench.parse_args()
    finally:
        sys.argv = old


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("workload,shape", [
    ("spheres", (1920, 1080, 12, 64)),
    ("sdf", (1920, 1080, 8, 64)),
    ("cornell", (512, 512, 8, 64)),
    ("ref", (512, 512, 8, 8)),
])
def test_workload_shapes(workload, shape):
    text, w, h, depth, spp, what = bench.workload(args("--workload", workload))
    assert (w, h, depth, spp) == shape
    assert text.lstrip().startswith("{") and f"{w}x{h}" in what


def test_default_run_is_the_whole_1024_spp_render():
    a = args()
    _, w, h, depth, spp, _ = bench.workload(a)
    assert a.workload == "spheres" and a.gpus == 1
    assert a.steps * spp == 1024


def launch_names():
    src = "".join(p.read_text() for p in (ROOT / "pathtracer_amd" / "csrc").glob("*.hip"))
    return set(re.findall(r'launch(?:_march<[^>]*>)?\(\s*ctx,\s*"(k_\w+)"', src))


def test_roofline_tables_name_launched_kernels():
    names = launch_names()
    assert "k_ext_intersect" in names and "k_ext_march_mandelbulb" in names
    for k in list(bench.BYTES_PER_UNIT) + list(bench.FLOPS_COUNTER):
        assert k in names, k


def test_roofline_units_are_counters():
    fields = {f for f, _ in pt.Counters._fields_}
    for terms in bench.BYTES_PER_UNIT.values():
        for bytes_per_unit, unit in terms:
            assert bytes_per_unit > 0 and unit in fields, unit
    for counter in bench.FLOPS_COUNTER.values():
        assert counter in fields, counter


def test_measured_peaks_read_from_the_committed_profile():
    mp = bench.measured_peaks()
    assert mp is not None, "profiles/peaks_copy_fma_mi355x.txt is missing its bench line"
    assert 1000.0 < mp["GB/s"] < bench.HBM_PEAK_GBS
    assert 10.0 < mp["TFLOP/s"] < bench.FP32_PEAK_TFLOPS


def test_peak_probe_flag():
    assert args("--peak-probe").peak_probe and not args().peak_probe


def test_roofline_of_an_hbm_kernel():
    """The headline's dominant kernel priced in algorithmic bytes against the
    published and the measured HBM peaks, traffic from its committed profile."""
    name = "k_ext_intersect"
    (bpu, unit), = bench.BYTES_PER_UNIT[name]
    prof = {name: (12, 36.0), "k_ext_shade_lambert": (12, 10.0)}
    pcnt = {unit: 10**9}
    r = bench.roofline_of(prof, pcnt, "spheres", "spheres4096_1920x1080")
    achieved = bpu * 1e9 / (36.0e-3) / 1e9
    assert r["bound"] == "hbm" and r["kernel"] == name and r["unit"] == "GB/s"
    assert r["achieved"] == round(achieved, 2) and r["frac"] == round(achieved / bench.HBM_PEAK_GBS, 5)
    mp = bench.measured_peaks()
    assert r["peak_measured"] == mp["GB/s"] and r["frac_measured"] == round(achieved / mp["GB/s"], 5)
    assert r["traffic"] == bench.traffic_from_profiles(name, "spheres4096_1920x1080") and r["traffic"] > 0
    assert r["avg_launch_ms"] == 3.0 and r["algorithmic_bytes_per_launch"] == round(bpu * 1e9 / 12)


def test_roofline_of_an_sdf_march():
    name = "k_ext_march_mandelbulb"
    counter = bench.FLOPS_COUNTER[name]
    r = bench.roofline_of({name: (8, 16.0)}, {counter: 2 * 10**11}, "sdf", "sdf_1920x1080")
    achieved = 2e11 / 16e-3 / 1e12
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s" and r["achieved"] == round(achieved, 3)
    assert r["peak"] == bench.FP32_PEAK_TFLOPS and r["frac_measured"] == round(achieved / bench.measured_peaks()["TFLOP/s"], 5)
    json.dumps(r)  # the bench line serialises it


def test_roofline_of_nothing_profiled():
    assert bench.roofline_of({}, {}, "spheres", "x") is None
