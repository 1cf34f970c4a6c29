"""The host BVH builder (pathtracer_amd/csrc/bvh.cpp) checked on the CPU by
tests/bvh_check.cpp: leaves cover every primitive reference once, every slot's
box holds its subtree (exact float containment), empty slots are inverted,
leaves respect max_leaf, a traversal that pushes every hit child but the
nearest stays within max_stack, and that traversal finds the same closest
sphere as testing every sphere. Sizes include the benchmark's 4096 spheres at
every width and leaf size the autotuner tries, and degenerate layouts."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "pathtracer_amd" / "csrc"


def build(tmp: Path, bvh_cpp: Path, flags=()) -> Path:
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp / "bvh_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", *flags, "-I", str(CSRC), str(ROOT / "tests" / "bvh_check.cpp"),
                    str(bvh_cpp), "-o", str(exe)], check=True, timeout=180)
    return exe


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    return build(tmp_path_factory.mktemp("bvh"), CSRC / "bvh.cpp")


def run(exe, n, width, max_leaf, layout, seed):
    r = subprocess.run([str(exe), str(n), str(width), str(max_leaf), str(layout), str(seed)],
                       capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("max_leaf", [1, 2, 4, 8])
def test_random_spheres_4096(checker, width, max_leaf):
    rc, out = run(checker, 4096, width, max_leaf, 0, 100 * width + max_leaf)
    assert rc == 0, out


@pytest.mark.parametrize("n", [0, 1, 2, 3, 17, 64, 65])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_small_inputs(checker, n, width):
    rc, out = run(checker, n, width, 1, 0, n + 7 * width)
    assert rc == 0, out


@pytest.mark.parametrize("layout", [1, 2, 3], ids=["coincident", "collinear", "clusters"])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_degenerate_layouts(checker, layout, width):
    rc, out = run(checker, 700, width, 2, layout, layout * 10 + width)
    assert rc == 0, out


@pytest.mark.parametrize("find,replace,message", [
    # the stack bound without the pushed siblings
    ("out.max_stack = std::max(out.max_stack, stack_here + 1);", "out.max_stack = std::max(out.max_stack, it.level);",
     "stack above max_stack"),
    # slot boxes written with the low corner for the high one
    ("&out.words[base + (size_t)((3 + k) * W) + j], &c.box.hi[k], 4);",
     "&out.words[base + (size_t)((3 + k) * W) + j], &c.box.lo[k], 4);", "misses"),
    # leaf ranges shifted by one primitive
    ("                child = c.first;", "                child = c.first + 1;", "leaf box misses a primitive"),
])
def test_checker_catches_planted_mistakes(tmp_path, find, replace, message):
    src = (CSRC / "bvh.cpp").read_text()
    planted = src.replace(find, replace)
    assert planted != src
    (tmp_path / "bvh.cpp").write_text(planted)
    exe = build(tmp_path, tmp_path / "bvh.cpp")
    rc, out = run(exe, 4096, 8, 4, 0, 1)
    assert rc == 1 and message in out, out


UBSAN = ("-fsanitize=undefined,float-cast-overflow", "-fno-sanitize-recover=all", "-g")


@pytest.mark.parametrize("width", [2, 4])
def test_overflowing_box_is_binned_in_range(tmp_path, width):
    """A primitive whose box overflows to +inf (a legal float centre 3e38 with
    radius 1e38) makes the centroid bounds infinite and the SAH bin offset NaN.
    The builder clamps the bin index instead of converting NaN to int (undefined
    behaviour, INT_MIN on x86, then an out-of-range bin write). Scene upload
    rejects such primitives before building; the builder stays defined anyway."""
    exe = build(tmp_path, CSRC / "bvh.cpp", UBSAN)
    rc, out = run(exe, 300, width, 1, 4, 5 + width)
    assert rc == 0 and "non-finite box" in out, out


def test_overflowing_box_fails_without_the_clamp(tmp_path):
    """The sanitized check above fails when the bin index is the plain cast."""
    src = (CSRC / "bvh.cpp").read_text()
    find = """        const float f = (c - lo) * scale;
        if (!(f > 0.0f)) return 0;"""
    planted = src.replace(find, """        const float f = (c - lo) * scale;
        if ((int)f < 0) return 0;""")
    assert planted != src
    (tmp_path / "bvh.cpp").write_text(planted)
    exe = build(tmp_path, tmp_path / "bvh.cpp", UBSAN)
    rc, out = run(exe, 300, 2, 1, 4, 7)
    assert rc != 0, out
