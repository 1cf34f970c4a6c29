"""The persistent kernels' queue shares (pathtracer_amd/csrc/wave_queue.h) run on
the CPU: tests/wave_queue_sim.cpp includes the same header with stand-ins for the
wave intrinsics, runs each lane of a grid of waves, and checks that every queue
entry reaches exactly one lane and that every wave leaves its loop, over grids
with fewer waves than segments, queues shorter than a wave, and both refill
rules the kernels use (intersect and shadow: 16 idle lanes or none busy; SDF
march: 8 idle lanes)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("wq") / "wave_queue_sim"
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I", str(ROOT / "pathtracer_amd" / "csrc"),
                    str(ROOT / "tests" / "wave_queue_sim.cpp"), "-o", str(exe)], check=True, timeout=120)
    return exe


def run(sim, blocks, threads, n, refill, all_idle, max_work, seed):
    r = subprocess.run([str(sim), str(blocks), str(threads), str(n), str(refill), str(int(all_idle)),
                        str(max_work), str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok"), r.stdout


GRIDS = [(1, 64), (1, 256), (3, 256), (8, 256), (40, 256), (9, 128)]
SIZES = [0, 1, 5, 63, 64, 65, 1000, 4097]
RULES = [(16, True), (8, False)]


@pytest.mark.parametrize("blocks,threads", GRIDS)
@pytest.mark.parametrize("refill,all_idle", RULES)
def test_every_entry_once_and_every_wave_leaves(sim, blocks, threads, refill, all_idle):
    for i, n in enumerate(SIZES):
        run(sim, blocks, threads, n, refill, all_idle, max_work=7, seed=1000 * blocks + i)


@pytest.mark.parametrize("max_work", [0, 1, 40])
def test_skipped_and_long_entries(sim, max_work):
    """0 steps: every entry is a ray the march skips; 40: long traversals."""
    for n in (0, 7, 20000):
        run(sim, 5, 256, n, 8, False, max_work, seed=n + max_work)


def test_large_queue_many_waves(sim):
    """More entries per segment than kChunkMax per wave, 1024 waves."""
    run(sim, 256, 256, 300007, 16, True, 3, seed=7)


@pytest.mark.parametrize("find,replace,grid,message", [
    # the cursor advances one entry less than it hands out
    ("atomicAdd(c.cur, c.chunk)", "atomicAdd(c.cur, c.chunk - 1)", ("4", "256"), "handed out twice"),
    # all kSeg segments even when the grid has fewer waves: some have no home wave
    ("const int nseg = min(kSeg, waves);", "const int nseg = kSeg;", ("2", "256"), "never handed out"),
])
def test_simulator_catches_planted_mistakes(sim, tmp_path, find, replace, grid, message):
    """A copy of the header with a planted mistake must fail the check."""
    src = (ROOT / "pathtracer_amd" / "csrc" / "wave_queue.h").read_text()
    planted = src.replace(find, replace)
    assert planted != src
    (tmp_path / "wave_queue.h").write_text(planted)
    exe = tmp_path / "sim_planted"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(tmp_path), str(ROOT / "tests" / "wave_queue_sim.cpp"),
                    "-o", str(exe)], check=True, timeout=120)
    r = subprocess.run([str(exe), *grid, "5000", "16", "1", "5", "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and message in r.stdout, r.stdout
