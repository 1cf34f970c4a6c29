"""The C++ host renderer (pathtracer_amd/host/pt_render.cpp) on the CPU: its
argument checks and file errors, which come before any device call."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "pathtracer_amd" / "host" / "pt_render"

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="pt_render not built (make -C pathtracer_amd/host)")


def run(*argv):
    return subprocess.run([str(EXE), *argv], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("argv", [
    [],
    ["--spp", "16"],
    ["scene.json", "--mode", "fast"],
    ["scene.json", "--spp", "0"],
    ["scene.json", "--spf", "-1"],
    ["scene.json", "--width", "0"],
    ["scene.json", "--height"],
    ["scene.json", "--unknown", "1"],
    ["a.json", "b.json"],
])
def test_bad_arguments_print_usage(argv):
    r = run(*argv)
    assert r.returncode == 2 and r.stderr.startswith("usage: pt_render"), (r.returncode, r.stderr)


def test_missing_scene_file(tmp_path):
    r = run(str(tmp_path / "nope.json"), "--spp", "4")
    assert r.returncode == 1 and "cannot open scene" in r.stderr
