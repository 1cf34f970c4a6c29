"""Mutation fuzzing of the scene JSON reader and writer, the scene loader,
packer and writer, and the SDF recogniser, on the CPU under AddressSanitizer
and UndefinedBehaviorSanitizer (tests/json_fuzz.cpp, built with g++ from
pathtracer_amd/csrc/json.cpp and scene.cpp): no memory or undefined-behaviour
error on mutants of the shipped scenes and their SDF texts, every rejection
carries a message, and for every accepted mutant the written text (JSON
document and scene) is a fixed point of read-then-write."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def fuzzer(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path_factory.mktemp("fuzz") / "json_fuzz"
    flags = ["-O1", "-g", "-std=c++17", "-Wall", "-Werror", "-fno-sanitize-recover=all",
             "-fsanitize=address,undefined,float-cast-overflow"]
    r = subprocess.run([gxx, *flags, "-I", str(ROOT / "pathtracer_amd" / "csrc"), str(ROOT / "tests" / "json_fuzz.cpp"),
                        str(ROOT / "pathtracer_amd" / "csrc" / "json.cpp"), str(ROOT / "pathtracer_amd" / "csrc" / "scene.cpp"),
                        "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mutated_scenes(fuzzer, seed):
    scenes = sorted(str(p) for p in (ROOT / "tests" / "golden" / "scenes").glob("*.json"))
    r = subprocess.run([str(fuzzer), "15000", str(seed), *scenes], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr[-2000:]
    parsed = int(r.stdout.split("parsed=")[1].split()[0])
    assert parsed > 500  # the mutants exercise the writer, not only the error paths
