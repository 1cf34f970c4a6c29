"""INTEGRATION.md's C++ binding compiles against include/pathtracer.h, so the
code a maintainer pastes into the reference host matches the ABI's names and
signatures. The reference's PushConstantValues (src/pathtracer.cpp:197-218) is
aliased to pt_push_constants, whose layout tests/test_capi.py checks."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def cpp_blocks(md: str) -> list:
    return re.findall(r"```cpp\n(.*?)```", md, re.S)


def test_binding_compiles_against_the_header(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    blocks = cpp_blocks((ROOT / "INTEGRATION.md").read_text())
    assert blocks, "no C++ block in INTEGRATION.md"
    src = tmp_path / "binding.cpp"
    # the first block is file-level code; the others are statements, compiled as
    # the body of a function given what they take from the caller
    body = "".join("static void example_%d(pt_ctx* ctx, const std::string& json, void (*load_weights)(std::vector<float>&)) {\n"
                   "%s\n}\n" % (i, b) for i, b in enumerate(blocks[1:]))
    src.write_text("#include <cstring>\n#include <stdexcept>\n#include <string>\n#include <vector>\n"
                   '#include "pathtracer.h"\nusing PushConstantValues = pt_push_constants;\n'
                   + blocks[0] + body)
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
