"""The C-ABI library loads and exports every entry point include/pathtracer.h
declares; context-free calls work without a device; structs have the
reference's byte layouts."""
import ctypes as C
import re
import subprocess

import pytest

import pathtracer_amd as pt
from tests.conftest import ROOT

HEADER = (ROOT / "include" / "pathtracer.h").read_text()
DECLARED = sorted(set(re.findall(r"^\s*(?:pt_status|void|double|int64_t|const char\*)\s+(pt_\w+)\s*\(", HEADER, re.M)))


def test_header_declarations_found():
    assert "pt_dispatch" in DECLARED and "pt_create" in DECLARED and len(DECLARED) >= 25


@pytest.mark.parametrize("name", DECLARED)
def test_symbol_exported(name):
    assert hasattr(pt.lib, name)
    assert name in pt.SIGNATURES, "binding table misses a declared entry point"


def test_exported_symbols_are_exactly_declared():
    out = subprocess.run(["nm", "-D", "--defined-only", str(pt.LIB_PATH)], capture_output=True, text=True).stdout
    exported = sorted({l.split()[-1] for l in out.splitlines() if re.search(r" T pt_\w+$", l)})
    assert exported == DECLARED


def test_struct_layouts():
    assert C.sizeof(pt.PackedScene) == 16388  # sizeof(UniformBufferObject)
    assert C.sizeof(pt.PushConstants) == 88   # sizeof(PushConstantValues)
    assert pt.PushConstants.cameraAngle.offset == 32
    assert pt.PushConstants.tonemap.offset == 84
    assert pt.PackedScene.cie.offset == 4 * (7 + 1024 + 768 + 783 + 128 + 64)
    # serving structs (no reference counterpart): C layouts on x86-64
    assert C.sizeof(pt.ServerConfig) == 24
    assert C.sizeof(pt.RequestDesc) == 40 and pt.RequestDesc.target_rel.offset == 32
    assert C.sizeof(pt.RequestStatus) == 88 and pt.RequestStatus.rel_error.offset == 24


def test_argument_validation_without_device():
    st = pt.lib.pt_create(None, 0, None)
    assert st == pt.PT_ERR_INVALID_ARG
    assert pt.lib.pt_dispatch(None, None, None, None, 0) == pt.PT_ERR_INVALID_ARG
    assert pt.lib.pt_sdf_recognize(None, 0, None) == pt.PT_ERR_INVALID_ARG
    assert pt.lib.pt_pack_json(None, 0, 1, None, None, None, 0, None) == pt.PT_ERR_INVALID_ARG
    assert pt.lib.pt_peak_probe(None, 0, 1024, 1, None) == pt.PT_ERR_INVALID_ARG


def test_create_reports_missing_device_or_succeeds():
    h = C.c_void_p()
    st = pt.lib.pt_create(C.byref(h), 0, None)
    if st == pt.PT_OK:
        pt.lib.pt_destroy(h)
    else:
        assert st == pt.PT_ERR_NO_DEVICE
        assert b"no HIP device" in pt.lib.pt_last_error_global()


def test_struct_layouts_match_a_c_compiler(tmp_path):
    """sizeof/offsetof from gcc on the header itself agree with the ctypes mirror."""
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stddef.h>
#include <stdio.h>
#include "pathtracer.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",
           sizeof(pt_packed_scene), sizeof(pt_push_constants), sizeof(pt_sdf_desc), sizeof(pt_config),
           sizeof(pt_counters), sizeof(pt_kernel_stat), sizeof(pt_server_config), sizeof(pt_request_desc),
           sizeof(pt_request_status), offsetof(pt_request_status, rel_error));
    return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(pt.PackedScene), C.sizeof(pt.PushConstants), C.sizeof(pt.SdfDesc), C.sizeof(pt.Config),
            C.sizeof(pt.Counters), C.sizeof(pt.KernelStat), C.sizeof(pt.ServerConfig), C.sizeof(pt.RequestDesc),
            C.sizeof(pt.RequestStatus), pt.RequestStatus.rel_error.offset]
    assert got == want
