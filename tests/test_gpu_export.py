"""Zero-copy display: an exportable accumulator's texels, exported as a dma-buf
file descriptor and mapped back through HIP's virtual-memory import (the part a
Vulkan renderer's VkImportMemoryFdInfoKHR plays in the reference's display
path, src/shader.frag:81-99), read the same bytes as pt_accum_read_xyz, and an
exportable accumulator renders exactly what an ordinary one does."""
import ctypes as C
import os

import numpy as np
import pytest

import pathtracer_amd as pt
from pathtracer_amd import scenes

pytestmark = pytest.mark.gpu


class MemLocation(C.Structure):
    _fields_ = [("type", C.c_int), ("id", C.c_int)]


class MemAccessDesc(C.Structure):
    _fields_ = [("location", MemLocation), ("flags", C.c_int)]


HIP_MEM_HANDLE_TYPE_POSIX_FD = 1
HIP_MEM_LOCATION_TYPE_DEVICE = 1
HIP_MEM_ACCESS_READ_WRITE = 3
HIP_MEMCPY_DEVICE_TO_HOST = 2


def hip():
    # the HIP runtime libpathtracer.so is bound to (by its soname: torch ships
    # its own copy, and a second runtime instance would see no device)
    lib = C.CDLL("libamdhip64.so.7")
    lib.hipMemImportFromShareableHandle.argtypes = [C.POINTER(C.c_uint64), C.c_void_p, C.c_int]
    lib.hipMemAddressReserve.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_size_t, C.c_void_p, C.c_ulonglong]
    lib.hipMemMap.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint64, C.c_ulonglong]
    lib.hipMemSetAccess.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(MemAccessDesc), C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemUnmap.argtypes = [C.c_void_p, C.c_size_t]
    lib.hipMemAddressFree.argtypes = [C.c_void_p, C.c_size_t]
    lib.hipMemRelease.argtypes = [C.c_uint64]
    return lib


def read_through_fd(fd: int, nbytes: int, want: int) -> np.ndarray:
    """The first `want` bytes of the exported memory, imported and mapped anew."""
    lib = hip()
    h = C.c_uint64(0)
    assert lib.hipMemImportFromShareableHandle(C.byref(h), C.c_void_p(fd), HIP_MEM_HANDLE_TYPE_POSIX_FD) == 0
    ptr = C.c_void_p()
    assert lib.hipMemAddressReserve(C.byref(ptr), nbytes, 0, None, 0) == 0
    try:
        assert lib.hipMemMap(ptr, nbytes, 0, h, 0) == 0
        try:
            desc = MemAccessDesc(MemLocation(HIP_MEM_LOCATION_TYPE_DEVICE, 0), HIP_MEM_ACCESS_READ_WRITE)
            assert lib.hipMemSetAccess(ptr, nbytes, C.byref(desc), 1) == 0
            out = np.empty(want // 4, np.float32)
            assert lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), ptr, want, HIP_MEMCPY_DEVICE_TO_HOST) == 0
            return out
        finally:
            lib.hipMemUnmap(ptr, nbytes)
    finally:
        lib.hipMemAddressFree(ptr, nbytes)
        lib.hipMemRelease(h)


def render_into(ctx, acc, text, W, H, frames, mode):
    sc = ctx.scene_from_json(text)
    pc = pt.push_constants(W, H, spf=2, path_length=6)
    sc.camera(1, pc)
    for f in range(frames):
        pc.frame = 2 * (f + 1)
        pc.currentSamples = 2 * (f + 1)
        ctx.dispatch(sc, acc, pc, mode)
    ctx.sync()
    sc.release()


@pytest.mark.parametrize("mode", [pt.MODE_REF, pt.MODE_EXT])
def test_exported_texels_are_the_accumulator(gpu_ctx, scene_texts, mode):
    W, H = 64, 48
    # the reference mode takes the reference's own scenes; the extended mode the Cornell box
    text = scenes.cornell(4, 3) if mode == pt.MODE_EXT else scene_texts["scene1"]
    plain = gpu_ctx.accum(W, H)
    exp = gpu_ctx.accum(W, H, exportable=True)
    try:
        render_into(gpu_ctx, plain, text, W, H, 3, mode)
        render_into(gpu_ctx, exp, text, W, H, 3, mode)
        a, b = plain.read(), exp.read()
        assert np.array_equal(a, b) and np.isfinite(b).all() and b[..., :3].max() > 0
        fd, nbytes = exp.export_fd()
        try:
            assert fd >= 0 and nbytes >= W * H * 16
            shared = read_through_fd(fd, nbytes, W * H * 16).reshape(H, W, 4)
        finally:
            os.close(fd)
        assert np.array_equal(shared, b)
        # the accumulator keeps working after the export
        exp.reset()
        assert not exp.read().any()
    finally:
        plain.release()
        exp.release()


def test_only_exportable_accumulators_export(gpu_ctx):
    acc = gpu_ctx.accum(16, 16)
    try:
        with pytest.raises(pt.PTError) as e:
            acc.export_fd()
        assert e.value.status == pt.PT_ERR_STATE
    finally:
        acc.release()
