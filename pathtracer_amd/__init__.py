"""pathtracer_amd — MI355X (gfx950) spectral path tracer behind a C ABI.

This module is a thin ctypes binding of ``include/pathtracer.h`` (the
library ``libpathtracer.so`` built in this directory) plus the reference
application's frame bookkeeping (``Renderer``), so Python tests and benches
drive exactly the C entry points a C++ host would.

The compute path is only the HIP library: importing this module fails loudly
when ``libpathtracer.so`` has not been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libpathtracer.so"

PT_OK = 0
PT_ERR_INVALID_ARG = -1
PT_ERR_JSON = -2
PT_ERR_CAPACITY = -3
PT_ERR_UNSUPPORTED_SDF = -4
PT_ERR_HIP = -5
PT_ERR_NO_DEVICE = -6
PT_ERR_BUDGET = -7
PT_ERR_STATE = -8

MODE_REF = 0
MODE_EXT = 1
FEATURES = 12  # floats per pixel of a feature image (PT_FEATURES)
DN_LOSS_L1_LOG, DN_LOSS_L2 = 0, 1

SDF_MANDELBULB, SDF_MENGER, SDF_BLOB, SDF_TERRAIN = 1, 2, 3, 4
SDFMAT_CONSTANT, SDFMAT_RADIAL, SDFMAT_BLOB = 0, 1, 2


class PackedScene(C.Structure):
    """UniformBufferObject (src/pathtracer.cpp:187-195), 16,388 bytes."""

    _fields_ = [
        ("numObjects", C.c_float * 7),
        ("objects", C.c_float * 1024),
        ("sdfs", C.c_float * 768),
        ("materials", C.c_float * 783),
        ("lights", C.c_float * 128),
        ("lightIDs", C.c_float * 64),
        ("cie", C.c_float * 1323),
    ]


class PushConstants(C.Structure):
    """PushConstantValues (src/pathtracer.cpp:197-218), 88 bytes."""

    _fields_ = [
        ("resolution", C.c_int32 * 2),
        ("frame", C.c_int32),
        ("currentSamples", C.c_int32),
        ("samplesPerFrame", C.c_int32),
        ("FPS", C.c_float),
        ("persistence", C.c_float),
        ("pathLength", C.c_int32),
        ("cameraAngle", C.c_float * 2),
        ("cameraPosX", C.c_float),
        ("cameraPosY", C.c_float),
        ("cameraPosZ", C.c_float),
        ("ISO", C.c_int32),
        ("cameraSize", C.c_float),
        ("apertureSize", C.c_float),
        ("apertureDist", C.c_float),
        ("lensRadius", C.c_float),
        ("lensFocalLength", C.c_float),
        ("lensThickness", C.c_float),
        ("lensDistance", C.c_float),
        ("tonemap", C.c_int32),
    ]


class SdfDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("iterations", C.c_int32),
        ("bailout", C.c_float),
        ("material_kind", C.c_int32),
        ("material", C.c_float * 4),
        ("pos", C.c_float * 3),
        ("size", C.c_float * 3),
    ]


class Config(C.Structure):
    _fields_ = [
        ("wave_paths", C.c_int32),
        ("block_size", C.c_int32),
        ("bvh_width", C.c_int32),
        ("sdf_max_steps", C.c_int32),
        ("hbm_budget", C.c_int64),
        ("traversal", C.c_int32),
        ("bvh_leaf", C.c_int32),
        ("graph_mode", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("march", C.c_int32),
        ("march_blocks_per_cu", C.c_int32),
    ]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class Counters(C.Structure):
    _fields_ = [
        ("segments", C.c_uint64),
        ("shadow_segments", C.c_uint64),
        ("sdf_steps", C.c_uint64),
        ("paths", C.c_uint64),
        ("nan_samples", C.c_uint64),
        ("dispatches", C.c_uint64),
        ("escaped", C.c_uint64),
        ("emitter_hits", C.c_uint64),
        ("lambert_hits", C.c_uint64),
        ("dielectric_hits", C.c_uint64),
        ("conductor_hits", C.c_uint64),
        ("bvh_nodes", C.c_uint64),
        ("shadow_bvh_nodes", C.c_uint64),
        ("sdf_exhausted", C.c_uint64),
        ("march_flops_mandelbulb", C.c_uint64),
        ("march_flops_menger", C.c_uint64),
        ("march_flops_blob", C.c_uint64),
        ("march_flops_terrain", C.c_uint64),
        ("shadow_march_flops", C.c_uint64),
        ("bad_refs", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


assert C.sizeof(PackedScene) == 16388
assert C.sizeof(PushConstants) == 88

class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_int64),
        ("max_active", C.c_int32),
        ("policy", C.c_int32),
        ("min_spp", C.c_int32),
        ("reserved", C.c_int32),
    ]


class RequestDesc(C.Structure):
    _fields_ = [
        ("json", C.c_char_p),
        ("len", C.c_size_t),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("spp_max", C.c_int32),
        ("path_length", C.c_int32),
        ("target_rel", C.c_float),
        ("denoise", C.c_int32),
    ]


class RequestStatus(C.Structure):
    _fields_ = [
        ("id", C.c_int32),
        ("state", C.c_int32),
        ("spp", C.c_int32),
        ("converged", C.c_int32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("rel_error", C.c_double),
        ("t_arrive", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
        ("gpu_seconds", C.c_double),
        ("segments", C.c_uint64),
        ("rel_plain", C.c_double),
        ("denoised", C.c_int32),
        ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


SERVE_SRPT, SERVE_FIFO, SERVE_HRRN, SERVE_IDLE = 0, 1, 2, 1
REQ_WAITING, REQ_ACTIVE, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

# Every symbol include/pathtracer.h declares, with (restype, argtypes).
_P = C.c_void_p
SIGNATURES = {
    "pt_create": (C.c_int, [C.POINTER(_P), C.c_int, C.POINTER(Config)]),
    "pt_destroy": (None, [_P]),
    "pt_last_error": (C.c_char_p, [_P]),
    "pt_last_error_global": (C.c_char_p, []),
    "pt_scene_from_json": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(_P)]),
    "pt_scene_from_packed": (C.c_int, [_P, C.POINTER(PackedScene), C.POINTER(SdfDesc), C.c_int, C.POINTER(_P)]),
    "pt_scene_get_packed": (C.c_int, [_P, C.POINTER(PackedScene)]),
    "pt_scene_get_sdfs": (C.c_int, [_P, C.POINTER(SdfDesc), C.c_int, C.POINTER(C.c_int)]),
    "pt_scene_camera": (C.c_int, [_P, C.c_int, C.POINTER(PushConstants)]),
    "pt_scene_device_bytes": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "pt_scene_release": (None, [_P]),
    "pt_sdf_recognize": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(SdfDesc)]),
    "pt_pack_json": (
        C.c_int,
        [C.c_char_p, C.c_size_t, C.c_int, C.POINTER(PackedScene), C.POINTER(PushConstants), C.POINTER(SdfDesc),
         C.c_int, C.POINTER(C.c_int)],
    ),
    "pt_scene_write_json": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pt_json_rewrite": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "pt_accum_create": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(_P)]),
    "pt_accum_reset": (C.c_int, [_P]),
    "pt_accum_read_xyz": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "pt_accum_write_xyz": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "pt_accum_release": (None, [_P]),
    "pt_accum_create_exportable": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(_P)]),
    "pt_accum_export_fd": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "pt_dispatch": (C.c_int, [_P, _P, _P, C.POINTER(PushConstants), C.c_int]),
    "pt_dispatch_noalloc": (C.c_int, [_P, _P, _P, C.POINTER(PushConstants)]),
    "pt_render_features": (C.c_int, [_P, _P, C.POINTER(PushConstants), C.POINTER(C.c_float), C.c_size_t]),
    "pt_sync": (C.c_int, [_P]),
    "pt_counters_get": (C.c_int, [_P, C.POINTER(Counters)]),
    "pt_counters_reset": (C.c_int, [_P]),
    "pt_profile_enable": (C.c_int, [_P, C.c_int]),
    "pt_profile_read": (C.c_int, [_P, C.POINTER(KernelStat), C.c_int, C.POINTER(C.c_int), C.c_int]),
    "pt_resolve_srgb8": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_uint8), C.c_size_t]),
    "pt_server_create": (C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    "pt_server_destroy": (None, [_P]),
    "pt_server_last_error": (C.c_char_p, [_P]),
    "pt_server_now": (C.c_double, [_P]),
    "pt_server_submit": (C.c_int, [_P, C.POINTER(RequestDesc), C.c_double, C.POINTER(C.c_int32)]),
    "pt_server_step": (C.c_int, [_P, C.POINTER(RequestStatus), C.c_int32, C.POINTER(C.c_int32)]),
    "pt_server_status": (C.c_int, [_P, C.c_int32, C.POINTER(RequestStatus)]),
    "pt_server_read": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_float), C.c_size_t]),
    "pt_server_release": (C.c_int, [_P, C.c_int32]),
    "pt_server_counts": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "pt_server_set_denoiser": (C.c_int, [_P, _P]),
    "pt_dn_create": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)]),
    "pt_dn_destroy": (None, [_P]),
    "pt_dn_last_error": (C.c_char_p, [_P]),
    "pt_dn_param_count": (C.c_int64, [_P]),
    "pt_dn_get_params": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "pt_dn_set_params": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "pt_dn_denoise": (C.c_int, [_P, C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "pt_dn_gradients": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t]),
    "pt_dn_set_loss": (C.c_int, [_P, C.c_int]),
    "pt_dn_train_step": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.c_int, C.c_int,
                                   C.c_float, C.POINTER(C.c_float)]),
    "pt_math_eval_host": (C.c_int, [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t]),
    "pt_math_eval_device": (
        C.c_int,
        [_P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t],
    ),
    "pt_peak_probe": (C.c_int, [_P, C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `make -C {_HERE}` (or __graft_entry__.build()) first; "
            "there is no CPU fallback"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class PTError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def _check(st: int, ctx=None) -> None:
    if st != PT_OK:
        if ctx is not None:
            msg = lib.pt_last_error(ctx)
        else:
            msg = lib.pt_last_error_global()
        raise PTError(st, (msg or b"").decode(errors="replace"))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def recognize_sdf(glsl: str) -> SdfDesc:
    d = SdfDesc()
    b = glsl.encode()
    _check(lib.pt_sdf_recognize(b, len(b), C.byref(d)))
    return d


def pack_json(text: str, shot: int = 1, pc: PushConstants | None = None):
    """Host-only UpdateFromJSON + UpdateUniformBuffer: returns (PackedScene,
    PushConstants with the camera of `shot`, [SdfDesc])."""
    b = text.encode()
    ubo = PackedScene()
    pc = pc if pc is not None else PushConstants()
    n = C.c_int()
    _check(lib.pt_pack_json(b, len(b), shot, C.byref(ubo), C.byref(pc), None, 0, C.byref(n)))
    arr = (SdfDesc * max(1, n.value))()
    _check(lib.pt_pack_json(b, len(b), shot, None, None, arr, n.value, C.byref(n)))
    return ubo, pc, [arr[i] for i in range(n.value)]


def _read_text(call) -> str:
    n = C.c_size_t()
    _check(call(None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    _check(call(buf, n.value + 1, C.byref(n)))
    return buf.raw[: n.value].decode()


def json_rewrite(text: str) -> str:
    """Host-only: parse a scene and write it back as the reference's SaveScene
    does (UpdateToJSON, src/pathtracer.cpp:2724-2858), extended keys exact."""
    b = text.encode()
    return _read_text(lambda out, cap, n: lib.pt_json_rewrite(b, len(b), out, cap, n))


def math_eval_host(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    yy = np.ascontiguousarray(y, dtype=np.float32) if y is not None else None
    out = np.empty_like(x)
    _check(lib.pt_math_eval_host(fn, _fptr(x), _fptr(yy) if yy is not None else None, _fptr(out), x.size))
    return out


class Context:
    """pt_ctx: one device, one stream."""

    def __init__(self, device: int = 0, wave_paths: int = 0, block_size: int = 0, bvh_width: int = 0,
                 sdf_max_steps: int = 0, hbm_budget: int = 0, traversal: int = 0, bvh_leaf: int = 0,
                 graph_mode: int = 0, blocks_per_cu: int = 0, march: int = 0, march_blocks_per_cu: int = 0):
        cfg = Config(wave_paths, block_size, bvh_width, sdf_max_steps, hbm_budget, traversal, bvh_leaf, graph_mode,
                     blocks_per_cu, march, march_blocks_per_cu)
        h = C.c_void_p()
        _check(lib.pt_create(C.byref(h), device, C.byref(cfg)))
        self.h = h

    def close(self) -> None:
        if self.h:
            lib.pt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int) -> None:
        _check(st, self.h)

    def scene_from_json(self, text: str) -> "Scene":
        b = text.encode()
        h = C.c_void_p()
        self.check(lib.pt_scene_from_json(self.h, b, len(b), C.byref(h)))
        return Scene(self, h)

    def scene_from_packed(self, ubo: PackedScene, sdfs: list[SdfDesc]) -> "Scene":
        arr = (SdfDesc * max(1, len(sdfs)))(*sdfs)
        h = C.c_void_p()
        self.check(lib.pt_scene_from_packed(self.h, C.byref(ubo), arr, len(sdfs), C.byref(h)))
        return Scene(self, h)

    def accum(self, w: int, h: int, exportable: bool = False) -> "Accum":
        """An accumulator of w x h texels; `exportable` ones can hand their
        device memory to another API as a dma-buf (Accum.export_fd)."""
        a = C.c_void_p()
        create = lib.pt_accum_create_exportable if exportable else lib.pt_accum_create
        self.check(create(self.h, w, h, C.byref(a)))
        return Accum(self, a, w, h)

    def dispatch(self, scene: "Scene", acc: "Accum", pc: PushConstants, mode: int = MODE_REF) -> None:
        self.check(lib.pt_dispatch(self.h, scene.h, acc.h, C.byref(pc), mode))

    def features(self, scene: "Scene", pc: PushConstants) -> np.ndarray:
        """Per-pixel means over the samples [currentSamples - samplesPerFrame,
        currentSamples) of XYZ, first-hit albedo, normal and distance, then the
        standard error of the mean of Y and the sample count:
        (height, width, FEATURES) float32."""
        w, h = pc.resolution[0], pc.resolution[1]
        out = np.empty((h, w, FEATURES), np.float32)
        self.check(lib.pt_render_features(self.h, scene.h, C.byref(pc),
                                          out.ctypes.data_as(C.POINTER(C.c_float)), out.size))
        return out

    def sync(self) -> None:
        self.check(lib.pt_sync(self.h))

    def counters(self) -> dict:
        c = Counters()
        self.check(lib.pt_counters_get(self.h, C.byref(c)))
        return c.as_dict()

    def reset_counters(self) -> None:
        self.check(lib.pt_counters_reset(self.h))

    def profile(self, enable: bool = True) -> None:
        self.check(lib.pt_profile_enable(self.h, 1 if enable else 0))

    def profile_read(self, reset: bool = True) -> dict:
        """{kernel name: (launches, total GPU ms)} since the last reset."""
        n = C.c_int()
        self.check(lib.pt_profile_read(self.h, None, 0, C.byref(n), 0))
        arr = (KernelStat * max(1, n.value))()
        self.check(lib.pt_profile_read(self.h, arr, n.value, C.byref(n), 1 if reset else 0))
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms)) for i in range(n.value)}

    def math_eval(self, fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        yy = np.ascontiguousarray(y, dtype=np.float32) if y is not None else None
        out = np.empty_like(x)
        self.check(lib.pt_math_eval_device(self.h, fn, _fptr(x), _fptr(yy) if yy is not None else None,
                                           _fptr(out), x.size))
        return out

    def peak_probe(self, what: int, n: int, reps: int) -> tuple:
        """Measured device peak: what 0 streams n floats (bytes/s), what 1 runs n
        rounds of the FP32 FMA loop (flop/s). Returns (rate per second, units per
        launch, average launch ms) from the HIP events around each timed launch."""
        units = C.c_double()
        self.profile(True)
        self.profile_read(reset=True)
        try:
            self.check(lib.pt_peak_probe(self.h, what, n, reps, C.byref(units)))
            stats = self.profile_read(reset=True)
        finally:
            self.profile(False)
        launches, total_ms = stats["k_peak_copy" if what == 0 else "k_peak_fma"]
        ms = total_ms / launches
        return units.value / (ms * 1e-3), units.value, ms


class Scene:
    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self.h = h

    def release(self) -> None:
        if self.h:
            lib.pt_scene_release(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.ctx.h:
                self.release()
        except Exception:
            pass

    def packed(self) -> PackedScene:
        u = PackedScene()
        _check(lib.pt_scene_get_packed(self.h, C.byref(u)))
        return u

    def sdfs(self) -> list[SdfDesc]:
        n = C.c_int()
        _check(lib.pt_scene_get_sdfs(self.h, None, 0, C.byref(n)))
        arr = (SdfDesc * max(1, n.value))()
        _check(lib.pt_scene_get_sdfs(self.h, arr, n.value, C.byref(n)))
        return [arr[i] for i in range(n.value)]

    def camera(self, shot: int, pc: PushConstants) -> PushConstants:
        _check(lib.pt_scene_camera(self.h, shot, C.byref(pc)))
        return pc

    def device_bytes(self) -> int:
        b = C.c_int64()
        _check(lib.pt_scene_device_bytes(self.h, C.byref(b)))
        return b.value

    def to_json(self) -> str:
        return _read_text(lambda out, cap, n: lib.pt_scene_write_json(self.h, out, cap, n))


class Accum:
    def __init__(self, ctx: Context, h, w: int, hh: int):
        self.ctx, self.h, self.w, self.hgt = ctx, h, w, hh

    def release(self) -> None:
        if self.h:
            lib.pt_accum_release(self.h)
            self.h = None

    def __del__(self):
        try:
            if self.ctx.h:
                self.release()
        except Exception:
            pass

    def reset(self) -> None:
        self.ctx.check(lib.pt_accum_reset(self.h))

    def read(self) -> np.ndarray:
        out = np.empty((self.hgt, self.w, 4), dtype=np.float32)
        self.ctx.check(lib.pt_accum_read_xyz(self.h, _fptr(out), out.size))
        return out

    def write(self, texels: np.ndarray) -> None:
        t = np.ascontiguousarray(texels, dtype=np.float32)
        self.ctx.check(lib.pt_accum_write_xyz(self.h, _fptr(t), t.size))

    def export_fd(self) -> tuple:
        """(file descriptor, bytes): a new dma-buf descriptor of the texels'
        device memory, owned by the caller (os.close it when done)."""
        fd, n = C.c_int(-1), C.c_uint64(0)
        self.ctx.check(lib.pt_accum_export_fd(self.h, C.byref(fd), C.byref(n)))
        return fd.value, n.value

    def resolve_srgb8(self, tonemap: int = 3) -> np.ndarray:
        out = np.empty((self.hgt, self.w, 3), dtype=np.uint8)
        self.ctx.check(lib.pt_resolve_srgb8(self.ctx.h, self.h, tonemap,
                                            out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size))
        return out


def push_constants(width: int, height: int, spf: int = 1, path_length: int = 5, persistence: float = 0.0625,
                   fps: float = 1.0 / 0.0166, tonemap: int = 3) -> PushConstants:
    """Push constants with the reference application's defaults
    (src/pathtracer.cpp:1163-1174); camera fields come from Scene.camera()."""
    pc = PushConstants()
    pc.resolution[0] = width
    pc.resolution[1] = height
    pc.samplesPerFrame = spf
    pc.pathLength = path_length
    pc.persistence = persistence
    pc.FPS = fps
    pc.tonemap = tonemap
    return pc


class Renderer:
    """Frame bookkeeping of the reference's MainLoop (src/pathtracer.cpp:4005-4050):
    `frame` counts every sample dispatched, `currentSamples` restarts at
    samplesPerFrame after a reset. Those two counters key the RNG and the
    running mean, so a sequence of render_frame() calls reproduces the
    reference's sequence of dispatches."""

    def __init__(self, ctx: Context, scene: Scene, width: int, height: int, spf: int = 1, path_length: int = 5,
                 shot: int = 1, mode: int = MODE_REF, tonemap: int = 3):
        self.ctx, self.scene, self.mode = ctx, scene, mode
        self.pc = push_constants(width, height, spf, path_length, tonemap=tonemap)
        scene.camera(shot, self.pc)
        self.acc = ctx.accum(width, height)
        self.frame = 0
        self.current = 0
        self.reset_pending = False

    def reset(self) -> None:
        self.reset_pending = True

    def render_frame(self) -> None:
        spf = self.pc.samplesPerFrame
        self.frame += spf
        if self.reset_pending:
            self.current = spf
            self.reset_pending = False
        else:
            self.current += spf
        self.pc.frame = self.frame
        self.pc.currentSamples = self.current
        self.ctx.dispatch(self.scene, self.acc, self.pc, self.mode)

    def render(self, total_samples: int) -> np.ndarray:
        while self.current < total_samples:
            self.render_frame()
        return self.acc.read()

    def save_ppm(self, path: str | os.PathLike, tonemap: int | None = None) -> None:
        rgb = self.acc.resolve_srgb8(self.pc.tonemap if tonemap is None else tonemap)
        with open(path, "wb") as f:
            f.write(b"P6\n%d\n%d\n255\n" % (self.acc.w, self.acc.hgt))
            f.write(rgb.tobytes())


class Server:
    """pt_server: render requests traced on one context until each reaches its
    relative-RMSE target (shortest estimated remaining time first, or FIFO)."""

    def __init__(self, ctx: "Context", policy: int = SERVE_SRPT, max_active: int = 0, hbm_budget: int = 0,
                 min_spp: int = 0):
        self.ctx = ctx
        cfg = ServerConfig(hbm_budget, max_active, policy, min_spp, 0)
        h = C.c_void_p()
        _check(lib.pt_server_create(ctx.h, C.byref(cfg), C.byref(h)), ctx.h)
        self.h = h

    def _ck(self, st: int) -> int:
        if st < 0:
            raise PTError(st, (lib.pt_server_last_error(self.h) or b"").decode(errors="replace"))
        return st

    def close(self) -> None:
        if self.h:
            lib.pt_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def now(self) -> float:
        return lib.pt_server_now(self.h)

    def submit(self, text: str, width: int, height: int, spp_max: int, path_length: int, target: float = 0.01,
               t_arrive: float | None = None, denoise: bool = False) -> int:
        b = text.encode()
        d = RequestDesc(b, len(b), width, height, spp_max, path_length, target, int(denoise))
        rid = C.c_int32()
        self._ck(lib.pt_server_submit(self.h, C.byref(d), self.now() if t_arrive is None else t_arrive,
                                      C.byref(rid)))
        return rid.value

    def step(self) -> list[RequestStatus] | None:
        """One wave; the statuses of the requests it traced, None when nothing is admitted."""
        arr = (RequestStatus * 64)()
        n = C.c_int32()
        rc = self._ck(lib.pt_server_step(self.h, arr, 64, C.byref(n)))
        return None if rc == SERVE_IDLE else [arr[i] for i in range(n.value)]

    def status(self, rid: int) -> RequestStatus:
        st = RequestStatus()
        self._ck(lib.pt_server_status(self.h, rid, C.byref(st)))
        return st

    def read(self, rid: int) -> np.ndarray:
        st = self.status(rid)
        out = np.empty((st.height, st.width, 4), np.float32)
        self._ck(lib.pt_server_read(self.h, rid, _fptr(out), out.size))
        return out

    def release(self, rid: int) -> None:
        self._ck(lib.pt_server_release(self.h, rid))

    def counts(self) -> tuple[int, int, int]:
        w, a, b = C.c_int32(), C.c_int32(), C.c_int64()
        self._ck(lib.pt_server_counts(self.h, C.byref(w), C.byref(a), C.byref(b)))
        return w.value, a.value, b.value

    def set_denoiser(self, dn: "Denoiser | None") -> None:
        """Requests submitted with denoise=True may finish with its image."""
        self._ck(lib.pt_server_set_denoiser(self.h, dn.h if dn is not None else None))
        self._dn = dn  # kept alive while attached


class Denoiser:
    """pt_denoiser: the U-Net denoiser (BF16 on the matrix cores, FP32 masters)."""

    def __init__(self, ctx: "Context", max_batch: int, max_height: int, max_width: int, seed: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib.pt_dn_create(ctx.h, max_batch, max_height, max_width, seed, C.byref(h)), ctx.h)
        self.h = h
        self.n = int(lib.pt_dn_param_count(h))

    def _ck(self, st: int) -> None:
        if st != PT_OK:
            raise PTError(st, (lib.pt_dn_last_error(self.h) or b"").decode(errors="replace"))

    def close(self) -> None:
        if self.h:
            lib.pt_dn_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self) -> np.ndarray:
        out = np.empty(self.n, np.float32)
        self._ck(lib.pt_dn_get_params(self.h, _fptr(out), self.n))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        self._ck(lib.pt_dn_set_params(self.h, _fptr(p), p.size))

    def denoise(self, feat: np.ndarray) -> np.ndarray:
        feat = np.ascontiguousarray(feat, np.float32)
        n, h, w, _ = feat.shape
        out = np.empty((n, h, w, 3), np.float32)
        self._ck(lib.pt_dn_denoise(self.h, _fptr(feat), n, h, w, _fptr(out)))
        return out

    def gradients(self, feat: np.ndarray, target: np.ndarray) -> tuple[float, np.ndarray]:
        feat = np.ascontiguousarray(feat, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        n, h, w, _ = feat.shape
        g = np.empty(self.n, np.float32)
        loss = C.c_float()
        self._ck(lib.pt_dn_gradients(self.h, _fptr(feat), _fptr(target), n, h, w, C.byref(loss), _fptr(g), self.n))
        return loss.value, g

    def set_loss(self, kind: int) -> None:
        self._ck(lib.pt_dn_set_loss(self.h, kind))

    def train_step(self, feat: np.ndarray, target: np.ndarray, lr: float) -> float:
        feat = np.ascontiguousarray(feat, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        n, h, w, _ = feat.shape
        loss = C.c_float()
        self._ck(lib.pt_dn_train_step(self.h, _fptr(feat), _fptr(target), n, h, w, lr, C.byref(loss)))
        return loss.value
