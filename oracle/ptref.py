"""ORACLE binding (test infrastructure only).

ctypes access to oracle/libptref.so, the FP32 CPU restatement of the
reference shader (oracle/ptref.c). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline use it, always as the checker or the CPU baseline,
never as the thing measured or shipped.

Arguments are passed as raw bytes in the C-ABI layouts of include/pathtracer.h
(pt_packed_scene 16,388 B, pt_push_constants 88 B, pt_sdf_desc 56 B each), so
this module does not import the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent


def _open(name: str) -> C.CDLL:
    path = _HERE / name
    if not path.exists():
        raise RuntimeError(f"{path} is not built: run `make -C {_HERE}`")
    lib = C.CDLL(str(path))
    lib.ptref_dispatch.restype = C.c_int
    lib.ptref_dispatch.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_float),
                                   C.POINTER(C.c_uint64), C.c_int]
    lib.ptref_pixel_samples.restype = C.c_int
    lib.ptref_pixel_samples.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_float)]
    lib.ptref_pcg.restype = C.c_uint32
    lib.ptref_pcg.argtypes = [C.c_uint32]
    for fn in ("ptref_bk7", "ptref_blackbody_peak"):
        getattr(lib, fn).restype = C.c_float
        getattr(lib, fn).argtypes = [C.c_float]
    lib.ptref_blackbody.restype = C.c_float
    lib.ptref_blackbody.argtypes = [C.c_float, C.c_float]
    lib.ptref_spd.restype = C.c_float
    lib.ptref_spd.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int]
    lib.ptref_wave_to_xyz.restype = None
    lib.ptref_wave_to_xyz.argtypes = [C.c_char_p, C.c_float, C.POINTER(C.c_float)]
    lib.ptref_sample_wavelengths.restype = None
    lib.ptref_sample_wavelengths.argtypes = [C.c_float, C.POINTER(C.c_float)]
    lib.ptref_rotation.restype = None
    lib.ptref_rotation.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.ptref_solve_quartic.restype = None
    lib.ptref_solve_quartic.argtypes = [C.c_float] * 5 + [C.POINTER(C.c_float), C.POINTER(C.c_int)]
    lib.ptref_sdf_eval.restype = C.c_float
    lib.ptref_sdf_eval.argtypes = [C.c_char_p, C.POINTER(C.c_float)]
    lib.ptref_sdfmat_eval.restype = C.c_float
    lib.ptref_sdfmat_eval.argtypes = [C.c_char_p, C.POINTER(C.c_float)]
    _probe_types(lib)
    return lib


def _probe_types(lib: C.CDLL) -> None:
    F = C.POINTER(C.c_float)
    lib.ptref_probe_isect.restype = C.c_int
    lib.ptref_probe_isect.argtypes = [C.c_int, F, F, F, F]
    lib.ptref_probe_cyclide.restype = None
    lib.ptref_probe_cyclide.argtypes = [F, F, F, F, F, F]
    lib.ptref_probe_sample.restype = None
    lib.ptref_probe_sample.argtypes = [C.c_int, C.c_uint32, F, C.c_float, F, F]
    lib.ptref_probe_light_pdf.restype = C.c_float
    lib.ptref_probe_light_pdf.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.ptref_probe_mis.restype = C.c_float
    lib.ptref_probe_mis.argtypes = [C.c_float, C.c_float]
    lib.ptref_probe_march.restype = C.c_int
    lib.ptref_probe_march.argtypes = [C.c_char_p, C.c_char_p, C.c_int, F, F, F, C.POINTER(C.c_uint64)]


def load(path: str | os.PathLike) -> C.CDLL:
    """Another build of ptref.c (a planted-mistake variant in a test), with the
    same argument types."""
    lib2 = C.CDLL(str(path))
    lib2.ptref_dispatch.restype = C.c_int
    lib2.ptref_dispatch.argtypes = lib.ptref_dispatch.argtypes
    lib2.ptref_pixel_samples.restype = C.c_int
    lib2.ptref_pixel_samples.argtypes = lib.ptref_pixel_samples.argtypes
    _probe_types(lib2)
    return lib2


lib = _open("libptref.so")
_libm = None


def libm_variant() -> C.CDLL:
    """The same restatement with glibc's transcendentals (divergence report)."""
    global _libm
    if _libm is None:
        _libm = _open("libptref_libm.so")
    return _libm


def dispatch(ubo: bytes, pc: bytes, sdfs: bytes, nsdf: int, texels: np.ndarray, threads: int = 0,
             use_libm: bool = False) -> dict:
    """One reference dispatch on the CPU; `texels` (H, W, 4) float32 is updated in place."""
    assert len(ubo) == 16388 and len(pc) == 88 and len(sdfs) >= 56 * nsdf
    assert texels.dtype == np.float32 and texels.flags.c_contiguous
    cnt = (C.c_uint64 * 6)()
    L = libm_variant() if use_libm else lib
    st = L.ptref_dispatch(ubo, pc, sdfs if nsdf else None, nsdf, texels.ctypes.data_as(C.POINTER(C.c_float)), cnt,
                          threads or (os.cpu_count() or 1))
    if st != 0:
        raise RuntimeError(f"ptref_dispatch failed with status {st}")
    keys = ("segments", "shadow_segments", "sdf_steps", "paths", "nan_samples", "dispatches")
    return {k: int(v) for k, v in zip(keys, cnt)}


def pixel_samples(ubo: bytes, pc: bytes, sdfs: bytes, nsdf: int, gx: int, gy: int, spf: int) -> np.ndarray:
    out = np.zeros((spf, 3), dtype=np.float32)
    st = lib.ptref_pixel_samples(ubo, pc, sdfs if nsdf else None, nsdf, gx, gy,
                                 out.ctypes.data_as(C.POINTER(C.c_float)))
    if st != 0:
        raise RuntimeError(f"ptref_pixel_samples failed with status {st}")
    return out


def pcg(x: int) -> int:
    return int(lib.ptref_pcg(x & 0xFFFFFFFF))


def wave_to_xyz(ubo: bytes, wave: float) -> np.ndarray:
    out = (C.c_float * 3)()
    lib.ptref_wave_to_xyz(ubo, wave, out)
    return np.array(out[:], dtype=np.float32)


def sample_wavelengths(lh: float) -> np.ndarray:
    out = (C.c_float * 4)()
    lib.ptref_sample_wavelengths(lh, out)
    return np.array(out[:], dtype=np.float32)


def rotation(angle_deg) -> np.ndarray:
    a = (C.c_float * 3)(*angle_deg)
    out = (C.c_float * 9)()
    lib.ptref_rotation(a, out)
    return np.array(out[:], dtype=np.float32).reshape(3, 3)  # rows = GLSL columns


def solve_quartic(a, b, c, d, e):
    roots = (C.c_float * 4)()
    real = (C.c_int * 4)()
    lib.ptref_solve_quartic(a, b, c, d, e, roots, real)
    return np.array(roots[:], dtype=np.float32), [bool(x) for x in real]


def sdf_eval(desc: bytes, p) -> float:
    return float(lib.ptref_sdf_eval(desc, (C.c_float * 3)(*p)))


def sdfmat_eval(desc: bytes, p) -> float:
    return float(lib.ptref_sdfmat_eval(desc, (C.c_float * 3)(*p)))


def _f(x, n=None) -> C.Array:
    a = np.ascontiguousarray(x, dtype=np.float32).ravel()
    return (C.c_float * (n or len(a)))(*a)


ISECT = {"sphere": 0, "plane": 1, "box": 2, "lens": 3, "cyclide": 4, "bound": 5}


def probe_isect(kind: str, params, o, d, L: C.CDLL | None = None):
    """One intersector on one ray: (hit, t, normal[3], isOutside)."""
    out = (C.c_float * 5)()
    hit = (L or lib).ptref_probe_isect(ISECT[kind], _f(params, 16), _f(o), _f(d), out)
    return bool(hit), float(out[0]), np.array(out[1:4], np.float32), int(out[4])


def probe_cyclide(params, o, d, L: C.CDLL | None = None):
    """(object-space origin, direction, quartic coefficients constant-first)."""
    lo, ld, cf = (C.c_float * 3)(), (C.c_float * 3)(), (C.c_float * 5)()
    (L or lib).ptref_probe_cyclide(_f(params, 16), _f(o), _f(d), lo, ld, cf)
    return np.array(lo[:], np.float32), np.array(ld[:], np.float32), np.array(cf[:], np.float32)


SAMPLE = {"sphere": 0, "bsdf": 1, "cone": 2, "to_world": 3}


def probe_samples(kind: str, seeds, n=(0.0, 0.0, 1.0), cos_max=0.0, v=(0.0, 0.0, 1.0), L: C.CDLL | None = None):
    L = L or lib
    out = np.zeros((len(seeds), 3), np.float32)
    nn, vv, o = _f(n), _f(v), (C.c_float * 3)()
    for i, s in enumerate(seeds):
        L.ptref_probe_sample(SAMPLE[kind], int(s) & 0xFFFFFFFF, nn, cos_max, vv, o)
        out[i] = o[:]
    return out


def probe_march(ubo: bytes, sdfs: bytes, nsdf: int, o, d, L: C.CDLL | None = None):
    """(hit, t, normal[3], material, distance evaluations)."""
    out = (C.c_float * 5)()
    steps = C.c_uint64()
    hit = (L or lib).ptref_probe_march(ubo, sdfs or None, nsdf, _f(o), _f(d), out, C.byref(steps))
    return bool(hit), float(out[0]), np.array(out[1:4], np.float32), float(out[4]), int(steps.value)
