"""ORACLE binding (test infrastructure only): the FP64 extended-mode tracer
oracle/ptext.c. The scene JSON is flattened here with Python's json module,
independently of the product's C++ loader."""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / "libptext.so"
if not _LIB.exists():
    raise RuntimeError(f"{_LIB} is not built: run `make -C {_HERE}`")
lib = C.CDLL(str(_LIB))

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


class Scene(C.Structure):
    _fields_ = [
        ("nsph", C.c_int), ("nquad", C.c_int), ("nplane", C.c_int), ("nsdf", C.c_int), ("nmat", C.c_int),
        ("nlight", C.c_int),
        ("sph", _D), ("sph_ml", _I), ("quad", _D), ("quad_ml", _I), ("plane_y", _D), ("plane_ml", _I),
        ("sdf", _D), ("mat", _D), ("light", _D), ("cie", C.POINTER(C.c_float)),
        ("has_env", C.c_int), ("env", C.c_double), ("n_lambda", C.c_int), ("lmin", C.c_double),
        ("lrange", C.c_double), ("rr_depth", C.c_int), ("seed", C.c_uint32),
        ("W", C.c_int), ("H", C.c_int), ("max_depth", C.c_int), ("sdf_steps", C.c_int),
        ("cam_pos", C.c_double * 3), ("cam_angle", C.c_double * 2), ("cam_size", C.c_double), ("ap_dist", C.c_double),
        ("camera", C.c_int), ("ap_size", C.c_double), ("lens_dist", C.c_double), ("lens_radius", C.c_double),
        ("lens_focal", C.c_double), ("lens_thick", C.c_double), ("nee", C.c_int), ("lens_missed", C.c_uint64),
        ("nlens", C.c_int), ("ncyc", C.c_int), ("lens", _D), ("cyc", _D), ("nlight_nee", C.c_int),
    ]


def _setup(L: C.CDLL) -> C.CDLL:
    L.ptext_render.restype = C.c_int
    L.ptext_render.argtypes = [C.POINTER(Scene), C.c_int, C.c_int, C.c_int, C.c_int, _D, _D,
                               C.POINTER(C.c_uint64), C.c_int]
    for fn in ("ptext_fresnel_dielectric", "ptext_fresnel_conductor", "ptext_planck"):
        getattr(L, fn).restype = C.c_double
        getattr(L, fn).argtypes = [C.c_double] * 3
    for fn in ("ptext_probe_quad", "ptext_probe_sphere"):
        getattr(L, fn).restype = C.c_int
        getattr(L, fn).argtypes = [_D, _D, _D, _D]
    for fn in ("ptext_probe_ior", "ptext_probe_reflectance"):
        getattr(L, fn).restype = C.c_double
        getattr(L, fn).argtypes = [_D, C.c_double]
    L.ptext_probe_march.restype = C.c_double
    L.ptext_probe_march.argtypes = [_D, _D, _D, C.c_int, C.POINTER(C.c_uint64)]
    L.ptext_probe_de.restype = C.c_double
    L.ptext_probe_de.argtypes = [_D, _D]
    for fn in ("ptext_probe_lens", "ptext_probe_cyclide"):
        getattr(L, fn).restype = C.c_int
        getattr(L, fn).argtypes = [_D, _D, _D, _D, _D]
    L.ptext_probe_poly_roots.restype = C.c_int
    L.ptext_probe_poly_roots.argtypes = [_D, C.c_int, C.c_double, C.c_double, _D]
    L.ptext_probe_camera.restype = C.c_int
    L.ptext_probe_camera.argtypes = [C.POINTER(Scene), C.c_uint32, C.c_double, C.c_double, C.c_double, _D, _D]
    return L


_setup(lib)


def load(path: str | os.PathLike) -> C.CDLL:
    """Another build of ptext.c (a planted-mistake variant in a test)."""
    return _setup(C.CDLL(str(path)))


SDF_KINDS = {"mandelbulb": 1, "menger": 2, "blob": 3, "terrain": 4}


def _rotation_glsl(angle_deg):
    """GLSL RotationMatrix: M = mX * mY * mZ with column-major constructors."""
    x, y, z = (a * math.pi / 180.0 for a in angle_deg)
    mX = np.array([[1, 0, 0], [0, math.cos(x), math.sin(x)], [0, -math.sin(x), math.cos(x)]])
    mY = np.array([[math.cos(y), 0, -math.sin(y)], [0, 1, 0], [math.sin(y), 0, math.cos(y)]])
    mZ = np.array([[math.cos(z), math.sin(z), 0], [-math.sin(z), math.cos(z), 0], [0, 0, 1]])
    return mX @ mY @ mZ  # as a row-major numpy matrix acting on column vectors: M v


class Flat:
    """A scene JSON flattened into the arrays of oracle/ptext.c."""

    def __init__(self, text: str, W: int, H: int, depth: int, cie: np.ndarray, sdf_steps: int = 256):
        sc = json.loads(text)
        rend = sc.get("render", {})
        lights_json = sc.get("light", [])
        sph, sph_ml, quad, quad_ml, lights = [], [], [], [], []

        def light_of(lid, typ, prim):
            if lid <= 0:
                return -1
            e = lights_json[lid - 1]["emission"]
            lights.append([typ, prim, float(e["temperature"]), float(e["luminosity"])])
            return len(lights) - 1

        for o in sc.get("sphere", []):
            sph_ml.append([o["materialID"] - 1, light_of(o["lightID"], 0, len(sph))])
            sph.append([*map(float, o["position"]), float(o["radius"])])

        def add_quad(Q, e1, e2, m, l):
            if np.linalg.norm(np.cross(e1, e2)) == 0:
                return
            quad_ml.append([m - 1, light_of(l, 1, len(quad))])
            quad.append([*Q, *e1, *e2])

        for q in sc.get("quad", []):
            add_quad(list(map(float, q["position"])), list(map(float, q["edge1"])), list(map(float, q["edge2"])),
                     q["materialID"], q["lightID"])
        for b in sc.get("box", []):
            M = _rotation_glsl(b["rotation"])
            hs = np.array(b["size"], float) / 2
            ax, ay, az = M[:, 0], M[:, 1], M[:, 2]
            X, Y, Z = ax * 2 * hs[0], ay * 2 * hs[1], az * 2 * hs[2]
            o = np.array(b["position"], float) - ax * hs[0] - ay * hs[1] - az * hs[2]
            o2 = o + X + Y + Z
            for Q, e1, e2 in [(o, Z, Y), (o, X, Z), (o, Y, X), (o2, -Y, -Z), (o2, -Z, -X), (o2, -X, -Y)]:
                add_quad(list(Q), list(e1), list(e2), b["materialID"], b["lightID"])
        nlight_nee = len(lights)  # spheres and parallelograms: sampled by NEE
        lens, cyc = [], []
        for o in sc.get("lens", []):
            li = light_of(o["lightID"], 2, len(lens))
            lens.append([*map(float, o["position"]), *map(float, o["rotation"]), float(o["radius"]),
                         float(o["focalLength"]), float(o["thickness"]), 1.0 if o["isConverging"] else 0.0,
                         o["materialID"] - 1, li])
        for o in sc.get("cyclide", []):
            li = light_of(o["lightID"], 3, len(cyc))
            cyc.append([*map(float, o["position"]), *map(float, o["rotation"]), *map(float, o["scale"]),
                        float(o["a"]), float(o["b"]), float(o["c"]), float(o["d"]), float(o["boundingRadius"]),
                        o["materialID"] - 1, li])
        planes = sc.get("plane", [])
        plane_y = [float(p["position"][1]) for p in planes]
        plane_ml = [[p["materialID"] - 1, -1] for p in planes]
        sdfs = []
        for s in sc.get("sdf", []):
            if "kind" not in s:
                raise ValueError("the FP64 oracle reads SDFs in the {'kind': ...} form")
            mat = s.get("material", 0.0)
            mat = mat[0] if isinstance(mat, list) else mat
            sdfs.append([SDF_KINDS[s["kind"]], s.get("iterations", 4 if s["kind"] == "mandelbulb" else 5),
                         s.get("bailout", 16.0), math.floor(mat), *map(float, s["position"]),
                         *map(float, s["boundingSize"])])
        mats = []
        for m in sc["material"]:
            t = m.get("type", "lambert")
            row = [0.0] * 12
            if t == "lambert" and "gaussians" in m:
                row[0], row[1] = 1, len(m["gaussians"])
                for i, g in enumerate(m["gaussians"]):
                    row[3 + 3 * i:6 + 3 * i] = g
            elif t == "lambert":
                r = m["reflection"]
                row[0] = 0
                row[3:6] = [r["peakWavelength"], r["sigma"], float(r["isInvert"])]
            elif t == "dielectric":
                row[0] = 2
                if "cauchy" in m:
                    row[1] = 0
                    row[3:5] = m["cauchy"]
                    row[2] = 1.0 if m["cauchy"][1] != 0 else 0.0
                else:
                    row[1] = 1
                    row[3:9] = m["sellmeier"]
                    row[2] = 1.0
            else:
                row[0] = 3
                row[3:5] = [m["n"], m["k"]]
            mats.append(row)
        cam = sc["camera"]
        self._keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(np.array(x, dtype=dt).reshape(-1) if len(x) else np.zeros(1, dt))
            self._keep.append(a)
            return a.ctypes.data_as(_D if dt == np.float64 else _I)

        s = Scene()
        s.nsph, s.nquad, s.nplane, s.nsdf, s.nmat, s.nlight = len(sph), len(quad), len(planes), len(sdfs), len(mats), len(lights)
        s.sph, s.sph_ml = arr(sph, np.float64), arr(sph_ml, np.int32)
        s.quad, s.quad_ml = arr(quad, np.float64), arr(quad_ml, np.int32)
        s.plane_y, s.plane_ml = arr(plane_y, np.float64), arr(plane_ml, np.int32)
        s.sdf, s.mat, s.light = arr(sdfs, np.float64), arr(mats, np.float64), arr(lights, np.float64)
        s.nlens, s.ncyc, s.nlight_nee = len(lens), len(cyc), nlight_nee
        s.lens, s.cyc = arr(lens, np.float64), arr(cyc, np.float64)
        self.cie = np.ascontiguousarray(cie, dtype=np.float32)
        s.cie = self.cie.ctypes.data_as(C.POINTER(C.c_float))
        env = sc.get("environment")
        s.has_env, s.env = (1, float(env["radiance"])) if env else (0, 0.0)
        s.n_lambda = int(rend.get("wavelengths", 4))
        s.lmin = float(rend.get("lambdaMin", 380.0))
        s.lrange = float(rend.get("lambdaMax", 780.0)) - s.lmin
        s.rr_depth = int(rend.get("russianRouletteDepth", 3))
        s.seed = int(rend.get("seed", 0)) & 0xFFFFFFFF
        s.W, s.H, s.max_depth, s.sdf_steps = W, H, depth, sdf_steps
        pos, ang = cam["position"][0], cam["angle"][0]
        s.cam_pos[:] = pos
        s.cam_angle[:] = [-ang[1], ang[0]]  # push-constant swizzle (src/pathtracer.cpp:3821)
        s.cam_size, s.ap_dist = float(cam["size"]), float(cam["apertureDistance"])
        s.camera = 1 if cam.get("model", "lens") == "pinhole" else 0
        s.ap_size, s.lens_dist = float(cam["apertureSize"]), float(cam["lensDistance"])
        s.lens_radius, s.lens_focal = float(cam["lensRadius"]), float(cam["lensFocalLength"])
        s.lens_thick = float(cam["lensThickness"])
        s.nee = 1
        self.exposure = float(cam["apertureSize"]) ** 2 * float(cam["ISO"])
        self.scene = s


def cie_table() -> np.ndarray:
    """The CIE 1931 table, read from the product's data file (data, not code)."""
    inc = (_HERE.parent / "pathtracer_amd" / "csrc" / "cie1931_2deg.inc").read_text()
    vals = [float(t.strip().rstrip("f")) for line in inc.splitlines() if "nm" in line and line.strip()[0].isdigit()
            for t in line.split("/*")[0].split(",") if t.strip()]
    return np.array(vals, np.float32)


def render(text: str, W: int, H: int, spp: int, depth: int, first_sample: int = 0, threads: int = 0,
           rows: tuple[int, int] | None = None, seed: int | None = None, nee: bool = True,
           L: C.CDLL | None = None) -> dict:
    """Renders rows [y0, y1) (all by default). Returns the per-pixel mean XYZ
    (exposure applied), the per-pixel sample variance of the mean, counters and
    seconds. nee=False renders with BSDF sampling alone."""
    if seed is not None:
        sc = json.loads(text)
        sc.setdefault("render", {})["seed"] = seed
        text = json.dumps(sc)
    f = Flat(text, W, H, depth, cie_table())
    f.scene.nee = 1 if nee else 0
    y0, y1 = rows if rows else (0, H)
    n = W * (y1 - y0) * 3
    sums = np.zeros(n, np.float64)
    sq = np.zeros(n, np.float64)
    cnt = (C.c_uint64 * 3)()
    t = time.perf_counter()
    st = (L or lib).ptext_render(C.byref(f.scene), first_sample, spp, y0, y1, sums.ctypes.data_as(_D),
                                 sq.ctypes.data_as(_D), cnt, threads or (os.cpu_count() or 1))
    dt = time.perf_counter() - t
    if st != 0:
        raise RuntimeError("ptext_render failed")
    mean = (sums / spp).reshape(y1 - y0, W, 3)
    var = np.maximum(sq / spp - (sums / spp) ** 2, 0.0).reshape(y1 - y0, W, 3) / spp
    return {"mean": mean * f.exposure, "var_of_mean": var * f.exposure ** 2, "segments": int(cnt[0]),
            "shadow_segments": int(cnt[1]), "sdf_steps": int(cnt[2]), "seconds": dt,
            "lens_missed": int(f.scene.lens_missed)}


def camera_ray(text: str, W: int, H: int, key: int, uvx: float, uvy: float, l0: float, L: C.CDLL | None = None):
    """The lens camera's ray for one (key, uv, hero wavelength): (ok, origin, direction)."""
    f = Flat(text, W, H, 1, cie_table())
    o, d = (C.c_double * 3)(), (C.c_double * 3)()
    ok = (L or lib).ptext_probe_camera(C.byref(f.scene), key & 0xFFFFFFFF, uvx, uvy, l0, o, d)
    return bool(ok), np.array(o[:]), np.array(d[:])


def material_row(m: dict) -> np.ndarray:
    """One material as the oracle's 12-double row (the Flat layout)."""
    f = Flat(json.dumps({"camera": {"position": [[0, 0, 0]], "angle": [[0, 0]], "size": 0.057,
                                    "apertureDistance": 0.049, "apertureSize": 0.0025, "ISO": 1,
                                    "lensDistance": 0.05, "lensRadius": 0.01, "lensFocalLength": 0.03,
                                    "lensThickness": 0.0},
                         "material": [m]}), 1, 1, 1, cie_table())
    return np.ctypeslib.as_array(f.scene.mat, shape=(12,)).copy()
