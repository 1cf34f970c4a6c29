"""Pins the FP64 extended-mode oracle (oracle/ptext.c) function by function
against independent mathematics: parallelogram and sphere hits against a
linear solve and the quadratic formula, dispersion against catalogue indices,
the 3-Gaussian reflectance against its closed forms, the lens camera against
paraxial ray-transfer optics, sphere tracing against brute-force root finding
of the distance estimator, and the light-sampling pdfs by rendering the same
image with next-event estimation and with BSDF sampling alone (two
estimators that share no light pdf must agree within the Monte Carlo bound).

Each group compiles ptext.c with a planted mistake and requires its check to
fail."""
import json
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import ptext
from pathtracer_amd import scenes

ROOT = Path(__file__).resolve().parent.parent
_D = ptext._D


def planted(tmp_path: Path, find: str, replace: str):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    src = (ROOT / "oracle" / "ptext.c").read_text()
    assert src.count(find) == 1, f"planted text must occur once: {find!r}"
    (tmp_path / "ptext.c").write_text(src.replace(find, replace))
    so = tmp_path / "libptext_planted.so"
    subprocess.run([gcc, "-O2", "-std=gnu11", "-fPIC", "-fopenmp", "-w", "-shared", "-o", str(so),
                    str(tmp_path / "ptext.c"), "-lm"], check=True, timeout=120)
    return ptext.load(so)


def fails(check, L) -> bool:
    try:
        check(L)
    except AssertionError:
        return True
    return False


def dv(x):
    a = np.ascontiguousarray(x, np.float64)
    return a, a.ctypes.data_as(_D)


# ---------------------------------------------------------------------------
# parallelogram and sphere hits

def check_quad(L):
    L = L or ptext.lib
    rng = np.random.default_rng(41)
    n_hit = 0
    for _ in range(3000):
        Q = rng.uniform(-1, 1, 3)
        e1, e2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        o = rng.uniform(-3, 3, 3)
        aim = Q + rng.uniform(-0.3, 1.3) * e1 + rng.uniform(-0.3, 1.3) * e2
        d = (aim - o) / np.linalg.norm(aim - o)
        q = np.concatenate([Q, e1, e2])
        t = np.zeros(1)
        _, pq = dv(q)
        _, po = dv(o)
        _, pd = dv(d)
        hit = L.ptext_probe_quad(pq, po, pd, t.ctypes.data_as(_D))
        # o + t d = Q + a e1 + b e2, solved as a 3x3 linear system
        M = np.stack([d, -e1, -e2], 1)
        if abs(np.linalg.det(M)) < 1e-6:
            continue
        tt, a, b = np.linalg.solve(M, Q - o)
        inside = 0 <= a <= 1 and 0 <= b <= 1 and tt > 1e-9
        if min(abs(a), abs(a - 1), abs(b), abs(b - 1)) < 1e-9:
            continue
        assert bool(hit) == inside, (a, b, tt)
        if inside:
            assert abs(t[0] - tt) < 1e-9 * (1 + abs(tt))
            n_hit += 1
    assert n_hit > 800


def check_sphere(L):
    L = L or ptext.lib
    rng = np.random.default_rng(42)
    n_hit = 0
    for _ in range(2000):
        c, r = rng.uniform(-1, 1, 3), rng.uniform(0.1, 1)
        o = rng.uniform(-3, 3, 3)
        d = c + rng.uniform(-1.3, 1.3, 3) * r - o
        d /= np.linalg.norm(d)
        roots = np.roots([1.0, 2 * d @ (o - c), (o - c) @ (o - c) - r * r])
        t = np.zeros(1)
        hit = L.ptext_probe_sphere(dv(np.concatenate([c, [r]]))[1], dv(o)[1], dv(d)[1], t.ctypes.data_as(_D))
        real = sorted(x.real for x in roots if abs(x.imag) < 1e-12 and x.real > 1e-9)
        assert bool(hit) == bool(real)
        if real:
            assert abs(t[0] - real[0]) < 1e-9 * (1 + real[0])
            n_hit += 1
    assert n_hit > 500


def test_parallelogram_and_sphere_hits():
    check_quad(None)
    check_sphere(None)


@pytest.mark.parametrize("find,replace,check", [
    ("    if (a < 0 || a > 1 || b < 0 || b > 1) return 0;", "    if (a < 0 || a > 1 || b < 0 || b > 2) return 0;", "quad"),
    ("    double a = dot(cross(hp, e2), n) / nn, b = dot(cross(e1, hp), n) / nn;",
     "    double a = dot(cross(hp, e2), n) / nn, b = dot(cross(hp, e1), n) / nn;", "quad"),
    ("    if (t0 > EPS && t0 < tmax) { *t = t0; return 1; }", "    if (t0 > EPS && t0 < tmax) { *t = t1; return 1; }",
     "sphere"),
], ids=["edge_range", "barycentric_sign", "near_root"])
def test_hit_checks_catch_planted_mistakes(tmp_path, find, replace, check):
    assert fails({"quad": check_quad, "sphere": check_sphere}[check], planted(tmp_path, find, replace))


# ---------------------------------------------------------------------------
# dispersion and reflectance

BK7 = [1.03961212, 0.231792344, 1.01046945, 6.00069867e-3, 2.00179144e-2, 103.560653]


def check_ior(L):
    L = L or ptext.lib
    cauchy = ptext.material_row({"type": "dielectric", "cauchy": [1.5, 0.01]})
    # Cauchy: A + B / lambda^2 with lambda in micrometres
    for lnm in (400.0, 500.0, 656.27):
        assert abs(L.ptext_probe_ior(dv(cauchy)[1], lnm) - (1.5 + 0.01 / (lnm * 1e-3) ** 2)) < 1e-12
    assert abs(L.ptext_probe_ior(dv(cauchy)[1], 500.0) - 1.54) < 1e-12
    # Sellmeier with Schott's N-BK7 coefficients reproduces the catalogue
    # indices n_F, n_d, n_C (486.13, 587.56, 656.27 nm)
    sell = ptext.material_row({"type": "dielectric", "sellmeier": BK7})
    for lnm, n in ((486.13, 1.5224), (587.56, 1.5168), (656.27, 1.5143)):
        assert abs(L.ptext_probe_ior(dv(sell)[1], lnm) - n) < 1e-4, lnm
    # Abbe number of N-BK7 (catalogue 64.17)
    nF, nd, nC = (L.ptext_probe_ior(dv(sell)[1], x) for x in (486.13, 587.56, 656.27))
    assert abs((nd - 1) / (nF - nC) - 64.17) < 0.2


def check_reflectance(L):
    L = L or ptext.lib
    g = [[0.3, 450.0, 20.0], [0.5, 560.0, 15.0], [0.15, 680.0, 25.0]]
    m = ptext.material_row({"type": "lambert", "gaussians": g})
    f = lambda x: L.ptext_probe_reflectance(dv(m)[1], x)
    for w, mu, s in g:
        assert abs(f(mu) - w) < 1e-5  # the other lobes are > 4.5 sigma away
        # one standard deviation off the mean: w exp(-1/2)
        assert abs(f(mu + s) - w * math.exp(-0.5)) < 5e-5
    # the area under each lobe is w sigma sqrt(2 pi)
    lam = np.linspace(300.0, 900.0, 60001)
    area = np.trapezoid([f(x) for x in lam], lam)
    assert abs(area - sum(w * s * math.sqrt(2 * math.pi) for w, mu, s in g)) < 1e-3
    # the sum is clamped at 1
    big = ptext.material_row({"type": "lambert", "gaussians": [[0.8, 550.0, 50.0], [0.8, 560.0, 50.0]]})
    assert L.ptext_probe_reflectance(dv(big)[1], 555.0) == 1.0
    # the reference's one-lobe spectrum exp(-((l - peak) / (2 sigma^2))^2), optionally inverted
    lobe = ptext.material_row({"reflection": {"peakWavelength": 550.0, "sigma": 6.5, "isInvert": True}})
    x = (600.0 - 550.0) / (2 * 6.5 ** 2)
    assert abs(L.ptext_probe_reflectance(dv(lobe)[1], 600.0) - (1 - math.exp(-x * x))) < 1e-12


def test_dispersion_matches_catalogue_indices():
    check_ior(None)


def test_three_gaussian_reflectance_closed_forms():
    check_reflectance(None)


@pytest.mark.parametrize("find,replace,check", [
    ("    if ((int)m[1] == 0) return m[3] + m[4] / l2;", "    if ((int)m[1] == 0) return m[3] + m[4] / sqrt(l2);", "ior"),
    ("    return sqrt(1.0 + m[3] * l2 / (l2 - m[6]) + m[4] * l2 / (l2 - m[7]) + m[5] * l2 / (l2 - m[8]));",
     "    return sqrt(1.0 + m[3] * l2 / (l2 - m[6]) + m[4] * l2 / (l2 - m[8]) + m[5] * l2 / (l2 - m[7]));", "ior"),
    ("        s += m[3 + 3 * i] * exp(-0.5 * d * d);", "        s += m[3 + 3 * i] * exp(-d * d);", "refl"),
], ids=["cauchy_lambda", "sellmeier_terms", "gaussian_width"])
def test_spectral_checks_catch_planted_mistakes(tmp_path, find, replace, check):
    assert fails({"ior": check_ior, "refl": check_reflectance}[check], planted(tmp_path, find, replace))


# ---------------------------------------------------------------------------
# the reference's lens camera

def bk7(lnm):
    l2 = (lnm * 1e-3) ** 2
    return math.sqrt(1 + BK7[0] * l2 / (l2 - BK7[3]) + BK7[1] * l2 / (l2 - BK7[4]) + BK7[2] * l2 / (l2 - BK7[5]))


def lens_scene(angle=(30.0, -20.0), thick=0.0):
    d = scenes.cornell_dict(4, 0)
    d["camera"] = scenes._camera([0.5, 0.5, -1.6], angle, model="lens")
    d["camera"]["lensThickness"] = thick
    return d


def fwd_of(angle):
    """The camera's forward axis: RotationMatrix(pitch, yaw, 0) applied to +z,
    with the push constants' (pitch, yaw) = (-angle[1], angle[0])."""
    x, y = math.radians(-angle[1]), math.radians(angle[0])
    r1 = np.array([0.0, -math.sin(x), math.cos(x)])
    return np.array([r1[0] * math.cos(y) + r1[2] * math.sin(y), r1[1], -r1[0] * math.sin(y) + r1[2] * math.cos(y)])


def check_lens_camera(L, thick=0.0):
    """Rays from the sensor's centre through the aperture leave the lens as if
    from one point on the axis. Its position follows from paraxial ray transfer
    through the two faces (radius 2f, vertices at -+(h + t/2)): entering
    n_a = n_BK7(l0), leaving with the inside index n_b = n_BK7(l0 / n_a) (the
    reference rescales the wavelength at the first face, src/shader.comp:1431-1437).
    Aberration grows as r^2 in the aperture radius; the fit's r -> 0 limit is
    compared."""
    angle = (30.0, -20.0)
    d = lens_scene(angle, thick)
    text = json.dumps(d)
    cam = d["camera"]
    fwd = fwd_of(angle)
    lp = np.array(cam["position"][0]) + cam["lensDistance"] * fwd
    l0 = 587.56
    rs, xs = [], []
    for key in range(1, 200):
        ok, o, dd = ptext.camera_ray(text, 32, 32, key * 2654435761, 0.0, 0.0, l0, L=L)
        assert ok
        A = dd - (dd @ fwd) * fwd
        b = (o - lp) - ((o - lp) @ fwd) * fwd
        if A @ A < 1e-20:
            continue
        s = -(A @ b) / (A @ A)
        lat = np.linalg.norm((o + s * dd - lp) - ((o + s * dd - lp) @ fwd) * fwd)
        assert lat < 1e-9  # the ray meets the axis (meridional)
        rs.append(np.linalg.norm(b))
        xs.append((o + s * dd - lp) @ fwd)
    rs, xs = np.array(rs), np.array(xs)
    k, x0 = np.polyfit(rs ** 2, xs, 1)
    assert np.abs(xs - (x0 + k * rs ** 2)).max() < 2e-4 * abs(x0)
    R = 2 * cam["lensFocalLength"]
    h = R - math.sqrt(R * R - cam["lensRadius"] ** 2)
    na = bk7(l0)
    nb = bk7(l0 / na)
    th = 1e-6
    v1 = -(h + thick / 2)  # first vertex, relative to the lens centre
    y1 = th * (cam["lensDistance"] + v1)
    t1 = (th - (na - 1) * y1 / R) / na
    y2 = y1 + t1 * (2 * h + thick)
    t2 = nb * t1 - (nb - 1) * y2 / R
    want = -v1 - y2 / t2
    assert abs(x0 - want) < 1e-3 * abs(want), (x0, want)


def test_lens_camera_images_like_paraxial_optics():
    check_lens_camera(None)
    check_lens_camera(None, thick=0.002)


@pytest.mark.parametrize("find,replace", [
    ("        l *= eta;\n", "\n"),
    ("                bn = entering ? rad : mul(rad, -1.0);", "                bn = rad;"),
    ("    double h = R - sqrt(R * R - s->lens_radius * s->lens_radius);\n    double half_t",
     "    double h = 0.0;\n    double half_t"),
], ids=["wavelength_rescale", "normal_orientation", "sagitta"])
def test_lens_check_catches_planted_mistakes(tmp_path, find, replace):
    assert fails(check_lens_camera, planted(tmp_path, find, replace))


# ---------------------------------------------------------------------------
# sphere tracing

def sdf_rec(kind, iters, pos, size):
    return np.array([ptext.SDF_KINDS[kind], iters, 16.0, 0.0, *pos, *size], np.float64)


def check_march(L):
    L = L or ptext.lib
    rng = np.random.default_rng(43)
    n_hit = n_miss = 0
    for kind, iters in (("blob", 0), ("menger", 3), ("mandelbulb", 8)):
        q = sdf_rec(kind, iters, (0.0, 1.0, 0.0), (3.0, 4.0, 3.0) if kind == "blob" else (2.0, 2.0, 2.0))
        qa, qp = dv(q)
        de = lambda p: L.ptext_probe_de(qp, dv(p)[1])
        for _ in range(60):
            o = np.array([rng.uniform(-4, 4), rng.uniform(-2, 4), -5.0])
            aim = q[4:7] + rng.uniform(-1.2, 1.2, 3)
            d = (aim - o) / np.linalg.norm(aim - o)
            steps = np.zeros(1, np.uint64)
            t = L.ptext_probe_march(qp, dv(o)[1], dv(d)[1], 256, steps.ctypes.data_as(ptext.C.POINTER(ptext.C.c_uint64)))
            ts = np.arange(0.0, 12.0, 1e-3)
            lo, hi = q[4:7] - q[7:10] / 2, q[4:7] + q[7:10] / 2
            pts = o[None, :] + ts[:, None] * d[None, :]
            inb = np.all((pts >= lo) & (pts <= hi), axis=1)
            vals = np.array([de(p) for p in pts[inb][::3]])
            tin = ts[inb][::3]
            below = np.nonzero(vals < 1e-4)[0]
            if t < 0:
                if len(below) == 0:
                    n_miss += 1
                else:
                    # a miss may only come from running out of steps
                    assert steps[0] >= 256 or kind == "mandelbulb", (kind, tin[below[0]])
                continue
            p = o + t * d
            assert de(p) < 1e-4
            # the first point of the ray within 1e-4 of the surface, to the sampling step
            if len(below):
                assert t <= tin[below[0]] + 3e-3, (kind, t, tin[below[0]])
                assert t >= tin[below[0]] - 3e-3 or kind == "mandelbulb"
            n_hit += 1
    assert n_hit > 60 and n_miss > 10, (n_hit, n_miss)


def test_sphere_tracing_finds_the_first_crossing():
    check_march(None)


@pytest.mark.parametrize("find,replace", [
    ("        if (dist < 1e-4) return t;", "        if (dist < 1e-2) return t;"),
    ("        t += dist;\n", "        t += 2.0 * dist;\n"),
], ids=["converge_eps", "overstep"])
def test_march_check_catches_planted_mistakes(tmp_path, find, replace):
    assert fails(check_march, planted(tmp_path, find, replace))


# ---------------------------------------------------------------------------
# light pdfs: NEE + MIS against BSDF sampling alone

def no_caustic_cornell(sphere_light=False):
    s = scenes.cornell_dict(4, 0)
    s["material"][4] = {"type": "lambert", "gaussians": [[0.6, 600.0, 80.0]]}
    if sphere_light:  # a second light: a small emitting sphere under the ceiling
        s["sphere"].append({"position": [0.25, 0.8, 0.6], "radius": 0.06, "materialID": 6, "lightID": 2})
        s["light"].append({"emission": {"temperature": 4000.0, "luminosity": 20.0}})
    return json.dumps(s)


def block(a, b):
    h, w = a.shape[0] // b * b, a.shape[1] // b * b
    return a[:h, :w].reshape(h // b, b, w // b, b, -1).mean(axis=(1, 3))


def check_nee_agrees_with_bsdf_sampling(L, sphere_light=False):
    text = no_caustic_cornell(sphere_light)
    W = H = 24
    on = ptext.render(text, W, H, 1024, 8, seed=101, L=L)
    off = ptext.render(text, W, H, 4096, 8, seed=202, nee=False, L=L)
    v = on["var_of_mean"] + off["var_of_mean"]
    z_img = (on["mean"].mean((0, 1)) - off["mean"].mean((0, 1))) / np.sqrt(v.sum((0, 1)) / (W * H) ** 2)
    assert np.all(np.abs(z_img) < 3.5), z_img
    zb = (block(on["mean"], 4) - block(off["mean"], 4)) / np.sqrt(block(v, 4) / 16)
    assert (np.abs(zb) <= 4).mean() >= 0.97, np.abs(zb).max()


def test_area_light_pdf_nee_equals_bsdf_sampling():
    check_nee_agrees_with_bsdf_sampling(None, sphere_light=False)


def test_sphere_light_pdf_nee_equals_bsdf_sampling():
    check_nee_agrees_with_bsdf_sampling(None, sphere_light=True)


@pytest.mark.parametrize("find,replace,sphere", [
    ("        pdf = dist * dist / (cl * area);", "        pdf = dist / (cl * area);", False),
    ("        double cl = fabs(dot(nn, dir)) / area;", "        double cl = 1.0;", False),
    ("        pdf = 1.0 / (2.0 * M_PI * (1.0 - cmax));", "        pdf = 1.0 / (M_PI * (1.0 - cmax));", True),
    ("    return pdf / s->nlight_nee;", "    return pdf;", True),
], ids=["quad_distance", "quad_cosine", "sphere_cone", "light_choice"])
def test_nee_check_catches_planted_mistakes(tmp_path, find, replace, sphere):
    L = planted(tmp_path, find, replace)
    assert fails(lambda L_: check_nee_agrees_with_bsdf_sampling(L_, sphere), L)
