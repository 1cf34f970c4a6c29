"""Pins the FP64 oracle's lenses and Dupin cyclides (oracle/ptext.c hit_lens,
hit_cyclide, poly_roots) against independent mathematics, and its handling of
emitters that next-event estimation does not sample:

- the polynomial root isolation against numpy's companion-matrix eigenvalues
  for quartics built from known roots;
- a converging lens of zero thickness is the intersection of two balls: the
  first change of that inside test along a finely sampled ray brackets the
  oracle's hit, and its normal points out of the solid;
- diverging lenses: the hit lies on a cap sphere, inside that cap, and the
  normal points into the sphere (out of the solid between the caps);
- a cyclide hit is a zero of the implicit quartic of src/shader.comp:633-679,
  no sign change of that function lies before it on a fine grid, numpy's
  roots of the same ray's polynomial find the same first root, and the
  normal is the finite-difference gradient of the world-space function;
- with an emitting lens and cyclide (found by BSDF sampling only) beside an
  area light, the image with next-event estimation equals the image with BSDF
  sampling alone within the Monte Carlo bound.

Each check fails on a planted mistake."""
import json
import math

import numpy as np
import pytest

from oracle import ptext
from pathtracer_amd import scenes
from tests.test_oracle_ext_pins import block, dv, fails, planted

_D = ptext._D


def rot_row(ang, v):
    """v * RotationMatrix(ang) (degrees), the reference's world-to-object map:
    rotations about x, then y, then z applied to the row vector."""
    x, y, z = (a * math.pi / 180.0 for a in ang)
    v = np.asarray(v, float)
    r1 = np.array([v[0], v[1] * math.cos(x) - v[2] * math.sin(x), v[1] * math.sin(x) + v[2] * math.cos(x)])
    r2 = np.array([r1[0] * math.cos(y) + r1[2] * math.sin(y), r1[1], -r1[0] * math.sin(y) + r1[2] * math.cos(y)])
    return np.array([r2[0] * math.cos(z) - r2[1] * math.sin(z), r2[0] * math.sin(z) + r2[1] * math.cos(z), r2[2]])


# ---------------------------------------------------------------------------
# polynomial roots

def check_poly_roots(L):
    L = L or ptext.lib
    rng = np.random.default_rng(5)
    tested = 0
    for _ in range(400):
        k = rng.integers(0, 5)  # real roots in the window
        r = np.sort(rng.uniform(-3.0, 3.0, k))
        if k > 1 and np.min(np.diff(r)) < 1e-3:
            continue
        poly = np.poly1d(r, r=True) if k else np.poly1d([1.0])
        if k % 2:  # odd: one more real root, outside the window
            poly = poly * np.poly1d([1.0, -rng.choice([-1.0, 1.0]) * rng.uniform(5.0, 8.0)])
        for _ in range((4 - k) // 2):  # fill to degree 4 with factors that have no real root
            poly = poly * np.poly1d([1.0, rng.uniform(-2, 2), rng.uniform(2.5, 6.0)])
        assert poly.order == 4
        poly = poly * rng.uniform(0.5, 2.0)
        c = poly.coeffs[::-1].copy()  # ascending
        out = np.zeros(4)
        _, pc = dv(c)
        n = L.ptext_probe_poly_roots(pc, 4, -4.0, 4.0, out.ctypes.data_as(_D))
        want = np.sort(np.real([z for z in np.roots(poly.coeffs) if abs(z.imag) < 1e-9 and -4 < z.real < 4]))
        assert n == len(want), (n, want)
        np.testing.assert_allclose(out[:n], want, atol=1e-9)
        tested += 1
    assert tested > 300


def test_poly_roots_match_companion_eigenvalues():
    check_poly_roots(None)


# ---------------------------------------------------------------------------
# lenses

def lens_rec(pos, ang, r, f, th, conv, mat=0, light=-1):
    return np.array([*pos, *ang, r, f, th, 1.0 if conv else 0.0, mat, light], float)


def probe(fn, rec, o, d):
    t, n = np.zeros(1), np.zeros(3)
    _, pq = dv(rec)
    _, po = dv(o)
    _, pd = dv(d)
    hit = fn(pq, po, pd, t.ctypes.data_as(_D), n.ctypes.data_as(_D))
    return hit, float(t[0]), n


def check_converging_lens(L):
    L = L or ptext.lib
    rng = np.random.default_rng(9)
    hits = 0
    for _ in range(300):
        pos, ang = rng.uniform(-1, 1, 3), rng.uniform(-60, 60, 3)
        f = rng.uniform(0.6, 2.0)
        r = rng.uniform(0.2, 0.95) * 2 * f
        R = 2 * f
        h = R - math.sqrt(R * R - r * r)
        rec = lens_rec(pos, ang, r, f, 0.0, True)
        o = pos + rng.normal(size=3) * 3.0
        aim = pos + rng.uniform(-0.8, 0.8, 3) * np.array([h, r, r])
        d = (aim - o) / np.linalg.norm(aim - o)
        hit, t, n = probe(L.ptext_probe_lens, rec, o, d)
        # inside = in both balls (centres on the object's x axis at +-(R - h))
        ts = np.linspace(0.0, 12.0, 240001)
        # rot_row is linear: apply it to o - pos and d once
        lo, ld = rot_row(ang, o - pos), rot_row(ang, d)
        Pl = lo[None, :] + ts[:, None] * ld[None, :]
        c0, c1 = np.array([R - h, 0, 0]), np.array([-(R - h), 0, 0])
        inside = (np.sum((Pl - c0) ** 2, 1) <= R * R) & (np.sum((Pl - c1) ** 2, 1) <= R * R)
        change = np.nonzero(inside[1:] != inside[:-1])[0]
        if len(change) == 0:
            assert not hit
            continue
        assert hit, "a ray through the lens solid must hit it"
        tb = ts[change[0]]
        assert abs(t - tb) <= 1e-4 + 1e-9, (t, tb)
        # the normal points out of the solid
        Pw = o + d * t
        outer = Pw + n * 1e-6
        ol = rot_row(ang, outer - pos)
        assert not (np.sum((ol - c0) ** 2) < R * R and np.sum((ol - c1) ** 2) < R * R)
        assert abs(np.linalg.norm(n) - 1) < 1e-12
        hits += 1
    assert hits > 100


def check_diverging_lens(L):
    L = L or ptext.lib
    rng = np.random.default_rng(10)
    hits = 0
    for _ in range(300):
        pos, ang = rng.uniform(-1, 1, 3), rng.uniform(-60, 60, 3)
        f = rng.uniform(0.6, 2.0)
        r = rng.uniform(0.2, 0.95) * 2 * f
        th = rng.uniform(0.0, 0.4)
        R = 2 * f
        h = R - math.sqrt(R * R - r * r)
        rec = lens_rec(pos, ang, r, f, th, False)
        o = pos + rng.normal(size=3) * 3.0
        aim = pos + rng.uniform(-0.8, 0.8, 3) * np.array([th + h, r, r])
        d = (aim - o) / np.linalg.norm(aim - o)
        hit, t, n = probe(L.ptext_probe_lens, rec, o, d)
        if not hit:
            continue
        pl = rot_row(ang, o + d * t - pos)
        nl = rot_row(ang, n)
        sp = -0.5 * th
        cs = [np.array([R - sp, 0, 0]), np.array([sp - R, 0, 0])]
        k = int(np.argmin([abs(np.linalg.norm(pl - c) - R) for c in cs]))
        assert abs(np.linalg.norm(pl - cs[k]) - R) < 1e-9 * max(1, R)
        xr = pl[0] - cs[k][0]
        assert (xr <= -(R - h) + 1e-9) if k == 0 else (xr >= R - h - 1e-9)
        # normal towards the cap sphere's centre
        np.testing.assert_allclose(nl, (cs[k] - pl) / R, atol=1e-9)
        hits += 1
    assert hits > 100


def test_lens_hits():
    check_converging_lens(None)
    check_diverging_lens(None)


@pytest.mark.parametrize("find,replace,check", [
    ("            if (k == 0 ? !(xr <= -(R - h)) : !(xr >= R - h)) continue;",
     "            if (k == 0 ? !(xr <= (R - h)) : !(xr >= R - h)) continue;", "converging"),
    ("        double cx = k == 0 ? R - sp : sp - R;", "        double cx = k == 0 ? R - sp : R - sp;", "converging"),
    ("            *n = rot_col(ang, conv ? nl : mul(nl, -1.0));", "            *n = rot_col(ang, nl);", "diverging"),
    ("    double R = 2.0 * f, h = R - sqrt(R * R - r * r), sp = conv ? 0.5 * th + h : -0.5 * th;",
     "    double R = 2.0 * f, h = R - sqrt(R * R - r * r), sp = conv ? 0.5 * th + h : 0.5 * th;", "diverging"),
], ids=["cap_side", "centre", "diverging_normal", "diverging_slice"])
def test_lens_checks_catch_planted_mistakes(tmp_path, find, replace, check):
    L = planted(tmp_path, find, replace)
    assert fails({"converging": check_converging_lens, "diverging": check_diverging_lens}[check], L)


# ---------------------------------------------------------------------------
# cyclides

def cyc_rec(pos, ang, scale, a, b, c, d, brad, mat=0, light=-1):
    return np.array([*pos, *ang, *scale, a, b, c, d, brad, mat, light], float)


def cyclide_F(rec, P):
    """The implicit function at world points P (n x 3)."""
    pos, ang, sc = rec[0:3], rec[3:6], rec[6:9]
    A, B, C, D = rec[9:13]
    # the object map is linear: rotate the unit axes once
    Mr = np.array([rot_row(ang, e) for e in np.eye(3)])  # row k = e_k * RotationMatrix
    loc = ((np.atleast_2d(P) - pos) @ Mr) / sc
    x, y, z = loc[:, 0], loc[:, 1], loc[:, 2]
    return (x * x + y * y + z * z + B * B - D * D) ** 2 - 4 * (A * x - C * D) ** 2 - 4 * B * B * z * z


# the reference scenes' cyclide parameters (tests/golden/scenes), with other placements
GOLDEN_CYC = dict(a=3.36, b=-3.17, c=-1.06, d=-1.5)


def check_cyclide(L):
    L = L or ptext.lib
    rng = np.random.default_rng(12)
    hits = 0
    for it in range(160):
        pos, ang = rng.uniform(-1, 1, 3), rng.uniform(-60, 60, 3)
        s = rng.uniform(0.15, 0.35)
        sc = np.array([s, s, s]) * (rng.uniform(0.8, 1.25, 3) if it % 2 else 1.0)
        rec = cyc_rec(pos, ang, sc, **GOLDEN_CYC, brad=6.0)
        o = pos + rng.normal(size=3) * 2.2
        aim = pos + rng.uniform(-0.6, 0.6, 3)
        d = (aim - o) / np.linalg.norm(aim - o)
        hit, t, n = probe(L.ptext_probe_cyclide, rec, o, d)
        # the ray inside the bounding sphere
        oc = o - pos
        bb, cc = oc @ d, oc @ oc - 6.0
        disc = bb * bb - cc
        if disc <= 0:
            assert not hit
            continue
        ta, tb = max(-bb - math.sqrt(disc), 1e-9), -bb + math.sqrt(disc)
        ts = np.linspace(ta, tb, 20001)
        F = cyclide_F(rec, o[None, :] + ts[:, None] * d[None, :])
        change = np.nonzero(np.sign(F[1:]) != np.sign(F[:-1]))[0]
        # numpy's roots of the polynomial fitted exactly through 5 samples
        tt = np.linspace(ta, tb, 5)
        coef = np.polyfit(tt, cyclide_F(rec, o[None, :] + tt[:, None] * d[None, :]), 4)
        rr = sorted(z.real for z in np.roots(coef) if abs(z.imag) < 1e-7 * max(1, abs(z.real)) and ta < z.real < tb)
        if len(change) == 0:
            continue  # tangent or no crossing at this sampling: nothing to bracket
        assert hit, "a sign change of F inside the bounding sphere must be a hit"
        assert ts[change[0]] - 1e-9 <= t <= ts[change[0] + 1] + 1e-9, (t, ts[change[0]])
        if rr:
            assert abs(t - rr[0]) < 1e-6 * max(1, abs(t)), (t, rr)
        # F vanishes at the hit (relative to its gradient) and n is the gradient
        P = o + d * t
        e = 1e-6
        g = np.array([(cyclide_F(rec, P + e * u) - cyclide_F(rec, P - e * u))[0] / (2 * e) for u in np.eye(3)])
        assert abs(cyclide_F(rec, P)[0]) <= 1e-7 * np.linalg.norm(g) * max(1, t)
        np.testing.assert_allclose(n, g / np.linalg.norm(g), atol=1e-5)
        hits += 1
    assert hits > 60


def test_cyclide_hits():
    check_cyclide(None)


@pytest.mark.parametrize("find,replace", [
    ("    c[2] = s1 * s1 + 2.0 * s2 * s0 - 4.0 * l1 * l1 - 4.0 * B * B * m1 * m1;",
     "    c[2] = s1 * s1 + 2.0 * s2 * s0 - 4.0 * l1 * l1;"),
    ("    double l1 = A * v.x, l0 = A * p.x - Cc * D, m1 = v.z, m0 = p.z;",
     "    double l1 = A * v.x, l0 = A * p.x - Cc * D, m1 = v.y, m0 = p.y;"),
    ("    d3 g = D3(4.0 * x.x * S - 8.0 * A * (A * x.x - Cc * D), 4.0 * x.y * S, 4.0 * x.z * S - 8.0 * B * B * x.z);",
     "    d3 g = D3(4.0 * x.x * S - 8.0 * A * (A * x.x - Cc * D), 4.0 * x.z * S, 4.0 * x.y * S - 8.0 * B * B * x.y);"),
    ("    *n = unit(rot_col(ang, D3(g.x / sc.x, g.y / sc.y, g.z / sc.z)));",
     "    *n = unit(rot_col(ang, g));"),
    ("    double tt = roots[0];", "    double tt = roots[poly_roots(c, 4, ta, tb, roots) - 1];"),
], ids=["coefficient", "swizzle", "normal_swizzle", "normal_scale", "last_root"])
def test_cyclide_check_catches_planted_mistakes(tmp_path, find, replace):
    L = planted(tmp_path, find, replace)
    assert fails(check_cyclide, L)


# ---------------------------------------------------------------------------
# emitters that NEE does not sample

def emitting_shapes_cornell():
    s = scenes.cornell_dict(4, 0)
    s["material"][4] = {"type": "lambert", "gaussians": [[0.6, 600.0, 80.0]]}
    s["sphere"] = s["sphere"][:1]
    s["light"].append({"emission": {"temperature": 3500.0, "luminosity": 6.0}})
    s["light"].append({"emission": {"temperature": 6500.0, "luminosity": 4.0}})
    nl = len(s["light"])
    s["lens"] = [{"position": [0.7, 0.45, 0.5], "rotation": [0.0, 20.0, 0.0], "radius": 0.12,
                  "focalLength": 0.15, "thickness": 0.0, "isConverging": True, "materialID": 1, "lightID": nl - 1}]
    s["cyclide"] = [{"position": [0.3, 0.3, 0.45], "rotation": [0.0, 0.0, 0.0], "scale": [0.04, 0.04, 0.04],
                     **GOLDEN_CYC, "boundingRadius": 0.09, "materialID": 2, "lightID": nl}]
    return json.dumps(s)


def check_unsampled_emitters(L):
    text = emitting_shapes_cornell()
    W = H = 24
    on = ptext.render(text, W, H, 1024, 8, seed=303, L=L)
    off = ptext.render(text, W, H, 4096, 8, seed=404, nee=False, L=L)
    v = on["var_of_mean"] + off["var_of_mean"]
    z_img = (on["mean"].mean((0, 1)) - off["mean"].mean((0, 1))) / np.sqrt(v.sum((0, 1)) / (W * H) ** 2)
    assert np.all(np.abs(z_img) < 3.5), z_img
    zb = (block(on["mean"], 4) - block(off["mean"], 4)) / np.sqrt(block(v, 4) / 16)
    assert (np.abs(zb) <= 4).mean() >= 0.97, np.abs(zb).max()


def test_emitting_lens_and_cyclide_nee_equals_bsdf_sampling():
    check_unsampled_emitters(None)


@pytest.mark.parametrize("find,replace", [
    ("            if (!specular && s->nee && li < s->nlight_nee) {", "            if (!specular && s->nee) {"),
    ("                int k = (int)(rnd(key, b, 3) * s->nlight_nee);\n                if (k >= s->nlight_nee) k = s->nlight_nee - 1;",
     "                int k = (int)(rnd(key, b, 3) * s->nlight);\n                if (k >= s->nlight) k = s->nlight - 1;"),
], ids=["mis_weight", "light_choice"])
def test_unsampled_emitter_check_catches_planted_mistakes(tmp_path, find, replace):
    L = planted(tmp_path, find, replace)
    assert fails(check_unsampled_emitters, L)
