"""Scene JSON loader, uniform-buffer packer, camera push constants and SDF
recogniser, against an independent restatement of UpdateFromJSON +
UpdateUniformBuffer (src/pathtracer.cpp:2576-2722, 3642-3834) written here
with Python's json module, on the reference's 11 shipped scenes."""
import json
import re

import numpy as np
import pytest

import pathtracer_amd as pt
from tests.conftest import SCENES

f32 = np.float32


def py_pack(scene: dict) -> np.ndarray:
    """UpdateUniformBuffer restated: returns the 4097 floats of the block."""
    sph, pln, box = scene.get("sphere", []), scene.get("plane", []), scene.get("box", [])
    lns, cyc, sdf = scene.get("lens", []), scene.get("cyclide", []), scene.get("sdf", [])
    objs, lids = [], []
    for i, o in enumerate(sph):
        objs += [*o["position"], o["radius"], o["materialID"], o["lightID"]]
        if o["lightID"] > 0:
            lids.append(i)
    for i, o in enumerate(pln):
        objs += [*o["position"], o["materialID"], o["lightID"]]
        if o["lightID"] > 0:
            lids.append(len(sph) + i)
    for i, o in enumerate(box):
        objs += [*o["position"], *o["rotation"], *o["size"], o["materialID"], o["lightID"]]
        if o["lightID"] > 0:
            lids.append(len(sph) + len(pln) + i)
    for i, o in enumerate(lns):
        objs += [*o["position"], *o["rotation"], o["radius"], o["focalLength"], o["thickness"],
                 float(o["isConverging"]), o["materialID"], o["lightID"]]
        # the reference tests planes[i].lightID here (src/pathtracer.cpp:3704)
        if i < len(pln) and pln[i]["lightID"] > 0:
            lids.append(len(sph) + len(pln) + len(box) + i)
    for i, o in enumerate(cyc):
        m = max(f32(s) for s in o["scale"])
        brad = f32(o["boundingRadius"])
        objs += [*o["position"], *o["rotation"], *o["scale"], o["a"], o["b"], o["c"], o["d"],
                 (brad * brad) * (m * m), o["materialID"], o["lightID"]]
        if i < len(pln) and pln[i]["lightID"] > 0:
            lids.append(len(sph) + len(pln) + len(box) + len(lns) + i)
    u = np.zeros(4097, np.float32)
    u[0:7] = [len(sph), len(pln), len(box), len(lns), len(cyc), len(sdf), len(lids)]
    u[7:7 + len(objs)] = np.array([f32(v) for v in objs], np.float32)
    s = []
    for o in sdf:
        s += [*o["position"], *o["boundingSize"]]
    u[1031:1031 + len(s)] = s
    mats = []
    for m in scene["material"]:
        r = m["reflection"]
        mats += [r["peakWavelength"], r["sigma"], float(r["isInvert"])]
    u[1799:1799 + len(mats)] = mats
    lts = []
    for l in scene["light"]:
        lts += [l["emission"]["temperature"], l["emission"]["luminosity"]]
    u[2582:2582 + len(lts)] = lts
    u[2710:2710 + len(lids)] = lids
    return u


@pytest.mark.parametrize("path", SCENES, ids=lambda p: p.stem)
def test_pack_matches_independent_restatement(path):
    text = path.read_text()
    ubo, pc, sdfs = pt.pack_json(text)
    got = np.frombuffer(bytes(ubo), np.float32)
    want = py_pack(json.loads(text))
    assert np.array_equal(got[:2774], want[:2774])  # everything but the CIE table
    # CIE 1931 table: 441 rows of (x, y, z) at 1 nm from 360 nm; y-bar(555) = 1
    cie = got[2774:].reshape(441, 3)
    assert cie[555 - 360, 1] == np.float32(1.0)
    assert cie[0, 0] == np.float32(0.0001299)


@pytest.mark.parametrize("path", SCENES, ids=lambda p: p.stem)
def test_camera_push_constants(path):
    scene = json.loads(path.read_text())
    cam = scene["camera"]
    for shot in range(1, cam["numShots"] + 1):
        _, pc, _ = pt.pack_json(path.read_text(), shot)
        pos, ang = cam["position"][shot - 1], cam["angle"][shot - 1]
        # cameraAngle = (-angle.y, angle.x) (src/pathtracer.cpp:3821)
        assert list(pc.cameraAngle) == [-f32(ang[1]), f32(ang[0])]
        assert [pc.cameraPosX, pc.cameraPosY, pc.cameraPosZ] == [f32(v) for v in pos]
        assert pc.ISO == cam["ISO"]
        assert pc.apertureDist == f32(cam["apertureDistance"])
        assert pc.lensFocalLength == f32(cam["lensFocalLength"])
    with pytest.raises(pt.PTError):
        pt.pack_json(path.read_text(), cam["numShots"] + 1)


EXPECTED_SDF = {
    "scene3": (pt.SDF_BLOB, pt.SDFMAT_BLOB, [3, 4, 5, 5]),
    "scene4": (pt.SDF_MANDELBULB, pt.SDFMAT_CONSTANT, [3]),
    "scene5": (pt.SDF_BLOB, pt.SDFMAT_BLOB, [3, 4, 5, 5]),
    "scene6": (pt.SDF_MANDELBULB, pt.SDFMAT_RADIAL, [3, 2, 0.8]),
    "scene7": (pt.SDF_MANDELBULB, pt.SDFMAT_RADIAL, [5, 4, 0.8]),
    "scene8": (pt.SDF_TERRAIN, pt.SDFMAT_CONSTANT, [0]),
    "scene9": (pt.SDF_MANDELBULB, pt.SDFMAT_RADIAL, [4, 3, 0.8]),
    "scene10": (pt.SDF_MENGER, pt.SDFMAT_CONSTANT, [0]),
}


@pytest.mark.parametrize("path", SCENES, ids=lambda p: p.stem)
def test_sdf_recognised(path):
    _, _, sdfs = pt.pack_json(path.read_text())
    want = EXPECTED_SDF.get(path.stem)
    if want is None:
        assert sdfs == []
        return
    (d,) = sdfs
    assert (d.kind, d.material_kind) == want[:2]
    np.testing.assert_allclose(list(d.material)[: len(want[2])], want[2], rtol=1e-7)
    if d.kind == pt.SDF_MANDELBULB:
        assert (d.iterations, d.bailout) == (4, 16.0)
    if d.kind == pt.SDF_MENGER:
        assert d.iterations == 5


def test_sdf_recognise_variants_and_rejects(scene_texts):
    mb = json.loads(scene_texts["scene4"])["sdf"][0]["glsl"]
    d = pt.recognize_sdf(mb.replace("i < 4", "i < 8").replace("r > 16.0", "r > 8.0"))
    assert (d.kind, d.iterations, d.bailout) == (pt.SDF_MANDELBULB, 8, 8.0)
    mg = json.loads(scene_texts["scene10"])["sdf"][0]["glsl"]
    assert pt.recognize_sdf(mg.replace("m < 5", "m < 4")).iterations == 4
    for bad in [mb.replace("7.0", "6.0"), mg.replace("3.0 * abs", "2.0 * abs"), "float sdf(in vec3 p) { return length(p) - 1.0; }",
                "not glsl at all {"]:
        with pytest.raises(pt.PTError) as e:
            pt.recognize_sdf(bad)
        assert e.value.status == pt.PT_ERR_UNSUPPORTED_SDF


def test_json_errors(scene_texts):
    with pytest.raises(pt.PTError) as e:
        pt.pack_json('{"camera": ')
    assert e.value.status == pt.PT_ERR_JSON and "line 1" in str(e.value)
    s = json.loads(scene_texts["scene0"])
    del s["camera"]["ISO"]
    with pytest.raises(pt.PTError) as e:
        pt.pack_json(json.dumps(s))
    assert e.value.status == pt.PT_ERR_JSON and "ISO" in str(e.value)
    s = json.loads(scene_texts["scene4"])
    s["sdf"][0]["glsl"] = "float sdf(in vec3 p) { return p.y; } float sdfmaterial(in vec3 p) { return 0.0; }"
    with pytest.raises(pt.PTError) as e:
        pt.pack_json(json.dumps(s))
    assert e.value.status == pt.PT_ERR_UNSUPPORTED_SDF


def test_capacity_overflow(scene_texts):
    s = json.loads(scene_texts["scene0"])
    s["sphere"] = [dict(s["sphere"][0]) for _ in range(171)]  # 171 * 6 > 1024 floats
    with pytest.raises(pt.PTError) as e:
        pt.pack_json(json.dumps(s))
    assert e.value.status == pt.PT_ERR_CAPACITY
    s["sphere"] = s["sphere"][:170]
    ubo, _, _ = pt.pack_json(json.dumps({**s, "plane": [], "box": [], "lens": [], "cyclide": []}))
    assert ubo.numObjects[0] == 170


def test_integer_and_real_literals_convert_like_nlohmann(scene_texts):
    s = json.loads(scene_texts["scene0"])
    text = json.dumps(s).replace('"radius": 1.0', '"radius": 1').replace('"ISO": 1600', '"ISO": 1600.7')
    ubo, pc, _ = pt.pack_json(text)
    assert ubo.objects[3] == 1.0 and pc.ISO == 1600
    # a real literal is parsed to binary64 and then rounded to binary32
    text2 = json.dumps(s).replace('"radius": 1.0', '"radius": 0.1000000000000000055511151231257827')
    ubo2, _, _ = pt.pack_json(text2)
    assert ubo2.objects[3] == np.float32(0.1)


SHAPE_CHANGES = [
    # valid GLSL computing another shape: each must be rejected, not rendered as the built-in
    ("scene3", "smin(s1, s2)", "smin(s1, s1)"),                  # a blob sphere dropped
    ("scene3", "), s3), s4)", "), s1), s4)"),                    # another one
    ("scene4", "cost2 = cost * cost", "cost2 = cost * cosp"),    # another operand
    ("scene4", "dot(vec2(z.x, z.z), vec2(z.x, z.z))", "dot(vec2(z.x, z.x), vec2(z.x, z.z))"),
    ("scene10", "vec3 q = abs(p)", "vec3 q = abs(-p)"),
    ("scene8", "float sdf(in vec3 p)", "float sdf(in vec3 q)"),  # the parameter renamed, the body still reads p
]


@pytest.mark.parametrize("scene,find,replace", SHAPE_CHANGES)
def test_recogniser_rejects_other_programs(scene_texts, scene, find, replace):
    glsl = json.loads(scene_texts[scene])["sdf"][0]["glsl"]
    assert find in glsl
    with pytest.raises(pt.PTError) as e:
        pt.recognize_sdf(glsl.replace(find, replace, 1))
    assert e.value.status == pt.PT_ERR_UNSUPPORTED_SDF


def test_recogniser_ignores_layout_comments_and_number_spelling(scene_texts):
    mb = json.loads(scene_texts["scene4"])["sdf"][0]["glsl"]
    d = pt.recognize_sdf(mb.replace("vec3 z = p;", "vec3   z=p; // the orbit\n/* block */"))
    assert (d.kind, d.iterations) == (pt.SDF_MANDELBULB, 4)
    mg = json.loads(scene_texts["scene10"])["sdf"][0]["glsl"]
    assert "vec3(1.0)" in mg
    assert pt.recognize_sdf(mg.replace("vec3(1.0)", "vec3(1.)")).kind == pt.SDF_MENGER


def code_positions(text: str) -> list:
    """Indices of the digits of `text` outside comments."""
    masked = re.sub(r"//[^\n]*|/\*.*?\*/", lambda m: " " * len(m.group(0)), text, flags=re.S)
    return [i for i, c in enumerate(masked) if c.isdigit()]


def desc_key(d):
    return (d.kind, d.material_kind, d.iterations, d.bailout, tuple(d.material))


@pytest.mark.parametrize("scene", ["scene3", "scene4", "scene6", "scene8", "scene10"])
def test_every_digit_change_is_a_parameter_or_rejected(scene_texts, scene):
    """Changing any one digit of the program (outside comments) either changes
    the recognised parameters (a free parameter) or is rejected: no change is
    accepted as the same built-in shape."""
    glsl = json.loads(scene_texts[scene])["sdf"][0]["glsl"]
    base = desc_key(pt.recognize_sdf(glsl))
    silent = []
    for i in code_positions(glsl):
        for c in "0123456789":
            if c == glsl[i]:
                continue
            try:
                if desc_key(pt.recognize_sdf(glsl[:i] + c + glsl[i + 1:])) == base:
                    silent.append((i, glsl[max(0, i - 20):i + 5]))
            except pt.PTError:
                pass
    assert not silent, silent[:5]
