"""The scene loader against the JSON library the reference itself uses.

oracle/_ref/json_probe is built (by oracle/Makefile, where the reference tree is
present) from a probe of this repository and the single header the reference
vendors, nlohmann-json 3.11.3. It prints, per number or bool leaf, the binary32
that get<float>() returns and the int that get<int>() returns: what
UpdateFromJSON (src/pathtracer.cpp:2576-2722) stores in its float and int
fields. The uniform block and the camera push constants packed by pt.pack_json
must hold exactly those values, on the 11 shipped scenes and on scenes whose
literals are hard for a decimal parser: long mantissas, binary32 ties reached
through the binary64 parse, denormals, and integers beyond int64 and uint64."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pathtracer_amd as pt
from tests.conftest import SCENES
from tests.test_scene import py_pack

ROOT = Path(__file__).resolve().parent.parent
PROBE = ROOT / "oracle" / "_ref" / "json_probe"

pytestmark = pytest.mark.skipif(not PROBE.exists(), reason="oracle/_ref/json_probe not built (needs the reference tree)")

# fields that UpdateFromJSON stores in ints (src/pathtracer.cpp:110-182)
INT_KEYS = {"materialID", "lightID", "ISO", "numShots"}


def nlohmann_values(text: str, tmp_path: Path) -> dict:
    f = tmp_path / "scene.json"
    f.write_text(text)
    r = subprocess.run([str(PROBE), str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        ptr, bits, i = line.split(" ")
        out[ptr] = (np.uint32(int(bits, 16)).view(np.float32), None if i == "x" else int(i))
    return out


def as_read_by_reference(node, vals: dict, ptr: str = ""):
    """The scene with each number or bool leaf replaced by what the reference's
    field of that name holds."""
    if isinstance(node, dict):
        return {k: as_read_by_reference(v, vals, f"{ptr}/{k}") for k, v in node.items()}
    if isinstance(node, list):
        return [as_read_by_reference(v, vals, f"{ptr}/{i}") for i, v in enumerate(node)]
    if isinstance(node, (bool, int, float)):
        f, i = vals[ptr]
        if ptr.rsplit("/", 1)[-1] in INT_KEYS:
            assert i is not None, ptr
            return i
        return f
    return node


def check_against_nlohmann(text: str, tmp_path: Path):
    vals = nlohmann_values(text, tmp_path)
    scene = as_read_by_reference(json.loads(text), vals)
    ubo, _, _ = pt.pack_json(text)
    got = np.frombuffer(bytes(ubo), np.float32)[:2774]
    want = py_pack(scene)[:2774]
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, [(int(k), float(got[k]), float(want[k])) for k in bad[:8]]
    cam = scene["camera"]
    for shot in range(1, cam["numShots"] + 1):
        _, pc, _ = pt.pack_json(text, shot)
        pos, ang = cam["position"][shot - 1], cam["angle"][shot - 1]
        # cameraAngle = (-angle.y, angle.x) (src/pathtracer.cpp:3821)
        got_cam = np.array([pc.cameraPosX, pc.cameraPosY, pc.cameraPosZ, *pc.cameraAngle], np.float32)
        want_cam = np.array([*pos, -ang[1], ang[0]], np.float32)
        assert np.array_equal(got_cam.view(np.uint32), want_cam.view(np.uint32)), (shot, got_cam, want_cam)
        assert pc.ISO == cam["ISO"]
        for name, key in (("apertureDist", "apertureDistance"), ("lensFocalLength", "lensFocalLength")):
            assert np.float32(getattr(pc, name)).view(np.uint32) == np.float32(cam[key]).view(np.uint32), name


@pytest.mark.parametrize("path", SCENES, ids=lambda p: p.stem)
def test_shipped_scenes_read_as_nlohmann_reads_them(path, tmp_path):
    check_against_nlohmann(path.read_text(), tmp_path)


HARD_REALS = [
    "0.1000000000000000055511151231257827",
    "0.333333333333333314829616256247390992939472198486328125",
    "2.5000000000000004",
    # just above the binary32 tie between 1 and 1 + 2^-23: binary64 rounds it onto
    # the tie, which binary32 then breaks to even (1.0)
    "1.00000005960464477539062500000000001",
    "1e-46",
    "7.006492321624085e-46",
    "1.401298464324817e-45",
    "1.1754942106924411e-38",
    "3.4028234663852886e38",
    "4.9406564584124654e-324",
    "-0.0",
    "-0",
    "0e0",
    "1E+2",
    "-2.5e-3",
    "123456789012345678901234567890",
    "16777217",
    "16777219",
    "-16777217",
    "9223372036854775807",
    "9223372036854775808",
    # 2^63 + 2^39 + 1: as an unsigned 64-bit integer it rounds up to 2^63 + 2^40 in
    # binary32; through binary64 it lands on the tie and rounds to 2^63
    "9223372586610589697",
    "18446744073709551615",
    "18446744073709551616",
    "-9223372036854775808",
    "-9223372036854775809",
    "-9223372586610589697",
]


def with_literals(scene: dict, slots: list, literals: list) -> str:
    """scene as JSON text with the leaves at `slots` (key paths) replaced by the
    given literals, verbatim."""
    for n, path in enumerate(slots):
        node = scene
        for k in path[:-1]:
            node = node[k]
        node[path[-1]] = f"@@{n}@@"
    text = json.dumps(scene, indent=1)
    for n, lit in enumerate(literals):
        text = text.replace(f'"@@{n}@@"', lit)
    return text


def float_slots(scene: dict) -> list:
    slots = []
    for i, s in enumerate(scene["sphere"]):
        slots += [("sphere", i, "position", k) for k in range(3)] + [("sphere", i, "radius")]
    for i, _ in enumerate(scene["plane"]):
        slots += [("plane", i, "position", k) for k in range(3)]
    for i, _ in enumerate(scene["material"]):
        slots += [("material", i, "reflection", "peakWavelength"), ("material", i, "reflection", "sigma")]
    for i, _ in enumerate(scene["light"]):
        slots += [("light", i, "emission", "temperature"), ("light", i, "emission", "luminosity")]
    for k in range(3):
        slots.append(("camera", "position", 0, k))
    slots += [("camera", "angle", 0, 0), ("camera", "angle", 0, 1), ("camera", "apertureDistance"),
              ("camera", "lensFocalLength")]
    return slots


@pytest.mark.parametrize("shift", range(4))
def test_hard_literals_read_as_nlohmann_reads_them(scene_texts, tmp_path, shift):
    base = json.loads(scene_texts["scene0"])
    slots = float_slots(base)
    lits = [HARD_REALS[(n + shift * 7) % len(HARD_REALS)] for n in range(len(slots))]
    # every literal lands somewhere across the four shifts
    check_against_nlohmann(with_literals(base, slots, lits), tmp_path)


def test_real_literals_in_int_fields(scene_texts, tmp_path):
    base = json.loads(scene_texts["scene0"])
    slots = [("sphere", 0, "materialID"), ("sphere", 1, "materialID"), ("sphere", 0, "lightID"),
             ("camera", "ISO"), ("plane", 0, "materialID")]
    lits = ["1.0", "0.99999", "0e0", "1600.7", "1E0"]
    check_against_nlohmann(with_literals(base, slots, lits), tmp_path)


def test_probe_sees_the_double_rounding_case(tmp_path):
    """The probe itself: an integer literal beyond int64 converts straight to
    binary32 (rounding up), while the same value written as a real literal goes
    through binary64 and lands on the tie (rounding to even)."""
    vals = nlohmann_values('{"a": 9223372586610589697, "b": 9223372586610589697.0}', tmp_path)
    assert vals["/a"][0] == np.float32(2.0 ** 63 + 2.0 ** 40)
    assert vals["/b"][0] == np.float32(2.0 ** 63)


def reference_save(text: str, tmp_path: Path) -> str:
    """text re-serialised as the reference's SaveScene writes a document:
    scene.dump(4, ' ', true) (src/pathtracer.cpp:3465-3470)."""
    f = tmp_path / "saved.json"
    f.write_bytes(text.encode("utf-8"))
    r = subprocess.run([str(PROBE), "--dump", str(f)], capture_output=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.decode("utf-8")


def generated_scenes():
    from pathtracer_amd import scenes
    return {"cornell": scenes.cornell(), "spheres": scenes.random_spheres(n=64), "sdf": scenes.sdf_scene()}


@pytest.mark.parametrize("path", SCENES, ids=lambda p: p.stem)
def test_saved_text_is_byte_identical_to_the_references(path, tmp_path):
    """The writer's text (UpdateToJSON's values) is the library's own layout of
    that document: numbers, escapes, indentation and empty containers."""
    ours = pt.json_rewrite(path.read_text())
    assert ours == reference_save(ours, tmp_path)


@pytest.mark.parametrize("name", ["cornell", "spheres", "sdf"])
def test_saved_extended_scenes_use_the_same_layout(name, tmp_path):
    ours = pt.json_rewrite(generated_scenes()[name])
    assert ours == reference_save(ours, tmp_path)


# binary32 values over every layout the serializer has: 0.000ddd down to 1e-4,
# d.ddde-XX below it, ddd.dd, digits with ".0" up to 15 places, d.ddde+XX above
LAYOUT_VALUES = [1.4e-45, 1.17549435e-38, 1e-30, 1.5e-5, 1e-5, 9.99e-5, 1e-4, 1.23e-4, 0.00123, 0.1, 0.5, 1.0,
                 1.0000001, 3.1415927, 100.0, 550.0, 16777216.0, 123456.79, 1e15, 1.2345678e15, 1e16, 9.9e37,
                 3.4028235e38]


def test_number_layouts(tmp_path):
    from pathtracer_amd import scenes
    s = scenes.cornell_dict()
    for a, b in zip(LAYOUT_VALUES[::2], LAYOUT_VALUES[1::2] + [2.0]):
        s["material"].append({"type": "conductor", "n": a, "k": b})
    s["environment"] = {"radiance": 6.0e-8}
    ours = pt.json_rewrite(json.dumps(s))
    assert ours == reference_save(ours, tmp_path)
    # every value was written, and reads back as the same binary32
    back = json.loads(ours)["material"][-len(LAYOUT_VALUES[::2]):]
    got = [v for m in back for v in (m["n"], m["k"])][:len(LAYOUT_VALUES)]
    assert np.array_equal(np.float32(got), np.float32(LAYOUT_VALUES))


def test_strings_escaped_as_the_reference_saves_them(scene_texts, tmp_path):
    """ensure_ascii: code points from 0x7F up as \\uXXXX (a surrogate pair above
    U+FFFF), control characters as \\u00XX or their two-character escapes."""
    s = json.loads(scene_texts["scene4"])
    s["sdf"][0]["glsl"] += "\n// café € \U0001d11e \x7f \x01 \x1f tab\there \"q\" back\\slash /"
    ours = pt.json_rewrite(json.dumps(s, ensure_ascii=False))
    assert ours.isascii() and "\\ud834\\udd1e" in ours and "\\u007f" in ours
    assert ours == reference_save(ours, tmp_path)
    assert json.loads(ours)["sdf"][0]["glsl"] == s["sdf"][0]["glsl"]


def wrap_extra_key(base: bytes, val: bytes) -> bytes:
    """A valid scene with one more top-level key holding `val` verbatim (both
    the reference and the loader ignore unknown keys)."""
    i = base.index(b"{")
    return base[:i + 1] + b'"x": ' + val + b"," + base[i + 1:]


# name -> a value placed under an extra top-level key of a valid scene
EDGE_VALUES = {
    "lone_low_surrogate": b'"\\udc00"',
    "lone_high_surrogate": b'"\\ud800x"',
    "high_then_ascii_escape": b'"\\ud800\\u0041"',
    "surrogate_pair": b'"\\ud834\\udd1e"',
    "raw_byte_ff": b'"a\xffb"',
    "overlong_nul": b'"\xc0\x80"',
    "encoded_surrogate": b'"\xed\xa0\x80"',
    "above_u10ffff": b'"\xf4\x90\x80\x80"',
    "four_byte_ok": '"\U0001d11e"'.encode(),
    "truncated_utf8": b'"\xe2\x82"',
    "stray_continuation": b'"\x80"',
    "del": b'"\x7f"',
    "escaped_nul": b'"\\u0000"',
    "raw_control": b'"\x01"',
    "leading_zero": b"01",
    "minus_only": b"-",
    "trailing_dot": b"1.",
    "leading_plus": b"+1",
    "trailing_comma": b"[1,2,]",
    "comment": b"/* c */ 1",
    "nan": b"NaN",
    "overflow": b"1e400",
    "negative_overflow": b"-1e400",
    "underflow": b"1e-400",
    "huge_integer": b"123456789012345678901234567890",
    "formfeed_whitespace": b"\x0c1",
    "upper_hex_escape": b'"\\u00E9"',
    "escaped_slash": b'"\\/"',
    "bad_escape": b'"\\a"',
    "capital_true": b"True",
    "null": b"null",
}


def probe_accepts(doc: bytes, tmp_path: Path) -> bool:
    f = tmp_path / "doc.json"
    f.write_bytes(doc)
    r = subprocess.run([str(PROBE), str(f)], capture_output=True, timeout=60)
    assert r.returncode in (0, 1), r.stderr
    return r.returncode == 0


def loader_accepts(doc: bytes) -> bool:
    import ctypes as C
    n = C.c_size_t()
    st = pt.lib.pt_json_rewrite(doc, len(doc), None, 0, C.byref(n))
    assert st in (pt.PT_OK, pt.PT_ERR_JSON), st
    return st == pt.PT_OK


def edge_documents(base: bytes) -> dict:
    docs = {k: wrap_extra_key(base, v) for k, v in EDGE_VALUES.items()}
    docs.update({"byte_order_mark": b"\xef\xbb\xbf" + base, "trailing_garbage": base + b" x",
                 "trailing_whitespace": base + b" \n\t\r ", "two_documents": base + base, "empty": b"",
                 "whitespace_only": b"  "})
    return docs


def test_accepts_and_rejects_what_nlohmann_does(tmp_path):
    """Strings (escapes, surrogates, RFC 3629 UTF-8), numbers (forms, overflow),
    a byte order mark and trailing content: the loader accepts a document
    exactly when the reference's parser does."""
    base = (ROOT / "tests" / "golden" / "scenes" / "scene0.json").read_bytes()
    diff = {name: (probe_accepts(doc, tmp_path), loader_accepts(doc))
            for name, doc in edge_documents(base).items()}
    assert all(a == b for a, b in diff.values()), {k: v for k, v in diff.items() if v[0] != v[1]}
    assert diff["byte_order_mark"] == (True, True) and diff["overflow"] == (False, False)


def test_repeated_key_takes_the_last_value_in_the_first_place(scene_texts, tmp_path):
    text = scene_texts["scene0"]
    assert '"ISO": 1600' in text
    doubled = text.replace('"ISO": 1600', '"ISO": 100, "ISO": 1600', 1)
    _, pc, _ = pt.pack_json(doubled)
    assert pc.ISO == 1600
    check_against_nlohmann(doubled, tmp_path)
    ours = pt.json_rewrite(doubled)
    assert ours == reference_save(ours, tmp_path)


def test_nesting_limit_is_the_one_deliberate_difference(tmp_path):
    """nlohmann parses any depth; the loader stops at 256 levels so that a
    hostile document cannot exhaust the calling thread's stack."""
    base = (ROOT / "tests" / "golden" / "scenes" / "scene0.json").read_bytes()
    ok = wrap_extra_key(base, b"[" * 200 + b"]" * 200)
    deep = wrap_extra_key(base, b"[" * 300 + b"]" * 300)
    assert probe_accepts(ok, tmp_path) and loader_accepts(ok)
    assert probe_accepts(deep, tmp_path) and not loader_accepts(deep)


def test_trailing_content_follows_the_strict_parse(tmp_path):
    """The reference reads its startup scene with ordered_json::parse
    (src/pathtracer.cpp:2573), which rejects content after the document, and a
    scene the user loads with `file >> json` (ReadJSON, :893-908), which stops
    after the first value. The loader follows the strict parse; the two reference
    paths differ on exactly these documents."""
    base = (ROOT / "tests" / "golden" / "scenes" / "scene0.json").read_bytes()
    f = tmp_path / "doc.json"
    for doc in (base + b" x", base + base, base + b"}"):
        f.write_bytes(doc)
        strict = subprocess.run([str(PROBE), str(f)], capture_output=True, timeout=60).returncode
        stream = subprocess.run([str(PROBE), "--stream", str(f)], capture_output=True, timeout=60).returncode
        assert (strict, stream) == (1, 0)
        assert not loader_accepts(doc)
    for name, doc in edge_documents(base).items():
        if name in ("trailing_garbage", "two_documents"):
            continue
        f.write_bytes(doc)
        strict = subprocess.run([str(PROBE), str(f)], capture_output=True, timeout=60).returncode
        stream = subprocess.run([str(PROBE), "--stream", str(f)], capture_output=True, timeout=60).returncode
        assert strict == stream, name


def json_leaves(node, ptr=""):
    if isinstance(node, dict):
        for k, v in node.items():
            yield from json_leaves(v, f"{ptr}/{k}")
    elif isinstance(node, list):
        for i, v in enumerate(node):
            yield from json_leaves(v, f"{ptr}/{i}")
    elif isinstance(node, (int, float)) and not isinstance(node, bool):
        yield ptr, node


# keys the writer saves exactly (extended schema); the reference's own float
# fields are saved on RoundDecimal's 1e-5 grid
EXACT_KEYS = {"quad", "gaussians", "cauchy", "sellmeier", "n", "k", "environment", "render", "bailout", "edge1",
              "edge2"}


@pytest.mark.parametrize("name", ["cornell", "spheres", "sdf"])
def test_extended_fields_read_as_nlohmann_reads_them(name, tmp_path):
    """The generated scenes the benchmarks render: every float the loader holds
    for an extended field (written back exactly by the writer) is nlohmann's
    get<float>() of the same literal; reference fields are within the 1e-5
    grid they are saved on."""
    text = generated_scenes()[name]
    vals = nlohmann_values(text, tmp_path)
    out = json.loads(pt.json_rewrite(text))
    exact = rounded = 0
    for ptr, v in json_leaves(out):
        if ptr not in vals:
            continue
        want = vals[ptr][0]
        if EXACT_KEYS & set(ptr.split("/")):
            assert np.float32(v).view(np.uint32) == want.view(np.uint32), ptr
            exact += 1
        else:
            assert abs(float(v) - float(want)) <= 0.5e-5 + abs(float(want)) * 2.0 ** -23, ptr
            rounded += 1
    assert exact > 10 and rounded > 10
