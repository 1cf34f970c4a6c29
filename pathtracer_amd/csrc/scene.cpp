// scene.cpp — see scene.hpp.
#include "scene.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "json.hpp"

namespace pt {

static const float kCie[PT_CIE_SIZE] = {
#include "cie1931_2deg.inc"
};

const float* cie_table() { return kCie; }
void copy_cie_table(float* dst) { std::memcpy(dst, kCie, sizeof kCie); }

using ptjson::Value;

namespace {

// nlohmann's operator[] on a missing key yields null, whose size() is 0; the
// reference relies on that for absent object lists (e.g. scene8 has no "plane").
const Value& member(const Value& v, const char* key) {
    static const Value null_value;
    const Value* m = v.get(key);
    return m ? *m : null_value;
}
const Value& need(const Value& v, const char* key) {
    const Value* m = v.get(key);
    if (!m) throw std::runtime_error(std::string("missing scene key '") + key + "'");
    return *m;
}
void read3(const Value& v, const char* what, float out[3]) {
    if (v.kind != ptjson::Kind::Array || v.size() < 3)
        throw std::runtime_error(std::string("scene key '") + what + "' must be an array of 3 numbers");
    for (int k = 0; k < 3; ++k) out[k] = ptjson::as_float(v.at(k), what);
}

void load_material(const Value& m, Material& mat) {
    const Value* refl = m.get("reflection");
    if (refl) {
        mat.reflection[0] = ptjson::as_float(need(*refl, "peakWavelength"), "peakWavelength");
        mat.reflection[1] = ptjson::as_float(need(*refl, "sigma"), "sigma");
        mat.reflection[2] = ptjson::as_float(need(*refl, "isInvert"), "isInvert");
    }
    const Value* type = m.get("type");
    std::string t = type ? ptjson::as_string(*type, "type") : std::string("lambert");
    if (t == "lambert") {
        const Value* g = m.get("gaussians");
        if (g) {
            if (g->kind != ptjson::Kind::Array || g->size() < 1 || g->size() > 3)
                throw std::runtime_error("material 'gaussians' must hold 1 to 3 [weight, mean, sigma] entries");
            mat.type = MAT_LAMBERT_GAUSS;
            mat.n_gauss = (int)g->size();
            for (int i = 0; i < mat.n_gauss; ++i) read3(g->at((size_t)i), "gaussians", &mat.gauss[3 * i]);
        } else {
            if (!refl) throw std::runtime_error("lambert material needs 'reflection' or 'gaussians'");
            mat.type = MAT_LAMBERT_LOBE;
        }
    } else if (t == "dielectric") {
        mat.type = MAT_DIELECTRIC;
        const Value* c = m.get("cauchy");
        const Value* s = m.get("sellmeier");
        if (c) {
            if (c->kind != ptjson::Kind::Array || c->size() != 2)
                throw std::runtime_error("dielectric 'cauchy' must be [A, B]");
            mat.dispersion = DISP_CAUCHY;
            mat.disp[0] = ptjson::as_float(c->at(0), "cauchy");
            mat.disp[1] = ptjson::as_float(c->at(1), "cauchy");
        } else if (s) {
            if (s->kind != ptjson::Kind::Array || s->size() != 6)
                throw std::runtime_error("dielectric 'sellmeier' must be [B1, B2, B3, C1, C2, C3]");
            mat.dispersion = DISP_SELLMEIER;
            for (int k = 0; k < 6; ++k) mat.disp[k] = ptjson::as_float(s->at((size_t)k), "sellmeier");
        } else {
            throw std::runtime_error("dielectric material needs 'cauchy' or 'sellmeier'");
        }
    } else if (t == "conductor") {
        mat.type = MAT_CONDUCTOR;
        mat.cond_n = ptjson::as_float(need(m, "n"), "n");
        mat.cond_k = ptjson::as_float(need(m, "k"), "k");
    } else {
        throw std::runtime_error("unknown material type '" + t + "'");
    }
}

int load_sdf(const Value& o, RSdf& sd, std::string& err) {
    read3(need(o, "position"), "position", sd.pos);
    read3(need(o, "boundingSize"), "boundingSize", sd.size);
    std::memset(&sd.desc, 0, sizeof sd.desc);
    const Value* g = o.get("glsl");
    const Value* kind = o.get("kind");
    if (g) {
        sd.glsl = ptjson::as_string(*g, "glsl");
        int st = recognize_sdf(sd.glsl, sd.desc, err);
        if (st != PT_OK) return st;
    } else if (kind) {
        // Extended form: {"kind": "mandelbulb", "iterations": 8, "bailout": 16,
        //                 "material": 0.0 | [a, b, d] | [m0, m1, m2, m3]}
        std::string k = ptjson::as_string(*kind, "kind");
        pt_sdf_desc& d = sd.desc;
        if (k == "mandelbulb") { d.kind = PT_SDF_MANDELBULB; d.iterations = 4; d.bailout = 16.0f; }
        else if (k == "menger") { d.kind = PT_SDF_MENGER; d.iterations = 5; }
        else if (k == "blob") { d.kind = PT_SDF_BLOB; d.material_kind = PT_SDFMAT_BLOB; d.material[0] = 3; d.material[1] = 4; d.material[2] = 5; d.material[3] = 5; }
        else if (k == "terrain") { d.kind = PT_SDF_TERRAIN; }
        else { err = "unknown sdf kind '" + k + "'"; return PT_ERR_UNSUPPORTED_SDF; }
        if (const Value* it = o.get("iterations")) d.iterations = ptjson::as_int(*it, "iterations");
        if (const Value* b = o.get("bailout")) d.bailout = ptjson::as_float(*b, "bailout");
        if (const Value* m = o.get("material")) {
            if (m->is_number()) { d.material_kind = PT_SDFMAT_CONSTANT; d.material[0] = ptjson::as_float(*m, "material"); }
            else if (m->kind == ptjson::Kind::Array && m->size() == 3 && d.kind != PT_SDF_BLOB) {
                d.material_kind = PT_SDFMAT_RADIAL;
                for (int q = 0; q < 3; ++q) d.material[q] = ptjson::as_float(m->at((size_t)q), "material");
            } else if (m->kind == ptjson::Kind::Array && m->size() == 4 && d.kind == PT_SDF_BLOB) {
                for (int q = 0; q < 4; ++q) d.material[q] = ptjson::as_float(m->at((size_t)q), "material");
            } else { err = "bad sdf 'material'"; return PT_ERR_JSON; }
        }
        if (d.iterations < 0 || d.iterations > 64) { err = "sdf iterations out of range [0, 64]"; return PT_ERR_JSON; }
    } else {
        err = "sdf entry needs 'glsl' or 'kind'";
        return PT_ERR_JSON;
    }
    for (int k = 0; k < 3; ++k) { sd.desc.pos[k] = sd.pos[k]; sd.desc.size[k] = sd.size[k]; }
    return PT_OK;
}

}  // namespace

int load_scene_json(const char* text, size_t len, SceneModel& s, std::string& err) {
    Value root;
    if (!ptjson::parse(text, len, root, err)) return PT_ERR_JSON;
    if (root.kind != ptjson::Kind::Object) { err = "scene JSON must be an object"; return PT_ERR_JSON; }
    s = SceneModel{};
    try {
        // Camera (UpdateFromJSON, src/pathtracer.cpp:2577-2606).
        const Value& cam = need(root, "camera");
        int nshots = ptjson::as_int(need(cam, "numShots"), "numShots");
        if (nshots < 1) throw std::runtime_error("camera numShots must be >= 1");
        const Value& pos = need(cam, "position");
        const Value& ang = need(cam, "angle");
        s.shots.resize((size_t)nshots);
        for (int i = 0; i < nshots; ++i) {
            read3(pos.at((size_t)i), "position", s.shots[(size_t)i].pos);
            const Value& a = ang.at((size_t)i);
            s.shots[(size_t)i].angle[0] = ptjson::as_float(a.at(0), "angle");
            s.shots[(size_t)i].angle[1] = ptjson::as_float(a.at(1), "angle");
        }
        s.cam.ISO = ptjson::as_int(need(cam, "ISO"), "ISO");
        s.cam.size = ptjson::as_float(need(cam, "size"), "size");
        s.cam.apertureSize = ptjson::as_float(need(cam, "apertureSize"), "apertureSize");
        s.cam.apertureDist = ptjson::as_float(need(cam, "apertureDistance"), "apertureDistance");
        s.cam.lensRadius = ptjson::as_float(need(cam, "lensRadius"), "lensRadius");
        s.cam.lensFocalLength = ptjson::as_float(need(cam, "lensFocalLength"), "lensFocalLength");
        s.cam.lensThickness = ptjson::as_float(need(cam, "lensThickness"), "lensThickness");
        s.cam.lensDistance = ptjson::as_float(need(cam, "lensDistance"), "lensDistance");
        if (const Value* model = cam.get("model")) {
            std::string m = ptjson::as_string(*model, "model");
            if (m == "lens") s.render.camera = 0;
            else if (m == "pinhole") s.render.camera = 1;
            else throw std::runtime_error("camera model must be 'lens' or 'pinhole'");
        }

        const Value& sph = member(root, "sphere");
        for (size_t i = 0; i < sph.size(); ++i) {
            const Value& o = sph.at(i);
            RSphere r{};
            read3(need(o, "position"), "position", r.pos);
            r.radius = ptjson::as_float(need(o, "radius"), "radius");
            r.materialID = ptjson::as_int(need(o, "materialID"), "materialID");
            r.lightID = ptjson::as_int(need(o, "lightID"), "lightID");
            s.spheres.push_back(r);
        }
        const Value& pln = member(root, "plane");
        for (size_t i = 0; i < pln.size(); ++i) {
            const Value& o = pln.at(i);
            RPlane r{};
            read3(need(o, "position"), "position", r.pos);
            r.materialID = ptjson::as_int(need(o, "materialID"), "materialID");
            r.lightID = ptjson::as_int(need(o, "lightID"), "lightID");
            s.planes.push_back(r);
        }
        const Value& bx = member(root, "box");
        for (size_t i = 0; i < bx.size(); ++i) {
            const Value& o = bx.at(i);
            RBox r{};
            read3(need(o, "position"), "position", r.pos);
            read3(need(o, "rotation"), "rotation", r.rotation);
            read3(need(o, "size"), "size", r.size);
            r.materialID = ptjson::as_int(need(o, "materialID"), "materialID");
            r.lightID = ptjson::as_int(need(o, "lightID"), "lightID");
            s.boxes.push_back(r);
        }
        const Value& ln = member(root, "lens");
        for (size_t i = 0; i < ln.size(); ++i) {
            const Value& o = ln.at(i);
            RLens r{};
            read3(need(o, "position"), "position", r.pos);
            read3(need(o, "rotation"), "rotation", r.rotation);
            r.radius = ptjson::as_float(need(o, "radius"), "radius");
            r.focalLength = ptjson::as_float(need(o, "focalLength"), "focalLength");
            r.thickness = ptjson::as_float(need(o, "thickness"), "thickness");
            r.isConverging = ptjson::as_bool(need(o, "isConverging"), "isConverging");
            r.materialID = ptjson::as_int(need(o, "materialID"), "materialID");
            r.lightID = ptjson::as_int(need(o, "lightID"), "lightID");
            s.lenses.push_back(r);
        }
        const Value& cy = member(root, "cyclide");
        for (size_t i = 0; i < cy.size(); ++i) {
            const Value& o = cy.at(i);
            RCyclide r{};
            read3(need(o, "position"), "position", r.pos);
            read3(need(o, "rotation"), "rotation", r.rotation);
            read3(need(o, "scale"), "scale", r.scale);
            r.a = ptjson::as_float(need(o, "a"), "a");
            r.b = ptjson::as_float(need(o, "b"), "b");
            r.c = ptjson::as_float(need(o, "c"), "c");
            r.d = ptjson::as_float(need(o, "d"), "d");
            r.brad = ptjson::as_float(need(o, "boundingRadius"), "boundingRadius");
            r.materialID = ptjson::as_int(need(o, "materialID"), "materialID");
            r.lightID = ptjson::as_int(need(o, "lightID"), "lightID");
            s.cyclides.push_back(r);
        }
        const Value& sd = member(root, "sdf");
        for (size_t i = 0; i < sd.size(); ++i) {
            RSdf r{};
            int st = load_sdf(sd.at(i), r, err);
            if (st != PT_OK) {
                err = "sdf " + std::to_string(i) + ": " + err;
                return st;
            }
            s.sdfs.push_back(r);
        }
        const Value& qd = member(root, "quad");
        for (size_t i = 0; i < qd.size(); ++i) {
            const Value& o = qd.at(i);
            Quad q{};
            read3(need(o, "position"), "position", q.corner);
            read3(need(o, "edge1"), "edge1", q.e1);
            read3(need(o, "edge2"), "edge2", q.e2);
            q.materialID = ptjson::as_int(need(o, "materialID"), "materialID");
            q.lightID = ptjson::as_int(need(o, "lightID"), "lightID");
            s.quads.push_back(q);
        }
        const Value& mt = member(root, "material");
        for (size_t i = 0; i < mt.size(); ++i) {
            Material m;
            load_material(mt.at(i), m);
            s.materials.push_back(m);
        }
        const Value& lt = member(root, "light");
        for (size_t i = 0; i < lt.size(); ++i) {
            const Value& em = need(lt.at(i), "emission");
            Light l{};
            l.emission[0] = ptjson::as_float(need(em, "temperature"), "temperature");
            l.emission[1] = ptjson::as_float(need(em, "luminosity"), "luminosity");
            s.lights.push_back(l);
        }
        if (const Value* env = root.get("environment")) {
            s.env_enabled = true;
            s.env_radiance = ptjson::as_float(need(*env, "radiance"), "radiance");
        }
        if (const Value* r = root.get("render")) {
            if (const Value* v = r->get("wavelengths")) s.render.n_lambda = ptjson::as_int(*v, "wavelengths");
            if (const Value* v = r->get("lambdaMin")) s.render.lambda_min = ptjson::as_float(*v, "lambdaMin");
            if (const Value* v = r->get("lambdaMax")) s.render.lambda_max = ptjson::as_float(*v, "lambdaMax");
            if (const Value* v = r->get("russianRouletteDepth")) s.render.rr_depth = ptjson::as_int(*v, "russianRouletteDepth");
            if (const Value* v = r->get("seed")) s.render.seed = (uint32_t)ptjson::as_double(*v, "seed");
            if (s.render.n_lambda != 1 && s.render.n_lambda != 4)
                throw std::runtime_error("render wavelengths must be 1 or 4");
            if (!(s.render.lambda_min >= 360.0f && s.render.lambda_max <= 800.0f &&
                  s.render.lambda_min < s.render.lambda_max))
                throw std::runtime_error("render wavelength range must lie in [360, 800] nm");
        }
    } catch (const std::exception& e) {
        err = e.what();
        return PT_ERR_JSON;
    }
    return PT_OK;
}

int pack_reference_ubo(const SceneModel& s, pt_packed_scene& ubo, std::string& err) {
    std::memset(&ubo, 0, sizeof ubo);
    std::vector<float> objects, lightIDs;
    for (size_t i = 0; i < s.spheres.size(); ++i) {
        const RSphere& o = s.spheres[i];
        objects.insert(objects.end(), {o.pos[0], o.pos[1], o.pos[2], o.radius, (float)o.materialID, (float)o.lightID});
        if (o.lightID > 0) lightIDs.push_back((float)i);
    }
    for (size_t i = 0; i < s.planes.size(); ++i) {
        const RPlane& o = s.planes[i];
        objects.insert(objects.end(), {o.pos[0], o.pos[1], o.pos[2], (float)o.materialID, (float)o.lightID});
        if (o.lightID > 0) lightIDs.push_back((float)(s.spheres.size() + i));
    }
    for (size_t i = 0; i < s.boxes.size(); ++i) {
        const RBox& o = s.boxes[i];
        objects.insert(objects.end(), {o.pos[0], o.pos[1], o.pos[2], o.rotation[0], o.rotation[1], o.rotation[2],
                                       o.size[0], o.size[1], o.size[2], (float)o.materialID, (float)o.lightID});
        if (o.lightID > 0) lightIDs.push_back((float)(s.spheres.size() + s.planes.size() + i));
    }
    // The reference tests planes[i].lightID for lenses and cyclides
    // (src/pathtracer.cpp:3704, 3728); an index past the plane list is read
    // here as "not a light".
    auto plane_light = [&](size_t i) { return i < s.planes.size() ? s.planes[i].lightID : 0; };
    for (size_t i = 0; i < s.lenses.size(); ++i) {
        const RLens& o = s.lenses[i];
        objects.insert(objects.end(), {o.pos[0], o.pos[1], o.pos[2], o.rotation[0], o.rotation[1], o.rotation[2],
                                       o.radius, o.focalLength, o.thickness, (float)o.isConverging,
                                       (float)o.materialID, (float)o.lightID});
        if (plane_light(i) > 0) lightIDs.push_back((float)(s.spheres.size() + s.planes.size() + s.boxes.size() + i));
    }
    for (size_t i = 0; i < s.cyclides.size(); ++i) {
        const RCyclide& o = s.cyclides[i];
        float m = std::max(std::max(o.scale[0], o.scale[1]), o.scale[2]);
        objects.insert(objects.end(), {o.pos[0], o.pos[1], o.pos[2], o.rotation[0], o.rotation[1], o.rotation[2],
                                       o.scale[0], o.scale[1], o.scale[2], o.a, o.b, o.c, o.d,
                                       o.brad * o.brad * (m * m), (float)o.materialID, (float)o.lightID});
        if (plane_light(i) > 0)
            lightIDs.push_back((float)(s.spheres.size() + s.planes.size() + s.boxes.size() + s.lenses.size() + i));
    }
    if (s.quads.size() > 0) {
        err = "quads have no reference packing";
        return PT_ERR_CAPACITY;
    }
    if (objects.size() > PT_MAX_OBJECTS_SIZE) {
        err = "objects need " + std::to_string(objects.size()) + " floats, the reference block holds " +
              std::to_string(PT_MAX_OBJECTS_SIZE);
        return PT_ERR_CAPACITY;
    }
    if (s.sdfs.size() > PT_MAX_SDF_OBJECTS) {
        err = "more than 32 SDFs (the reference's set mask holds 32)";
        return PT_ERR_CAPACITY;
    }
    if (s.materials.size() * 3 > PT_MAX_MATERIALS_SIZE) { err = "too many materials"; return PT_ERR_CAPACITY; }
    if (s.lights.size() * 2 > PT_MAX_LIGHTS_SIZE) { err = "too many lights"; return PT_ERR_CAPACITY; }
    if (lightIDs.size() > PT_MAX_LIGHTIDS_SIZE) { err = "too many emissive objects"; return PT_ERR_CAPACITY; }

    ubo.numObjects[0] = (float)s.spheres.size();
    ubo.numObjects[1] = (float)s.planes.size();
    ubo.numObjects[2] = (float)s.boxes.size();
    ubo.numObjects[3] = (float)s.lenses.size();
    ubo.numObjects[4] = (float)s.cyclides.size();
    ubo.numObjects[5] = (float)s.sdfs.size();
    ubo.numObjects[6] = (float)lightIDs.size();
    std::copy(objects.begin(), objects.end(), ubo.objects);
    for (size_t i = 0; i < s.sdfs.size(); ++i)
        for (int k = 0; k < 3; ++k) {
            ubo.sdfs[6 * i + k] = s.sdfs[i].pos[k];
            ubo.sdfs[6 * i + 3 + k] = s.sdfs[i].size[k];
        }
    for (size_t i = 0; i < s.materials.size(); ++i)
        for (int k = 0; k < 3; ++k) ubo.materials[3 * i + k] = s.materials[i].reflection[k];
    for (size_t i = 0; i < s.lights.size(); ++i) {
        ubo.lights[2 * i] = s.lights[i].emission[0];
        ubo.lights[2 * i + 1] = s.lights[i].emission[1];
    }
    std::copy(lightIDs.begin(), lightIDs.end(), ubo.lightIDs);
    copy_cie_table(ubo.cie);
    return PT_OK;
}

int camera_to_push(const SceneModel& s, int shot, pt_push_constants& pc, std::string& err) {
    if (shot < 1 || (size_t)shot > s.shots.size()) {
        err = "camera shot " + std::to_string(shot) + " out of range [1, " + std::to_string(s.shots.size()) + "]";
        return PT_ERR_INVALID_ARG;
    }
    const CameraShot& c = s.shots[(size_t)shot - 1];
    pc.cameraAngle[0] = -c.angle[1];
    pc.cameraAngle[1] = c.angle[0];
    pc.cameraPosX = c.pos[0];
    pc.cameraPosY = c.pos[1];
    pc.cameraPosZ = c.pos[2];
    pc.ISO = s.cam.ISO;
    pc.cameraSize = s.cam.size;
    pc.apertureSize = s.cam.apertureSize;
    pc.apertureDist = s.cam.apertureDist;
    pc.lensRadius = s.cam.lensRadius;
    pc.lensFocalLength = s.cam.lensFocalLength;
    pc.lensThickness = s.cam.lensThickness;
    pc.lensDistance = s.cam.lensDistance;
    return PT_OK;
}

// ---------------------------------------------------------------------------
// SDF recogniser.
//
// The GLSL text is tokenised (comments dropped) and split into functions. The
// `sdf` body, its helpers and the `sdfmaterial` body are classified by the
// built-in functions they call and verified against the exact sequence of
// numeric literals each built-in kernel implements; the free parameters
// (iterations, escape radius, material values) are read from that sequence.
// Then the whole program must be the built-in one token for token (a 64-bit
// fingerprint of its functions with the free parameters masked), so that a
// valid program computing another shape is not rendered as the built-in one.
// Whitespace, comments and the spelling of numbers do not count. Anything else
// is rejected.
namespace {

struct Tok { std::string t; bool num; };

std::vector<Tok> tokenize(const std::string& src) {
    std::vector<Tok> out;
    size_t i = 0, n = src.size();
    while (i < n) {
        char c = src[i];
        if (std::isspace((unsigned char)c)) { ++i; continue; }
        if (c == '/' && i + 1 < n && src[i + 1] == '/') {
            while (i < n && src[i] != '\n') ++i;
            continue;
        }
        if (c == '/' && i + 1 < n && src[i + 1] == '*') {
            size_t e = src.find("*/", i + 2);
            i = (e == std::string::npos) ? n : e + 2;
            continue;
        }
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t j = i;
            while (j < n && (std::isalnum((unsigned char)src[j]) || src[j] == '_')) ++j;
            out.push_back({src.substr(i, j - i), false});
            i = j;
            continue;
        }
        if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < n && std::isdigit((unsigned char)src[i + 1]))) {
            size_t j = i;
            while (j < n && (std::isdigit((unsigned char)src[j]) || src[j] == '.')) ++j;
            if (j < n && (src[j] == 'e' || src[j] == 'E')) {
                ++j;
                if (j < n && (src[j] == '+' || src[j] == '-')) ++j;
                while (j < n && std::isdigit((unsigned char)src[j])) ++j;
            }
            if (j < n && (src[j] == 'f' || src[j] == 'F')) ++j;
            out.push_back({src.substr(i, j - i), true});
            i = j;
            continue;
        }
        static const char* two[] = {"++", "--", "<=", ">=", "==", "!=", "&&", "||", "+=", "-=", "*=", "/="};
        bool matched = false;
        for (const char* t : two)
            if (i + 1 < n && src[i] == t[0] && src[i + 1] == t[1]) {
                out.push_back({std::string(t), false});
                i += 2;
                matched = true;
                break;
            }
        if (matched) continue;
        out.push_back({std::string(1, c), false});
        ++i;
    }
    return out;
}

struct Func { std::string name; std::vector<Tok> head, body; };

// Splits `float name(...) { ... }` definitions.
bool split_functions(const std::vector<Tok>& t, std::vector<Func>& out) {
    size_t i = 0;
    while (i < t.size()) {
        if (i + 2 < t.size() && !t[i].num && !t[i + 1].num && t[i + 2].t == "(") {
            Func f;
            f.name = t[i + 1].t;
            size_t j = i + 3;
            while (j < t.size() && t[j].t != ")") ++j;
            ++j;
            if (j >= t.size() || t[j].t != "{") return false;
            int depth = 0;
            size_t k = j;
            for (; k < t.size(); ++k) {
                if (t[k].t == "{") ++depth;
                else if (t[k].t == "}") { if (--depth == 0) break; }
            }
            if (k >= t.size()) return false;
            f.head.assign(t.begin() + (long)i, t.begin() + (long)j);
            f.body.assign(t.begin() + (long)j + 1, t.begin() + (long)k);
            out.push_back(std::move(f));
            i = k + 1;
        } else {
            return false;
        }
    }
    return true;
}

std::vector<double> literals(const std::vector<Tok>& body) {
    std::vector<double> v;
    for (const Tok& k : body)
        if (k.num) v.push_back(std::strtod(k.t.c_str(), nullptr));
    return v;
}
bool calls(const std::vector<Tok>& body, const char* name) {
    for (size_t i = 0; i + 1 < body.size(); ++i)
        if (body[i].t == name && body[i + 1].t == "(") return true;
    return false;
}
int count_calls(const std::vector<Tok>& body, const char* name) {
    int n = 0;
    for (size_t i = 0; i + 1 < body.size(); ++i)
        if (body[i].t == name && body[i + 1].t == "(") ++n;
    return n;
}
// NaN in `want` marks a free parameter, written to `got` in order.
bool match_literals(const std::vector<double>& have, const std::vector<double>& want, std::vector<double>& got) {
    if (have.size() != want.size()) return false;
    got.clear();
    for (size_t i = 0; i < want.size(); ++i) {
        if (want[i] != want[i]) got.push_back(have[i]);
        else if (have[i] != want[i]) return false;
    }
    return true;
}

const double P = std::nan("");  // placeholder

// Token fingerprints (see fingerprint() below) of the reference's four SDF
// programs and three kinds of sdfmaterial, as shipped in sdfs/*.glsl and in
// the scenes that embed them (scenes/scene3..10.json): each kind has one.
#define SDF_FP_MANDELBULB 0x2d1eecbac9e0ac6dull
#define SDF_FP_MENGER 0x087bb4cf520f90bdull
#define SDF_FP_BLOB 0x788e0e00fa30ff89ull
#define SDF_FP_TERRAIN 0x25b927ec8dc4614cull
#define SDFMAT_FP_CONSTANT 0x23a4b3641549a38eull
#define SDFMAT_FP_RADIAL 0x0c5b2fb52fbaf614ull
#define SDFMAT_FP_BLOB 0xfa8a08568569156dull

// FNV-1a (64-bit) of a function's tokens: its header, then its body, numbers
// written as their value ("1.0" and "1." are one token) and the literals at the
// NaN positions of `pattern` (the free parameters) as "#". Literal positions
// count the body's numeric tokens in order, as match_literals does.
void fingerprint(const Func& f, const std::vector<double>& pattern, uint64_t& h) {
    auto mix = [&](const std::string& t) {
        for (unsigned char c : t) h = (h ^ c) * 1099511628211ull;
        h = (h ^ 0x1Fu) * 1099511628211ull;  // token separator
    };
    char b[40];
    for (const Tok& k : f.head) {
        if (!k.num) { mix(k.t); continue; }
        std::snprintf(b, sizeof b, "%.17g", std::strtod(k.t.c_str(), nullptr));
        mix(b);
    }
    mix("{");
    size_t lit = 0;
    for (const Tok& k : f.body) {
        if (!k.num) { mix(k.t); continue; }
        const bool free = lit < pattern.size() && pattern[lit] != pattern[lit];
        ++lit;
        if (free) { mix("#"); continue; }
        std::snprintf(b, sizeof b, "%.17g", std::strtod(k.t.c_str(), nullptr));
        mix(b);
    }
    mix("}");
}

}  // namespace

int recognize_sdf(const std::string& glsl, pt_sdf_desc& d, std::string& err) {
    std::memset(&d, 0, sizeof d);
    std::vector<Func> fns;
    if (!split_functions(tokenize(glsl), fns)) {
        err = "SDF text is not a list of function definitions";
        return PT_ERR_UNSUPPORTED_SDF;
    }
    const Func* sdf = nullptr;
    const Func* mat = nullptr;
    std::vector<const Func*> helpers;
    for (const Func& f : fns) {
        if (f.name == "sdf") sdf = &f;
        else if (f.name == "sdfmaterial") mat = &f;
        else helpers.push_back(&f);
    }
    if (!sdf || !mat) {
        err = "SDF text must define sdf() and sdfmaterial()";
        return PT_ERR_UNSUPPORTED_SDF;
    }
    std::vector<double> lit = literals(sdf->body), got;

    // Mandelbulb, power 8 via Chebyshev polynomials (sdfs/mandelbulb.glsl).
    static const std::vector<double> kMandel = {1.0, 0, P, 8.0, 7.0, 1.0, 1.0, 128.0, 256.0, 160.0, 32.0, 1.0,
                                                128.0, 256.0, 160.0, 32.0, 1.0, 128.0, 192.0, 80.0, 8.0,
                                                128.0, 192.0, 80.0, 8.0, P, 0.5};
    // Menger sponge (sdfs/menger.glsl).
    static const std::vector<double> kMenger = {1.0, 0.0, 0.0, 1.0, 0, P, 2.0, 1.0, 3.0, 1.0, 3.0, 1.0};
    // Blob of four smooth-min spheres (sdfs/blob.glsl).
    static const std::vector<double> kBlob = {1.0, 1.0, 0.7, 0.25, 1.25, 0.5, 0.25, 0.25, 1.25, 0.5, 0.25};
    static const std::vector<double> kBlobMat = {1.0, P, 1.0, 0.7, P, 0.25, 1.25, 0.5, 0.25, P, 0.25, 1.25, 0.5, 0.25, P};
    // Terrain height h(x) (sdfs/terrain.glsl).
    static const std::vector<double> kTerrain = {0.0625, 0.125, 10.0, 0.25, 0.25, 0.125, 0.5, 0.25, 0.25, 2.0, 1.0,
                                                 0.125, 4.0, 2.0, 0.0625, 8.0, 1.0, 0.03125, 16.0, 5.0, 0.015625,
                                                 32.0, 0.0078125, 64.0};

    if (helpers.empty() && calls(sdf->body, "pow") && calls(sdf->body, "inversesqrt") && calls(sdf->body, "log") &&
        match_literals(lit, kMandel, got)) {
        d.kind = PT_SDF_MANDELBULB;
        if (!(got[0] >= 0.0 && got[0] <= 64.0) || got[0] != std::floor(got[0])) {
            err = "mandelbulb iteration count must be an integer in [0, 64]";
            return PT_ERR_UNSUPPORTED_SDF;
        }
        if (!(got[1] > 0.0 && got[1] <= 3.4e38)) {
            err = "mandelbulb escape radius must be a positive number in float's range";
            return PT_ERR_UNSUPPORTED_SDF;
        }
        d.iterations = (int)got[0];
        d.bailout = (float)got[1];
    } else if (helpers.empty() && calls(sdf->body, "mod") && match_literals(lit, kMenger, got)) {
        d.kind = PT_SDF_MENGER;
        if (!(got[0] >= 0.0 && got[0] <= 16.0) || got[0] != std::floor(got[0])) {
            err = "menger level count must be an integer in [0, 16]";
            return PT_ERR_UNSUPPORTED_SDF;
        }
        d.iterations = (int)got[0];
    } else if (helpers.empty() && count_calls(sdf->body, "smin") == 3 && match_literals(lit, kBlob, got)) {
        d.kind = PT_SDF_BLOB;
    } else if (helpers.size() == 1 && count_calls(helpers[0]->body, "sin") == 11 &&
               match_literals(literals(helpers[0]->body), kTerrain, got) && lit.empty() &&
               count_calls(sdf->body, helpers[0]->name.c_str()) == 2) {
        d.kind = PT_SDF_TERRAIN;
    } else {
        err = "unrecognised SDF program (supported: mandelbulb, menger, blob, terrain)";
        return PT_ERR_UNSUPPORTED_SDF;
    }

    // The whole program, not only its literals and calls, must be the built-in
    // one: every function but sdfmaterial, tokens in order, free parameters
    // masked. A valid program that computes another shape (a sphere dropped from
    // the blob, another operand somewhere) is rejected, not rendered as the
    // built-in shape.
    {
        const std::vector<double> none;
        const std::vector<double>& pat = d.kind == PT_SDF_MANDELBULB ? kMandel : d.kind == PT_SDF_MENGER ? kMenger : none;
        uint64_t h = 14695981039346656037ull;
        for (const Func& f : fns)
            if (&f != mat) fingerprint(f, &f == sdf ? pat : none, h);
        // one per kind: the reference's sdfs/*.glsl and the scenes that embed them
        static const uint64_t kProgram[5] = {0, SDF_FP_MANDELBULB, SDF_FP_MENGER, SDF_FP_BLOB, SDF_FP_TERRAIN};
        if (h != kProgram[d.kind]) {
            err = "SDF program differs from the built-in one it resembles (its tokens, not only its constants, must match)";
            return PT_ERR_UNSUPPORTED_SDF;
        }
    }

    std::vector<double> mlit = literals(mat->body);
    for (double v : mlit)
        if (!(std::fabs(v) <= 3.4e38)) {
            err = "sdfmaterial() values must be numbers in float's range";
            return PT_ERR_UNSUPPORTED_SDF;
        }
    if (d.kind == PT_SDF_BLOB) {
        if (count_calls(mat->body, "smin") != 3 || !match_literals(mlit, kBlobMat, got)) {
            err = "unrecognised blob sdfmaterial()";
            return PT_ERR_UNSUPPORTED_SDF;
        }
        d.material_kind = PT_SDFMAT_BLOB;
        for (int k = 0; k < 4; ++k) d.material[k] = (float)got[(size_t)k];
    } else if (mlit.size() == 1 && mat->body.size() == 3 && mat->body[0].t == "return") {
        d.material_kind = PT_SDFMAT_CONSTANT;
        d.material[0] = (float)mlit[0];
    } else if (mlit.size() == 3 && calls(mat->body, "mix") && calls(mat->body, "dot")) {
        d.material_kind = PT_SDFMAT_RADIAL;
        for (int k = 0; k < 3; ++k) d.material[k] = (float)mlit[(size_t)k];
    } else {
        err = "unrecognised sdfmaterial() (supported: constant, radial mix, blob)";
        return PT_ERR_UNSUPPORTED_SDF;
    }
    // sdfmaterial likewise, its material values masked
    {
        static const std::vector<double> kConst = {P}, kRadial = {P, P, P};
        const std::vector<double>& pat =
            d.material_kind == PT_SDFMAT_BLOB ? kBlobMat : d.material_kind == PT_SDFMAT_RADIAL ? kRadial : kConst;
        uint64_t h = 14695981039346656037ull;
        fingerprint(*mat, pat, h);
        static const uint64_t kMaterial[3] = {SDFMAT_FP_CONSTANT, SDFMAT_FP_RADIAL, SDFMAT_FP_BLOB};
        if (h != kMaterial[d.material_kind]) {
            err = "sdfmaterial() differs from the built-in one it resembles (its tokens, not only its constants, must match)";
            return PT_ERR_UNSUPPORTED_SDF;
        }
    }
    return PT_OK;
}

// ---------------------------------------------------------------------------------
// Scene writer: UpdateToJSON + SaveScene (src/pathtracer.cpp:2724-2858, 3465-3470).

namespace {

// RoundDecimal(x, 1e5) (src/pathtracer.cpp:935-943): round half away from zero
// on the 1e-5 grid, through a truncating integer cast (64-bit here, where the
// reference's 32-bit cast overflows above |x| = 21474.8).
double round_decimal(double x) {
    const double p = 1e5;
    // beyond int64 (and for NaN or infinity) there is nothing to round: the
    // value is written as it is
    if (!(std::fabs(x) * p < 9.2e18)) return x;
    double n = x >= 0.0 ? (double)(int64_t)(x * p + 0.5) : (double)(int64_t)(x * p - 0.5);
    return n / p;
}
Value jnum(float x) {  // a reference field: rounded as the reference saves it
    Value v;
    v.kind = ptjson::Kind::Real;
    v.d = round_decimal((double)x);
    return v;
}
Value jf32(float x) {  // an extended field: the float itself, shortest text
    Value v;
    v.kind = ptjson::Kind::Real;
    v.d = (double)x;
    v.f32 = true;
    return v;
}
Value jint(int64_t x) {
    Value v;
    v.kind = ptjson::Kind::Int;
    v.i = x;
    return v;
}
Value jbool(bool x) {
    Value v;
    v.kind = ptjson::Kind::Bool;
    v.b = x;
    return v;
}
Value jstr(const std::string& x) {
    Value v;
    v.kind = ptjson::Kind::String;
    v.s = x;
    return v;
}
Value jarr(std::initializer_list<Value> xs) {
    Value v;
    v.kind = ptjson::Kind::Array;
    v.arr.assign(xs.begin(), xs.end());
    return v;
}
Value jobj() {
    Value v;
    v.kind = ptjson::Kind::Object;
    return v;
}
Value jlist() {
    Value v;
    v.kind = ptjson::Kind::Array;
    return v;
}
void put(Value& o, const char* key, Value v) { o.obj.emplace_back(key, std::move(v)); }
Value vec3(const float* a) { return jarr({jnum(a[0]), jnum(a[1]), jnum(a[2])}); }
Value vec3f(const float* a) { return jarr({jf32(a[0]), jf32(a[1]), jf32(a[2])}); }

}  // namespace

std::string write_scene_json(const SceneModel& s) {
    Value root = jobj();
    Value cam = jobj();
    put(cam, "numShots", jint((int64_t)s.shots.size()));
    Value pos = jlist(), ang = jlist();
    for (const CameraShot& c : s.shots) {
        pos.arr.push_back(vec3(c.pos));
        ang.arr.push_back(jarr({jnum(c.angle[0]), jnum(c.angle[1])}));
    }
    if (!s.shots.empty()) {
        put(cam, "position", pos);
        put(cam, "angle", ang);
    }
    put(cam, "ISO", jint(s.cam.ISO));
    put(cam, "size", jnum(s.cam.size));
    put(cam, "apertureSize", jnum(s.cam.apertureSize));
    put(cam, "apertureDistance", jnum(s.cam.apertureDist));
    put(cam, "lensRadius", jnum(s.cam.lensRadius));
    put(cam, "lensFocalLength", jnum(s.cam.lensFocalLength));
    put(cam, "lensThickness", jnum(s.cam.lensThickness));
    put(cam, "lensDistance", jnum(s.cam.lensDistance));
    if (s.render.camera == 1) put(cam, "model", jstr("pinhole"));  // extended key
    put(root, "camera", cam);

    auto ids = [](Value& o, int m, int l) {
        put(o, "materialID", jint(m));
        put(o, "lightID", jint(l));
    };
    if (!s.spheres.empty()) {
        Value a = jlist();
        for (const RSphere& r : s.spheres) {
            Value o = jobj();
            put(o, "position", vec3(r.pos));
            put(o, "radius", jnum(r.radius));
            ids(o, r.materialID, r.lightID);
            a.arr.push_back(o);
        }
        put(root, "sphere", a);
    }
    if (!s.planes.empty()) {
        Value a = jlist();
        for (const RPlane& r : s.planes) {
            Value o = jobj();
            put(o, "position", vec3(r.pos));
            ids(o, r.materialID, r.lightID);
            a.arr.push_back(o);
        }
        put(root, "plane", a);
    }
    if (!s.boxes.empty()) {
        Value a = jlist();
        for (const RBox& r : s.boxes) {
            Value o = jobj();
            put(o, "position", vec3(r.pos));
            put(o, "rotation", vec3(r.rotation));
            put(o, "size", vec3(r.size));
            ids(o, r.materialID, r.lightID);
            a.arr.push_back(o);
        }
        put(root, "box", a);
    }
    if (!s.lenses.empty()) {
        Value a = jlist();
        for (const RLens& r : s.lenses) {
            Value o = jobj();
            put(o, "position", vec3(r.pos));
            put(o, "rotation", vec3(r.rotation));
            put(o, "radius", jnum(r.radius));
            put(o, "focalLength", jnum(r.focalLength));
            put(o, "thickness", jnum(r.thickness));
            put(o, "isConverging", jbool(r.isConverging));
            ids(o, r.materialID, r.lightID);
            a.arr.push_back(o);
        }
        put(root, "lens", a);
    }
    if (!s.cyclides.empty()) {
        Value a = jlist();
        for (const RCyclide& r : s.cyclides) {
            Value o = jobj();
            put(o, "position", vec3(r.pos));
            put(o, "rotation", vec3(r.rotation));
            put(o, "scale", vec3(r.scale));
            put(o, "a", jnum(r.a));
            put(o, "b", jnum(r.b));
            put(o, "c", jnum(r.c));
            put(o, "d", jnum(r.d));
            put(o, "boundingRadius", jnum(r.brad));
            ids(o, r.materialID, r.lightID);
            a.arr.push_back(o);
        }
        put(root, "cyclide", a);
    }
    if (!s.sdfs.empty()) {
        Value a = jlist();
        for (const RSdf& r : s.sdfs) {
            Value o = jobj();
            put(o, "position", vec3(r.pos));
            put(o, "boundingSize", vec3(r.size));
            if (!r.glsl.empty()) {
                put(o, "glsl", jstr(r.glsl));
            } else {  // extended form
                const pt_sdf_desc& d = r.desc;
                const char* kinds[] = {"none", "mandelbulb", "menger", "blob", "terrain"};
                put(o, "kind", jstr(kinds[d.kind >= 0 && d.kind <= 4 ? d.kind : 0]));
                put(o, "iterations", jint(d.iterations));
                if (d.kind == PT_SDF_MANDELBULB) put(o, "bailout", jf32(d.bailout));
                if (d.material_kind == PT_SDFMAT_CONSTANT) put(o, "material", jf32(d.material[0]));
                else if (d.material_kind == PT_SDFMAT_RADIAL)
                    put(o, "material", jarr({jf32(d.material[0]), jf32(d.material[1]), jf32(d.material[2])}));
                else
                    put(o, "material", jarr({jf32(d.material[0]), jf32(d.material[1]), jf32(d.material[2]), jf32(d.material[3])}));
            }
            a.arr.push_back(o);
        }
        put(root, "sdf", a);
    }
    if (!s.materials.empty()) {
        Value a = jlist();
        for (const Material& m : s.materials) {
            Value o = jobj();
            if (m.type == MAT_LAMBERT_LOBE || m.reflection[0] != 0.0f || m.reflection[1] != 0.0f || m.reflection[2] != 0.0f) {
                Value r = jobj();
                put(r, "peakWavelength", jnum(m.reflection[0]));
                put(r, "sigma", jnum(m.reflection[1]));
                put(r, "isInvert", jbool(m.reflection[2] != 0.0f));
                put(o, "reflection", r);
            }
            switch (m.type) {  // extended keys
                case MAT_LAMBERT_GAUSS: {
                    put(o, "type", jstr("lambert"));
                    Value g = jlist();
                    for (int i = 0; i < m.n_gauss; ++i) g.arr.push_back(vec3f(&m.gauss[3 * i]));
                    put(o, "gaussians", g);
                    break;
                }
                case MAT_DIELECTRIC:
                    put(o, "type", jstr("dielectric"));
                    if (m.dispersion == DISP_SELLMEIER) {
                        Value c = jlist();
                        for (int k = 0; k < 6; ++k) c.arr.push_back(jf32(m.disp[k]));
                        put(o, "sellmeier", c);
                    } else {
                        put(o, "cauchy", jarr({jf32(m.disp[0]), jf32(m.disp[1])}));
                    }
                    break;
                case MAT_CONDUCTOR:
                    put(o, "type", jstr("conductor"));
                    put(o, "n", jf32(m.cond_n));
                    put(o, "k", jf32(m.cond_k));
                    break;
                default: break;
            }
            a.arr.push_back(o);
        }
        put(root, "material", a);
    }
    if (!s.lights.empty()) {
        Value a = jlist();
        for (const Light& l : s.lights) {
            Value o = jobj(), e = jobj();
            put(e, "temperature", jnum(l.emission[0]));
            put(e, "luminosity", jnum(l.emission[1]));
            put(o, "emission", e);
            a.arr.push_back(o);
        }
        put(root, "light", a);
    }
    // extended top-level keys
    if (!s.quads.empty()) {
        Value a = jlist();
        for (const Quad& q : s.quads) {
            Value o = jobj();
            put(o, "position", vec3f(q.corner));
            put(o, "edge1", vec3f(q.e1));
            put(o, "edge2", vec3f(q.e2));
            ids(o, q.materialID, q.lightID);
            a.arr.push_back(o);
        }
        put(root, "quad", a);
    }
    if (s.env_enabled) {
        Value e = jobj();
        put(e, "radiance", jf32(s.env_radiance));
        put(root, "environment", e);
    }
    const RenderSettings def;
    const RenderSettings& r = s.render;
    if (r.n_lambda != def.n_lambda || r.lambda_min != def.lambda_min || r.lambda_max != def.lambda_max ||
        r.rr_depth != def.rr_depth || r.seed != def.seed) {
        Value o = jobj();
        put(o, "wavelengths", jint(r.n_lambda));
        put(o, "lambdaMin", jf32(r.lambda_min));
        put(o, "lambdaMax", jf32(r.lambda_max));
        put(o, "russianRouletteDepth", jint(r.rr_depth));
        put(o, "seed", jint((int64_t)r.seed));
        put(root, "render", o);
    }
    return ptjson::dump(root, 4);
}

}  // namespace pt
