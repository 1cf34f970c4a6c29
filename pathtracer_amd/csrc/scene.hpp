// scene.hpp — host scene model, reference JSON loader, uniform-buffer packer,
// SDF recogniser and extended-scene schema.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pathtracer.h"

namespace pt {

// Host mirrors of the reference's scene structs (src/pathtracer.cpp:110-185).
struct CameraShot { float pos[3]; float angle[2]; };
struct CameraParams {
    int ISO = 0;
    float size = 0, apertureSize = 0, apertureDist = 0;
    float lensRadius = 0, lensFocalLength = 0, lensThickness = 0, lensDistance = 0;
};
struct RSphere { float pos[3]; float radius; int materialID; int lightID; };
struct RPlane { float pos[3]; int materialID; int lightID; };
struct RBox { float pos[3]; float rotation[3]; float size[3]; int materialID; int lightID; };
struct RLens {
    float pos[3]; float rotation[3]; float radius, focalLength, thickness; bool isConverging;
    int materialID; int lightID;
};
struct RCyclide {
    float pos[3]; float rotation[3]; float scale[3]; float a, b, c, d, brad;
    int materialID; int lightID;
};
struct RSdf { float pos[3]; float size[3]; std::string glsl; pt_sdf_desc desc; };
// Extended: a parallelogram with corner `corner` and edges e1, e2 (front face e1 x e2).
struct Quad { float corner[3]; float e1[3]; float e2[3]; int materialID; int lightID; };

enum MaterialType : int32_t {
    MAT_LAMBERT_LOBE = 0,   // the reference's one-lobe spectrum (src/shader.comp:1030-1038)
    MAT_LAMBERT_GAUSS = 1,  // sum of up to 3 Gaussians, clamped to [0, 1]
    MAT_DIELECTRIC = 2,     // smooth dielectric, Cauchy or Sellmeier dispersion
    MAT_CONDUCTOR = 3,      // smooth conductor, complex index n + ik
};
enum DispersionModel : int32_t { DISP_CAUCHY = 0, DISP_SELLMEIER = 1 };

struct Material {
    float reflection[3] = {0, 0, 0};  // (peak nm, sigma, invert) — the packed reference form
    int32_t type = MAT_LAMBERT_LOBE;
    int32_t n_gauss = 0;
    float gauss[9] = {0};             // (weight, mean nm, sigma nm) x 3
    int32_t dispersion = DISP_CAUCHY;
    float disp[6] = {0};              // Cauchy (A, B[um^2]) or Sellmeier (B1,B2,B3,C1,C2,C3[um^2])
    float cond_n = 0, cond_k = 0;
};
struct Light { float emission[2]; };  // (temperature K, luminosity)

struct RenderSettings {
    int n_lambda = 4;            // wavelengths carried per path in PT_MODE_EXT (1 or 4)
    float lambda_min = 380.0f;   // nm
    float lambda_max = 780.0f;
    int rr_depth = 3;            // first bounce at which Russian roulette may end a path
    uint32_t seed = 0;           // scene seed mixed into the counter-based RNG
    int camera = 0;              // 0 = reference lens camera, 1 = pinhole
};

struct SceneModel {
    std::vector<CameraShot> shots;
    CameraParams cam;
    std::vector<RSphere> spheres;
    std::vector<RPlane> planes;
    std::vector<RBox> boxes;
    std::vector<RLens> lenses;
    std::vector<RCyclide> cyclides;
    std::vector<RSdf> sdfs;
    std::vector<Quad> quads;
    std::vector<Material> materials;
    std::vector<Light> lights;
    bool env_enabled = false;
    float env_radiance = 0.0f;   // uniform spectral radiance of the environment
    RenderSettings render;
};

// Parses the reference scene JSON (plus extended keys). Returns PT_OK,
// PT_ERR_JSON or PT_ERR_UNSUPPORTED_SDF; `err` explains a failure.
int load_scene_json(const char* text, size_t len, SceneModel& out, std::string& err);

// UpdateUniformBuffer (src/pathtracer.cpp:3642-3811). Returns PT_OK or
// PT_ERR_CAPACITY when an array would exceed the uniform block's caps.
int pack_reference_ubo(const SceneModel& s, pt_packed_scene& ubo, std::string& err);

// Copies the CIE 1931 table into `dst` (1323 floats, src/pathtracer.cpp:2231-2233).
void copy_cie_table(float* dst);
const float* cie_table();

// Camera part of UpdateFromJSON + UpdatePushConstant (src/pathtracer.cpp:2577-2606, 3813-3834).
int camera_to_push(const SceneModel& s, int shot, pt_push_constants& pc, std::string& err);

// UpdateToJSON + SaveScene (src/pathtracer.cpp:2724-2858, 3465-3470): the
// reference's keys in its order, each float rounded to the 1e-5 grid as the
// reference saves it; extended keys follow with their floats exact (the
// shortest text that reads back as the same binary32 value).
std::string write_scene_json(const SceneModel& s);

// InsertSDF's input recognised as a built-in kind (src/pathtracer.cpp:2004-2054).
int recognize_sdf(const std::string& glsl, pt_sdf_desc& desc, std::string& err);

}  // namespace pt
