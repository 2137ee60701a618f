// host_scene.h -- scene input: JSON scene file, OBJ meshes, config.json paths,
// transform composition, and flattening into the camera-space triangle soup
// the tracer consumes.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/aurora_hip.h"

namespace aur {

// 4x4 row-major float matrix with the reference's arithmetic
// (src/core/matrix4x4.cpp, src/core/transform.cpp).
struct Mat4 {
    float m[4][4];
    static Mat4 identity();
    Mat4 mul(const Mat4 &b) const;  // this * b, matrix4x4.cpp operator*
    Mat4 inverse() const;           // Gauss-Jordan with full pivoting, matrix4x4.cpp inverse()
};

struct Xform {
    Mat4 m, inv;
    Xform() : m(Mat4::identity()), inv(Mat4::identity()) {}
    void translate(float x, float y, float z);
    void scale(float x, float y, float z);
    void rotate_x(float deg);
    void rotate_y(float deg);
    void rotate_z(float deg);
};

struct Options {
    int width = 512, height = 512;  // options.h:49-61 defaults
    int pixelsamples = 64, lightsamples = 64, mindepth = 1, maxdepth = 16;
    float fov = 0.78539816f;
    std::string filename;
};

struct Displace {
    bool on = false;
    float amplitude = 0.f, frequency = 1.f;
    int octaves = 0;
    uint32_t seed = 0;
};

// One geometry object after OBJ ingest: de-indexed per-corner arrays, as
// ObjTriangleMesh builds them (objTriangleMesh.cpp:26-92).
struct HostObject {
    std::string name, material;
    int attr = 0;
    int subdlevels = 0;
    float o2c[16], c2o[16];
    std::vector<float> P;   // 9 per triangle, object space
    std::vector<float> N;   // 9 per triangle, object space
    std::vector<float> UV;  // 6 per triangle, or empty when the OBJ has no `vt`
    Displace displace;
};

// A texture as the reference holds it: FP16 RGBA (Imf::Rgba) widened to float.
struct HostTexture {
    int width = 0, height = 0;
    std::vector<float> rgba;
};

struct HostScene {
    Options opt;
    std::vector<HostObject> objects;
    std::vector<aur_material> attr_mats;  // objects then lights
    std::vector<aur_light> lights;        // JSON (sorted) order
    std::vector<int> light_attr;          // attribute index of each light
    std::vector<aur_shader> shaders;
    std::vector<HostTexture> textures;
    float cam_m[16], cam_inv[16];         // camera-to-world and its inverse
    // flattened camera-space scene (filled by flatten or by upload)
    bool flat = false;
    std::vector<float> verts, normals, uvs;
    std::vector<int> tris, attr;
    HostScene();
};

// Texel pool layout of the scene's textures: offsets in texels.
std::vector<aur_texture> texture_table(const HostScene &s, int64_t *n_texels);

// Parses scene + OBJ files.  Throws std::runtime_error with a message; the
// C ABI turns that into an error code.
void load_scene(const std::string &scene_path, const std::string &config_path, HostScene *out);

// Light triangles (renderer.cpp:127-134, squareLight.cpp shape()) appended
// after all objects; objects that need no subdivision are flattened on the
// host.  Returns the list of object indices that still need GPU subdivision.
void light_geometry(const aur_light &l, float verts[12], float n[3]);

// Applies the reference's path templates (stringUtils.h:37-122).
std::string string_template(const std::string &in, const std::string &config_path);

// Derived light data (squareLight.cpp:20-52): corners, normal, pdf triangles.
void light_corners(const aur_light &l, float pc[4][3], float N[3]);

// Row-major 4x4 point / vector transforms (transform.cpp:206-232).
void transform_point(const float *m16, const float *p, float *out);
void transform_vector(const float *m16, const float *v, float *out);

}  // namespace aur
