// host_scene.cpp -- scene input for the MI355X tracer (host side).
//
// Follows src/translator/jsonParser.cpp (scene keys and composition order),
// src/core/transform.cpp + matrix4x4.cpp (transform arithmetic),
// src/geometry/objParser.cpp + objTriangleMesh.cpp (mesh ingest) and
// src/core/stringUtils.h (path templates).
#include "host_scene.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <utility>

#include "exr_io.h"
#include "json_lite.h"

namespace aur {

// ---------------------------------------------------------------------------
// matrices: matrix4x4.cpp
// ---------------------------------------------------------------------------
Mat4 Mat4::identity() {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r.m[i][j] = (i == j) ? 1.f : 0.f;
    return r;
}

Mat4 Mat4::mul(const Mat4 &b) const {
    Mat4 r;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i][j] = m[i][0] * b.m[0][j] + m[i][1] * b.m[1][j] + m[i][2] * b.m[2][j] + m[i][3] * b.m[3][j];
    return r;
}

// Gauss-Jordan elimination with full pivoting, the same pivot order and
// float arithmetic as matrix4x4.cpp inverse().
Mat4 Mat4::inverse() const {
    int indxc[4], indxr[4];
    int ipiv[4] = {0, 0, 0, 0};
    float minv[4][4];
    std::memcpy(minv, m, sizeof minv);
    for (int i = 0; i < 4; i++) {
        int irow = -1, icol = -1;
        float big = 0.f;
        for (int j = 0; j < 4; j++) {
            if (ipiv[j] != 1) {
                for (int k = 0; k < 4; k++) {
                    if (ipiv[k] == 0) {
                        if (std::fabs(minv[j][k]) >= big) {
                            big = std::fabs(minv[j][k]);
                            irow = j;
                            icol = k;
                        }
                    } else if (ipiv[k] > 1) {
                        throw std::runtime_error("singular matrix in transform inverse");
                    }
                }
            }
        }
        if (icol < 0) throw std::runtime_error("singular matrix in transform inverse");
        ++ipiv[icol];
        if (irow != icol)
            for (int k = 0; k < 4; ++k) std::swap(minv[irow][k], minv[icol][k]);
        indxr[i] = irow;
        indxc[i] = icol;
        if (minv[icol][icol] == 0.f) throw std::runtime_error("singular matrix in transform inverse");
        float pivinv = 1.f / minv[icol][icol];
        minv[icol][icol] = 1.f;
        for (int j = 0; j < 4; j++) minv[icol][j] *= pivinv;
        for (int j = 0; j < 4; j++) {
            if (j != icol) {
                float save = minv[j][icol];
                minv[j][icol] = 0;
                for (int k = 0; k < 4; k++) minv[j][k] -= minv[icol][k] * save;
            }
        }
    }
    for (int j = 3; j >= 0; j--) {
        if (indxr[j] != indxc[j])
            for (int k = 0; k < 4; k++) std::swap(minv[k][indxr[j]], minv[k][indxc[j]]);
    }
    Mat4 r;
    std::memcpy(r.m, minv, sizeof minv);
    return r;
}

// ---------------------------------------------------------------------------
// transform.cpp:44-104: post-multiplication, inverse recomputed each time
// ---------------------------------------------------------------------------
static Mat4 make(float a00, float a01, float a02, float a03, float a10, float a11, float a12, float a13,
                 float a20, float a21, float a22, float a23, float a30, float a31, float a32, float a33) {
    Mat4 r;
    float v[16] = {a00, a01, a02, a03, a10, a11, a12, a13, a20, a21, a22, a23, a30, a31, a32, a33};
    std::memcpy(r.m, v, sizeof v);
    return r;
}
// math_utils.h:212-214: degrees * M_PI / 180.f evaluated in double
static float radians(float deg) { return (float)(deg * M_PI / 180.f); }

void Xform::translate(float x, float y, float z) {
    m = m.mul(make(1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z, 0, 0, 0, 1));
    inv = m.inverse();
}
void Xform::scale(float x, float y, float z) {
    m = m.mul(make(x, 0, 0, 0, 0, y, 0, 0, 0, 0, z, 0, 0, 0, 0, 1));
    inv = m.inverse();
}
void Xform::rotate_x(float a) {
    float s = sinf(radians(a)), c = cosf(radians(a));
    m = m.mul(make(1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0, 0, 0, 0, 1));
    inv = m.inverse();
}
void Xform::rotate_y(float a) {
    float s = sinf(radians(a)), c = cosf(radians(a));
    m = m.mul(make(c, 0, s, 0, 0, 1, 0, 0, -s, 0, c, 0, 0, 0, 0, 1));
    inv = m.inverse();
}
void Xform::rotate_z(float a) {
    float s = sinf(radians(a)), c = cosf(radians(a));
    m = m.mul(make(c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1));
    inv = m.inverse();
}

// transform.cpp:206-214 (points) and 226-232 (vectors)
static void xpoint(const Mat4 &M, const float *p, float *o) {
    const float(*m)[4] = M.m;
    float x = p[0], y = p[1], z = p[2];
    float xp = m[0][0] * x + m[0][1] * y + m[0][2] * z + m[0][3];
    float yp = m[1][0] * x + m[1][1] * y + m[1][2] * z + m[1][3];
    float zp = m[2][0] * x + m[2][1] * y + m[2][2] * z + m[2][3];
    float wp = m[3][0] * x + m[3][1] * y + m[3][2] * z + m[3][3];
    if (wp == 1.f) { o[0] = xp; o[1] = yp; o[2] = zp; return; }
    float inv = 1.f / wp;
    o[0] = xp * inv; o[1] = yp * inv; o[2] = zp * inv;
}
static void xvector(const Mat4 &M, const float *v, float *o) {
    const float(*m)[4] = M.m;
    float x = v[0], y = v[1], z = v[2];
    o[0] = m[0][0] * x + m[0][1] * y + m[0][2] * z;
    o[1] = m[1][0] * x + m[1][1] * y + m[1][2] * z;
    o[2] = m[2][0] * x + m[2][1] * y + m[2][2] * z;
}

void transform_point(const float *m16, const float *p, float *out) {
    Mat4 M;
    std::memcpy(M.m, m16, sizeof M.m);
    xpoint(M, p, out);
}
void transform_vector(const float *m16, const float *v, float *out) {
    Mat4 M;
    std::memcpy(M.m, m16, sizeof M.m);
    xvector(M, v, out);
}

// ---------------------------------------------------------------------------
// config.json templates: stringUtils.h:37-122
// ---------------------------------------------------------------------------
static std::string read_file(const std::string &path) {
    std::ifstream f(path.c_str(), std::ios::in | std::ios::binary);
    if (!f.good()) throw std::runtime_error("cannot open file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string string_template(const std::string &input, const std::string &config_path) {
    if (input.find('$') == std::string::npos && input.find("####") == std::string::npos) return input;
    std::map<std::string, std::string> rep;
    if (!config_path.empty()) {
        JsonValue cfg = JsonReader(read_file(config_path)).parse();
        auto get = [&](const char *k) { return cfg.has(k) ? cfg.at(k).as_string() : std::string(); };
        rep["$BASE"] = get("basepath");
        rep["$CACHE"] = get("cachepath");
        rep["$RENDER"] = get("renderpath");
        rep["$TEXTURE"] = get("texturepath");
        rep["$HDRI"] = get("hdripath");
    }
    std::string out = input;
    while (out.find('$') != std::string::npos) {
        bool did = false;
        for (auto &kv : rep) {
            size_t pos = out.find(kv.first);
            if (pos != std::string::npos) { out.replace(pos, kv.first.size(), kv.second); did = true; }
        }
        if (!did) throw std::runtime_error("cannot replace unknown keywords in string: " + out);
    }
    size_t dbl = out.find("//");
    while (dbl != std::string::npos) { out.replace(dbl, 2, "/"); dbl = out.find("//"); }
    size_t hash = out.find("####");
    if (hash != std::string::npos) {
        for (int frame = 1;; ++frame) {
            char num[32];
            std::snprintf(num, sizeof num, "%04d", frame);
            std::string cand = out;
            cand.replace(hash, 4, num);
            std::ifstream f(cand.c_str());
            if (!f.good()) { out = cand; break; }
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// OBJ: objParser.cpp:58-202, objTriangleMesh.cpp:26-92
// ---------------------------------------------------------------------------
// Reads the whole file and scans it line by line with the semantics of the
// stream parser it replaces (objParser.cpp: `ss >> op` then values): on a
// 'v' line the first token is skipped and three floats read ("vn"), two
// ("vt") or three (positions), a value that does not parse reads 0 and ends
// the line; an 'f' line contributes its first three corner tokens.  Floats
// convert with strtof, as the stream extraction does.
static const char *skip_ws(const char *p, const char *e) {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    return p;
}
static const char *skip_token(const char *p, const char *e) {
    while (p < e && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    return p;
}
static void read_floats(const char *p, const char *e, int n, std::vector<float> *out) {
    bool ok = true;
    char buf[64];
    for (int k = 0; k < n; ++k) {
        float v = 0.f;
        if (ok) {
            p = skip_ws(p, e);
            const char *q = skip_token(p, e);
            const size_t len = std::min<size_t>((size_t)(q - p), sizeof buf - 1);
            std::memcpy(buf, p, len);
            buf[len] = 0;
            char *end = nullptr;
            v = std::strtof(buf, &end);
            if (end == buf) { v = 0.f; ok = false; }
            p = q;
        }
        out->push_back(v);
    }
}
static int corner_int(const char *p, const char *e) {
    // atoi of the characters [p, e)
    char buf[32];
    const size_t len = std::min<size_t>((size_t)(e - p), sizeof buf - 1);
    std::memcpy(buf, p, len);
    buf[len] = 0;
    return std::atoi(buf);
}

static void load_obj(const std::string &path, bool need_normals, HostObject *obj) {
    std::FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("cannot open OBJ file: " + path);
    std::string data;
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (size > 0) {
        data.resize((size_t)size);
        const size_t got = std::fread(&data[0], 1, (size_t)size, fp);
        data.resize(got);
    }
    std::fclose(fp);
    std::vector<float> pts, nrm, tex;
    std::vector<int> vi, ni, ti;
    const char *p = data.data(), *end = data.data() + data.size();
    while (p < end) {
        const char *eol = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!eol) eol = end;
        if (eol > p) {
            if (p[0] == 'v') {
                const char *q = skip_token(p, eol);
                if (eol - p > 1 && p[1] == 'n') read_floats(q, eol, 3, &nrm);
                else if (eol - p > 1 && p[1] == 't') read_floats(q, eol, 2, &tex);
                else read_floats(q, eol, 3, &pts);
            } else if (p[0] == 'f') {
                const char *q = skip_token(p, eol);
                for (int k = 0; k < 3; ++k) {
                    q = skip_ws(q, eol);
                    const char *t = q, *te = skip_token(q, eol);
                    q = te;
                    // vertex: digits before the first '/'; uv: between the first
                    // and second '/' (objParser.cpp:140-176); normal: after the
                    // last '/', or the whole token without one
                    const char *s1 = static_cast<const char *>(std::memchr(t, '/', (size_t)(te - t)));
                    vi.push_back(corner_int(t, s1 ? s1 : te));
                    int uv = 0;
                    const char *last = s1;
                    if (s1) {
                        const char *s2 = static_cast<const char *>(std::memchr(s1 + 1, '/', (size_t)(te - s1 - 1)));
                        uv = corner_int(s1 + 1, s2 ? s2 : te);
                        for (const char *c = s1; c < te; ++c)
                            if (*c == '/') last = c;
                    }
                    ti.push_back(uv);
                    ni.push_back(corner_int(last ? last + 1 : t, te));
                }
            }
        }
        p = eol + 1;
    }
    size_t nverts = vi.size();
    obj->P.resize(nverts * 3);
    obj->N.resize(nverts * 3);
    // objTriangleMesh.cpp:43-72: per-corner uvs only when the file has any
    if (!tex.empty()) obj->UV.resize(nverts * 2);
    else obj->UV.clear();
    for (size_t i = 0; i < nverts; ++i) {
        int v = vi[i] - 1, n = ni[i] - 1;
        if (v < 0 || (size_t)v * 3 + 2 >= pts.size() + 0 || (size_t)(v * 3 + 2) >= pts.size())
            throw std::runtime_error("OBJ vertex index out of range in " + path);
        for (int k = 0; k < 3; ++k) obj->P[3 * i + k] = pts[3 * v + k];
        if (n >= 0 && (size_t)(n * 3 + 2) < nrm.size()) {
            float x = nrm[3 * n], y = nrm[3 * n + 1], z = nrm[3 * n + 2];
            if (x == 0.f && y == 0.f && z == 0.f) { x = 0.f; y = 1.f; z = 0.f; }  // objTriangleMesh.cpp:232-235
            obj->N[3 * i] = x; obj->N[3 * i + 1] = y; obj->N[3 * i + 2] = z;
        } else if (need_normals) {
            throw std::runtime_error("OBJ normal index out of range in " + path);
        } else {
            obj->N[3 * i] = 0.f; obj->N[3 * i + 1] = 1.f; obj->N[3 * i + 2] = 0.f;
        }
        if (!tex.empty()) {
            int t = ti[i] - 1;
            if (t < 0 || (size_t)(2 * t + 1) >= tex.size()) throw std::runtime_error("OBJ uv index out of range in " + path);
            obj->UV[2 * i] = tex[2 * t];
            obj->UV[2 * i + 1] = tex[2 * t + 1];
        }
    }
    if (nverts % 3) throw std::runtime_error("OBJ face list is not triangles: " + path);
}

// ---------------------------------------------------------------------------
// scene: jsonParser.cpp
// ---------------------------------------------------------------------------
static Xform transforms_attr(const JsonValue &root) {
    // jsonParser.cpp:168-223
    Xform t;
    const JsonValue &list = root.at("transforms");
    if (list.kind != JsonValue::Array && list.kind != JsonValue::Object) throw std::runtime_error("transforms must be a list");
    for (const JsonValue &item : list.arr) {
        if (item.kind != JsonValue::Object || item.obj.empty()) continue;
        const std::string &k = item.obj.begin()->first;
        const JsonValue &v = item.obj.begin()->second;
        if (k == "translate") t.translate((float)v.arr.at(0).as_double(), (float)v.arr.at(1).as_double(), (float)v.arr.at(2).as_double());
        else if (k == "scale") t.scale((float)v.arr.at(0).as_double(), (float)v.arr.at(1).as_double(), (float)v.arr.at(2).as_double());
        else if (k == "rotateX") t.rotate_x((float)v.as_double());
        else if (k == "rotateY") t.rotate_y((float)v.as_double());
        else if (k == "rotateZ") t.rotate_z((float)v.as_double());
    }
    return t;
}

static void color_attr(const JsonValue &root, const char *name, float *out) {
    const JsonValue &v = root.at(name);
    if (v.kind == JsonValue::String) throw std::runtime_error(std::string("textured parameter not supported: ") + name);
    if (v.kind != JsonValue::Array || v.arr.size() != 3) throw std::runtime_error(std::string("expected 3 values for ") + name);
    for (int i = 0; i < 3; ++i) out[i] = (float)v.arr[i].as_double();
}
static float float_attr(const JsonValue &root, const char *name) {
    const JsonValue &v = root.at(name);
    if (v.kind == JsonValue::String) throw std::runtime_error(std::string("textured parameter not supported: ") + name);
    return (float)v.as_double();
}
static int distribution_attr(const JsonValue &root) {
    if (!root.has("distribution")) return AUR_DIST_BECKMANN;
    const std::string &d = root.at("distribution").as_string();
    if (d == "ggx" || d == "GGX") return AUR_DIST_GGX;
    if (d == "beckmann" || d == "Beckmann") return AUR_DIST_BECKMANN;
    throw std::runtime_error("unknown microfacet distribution: " + d);
}

// Shaders of the "shaders" section by name (ShadingEngine::m_index_map).
typedef std::map<std::string, int> ShaderNames;

// getShaderAttr + initShader (jsonParser.cpp:145-165, 284-310): a string names
// a shader of the "shaders" section, anything else is the constant value.
// Returns the shader index or -1 for a constant written to out.
static int color_slot(const JsonValue &root, const char *name, float *out, const ShaderNames &sn,
                      const std::vector<aur_shader> &shaders) {
    const JsonValue &v = root.at(name);
    if (v.kind != JsonValue::String) { color_attr(root, name, out); return -1; }
    auto it = sn.find(v.as_string());
    if (it == sn.end()) throw std::runtime_error("unknown shader '" + v.as_string() + "' for " + name);
    if (shaders[it->second].kind != AUR_SHADER_COLOR)
        throw std::runtime_error("shader '" + v.as_string() + "' is a float shader, " + name + " needs a color");
    return it->second;
}
static int float_slot(const JsonValue &root, const char *name, float *out, const ShaderNames &sn,
                      const std::vector<aur_shader> &shaders) {
    const JsonValue &v = root.at(name);
    if (v.kind != JsonValue::String) { *out = float_attr(root, name); return -1; }
    auto it = sn.find(v.as_string());
    if (it == sn.end()) throw std::runtime_error("unknown shader '" + v.as_string() + "' for " + name);
    if (shaders[it->second].kind != AUR_SHADER_FLOAT)
        throw std::runtime_error("shader '" + v.as_string() + "' is a color shader, " + name + " needs a float");
    return it->second;
}

// jsonParser.cpp:313-429
static aur_material parse_material(const JsonValue &root, const ShaderNames &sn, std::vector<aur_shader> *shaders) {
    aur_material m;
    std::memset(&m, 0, sizeof m);
    for (int k = 0; k < AUR_N_SLOTS; ++k) m.shader[k] = -1;
    const std::vector<aur_shader> &sh = *shaders;
    const std::string &type = root.at("type").as_string();
    if (type == "kelemenMaterial") {
        m.type = AUR_MAT_KELEMEN;
        m.shader[AUR_SLOT_DIFFUSE] = color_slot(root, "diffusecolor", m.diffuse, sn, sh);
        m.shader[AUR_SLOT_SPEC] = color_slot(root, "speccolor", m.spec, sn, sh);
        m.reflectance = float_attr(root, "reflectance");  // a constant only (jsonParser.cpp:331-333)
        m.shader[AUR_SLOT_ROUGH] = float_slot(root, "roughness", &m.roughness, sn, sh);
        m.distribution = distribution_attr(root);
    } else if (type == "matteMaterial") {
        m.type = AUR_MAT_MATTE;
        m.shader[AUR_SLOT_DIFFUSE] = color_slot(root, "color", m.diffuse, sn, sh);
    } else if (type == "metalMaterial") {
        m.type = AUR_MAT_METAL;
        m.shader[AUR_SLOT_SPEC] = color_slot(root, "reflectance", m.spec, sn, sh);
        m.shader[AUR_SLOT_ROUGH] = float_slot(root, "roughnessA", &m.roughness, sn, sh);
        m.shader[AUR_SLOT_ROUGHB] = float_slot(root, "roughnessB", &m.roughnessB, sn, sh);
        m.shader[AUR_SLOT_MIX] = float_slot(root, "mix", &m.mix, sn, sh);
        m.shader[AUR_SLOT_GAIN] = color_slot(root, "gain", m.gain, sn, sh);
        m.distribution = distribution_attr(root);
    } else if (type == "glassMaterial") {
        // jsonParser.cpp:412-419: constants only (the attributes are read with getColor / getFloat)
        m.type = AUR_MAT_GLASS;
        color_attr(root, "specColor", m.spec);
        m.reflectance = float_attr(root, "reflectance");
        m.ior = float_attr(root, "ior");
        color_attr(root, "transmitColor", m.transmit);
    } else {
        throw std::runtime_error("unsupported material type: " + type);
    }
    // jsonParser.cpp:316-321: an optional normal map, a shader or a constant colour
    if (root.has("normalmap")) {
        float c[3];
        int idx = color_slot(root, "normalmap", c, sn, sh);
        if (idx < 0) {
            aur_shader k;
            std::memset(&k, 0, sizeof k);
            k.type = AUR_SHADER_CONSTANT;
            k.kind = AUR_SHADER_COLOR;
            k.texture = -1;
            for (int i = 0; i < 3; ++i) k.value[i] = c[i];
            idx = (int)shaders->size();
            shaders->push_back(k);
        }
        m.shader[AUR_SLOT_NORMAL] = idx;
    }
    return m;
}

// Texture2D::Texture2D (texture2D.cpp:23-46): the file through Imf::RgbaInputFile,
// i.e. every channel rounded to FP16.  One texture per file.
static int load_texture(const std::string &name, const std::string &scene_dir, std::map<std::string, int> *by_path,
                        std::vector<HostTexture> *textures) {
    std::string path = name;
    if (!path.empty() && path[0] != '/') {
        std::ifstream probe(path.c_str());
        if (!probe.good()) path = scene_dir + path;
    }
    auto it = by_path->find(path);
    if (it != by_path->end()) return it->second;
    ExrImage img;
    try {
        img = exr_read(path);
    } catch (const std::runtime_error &e) {
        throw std::runtime_error(std::string("texture: ") + e.what());
    }
    HostTexture t;
    t.width = img.width;
    t.height = img.height;
    t.rgba.resize(img.rgba.size());
    for (size_t i = 0; i < img.rgba.size(); ++i) t.rgba[i] = half_to_float(float_to_half(img.rgba[i]));
    const int idx = (int)textures->size();
    textures->push_back(std::move(t));
    (*by_path)[path] = idx;
    return idx;
}

// jsonParser.cpp:537-617
static void parse_shaders(const JsonValue &list, const std::string &scene_dir, ShaderNames *names,
                          std::vector<aur_shader> *shaders, std::map<std::string, int> *tex_by_path,
                          std::vector<HostTexture> *textures) {
    for (auto &kv : list.obj) {
        const JsonValue &o = kv.second;
        aur_shader k;
        std::memset(&k, 0, sizeof k);
        const std::string &at = o.at("attributetype").as_string();
        if (at == "color") k.kind = AUR_SHADER_COLOR;
        else if (at == "float") k.kind = AUR_SHADER_FLOAT;
        else throw std::runtime_error("shader " + kv.first + ": unknown attribute type " + at);
        const std::string &type = o.at("type").as_string();
        if (type == "texture2DShader") {
            k.type = AUR_SHADER_TEXTURE2D;
        } else if (type == "triplanarTextureShader") {
            k.type = AUR_SHADER_TRIPLANAR;
            k.scale = (float)o.at("scale").as_double();
            if (k.scale == 0.f) throw std::runtime_error("shader " + kv.first + ": scale must not be zero");
        } else {
            throw std::runtime_error("shader " + kv.first + ": unknown shader type " + type);
        }
        k.texture = load_texture(o.at("texturename").as_string(), scene_dir, tex_by_path, textures);
        (*names)[kv.first] = (int)shaders->size();
        shaders->push_back(k);
    }
}

void light_corners(const aur_light &l, float pc[4][3], float N[3]) {
    Mat4 M;
    std::memcpy(M.m, l.m, sizeof M.m);
    float c[4][3] = {{-l.sx, -l.sy, 0}, {-l.sx, l.sy, 0}, {l.sx, l.sy, 0}, {l.sx, -l.sy, 0}};
    for (int i = 0; i < 4; ++i) xpoint(M, c[i], pc[i]);
    // squareLight.cpp:35-38: normalize(cross(p1 - p2, p2 - p3))
    float v1[3] = {pc[0][0] - pc[1][0], pc[0][1] - pc[1][1], pc[0][2] - pc[1][2]};
    float v2[3] = {pc[1][0] - pc[2][0], pc[1][1] - pc[2][1], pc[1][2] - pc[2][2]};
    float n[3] = {(v1[1] * v2[2]) - (v1[2] * v2[1]), (v1[2] * v2[0]) - (v1[0] * v2[2]), (v1[0] * v2[1]) - (v1[1] * v2[0])};
    float len = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len == 0.f) len = 0.0001f;
    float inv = 1.f / len;
    N[0] = n[0] * inv; N[1] = n[1] * inv; N[2] = n[2] * inv;
}

// Light quad as two triangles, object vertex order {0,1,3, 2,1,3}
// (squareLight.cpp shape()), camera-space corners.
void light_geometry(const aur_light &l, float verts[12], float n[3]) {
    float pc[4][3];
    light_corners(l, pc, n);
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) verts[3 * i + k] = pc[i][k];
}

HostScene::HostScene() {
    Mat4 I = Mat4::identity();
    std::memcpy(cam_m, I.m, sizeof cam_m);
    std::memcpy(cam_inv, I.m, sizeof cam_inv);
}

std::vector<aur_texture> texture_table(const HostScene &s, int64_t *n_texels) {
    std::vector<aur_texture> out(s.textures.size());
    int64_t off = 0;
    for (size_t i = 0; i < s.textures.size(); ++i) {
        out[i].width = s.textures[i].width;
        out[i].height = s.textures[i].height;
        out[i].offset = off;
        off += (int64_t)s.textures[i].width * s.textures[i].height;
    }
    if (n_texels) *n_texels = off;
    return out;
}

void load_scene(const std::string &scene_path, const std::string &config_path, HostScene *out) {
    HostScene s;
    JsonValue root = JsonReader(read_file(scene_path)).parse();
    if (root.kind != JsonValue::Object || root.obj.empty()) throw std::runtime_error("no keys in json object: " + scene_path);
    std::string scene_dir;
    size_t slash = scene_path.find_last_of('/');
    if (slash != std::string::npos) scene_dir = scene_path.substr(0, slash + 1);

    // jsonParser.cpp:431-475 (all six keys are required)
    if (root.has("options")) {
        const JsonValue &o = root.at("options");
        s.opt.pixelsamples = o.at("pixelsamples").as_int();
        s.opt.lightsamples = o.at("lightsamples").as_int();
        s.opt.mindepth = o.at("mindepth").as_int();
        s.opt.maxdepth = o.at("maxdepth").as_int();
        const JsonValue &r = o.at("resolution");
        if (r.kind != JsonValue::Array || r.arr.size() != 2) throw std::runtime_error("resolution needs two values");
        s.opt.width = r.arr[0].as_int();
        s.opt.height = r.arr[1].as_int();
        s.opt.filename = string_template(o.at("filename").as_string(), config_path);
    }
    // jsonParser.cpp:477-505
    Xform cam;
    if (root.has("renderCam")) {
        const JsonValue &c = root.at("renderCam");
        s.opt.fov = (float)c.at("fov").as_double();
        cam = transforms_attr(c);
    }
    Mat4 inv_cam = cam.inv;  // Transform::inverse swaps m and mInv (transform.cpp:36-38)
    std::memcpy(s.cam_m, cam.m.m, sizeof s.cam_m);
    std::memcpy(s.cam_inv, cam.inv.m, sizeof s.cam_inv);

    // shaders, then materials (jsonParser.cpp:49-50, 507-535, 580-617)
    ShaderNames shader_names;
    std::map<std::string, int> tex_by_path;
    if (root.has("shaders")) parse_shaders(root.at("shaders"), scene_dir, &shader_names, &s.shaders, &tex_by_path, &s.textures);
    std::map<std::string, aur_material> mats;
    if (root.has("materials"))
        for (auto &kv : root.at("materials").obj) mats[kv.first] = parse_material(kv.second, shader_names, &s.shaders);

    // objects in sorted key order (jsonParser.cpp:619-661)
    std::vector<aur_light> lights;
    std::vector<std::string> obj_paths;  // OBJ file of each geometry object, read below
    if (root.has("objects")) {
        for (auto &kv : root.at("objects").obj) {
            const JsonValue &o = kv.second;
            const std::string &type = o.at("type").as_string();
            if (type == "geometry") {
                // jsonParser.cpp:225-241
                HostObject obj;
                obj.name = o.at("name").as_string();
                obj.material = o.at("material").as_string();
                obj.subdlevels = o.at("subdlevels").as_int();
                Xform t = transforms_attr(o);
                Mat4 o2c = inv_cam.mul(t.m);
                std::memcpy(obj.o2c, o2c.m, sizeof obj.o2c);
                // the shape's cameraToObject: (invCam * objTransform)'s inverse
                // matrix, objTransform.mInv * cam.m (transform.cpp:189-194)
                Mat4 c2o = t.inv.mul(cam.m);
                std::memcpy(obj.c2o, c2o.m, sizeof obj.c2o);
                std::string path = string_template(o.at("path").as_string(), config_path);
                if (!path.empty() && path[0] != '/') {
                    std::ifstream probe(path.c_str());
                    if (!probe.good()) path = scene_dir + path;
                }
                if (o.has("displace")) {
                    const JsonValue &d = o.at("displace");
                    obj.displace.on = true;
                    obj.displace.amplitude = (float)d.at("amplitude").as_double();
                    obj.displace.octaves = d.has("octaves") ? d.at("octaves").as_int() : 3;
                    obj.displace.frequency = d.has("frequency") ? (float)d.at("frequency").as_double() : 1.f;
                    obj.displace.seed = d.has("seed") ? (uint32_t)d.at("seed").as_double() : 0u;
                }
                if (mats.find(obj.material) == mats.end()) throw std::runtime_error("unknown material: " + obj.material);
                obj_paths.push_back(path);
                s.objects.push_back(std::move(obj));
            } else if (type == "light") {
                // jsonParser.cpp:243-263
                aur_light l;
                std::memset(&l, 0, sizeof l);
                l.type = AUR_LIGHT_AREA;
                Xform t = transforms_attr(o);
                Mat4 o2c = inv_cam.mul(t.m);
                std::memcpy(l.m, o2c.m, sizeof l.m);
                l.envmap = -1;
                l.sx = (float)o.at("scaleX").as_double();
                l.sy = (float)o.at("scaleY").as_double();
                l.exposure = (float)o.at("exposure").as_double();
                color_attr(o, "color", l.color);
                o.at("name").as_string();
                lights.push_back(l);
            } else if (type == "envlight") {
                // jsonParser.cpp:265-282; the map is oriented by the camera
                // transform alone (c2w / w2c), the light's own transform is unused
                aur_light l;
                std::memset(&l, 0, sizeof l);
                l.type = AUR_LIGHT_ENV;
                Xform t = transforms_attr(o);
                std::memcpy(l.m, t.m.m, sizeof l.m);
                l.exposure = (float)o.at("exposure").as_double();
                color_attr(o, "color", l.color);
                const std::string &map = o.at("envmap").as_string();
                l.envmap = map.empty() ? -1 : load_texture(map, scene_dir, &tex_by_path, &s.textures);
                lights.push_back(l);
            }
            // unknown object types are skipped with a warning in the reference
        }
    }
    // the meshes, read concurrently (one file per task)
    {
        const size_t n = s.objects.size();
        std::vector<std::exception_ptr> err(n);
        std::atomic<size_t> next(0);
        auto work = [&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) {
                try {
                    load_obj(obj_paths[i], s.objects[i].subdlevels == 0, &s.objects[i]);
                } catch (...) {
                    err[i] = std::current_exception();
                }
            }
        };
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t n_threads = std::min<size_t>(std::min<size_t>(n, hw), 16);
        std::vector<std::thread> pool;
        for (size_t t = 1; t < n_threads; ++t) pool.emplace_back(work);
        work();
        for (std::thread &t : pool) t.join();
        for (size_t i = 0; i < n; ++i)
            if (err[i]) std::rethrow_exception(err[i]);
    }
    // attributes: objects first, then lights (renderer.cpp:118-134)
    for (size_t i = 0; i < s.objects.size(); ++i) {
        s.objects[i].attr = (int)i;
        aur_material m = mats[s.objects[i].material];
        std::memcpy(m.c2o, s.objects[i].c2o, sizeof m.c2o);
        s.attr_mats.push_back(m);
    }
    for (size_t i = 0; i < lights.size(); ++i) {
        aur_material m;
        std::memset(&m, 0, sizeof m);
        m.type = AUR_MAT_LIGHT;
        for (int k = 0; k < AUR_N_SLOTS; ++k) m.shader[k] = -1;
        if (lights[i].type == AUR_LIGHT_AREA) {
            // SquareLight::emission(): color * powf(2, exposure)
            float e = powf(2, lights[i].exposure);
            for (int k = 0; k < 3; ++k) m.emission[k] = lights[i].color[k] * e;
        }
        s.light_attr.push_back((int)(s.objects.size() + i));
        s.attr_mats.push_back(m);
    }
    s.lights = lights;
    *out = std::move(s);
}

}  // namespace aur
