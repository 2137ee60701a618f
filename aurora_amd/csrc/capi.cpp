// capi.cpp -- the extern "C" boundary (include/aurora_hip.h).
//
// Every entry point catches C++ exceptions and converts them to error codes
// with a thread-local message; nothing throws or aborts across the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "context.h"
#include "exr_io.h"
#include "subdiv.h"

using namespace aur;

struct aur_ctx {
    Ctx c;
};

static thread_local std::string g_last_error;

template <class F>
static int guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return AUR_OK;
    } catch (const HipError &e) {
        g_last_error = e.what();
        return AUR_EHIP;
    } catch (const StateError &e) {
        g_last_error = e.what();
        return AUR_ESTATE;
    } catch (const CapacityError &e) {
        g_last_error = e.what();
        return AUR_ECAPACITY;
    } catch (const IoError &e) {
        g_last_error = e.what();
        return AUR_EIO;
    } catch (const std::invalid_argument &e) {
        g_last_error = e.what();
        return AUR_EINVAL;
    } catch (const std::ios_base::failure &e) {
        g_last_error = e.what();
        return AUR_EIO;
    } catch (const std::bad_alloc &) {
        g_last_error = "out of host memory";
        return AUR_ECAPACITY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        std::string m = e.what();
        if (m.rfind("cannot open", 0) == 0) return AUR_EIO;
        return AUR_EPARSE;
    } catch (...) {
        g_last_error = "unknown error";
        return AUR_EPARSE;
    }
}

#define NEED(ptr)                                                      \
    do {                                                               \
        if (!(ptr)) throw std::invalid_argument("null argument: " #ptr); \
    } while (0)

static void set_device(Ctx *c) {
    if (c->device < 0) throw StateError("host-only context (created with device -1) cannot run GPU work");
    HIP_CHECK(hipSetDevice(c->device));
}

static bool needs_gpu_flatten(const HostScene &s) {
    for (const HostObject &o : s.objects)
        if (o.subdlevels > 0 || o.displace.on) return true;
    return false;
}

// Radiance sources must be finite.  The reference takes color * 2^exposure
// in float for square lights (squareLight.h:84-86) and in double for the
// environment (infiniteAreaLight.cpp:61-78), and stores textures as FP16
// (Imf::Rgba); an exposure of 128 or a texel above 65504 turns into +inf
// there, and an infinite radiance makes the firefly clamp of
// integrator.cpp:277-281 loop forever.  Such scenes are refused here.
static void check_radiance(const HostScene &s) {
    for (const aur_light &l : s.lights) {
        if (!std::isfinite(l.exposure)) throw std::invalid_argument("light exposure is not finite");
        for (int k = 0; k < 3; ++k) {
            const float le = l.type == AUR_LIGHT_AREA ? l.color[k] * powf(2, l.exposure)
                                                      : (float)(l.color[k] * std::pow(2.0, (double)l.exposure));
            if (!std::isfinite(le)) throw std::invalid_argument("light radiance color * 2^exposure is not finite");
        }
    }
    for (const aur_material &m : s.attr_mats)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(m.emission[k])) throw std::invalid_argument("material emission is not finite");
    for (const HostTexture &t : s.textures)
        for (float v : t.rgba)
            if (!std::isfinite(half_to_float(float_to_half(v))))
                throw std::invalid_argument("texture holds a texel that is not finite as FP16");
}

// ---------------------------------------------------------------------------
// device scene upload
// ---------------------------------------------------------------------------
static DLight make_dlight(const aur_light &l) {
    DLight d;
    std::memset(&d, 0, sizeof d);
    d.type = l.type == AUR_LIGHT_AREA ? LIGHT_AREA : LIGHT_ENV;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k) d.m[4 * r + k] = l.m[4 * r + k];
    d.sx = l.sx;
    d.sy = l.sy;
    d.intensity = powf(2, l.exposure);  // light.h:42-45
    if (d.type == LIGHT_AREA) {
        for (int k = 0; k < 3; ++k) d.Le[k] = l.color[k] * d.intensity;  // squareLight.h:84-86
        float pc[4][3], N[3];
        light_corners(l, pc, N);
        for (int k = 0; k < 3; ++k) d.N[k] = N[k];
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 3; ++k) d.pcam[i][k] = pc[i][k];
        // RenderableTriangle(p1, p2, p4) and (p3, p2, p4): e1 = p2 - p, e2 = p4 - p
        const int base[2] = {0, 2};
        for (int t = 0; t < 2; ++t) {
            const float *p = pc[base[t]];
            for (int k = 0; k < 3; ++k) {
                d.tri[t][k] = p[k];
                d.tri[t][4 + k] = pc[1][k] - p[k];
                d.tri[t][8 + k] = pc[3][k] - p[k];
            }
        }
    } else {
        double e = std::pow(2.0, (double)l.exposure);  // infiniteAreaLight.cpp:61-78 (double pow)
        for (int k = 0; k < 3; ++k) d.Le[k] = (float)(l.color[k] * e);
        d.escale = e;
    }
    d.tex = (d.type == LIGHT_ENV) ? l.envmap : -1;
    return d;
}

// Texture2D::initPdf (texture2D.cpp:120-155) for env maps: the alpha channel
// becomes the FP16-rounded mean of r, g, b; the normalisation is the float sum
// of the unrounded means (row by row) over the texel count.
static float env_pdf_alpha(HostTexture &t) {
    float lum = 0.f;
    const size_t n = (size_t)t.width * t.height;
    for (size_t i = 0; i < n; ++i) {
        float *p = &t.rgba[4 * i];
        const float pdf = (p[0] + p[1] + p[2]) / 3.f;
        p[3] = half_to_float(float_to_half(pdf));
        lum += pdf;
    }
    return lum / (float)(t.width * t.height);
}

// The Halton source's tables (once per context) and its enumeration for the
// scene's resolution (device memory, read through SceneView::qmc).
static void ensure_halton(Ctx *c) {
    if (c->sampler != AUR_SAMPLER_HALTON || c->device < 0) return;
    if (!c->d_halton_dims.p) {
        std::vector<DHaltonDim> dims;
        std::vector<uint16_t> tables;
        halton_tables(&dims, &tables);
        c->d_halton_dims.reserve(sizeof(DHaltonDim) * dims.size());
        c->d_halton_tables.reserve(sizeof(uint16_t) * tables.size());
        HIP_CHECK(hipMemcpy(c->d_halton_dims.p, dims.data(), sizeof(DHaltonDim) * dims.size(), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(c->d_halton_tables.p, tables.data(), sizeof(uint16_t) * tables.size(), hipMemcpyHostToDevice));
        c->halton_w = c->halton_h = 0;
    }
    const int w = c->scene.opt.width, h = c->scene.opt.height;
    if (w == c->halton_w && h == c->halton_h && c->d_halton_view.p) return;
    HaltonView v;
    halton_enum(w, h, &v);
    v.dims = c->d_halton_dims.as<DHaltonDim>();
    v.tables = c->d_halton_tables.as<uint16_t>();
    c->d_halton_view.reserve(sizeof(HaltonView));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipMemcpy(c->d_halton_view.p, &v, sizeof v, hipMemcpyHostToDevice));
    c->halton_w = w;
    c->halton_h = h;
}

static void upload_shading(Ctx *c) {
    HostScene &s = c->scene;
    const int nt = (int)(s.tris.size() / 3);
    // texture coordinates, only when some material reads them
    c->has_uv = !c->ptex.empty();
    for (const aur_shader &k : s.shaders) c->has_uv = c->has_uv || k.type == AUR_SHADER_TEXTURE2D;
    if (c->has_uv && nt > 0) {
        if (s.uvs.size() != (size_t)6 * nt) throw std::invalid_argument("uv array has the wrong size");
        c->d_uv.reserve(sizeof(float) * s.uvs.size());
        HIP_CHECK(hipMemcpyAsync(c->d_uv.p, s.uvs.data(), sizeof(float) * s.uvs.size(), hipMemcpyHostToDevice, c->stream));
    } else {
        c->has_uv = false;
    }
    if (!c->ptex.empty()) {
        c->d_ptex.reserve(sizeof(float) * c->ptex.size());
        HIP_CHECK(hipMemcpyAsync(c->d_ptex.p, c->ptex.data(), sizeof(float) * c->ptex.size(), hipMemcpyHostToDevice,
                                 c->stream));
    }
    // per-attribute slots and transforms
    std::vector<DAttrX> ax(s.attr_mats.size());
    c->has_attrx = false;
    for (size_t a = 0; a < s.attr_mats.size(); ++a) {
        const aur_material &m = s.attr_mats[a];
        DAttrX &x = ax[a];
        std::memset(&x, 0, sizeof x);
        for (int k = 0; k < 8; ++k) x.slot[k] = -1;
        int mask = 0;
        if (m.type != AUR_MAT_LIGHT)
            for (int k = 0; k < AUR_N_SLOTS; ++k) {
                const int sh = m.shader[k];
                if (sh < 0) continue;
                if (sh >= (int)s.shaders.size()) throw std::invalid_argument("material shader index out of range");
                x.slot[k] = sh;
                mask |= 1 << k;
            }
        std::memcpy(x.c2o, m.c2o, sizeof x.c2o);
        // the albedo texture replaces the diffuse colour (and its shader)
        if (std::find(c->ptex_attrs.begin(), c->ptex_attrs.end(), (int)a) != c->ptex_attrs.end() && !c->ptex.empty())
            mask = (mask & ~(1 << AUR_SLOT_DIFFUSE)) | TEXMASK_PTEX;
        c->dmats[a].texmask = mask;
        c->has_attrx = c->has_attrx || (mask & ~TEXMASK_PTEX) != 0;
    }
    if (c->has_attrx) {
        c->d_attrx.reserve(sizeof(DAttrX) * ax.size());
        HIP_CHECK(hipMemcpyAsync(c->d_attrx.p, ax.data(), sizeof(DAttrX) * ax.size(), hipMemcpyHostToDevice, c->stream));
    }
    // shaders
    std::vector<DShader> ds(s.shaders.size());
    for (size_t i = 0; i < s.shaders.size(); ++i) {
        const aur_shader &k = s.shaders[i];
        DShader &d = ds[i];
        std::memset(&d, 0, sizeof d);
        if (k.type < AUR_SHADER_TEXTURE2D || k.type > AUR_SHADER_CONSTANT) throw std::invalid_argument("unknown shader type");
        if (k.kind != AUR_SHADER_COLOR && k.kind != AUR_SHADER_FLOAT) throw std::invalid_argument("unknown shader kind");
        if (k.type != AUR_SHADER_CONSTANT && (k.texture < 0 || k.texture >= (int)s.textures.size()))
            throw std::invalid_argument("shader texture index out of range");
        if (k.type == AUR_SHADER_TRIPLANAR && k.scale == 0.f) throw std::invalid_argument("triplanar scale must not be zero");
        d.type = k.type;
        d.kind = k.kind;
        d.tex = k.type == AUR_SHADER_CONSTANT ? -1 : k.texture;
        d.scale = k.scale;
        for (int j = 0; j < 3; ++j) d.value[j] = k.value[j];
    }
    if (!ds.empty()) {
        c->d_shaders.reserve(sizeof(DShader) * ds.size());
        HIP_CHECK(hipMemcpyAsync(c->d_shaders.p, ds.data(), sizeof(DShader) * ds.size(), hipMemcpyHostToDevice, c->stream));
    }
    // textures: FP16 texel pool; env maps get their pdf in alpha
    std::vector<char> env_use(s.textures.size(), 0);
    for (DLight &l : c->dlights) {
        if (l.type != LIGHT_ENV || l.tex < 0) continue;
        if (l.tex >= (int)s.textures.size()) throw std::invalid_argument("env map texture index out of range");
        env_use[l.tex] = 1;
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) {
                l.c2w[4 * r + k] = s.cam_m[4 * r + k];
                l.w2c[4 * r + k] = s.cam_inv[4 * r + k];
            }
    }
    int64_t n_texels = 0;
    std::vector<aur_texture> table = texture_table(s, &n_texels);
    std::vector<DTex> dt(s.textures.size());
    std::vector<ushort4> texels((size_t)n_texels);
    for (size_t i = 0; i < s.textures.size(); ++i) {
        HostTexture t = s.textures[i];
        if (t.width <= 0 || t.height <= 0 || t.rgba.size() != (size_t)4 * t.width * t.height)
            throw std::invalid_argument("bad texture");
        DTex &d = dt[i];
        std::memset(&d, 0, sizeof d);
        d.off = table[i].offset;
        d.w = t.width;
        d.h = t.height;
        d.pdf_norm = env_use[i] ? env_pdf_alpha(t) : 1.f;
        for (size_t p = 0; p < (size_t)t.width * t.height; ++p) {
            ushort4 &q = texels[(size_t)d.off + p];
            q.x = float_to_half(t.rgba[4 * p]);
            q.y = float_to_half(t.rgba[4 * p + 1]);
            q.z = float_to_half(t.rgba[4 * p + 2]);
            q.w = float_to_half(t.rgba[4 * p + 3]);
        }
    }
    if (!dt.empty()) {
        c->d_texs.reserve(sizeof(DTex) * dt.size());
        HIP_CHECK(hipMemcpyAsync(c->d_texs.p, dt.data(), sizeof(DTex) * dt.size(), hipMemcpyHostToDevice, c->stream));
    }
    if (!texels.empty()) {
        c->d_texels.reserve(sizeof(ushort4) * texels.size());
        HIP_CHECK(hipMemcpyAsync(c->d_texels.p, texels.data(), sizeof(ushort4) * texels.size(), hipMemcpyHostToDevice, c->stream));
    }
    // pageable sources: the copies above complete before this returns
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

static void upload(Ctx *c) {
    HostScene &s = c->scene;
    c->n_verts = (int)(s.verts.size() / 3);
    c->n_tris = (int)(s.tris.size() / 3);
    const int nt = c->n_tris;
    c->d_verts.reserve(sizeof(float) * std::max<size_t>(s.verts.size(), 1));
    c->d_tris.reserve(sizeof(int) * std::max<size_t>(s.tris.size(), 1));
    if (!s.verts.empty())
        HIP_CHECK(hipMemcpyAsync(c->d_verts.p, s.verts.data(), sizeof(float) * s.verts.size(), hipMemcpyHostToDevice, c->stream));
    if (!s.tris.empty())
        HIP_CHECK(hipMemcpyAsync(c->d_tris.p, s.tris.data(), sizeof(int) * s.tris.size(), hipMemcpyHostToDevice, c->stream));
    for (int i = 0; i < nt; ++i)
        for (int k = 0; k < 3; ++k) {
            int v = s.tris[3 * i + k];
            if (v < 0 || v >= c->n_verts) throw std::invalid_argument("triangle index out of range");
        }
    for (int a : s.attr)
        if (a < 0 || a >= (int)s.attr_mats.size()) throw std::invalid_argument("attribute index out of range");
    // shading records: (n0, attr), (n1, 0), (n2, 0)
    std::vector<float> shade((size_t)12 * std::max(nt, 1), 0.f);
    for (int i = 0; i < nt; ++i) {
        for (int cidx = 0; cidx < 3; ++cidx)
            for (int k = 0; k < 3; ++k) shade[12 * (size_t)i + 4 * cidx + k] = s.normals[9 * (size_t)i + 3 * cidx + k];
        int a = s.attr[i];
        std::memcpy(&shade[12 * (size_t)i + 3], &a, sizeof(int));
    }
    c->d_shade.reserve(sizeof(float) * shade.size());
    HIP_CHECK(hipMemcpyAsync(c->d_shade.p, shade.data(), sizeof(float) * shade.size(), hipMemcpyHostToDevice, c->stream));

    // materials and Kelemen tables (one per rounded reflectance and distribution,
    // as the reference loads one table file per reflectance, kelemenMaterial.cpp:47-81)
    c->dmats.clear();
    c->table_ids.clear();
    std::vector<std::pair<float, int>> tables;
    for (const aur_material &m : s.attr_mats) {
        DMat d;
        std::memset(&d, 0, sizeof d);
        d.type = m.type;
        d.dist = m.distribution;
        for (int k = 0; k < 3; ++k) {
            d.diffuse[k] = m.diffuse[k];
            d.spec[k] = m.spec[k];
            d.gain[k] = m.gain[k];
            d.emission[k] = m.emission[k];
        }
        d.rough = m.roughness;
        d.roughB = m.roughnessB;
        d.mix = m.mix;
        d.refl = m.reflectance;
        for (int k = 0; k < 3; ++k) d.transmit[k] = m.transmit[k];
        d.transmit[3] = m.ior;
        d.table = 0;
        if (m.type == AUR_MAT_KELEMEN) {
            int q = (int)std::lround(m.reflectance * 20.f);
            q = std::min(std::max(q, 0), 20);
            auto key = std::make_pair(q, m.distribution);
            auto it = c->table_ids.find(key);
            if (it == c->table_ids.end()) {
                int id = (int)tables.size();
                c->table_ids[key] = id;
                tables.push_back({q / 20.f, m.distribution});
                d.table = id;
            } else {
                d.table = it->second;
            }
        }
        c->dmats.push_back(d);
    }
    c->n_mats = (int)c->dmats.size();
    c->d_mats.reserve(sizeof(DMat) * std::max<size_t>(c->dmats.size(), 1));
    if (!c->dmats.empty())
        HIP_CHECK(hipMemcpyAsync(c->d_mats.p, c->dmats.data(), sizeof(DMat) * c->dmats.size(), hipMemcpyHostToDevice, c->stream));
    const size_t tsz = AUR_KEL_COS * AUR_KEL_ROUGH;
    c->d_tables.reserve(sizeof(float) * tsz * std::max<size_t>(tables.size(), 1));
    for (size_t t = 0; t < tables.size(); ++t)
        // the reference's table: SPECALBEDO_SAMPLES samples per entry, stored as FP16
        kelemen_table(c, tables[t].first, tables[t].second, AUR_KEL_SAMPLES, 0x4B454Cu, c->d_tables.as<float>() + t * tsz,
                      true);

    c->dlights.clear();
    c->n_env = 0;
    c->env_le[0] = c->env_le[1] = c->env_le[2] = 0.f;
    for (const aur_light &l : s.lights) {
        DLight d = make_dlight(l);
        if (d.type == LIGHT_ENV) {
            ++c->n_env;
            for (int k = 0; k < 3; ++k) c->env_le[k] += d.Le[k];
        }
        c->dlights.push_back(d);
    }
    upload_shading(c);  // sets DMat::texmask and the env maps' transforms
    ensure_halton(c);
    if (!c->dmats.empty())
        HIP_CHECK(hipMemcpyAsync(c->d_mats.p, c->dmats.data(), sizeof(DMat) * c->dmats.size(), hipMemcpyHostToDevice, c->stream));
    c->n_lights = (int)c->dlights.size();
    c->d_lights.reserve(sizeof(DLight) * std::max<size_t>(c->dlights.size(), 1));
    if (!c->dlights.empty())
        HIP_CHECK(hipMemcpyAsync(c->d_lights.p, c->dlights.data(), sizeof(DLight) * c->dlights.size(), hipMemcpyHostToDevice, c->stream));
    if (!c->tri_albedo.empty()) {
        if (c->tri_albedo.size() != (size_t)3 * nt) throw std::invalid_argument("per-triangle albedo has the wrong size");
        c->d_albedo.reserve(sizeof(float) * c->tri_albedo.size());
        HIP_CHECK(hipMemcpyAsync(c->d_albedo.p, c->tri_albedo.data(), sizeof(float) * c->tri_albedo.size(), hipMemcpyHostToDevice, c->stream));
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
}

// objects (sorted JSON order), then area-light quads (renderer.cpp:118-134)
static void flatten(Ctx *c) {
    HostScene &s = c->scene;
    s.verts.clear(); s.normals.clear(); s.tris.clear(); s.attr.clear(); s.uvs.clear();
    for (const HostObject &o : s.objects) {
        if (o.subdlevels > 0 || o.displace.on) {
            SubdivResult r = subdivide_object(c, o);
            int base = (int)(s.verts.size() / 3);
            s.verts.insert(s.verts.end(), r.verts.begin(), r.verts.end());
            for (size_t i = 0; i < r.tris.size(); ++i) s.tris.push_back(base + r.tris[i]);
            s.normals.insert(s.normals.end(), r.corner_normals.begin(), r.corner_normals.end());
            for (size_t i = 0; i < r.tris.size() / 3; ++i) s.attr.push_back(o.attr);
            s.uvs.insert(s.uvs.end(), r.corner_uvs.begin(), r.corner_uvs.end());
            continue;
        }
        // TriangleMesh::makeEmbree, triangleMesh.cpp:57-81
        int base = (int)(s.verts.size() / 3);
        size_t ncorner = o.P.size() / 3;
        for (size_t k = 0; k < ncorner; ++k) {
            float p[3];
            transform_point(o.o2c, &o.P[3 * k], p);
            s.verts.insert(s.verts.end(), p, p + 3);
            float n[3];
            transform_vector(o.o2c, &o.N[3 * k], n);
            s.normals.insert(s.normals.end(), n, n + 3);
            // triangleMesh.cpp:40-46, 73-75: uv() = (0, 0) without `vt`
            s.uvs.push_back(o.UV.empty() ? 0.f : o.UV[2 * k]);
            s.uvs.push_back(o.UV.empty() ? 0.f : o.UV[2 * k + 1]);
        }
        for (size_t t = 0; t < ncorner / 3; ++t) {
            s.tris.push_back(base + (int)(3 * t));
            s.tris.push_back(base + (int)(3 * t + 1));
            s.tris.push_back(base + (int)(3 * t + 2));
            s.attr.push_back(o.attr);
        }
    }
    for (size_t i = 0; i < s.lights.size(); ++i) {
        const aur_light &l = s.lights[i];
        if (l.type != AUR_LIGHT_AREA) continue;
        float v[12], n[3], nt[3];
        light_geometry(l, v, n);
        transform_vector(l.m, n, nt);
        int base = (int)(s.verts.size() / 3);
        s.verts.insert(s.verts.end(), v, v + 12);
        const int idx[6] = {0, 1, 3, 2, 1, 3};
        for (int t = 0; t < 2; ++t) {
            for (int k = 0; k < 3; ++k) {
                s.tris.push_back(base + idx[3 * t + k]);
                s.normals.insert(s.normals.end(), nt, nt + 3);
                s.uvs.push_back(0.f);
                s.uvs.push_back(0.f);
            }
            s.attr.push_back(s.light_attr[i]);
        }
    }
    s.flat = true;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *aur_last_error(void) { return g_last_error.c_str(); }

int aur_create(aur_ctx **out, int hip_device) {
    return guarded([&] {
        NEED(out);
        *out = nullptr;
        if (hip_device == -1) {
            // host-only context: scene parsing and flattening, no GPU work
            std::unique_ptr<aur_ctx> c(new aur_ctx());
            c->c.device = -1;
            *out = c.release();
            return;
        }
        int n = 0;
        HIP_CHECK(hipGetDeviceCount(&n));
        if (hip_device < 0 || hip_device >= n) throw std::invalid_argument("no such HIP device");
        std::unique_ptr<aur_ctx> c(new aur_ctx());
        c->c.device = hip_device;
        set_device(&c->c);
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, hip_device));
        c->c.n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        c->c.launch.shade_list_tiles = 16 * c->c.n_cu;  // four tiles per shade workgroup
        noalloc_init(hip_device);  // the device is current (set_device above)
        HIP_CHECK(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
        c->c.own_stream = true;
        ensure_lane_streams(&c->c);
        *out = c.release();
    });
}

void aur_destroy(aur_ctx *ctx) {
    if (!ctx) return;
    try {
        if (ctx->c.device < 0) {
            delete ctx;
            return;
        }
        (void)hipSetDevice(ctx->c.device);
        (void)hipStreamSynchronize(ctx->c.stream);
        try {
            timer_release(&ctx->c);
        } catch (...) {
        }
        hipStream_t s = ctx->c.stream;
        bool own = ctx->c.own_stream;
        if (ctx->c.lane_stream[0]) {
            for (hipStream_t ls : ctx->c.lane_stream) {
                (void)hipStreamSynchronize(ls);
                (void)hipStreamDestroy(ls);
            }
            for (int l = 0; l < 2; ++l) {
                (void)hipEventDestroy(ctx->c.lane_fork[l]);
                (void)hipEventDestroy(ctx->c.lane_shaded[l]);
                (void)hipEventDestroy(ctx->c.lane_done[l]);
                (void)hipEventDestroy(ctx->c.ev_resolved[l]);
            }
            (void)hipEventDestroy(ctx->c.ev_start);
            (void)hipEventDestroy(ctx->c.ev_join);
        }
        delete ctx;
        if (own && s) (void)hipStreamDestroy(s);
    } catch (...) {
    }
}

int aur_set_stream(aur_ctx *ctx, void *hip_stream) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        set_device(c);
        HIP_CHECK(hipStreamSynchronize(c->stream));
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        if (hip_stream) {
            c->stream = (hipStream_t)hip_stream;
            c->own_stream = false;
        } else {
            HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
    });
}

int aur_set_seed(aur_ctx *ctx, uint32_t seed) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        ctx->c.seed = seed;
    });
}

int aur_scene_load(aur_ctx *ctx, const char *scene_path, const char *config_json_path) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        NEED(scene_path);
        HostScene s;
        std::string cfg = config_json_path ? config_json_path : "";
        if (cfg.empty()) {
            // the reference reads one fixed config.json (constants.h:46); without an
            // explicit path, use the one next to the scene file when present
            std::string sp = scene_path;
            size_t slash = sp.find_last_of('/');
            std::string cand = (slash == std::string::npos ? std::string() : sp.substr(0, slash + 1)) + "config.json";
            FILE *f = std::fopen(cand.c_str(), "r");
            if (f) { std::fclose(f); cfg = cand; }
        }
        load_scene(scene_path, cfg, &s);
        check_radiance(s);
        ctx->c.scene = std::move(s);
        ctx->c.have_scene = true;
        ctx->c.built = false;
        ctx->c.tri_albedo.clear();
        ctx->c.ptex.clear();
        ctx->c.ptex_attrs.clear();
    });
}

int aur_scene_upload(aur_ctx *ctx, const aur_scene_desc *d) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        NEED(d);
        if (d->n_tris < 0 || d->n_verts < 0 || d->n_attrs < 0 || d->n_lights < 0) throw std::invalid_argument("negative count");
        if (d->n_tris > 0) { NEED(d->verts); NEED(d->tris); NEED(d->normals); NEED(d->attr); }
        if (d->n_attrs > 0) NEED(d->mats);
        if (d->n_lights > 0) NEED(d->lights);
        if (d->width <= 0 || d->height <= 0) throw std::invalid_argument("resolution must be positive");
        if (d->maxdepth < 0 || d->maxdepth > 200) throw std::invalid_argument("maxdepth out of range");
        HostScene s;
        s.opt.width = d->width;
        s.opt.height = d->height;
        s.opt.pixelsamples = d->pixelsamples;
        s.opt.lightsamples = d->lightsamples;
        s.opt.mindepth = d->mindepth;
        s.opt.maxdepth = d->maxdepth;
        s.opt.fov = d->fov;
        s.verts.assign(d->verts, d->verts + 3 * (size_t)d->n_verts);
        s.tris.assign(d->tris, d->tris + 3 * (size_t)d->n_tris);
        s.normals.assign(d->normals, d->normals + 9 * (size_t)d->n_tris);
        s.attr.assign(d->attr, d->attr + d->n_tris);
        s.attr_mats.assign(d->mats, d->mats + d->n_attrs);
        s.lights.assign(d->lights, d->lights + d->n_lights);
        if (d->uvs) s.uvs.assign(d->uvs, d->uvs + 6 * (size_t)d->n_tris);
        else s.uvs.assign(6 * (size_t)d->n_tris, 0.f);
        if (d->n_shaders < 0 || d->n_textures < 0 || d->n_texels < 0) throw std::invalid_argument("negative count");
        if (d->n_shaders > 0) { NEED(d->shaders); s.shaders.assign(d->shaders, d->shaders + d->n_shaders); }
        if (d->n_textures > 0) {
            NEED(d->textures);
            NEED(d->texels);
            for (int i = 0; i < d->n_textures; ++i) {
                const aur_texture &t = d->textures[i];
                if (t.width <= 0 || t.height <= 0 || t.offset < 0 || t.offset + (int64_t)t.width * t.height > d->n_texels)
                    throw std::invalid_argument("texture outside the texel pool");
                HostTexture h;
                h.width = t.width;
                h.height = t.height;
                const float *p = d->texels + 4 * t.offset;
                h.rgba.resize((size_t)4 * t.width * t.height);
                // the reference holds FP16 texels (Imf::Rgba)
                for (size_t k = 0; k < h.rgba.size(); ++k) h.rgba[k] = half_to_float(float_to_half(p[k]));
                s.textures.push_back(std::move(h));
            }
        }
        std::memcpy(s.cam_m, d->cam_to_world, sizeof s.cam_m);
        std::memcpy(s.cam_inv, d->world_to_cam, sizeof s.cam_inv);
        for (const aur_shader &k : s.shaders)
            if (k.type != AUR_SHADER_CONSTANT && (k.texture < 0 || k.texture >= d->n_textures))
                throw std::invalid_argument("shader texture index out of range");
        for (const aur_material &m : s.attr_mats)
            for (int k = 0; k < AUR_N_SLOTS; ++k)
                if (m.shader[k] >= d->n_shaders) throw std::invalid_argument("material shader index out of range");
        for (const aur_light &l : s.lights)
            if (l.type == AUR_LIGHT_ENV && l.envmap >= d->n_textures) throw std::invalid_argument("env map index out of range");
        for (int v : s.tris)
            if (v < 0 || v >= d->n_verts) throw std::invalid_argument("triangle index out of range");
        for (int a : s.attr)
            if (a < 0 || a >= d->n_attrs) throw std::invalid_argument("attribute index out of range");
        for (const aur_material &m : s.attr_mats)
            if (m.type < AUR_MAT_MATTE || m.type > AUR_MAT_GLASS) throw std::invalid_argument("unknown material type");
        for (const aur_light &l : s.lights)
            if (l.type != AUR_LIGHT_AREA && l.type != AUR_LIGHT_ENV) throw std::invalid_argument("unknown light type");
        check_radiance(s);
        s.flat = true;
        ctx->c.scene = std::move(s);
        ctx->c.have_scene = true;
        ctx->c.built = false;
        ctx->c.tri_albedo.clear();
        ctx->c.ptex.clear();
        ctx->c.ptex_attrs.clear();
    });
}

int aur_set_albedo_texture(aur_ctx *ctx, int width, int height, const float *rgb, const int *attrs, int n_attrs) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        Ctx *c = &ctx->c;
        if (!c->built) throw StateError("the albedo texture needs a built scene (call aur_build_accel first)");
        if (!rgb) {
            c->ptex.clear();
            c->ptex_attrs.clear();
        } else {
            if (width < 1 || height < 1 || (int64_t)width * height > ((int64_t)1 << 28))
                throw std::invalid_argument("albedo texture size out of range");
            if (n_attrs < 1 || !attrs) throw std::invalid_argument("the albedo texture needs at least one material");
            const HostScene &s = c->scene;
            const int nt = (int)(s.tris.size() / 3);
            if (s.uvs.size() != (size_t)6 * nt) throw std::invalid_argument("the albedo texture needs per-corner uvs");
            std::vector<int> as(attrs, attrs + n_attrs);
            for (int a : as) {
                if (a < 0 || a >= (int)s.attr_mats.size()) throw std::invalid_argument("material index out of range");
                const int t = s.attr_mats[a].type;
                if (t != AUR_MAT_MATTE && t != AUR_MAT_KELEMEN)
                    throw std::invalid_argument("the albedo texture drives matte and Kelemen diffuse colours only");
            }
            const bool same = !c->ptex.empty() && width == c->ptex_w && height == c->ptex_h && as == c->ptex_attrs;
            c->ptex.assign(rgb, rgb + (size_t)3 * width * height);
            c->ptex_w = width;
            c->ptex_h = height;
            c->ptex_attrs = as;
            if (same) {
                // new texel values only (an optimiser's step)
                set_device(c);
                HIP_CHECK(hipMemcpyAsync(c->d_ptex.p, c->ptex.data(), sizeof(float) * c->ptex.size(),
                                         hipMemcpyHostToDevice, c->stream));
                HIP_CHECK(hipStreamSynchronize(c->stream));
                return;
            }
        }
        set_device(c);
        upload_shading(c);
        if (!c->dmats.empty())
            HIP_CHECK(hipMemcpyAsync(c->d_mats.p, c->dmats.data(), sizeof(DMat) * c->dmats.size(),
                                     hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int aur_set_tri_albedo(aur_ctx *ctx, const float *albedo_rgb) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        Ctx *c = &ctx->c;
        if (!albedo_rgb) { c->tri_albedo.clear(); return; }
        if (!c->scene.flat) throw StateError("per-triangle albedo needs a flattened scene (call aur_build_accel first)");
        size_t n = c->scene.tris.size();
        c->tri_albedo.assign(albedo_rgb, albedo_rgb + n);
        if (c->built) {
            set_device(c);
            c->d_albedo.reserve(sizeof(float) * n);
            HIP_CHECK(hipMemcpyAsync(c->d_albedo.p, c->tri_albedo.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
        }
    });
}

int aur_build_accel(aur_ctx *ctx, const aur_accel_opts *opts) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        Ctx *c = &ctx->c;
        if (!c->have_scene) throw StateError("no scene loaded");
        set_device(c);
        int width = opts && opts->width ? opts->width : 4;
        int leaf = opts && opts->max_leaf ? opts->max_leaf : 2;
        if (width != 4 && width != 8) throw std::invalid_argument("BVH width must be 4 or 8");
        if (leaf < 1 || leaf > 16) throw std::invalid_argument("max_leaf must be in 1..16");
        if (!c->scene.flat) flatten(c);
        upload(c);
        build_bvh(c, width, leaf);
        c->built = true;
    });
}

int aur_scene_info_get(aur_ctx *ctx, aur_scene_info *out) {
    return guarded([&] {
        NEED(ctx);
        NEED(out);
        Ctx *c = &ctx->c;
        if (!c->have_scene) throw StateError("no scene loaded");
        if (!c->scene.flat && !needs_gpu_flatten(c->scene)) flatten(c);
        const HostScene &s = c->scene;
        std::memset(out, 0, sizeof *out);
        out->n_verts = (int)(s.verts.size() / 3);
        out->n_tris = (int)(s.tris.size() / 3);
        out->n_attrs = (int)s.attr_mats.size();
        out->n_lights = (int)s.lights.size();
        out->n_objects = (int)s.objects.size();
        out->width = s.opt.width;
        out->height = s.opt.height;
        out->pixelsamples = s.opt.pixelsamples;
        out->lightsamples = s.opt.lightsamples;
        out->mindepth = s.opt.mindepth;
        out->maxdepth = s.opt.maxdepth;
        out->fov = s.opt.fov;
        out->bvh_width = c->bvh.width;
        out->bvh_nodes = c->bvh.n_nodes;
        out->bvh_depth = c->bvh.depth;
        out->bvh_max_leaf = c->bvh.max_leaf;
        out->n_shaders = (int)s.shaders.size();
        out->n_textures = (int)s.textures.size();
        int64_t n_texels = 0;
        texture_table(s, &n_texels);
        out->n_texels = n_texels;
    });
}

int aur_bvh_export(aur_ctx *ctx, void *nodes, float *tri_records) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (!c->built) throw StateError("no acceleration structure (call aur_build_accel first)");
        set_device(c);
        HIP_CHECK(hipStreamSynchronize(c->stream));
        if (nodes)
            HIP_CHECK(hipMemcpy(nodes, c->bvh.nodes.p, (size_t)c->bvh.n_nodes * 32 * c->bvh.width, hipMemcpyDeviceToHost));
        if (tri_records)
            HIP_CHECK(hipMemcpy(tri_records, c->bvh.tri_sorted.p, sizeof(float) * 12 * (size_t)c->n_tris,
                                hipMemcpyDeviceToHost));
    });
}

int aur_scene_export(aur_ctx *ctx, float *verts, int *tris, float *normals, int *attr, aur_material *mats, aur_light *lights) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (!c->have_scene) throw StateError("no scene loaded");
        if (!c->scene.flat) {
            if (needs_gpu_flatten(c->scene)) throw StateError("scene needs GPU subdivision (call aur_build_accel first)");
            flatten(c);
        }
        const HostScene &s = c->scene;
        if (verts) std::memcpy(verts, s.verts.data(), sizeof(float) * s.verts.size());
        if (tris) std::memcpy(tris, s.tris.data(), sizeof(int) * s.tris.size());
        if (normals) std::memcpy(normals, s.normals.data(), sizeof(float) * s.normals.size());
        if (attr) std::memcpy(attr, s.attr.data(), sizeof(int) * s.attr.size());
        if (mats) std::memcpy(mats, s.attr_mats.data(), sizeof(aur_material) * s.attr_mats.size());
        if (lights) std::memcpy(lights, s.lights.data(), sizeof(aur_light) * s.lights.size());
    });
}

int aur_scene_export_shading(aur_ctx *ctx, float *uvs, aur_shader *shaders, aur_texture *textures, float *texels,
                             float *cam_to_world_and_inverse) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (!c->have_scene) throw StateError("no scene loaded");
        if (!c->scene.flat) {
            if (needs_gpu_flatten(c->scene)) throw StateError("scene needs GPU subdivision (call aur_build_accel first)");
            flatten(c);
        }
        const HostScene &s = c->scene;
        if (uvs) {
            const size_t n = 6 * (s.tris.size() / 3);
            if (s.uvs.size() == n) std::memcpy(uvs, s.uvs.data(), sizeof(float) * n);
            else std::memset(uvs, 0, sizeof(float) * n);
        }
        if (shaders && !s.shaders.empty()) std::memcpy(shaders, s.shaders.data(), sizeof(aur_shader) * s.shaders.size());
        int64_t n_texels = 0;
        std::vector<aur_texture> table = texture_table(s, &n_texels);
        if (textures && !table.empty()) std::memcpy(textures, table.data(), sizeof(aur_texture) * table.size());
        if (texels)
            for (size_t i = 0; i < s.textures.size(); ++i)
                std::memcpy(texels + 4 * table[i].offset, s.textures[i].rgba.data(), sizeof(float) * s.textures[i].rgba.size());
        if (cam_to_world_and_inverse) {
            std::memcpy(cam_to_world_and_inverse, s.cam_m, sizeof(float) * 16);
            std::memcpy(cam_to_world_and_inverse + 16, s.cam_inv, sizeof(float) * 16);
        }
    });
}

// Passes [first_pass, first_pass + n_passes) into the accumulation buffer and
// the FP16 display; first_pass == 0 (or a new resolution) clears both.
static void render_range(Ctx *c, int first_pass, int n_passes, aur_stats *st) {
    if (!c->built) throw StateError("render before aur_build_accel");
    if (first_pass < 0 || n_passes < 0) throw std::invalid_argument("negative pass range");
    set_device(c);
    const int w = c->scene.opt.width, h = c->scene.opt.height;
    const size_t wh = (size_t)w * h;
    bool resized = c->fb_w != w || c->fb_h != h;
    c->accum.reserve(sizeof(float4) * wh);
    c->display.reserve(sizeof(ushort4) * wh);
    if (first_pass == 0 || resized) {
        fill_zero(c, c->accum.p, sizeof(float4) * wh);
        fill_zero(c, c->display.p, sizeof(ushort4) * wh);
        c->display_passes = 0;
    }
    c->fb_w = w;
    c->fb_h = h;
    std::memset(st, 0, sizeof *st);
    if (n_passes > 0) {
        PathSource src;
        src.mode = 0;
        src.first_pass = first_pass;
        src.pixel = src.sample = nullptr;
        src.slots = nullptr;
        src.ls = c->scene.opt.lightsamples;
        src.display = c->display.as<ushort4>();
        src.display_base = c->display_passes;
        if (src.ls < 1) throw std::invalid_argument("lightsamples must be at least 1");
        // kept records: the passes of this call inside the kept range also
        // write their per-path records (one path per pixel and pass)
        const int r0 = std::max(first_pass, c->rec_first), r1 = std::min(first_pass + n_passes, c->rec_first + c->rec_n);
        const bool keep = c->rec_n > 0 && c->path_rec.p && c->rec_wh == wh && src.ls == 1 && r0 < r1;
        if (keep) {
            src.adj_rec = c->path_rec.as<float4>();
            src.rec_base = (long long)(first_pass - c->rec_first) * (long long)wh;
            src.rec_count = (long long)c->rec_n * (long long)wh;
            src.rec_only = false;
        }
        run_paths(c, src, pass_paths(wh, first_pass, n_passes, src.ls), c->accum.as<float4>(), st);
        c->display_passes += n_passes;
        if (keep)
            for (int p = r0; p < r1; ++p) c->rec_version[p - c->rec_first] = c->version;
    }
}

int aur_render(aur_ctx *ctx, int first_pass, int n_passes, float *rgba_sum_host, aur_stats *out) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        aur_stats st;
        render_range(c, first_pass, n_passes, &st);
        if (rgba_sum_host) {
            const size_t wh = (size_t)c->fb_w * c->fb_h;
            HIP_CHECK(hipMemcpyAsync(rgba_sum_host, c->accum.p, sizeof(float4) * wh, hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
        }
        if (out) *out = st;
    });
}

// The display written by a host thread while the device traces on: one write
// in flight at a time (the reference waits for its draw task before the next
// pass, renderer.cpp:238), every exception carried back to the caller.
struct DisplayWriter {
    std::thread t;
    std::exception_ptr err;
    void wait() {
        if (t.joinable()) t.join();
        if (err) {
            std::exception_ptr e = err;
            err = nullptr;
            std::rethrow_exception(e);
        }
    }
    ~DisplayWriter() {
        if (t.joinable()) t.join();
    }
};

// Renderer::renderImage, outputStats and postRenderCleanup (renderer.cpp:200-304).
int aur_render_image(aur_ctx *ctx, const char *path, double draw_delay_s, aur_stats *out, int *n_draws) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (!c->built) throw StateError("render before aur_build_accel");
        const std::string fn = path ? std::string(path) : c->scene.opt.filename;
        if (fn.empty()) throw std::invalid_argument("render output filename is blank");  // renderer.cpp:217-219
        if (!(draw_delay_s <= 1e9)) throw std::invalid_argument("draw delay must be a number of seconds");
        const double delay = draw_delay_s > 0 ? draw_delay_s : 1.0;  // DRAW_DELAY (constants.h:20)
        const int total = c->scene.opt.pixelsamples;
        const int w = c->scene.opt.width, h = c->scene.opt.height;
        const size_t n = (size_t)w * h;
        using clock = std::chrono::steady_clock;
        const clock::time_point t0 = clock::now();
        clock::time_point last_draw = t0;
        std::vector<uint16_t> shown(4 * n);
        DisplayWriter writer;
        int draws = 0;
        auto draw = [&](std::vector<std::pair<std::string, std::string>> meta) {
            writer.wait();
            HIP_CHECK(hipMemcpyAsync(shown.data(), c->display.p, sizeof(uint16_t) * 4 * n, hipMemcpyDeviceToHost,
                                     c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
            meta.insert(meta.begin(), {"renderer", "Aurora (MI355X wavefront tracer)"});
            writer.t = std::thread([&writer, &shown, fn, w, h, n, meta] {
                try {
                    std::vector<float> img(4 * n);
                    for (size_t i = 0; i < 4 * n; ++i) img[i] = half_to_float(shown[i]);
                    exr_write(fn, w, h, img.data(), EXR_ZIP, true, meta);
                } catch (const std::exception &e) {
                    writer.err = std::make_exception_ptr(IoError(e.what()));
                }
            });
            ++draws;
        };
        aur_stats sum, st;
        std::memset(&sum, 0, sizeof sum);
        // Passes are traced in chunks sized to a quarter of the draw delay
        // (pass 0 alone: it is always drawn); the reference's per-pass check
        // then falls on the end of the first chunk that ends past the delay.
        int done = 0;
        double per_pass_s = 0.0;
        render_range(c, 0, 0, &st);
        while (done < total) {
            int chunk = 1;
            if (done > 0 && per_pass_s > 0.0)
                chunk = (int)std::min<double>(total - done, std::max(1.0, std::floor(0.25 * delay / per_pass_s)));
            const clock::time_point ts = clock::now();
            render_range(c, done, chunk, &st);
            HIP_CHECK(hipStreamSynchronize(c->stream));
            const clock::time_point te = clock::now();
            per_pass_s = std::chrono::duration<double>(te - ts).count() / chunk;
            sum.paths += st.paths;
            sum.ext_rays += st.ext_rays;
            sum.hit_rays += st.hit_rays;
            sum.shadow_rays += st.shadow_rays;
            sum.ms += st.ms;
            const bool first = done == 0;
            done += chunk;
            if (first || std::chrono::duration<double>(te - last_draw).count() > delay) {
                draw({});
                last_draw = clock::now();
            }
        }
        // outputStats (renderer.cpp:276-298): whole seconds since the render
        // began, and the traced rays (camera, extension and shadow rays, as
        // integrator.cpp counts them) per second of tracing, in thousands
        const long long secs = (long long)std::chrono::duration<double>(clock::now() - t0).count();
        const double trace_s = sum.ms * 1e-3;
        const long long krays =
            trace_s > 0.0 ? (long long)((double)(sum.ext_rays + sum.shadow_rays) * 1e-3 / trace_s) : 0;
        draw({{"renderTime", std::to_string(secs / 3600) + " h " + std::to_string((secs / 60) % 60) + " min " +
                                 std::to_string(secs % 60) + " sec."},
              {"RaySpeed", std::to_string(krays) + "K rays/sec."}});
        writer.wait();
        sum.ms = std::chrono::duration<double>(clock::now() - t0).count() * 1e3;
        if (out) *out = sum;
        if (n_draws) *n_draws = draws;
    });
}

int aur_set_sampler(aur_ctx *ctx, int sampler) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        Ctx *c = &ctx->c;
        if (sampler != AUR_SAMPLER_HASH && sampler != AUR_SAMPLER_HALTON) throw std::invalid_argument("unknown sampler");
        c->sampler = sampler;
        if (c->built) {
            set_device(c);
            ensure_halton(c);
        }
    });
}

int aur_framebuffer(aur_ctx *ctx, float *rgba_sum_host) {
    return guarded([&] {
        NEED(ctx);
        NEED(rgba_sum_host);
        Ctx *c = &ctx->c;
        set_device(c);
        if (!c->accum.p || c->fb_w <= 0) throw StateError("no framebuffer yet (render first)");
        HIP_CHECK(hipStreamSynchronize(c->stream));
        HIP_CHECK(hipMemcpy(rgba_sum_host, c->accum.p, sizeof(float) * 4 * (size_t)c->fb_w * c->fb_h, hipMemcpyDeviceToHost));
    });
}

int aur_display(aur_ctx *ctx, uint16_t *rgba_half_host) {
    return guarded([&] {
        NEED(ctx);
        NEED(rgba_half_host);
        Ctx *c = &ctx->c;
        set_device(c);
        if (!c->display.p || c->fb_w <= 0) throw StateError("no framebuffer yet (render first)");
        HIP_CHECK(hipStreamSynchronize(c->stream));
        HIP_CHECK(hipMemcpy(rgba_half_host, c->display.p, sizeof(uint16_t) * 4 * (size_t)c->fb_w * c->fb_h,
                            hipMemcpyDeviceToHost));
    });
}

static std::vector<std::pair<std::string, std::string>> meta_list(const char *const *keys, const char *const *values, int n) {
    if (n < 0) throw std::invalid_argument("negative metadata count");
    if (n > 0 && (!keys || !values)) throw std::invalid_argument("null metadata arrays");
    std::vector<std::pair<std::string, std::string>> m;
    for (int k = 0; k < n; ++k) {
        if (!keys[k] || !values[k] || !*keys[k]) throw std::invalid_argument("empty metadata key or value");
        m.push_back({keys[k], values[k]});
    }
    return m;
}

int aur_write_image(aur_ctx *ctx, const char *path, int n_passes, const char *const *meta_keys,
                    const char *const *meta_values, int n_meta) {
    return guarded([&] {
        NEED(ctx);
        NEED(path);
        if (n_passes < 0) throw std::invalid_argument("n_passes must not be negative");
        Ctx *c = &ctx->c;
        set_device(c);
        if (!c->accum.p || c->fb_w <= 0) throw StateError("no framebuffer yet (render first)");
        auto meta = meta_list(meta_keys, meta_values, n_meta);
        bool has_renderer = false;
        for (auto &kv : meta) has_renderer |= kv.first == "renderer";
        if (!has_renderer) meta.insert(meta.begin(), {"renderer", "Aurora (MI355X wavefront tracer)"});
        const size_t n = (size_t)c->fb_w * c->fb_h;
        std::vector<float> img(4 * n);
        HIP_CHECK(hipStreamSynchronize(c->stream));
        if (n_passes > 0) {
            HIP_CHECK(hipMemcpy(img.data(), c->accum.p, sizeof(float) * 4 * n, hipMemcpyDeviceToHost));
            const float inv = 1.f / (float)n_passes;
            for (float &v : img) v *= inv;
        } else {
            // the FP16 running mean as the reference's display holds it (the
            // half -> float -> half round trip of the writer is exact)
            std::vector<uint16_t> h(4 * n);
            HIP_CHECK(hipMemcpy(h.data(), c->display.p, sizeof(uint16_t) * 4 * n, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < 4 * n; ++i) img[i] = half_to_float(h[i]);
        }
        try {
            exr_write(path, c->fb_w, c->fb_h, img.data(), EXR_ZIP, true, meta);
        } catch (const std::runtime_error &e) {
            throw IoError(e.what());
        }
    });
}

int aur_exr_write(const char *path, int width, int height, const float *rgba, int compression, int half,
                  const char *const *meta_keys, const char *const *meta_values, int n_meta) {
    return guarded([&] {
        NEED(path);
        NEED(rgba);
        if (width <= 0 || height <= 0) throw std::invalid_argument("image size must be positive");
        try {
            exr_write(path, width, height, rgba, compression, half != 0, meta_list(meta_keys, meta_values, n_meta));
        } catch (const std::runtime_error &e) {
            throw IoError(e.what());
        }
    });
}

int aur_exr_read(const char *path, int *width, int *height, float *rgba_out, int64_t capacity_floats) {
    return guarded([&] {
        NEED(path);
        NEED(width);
        NEED(height);
        ExrImage img;
        try {
            img = exr_read(path);
        } catch (const std::runtime_error &e) {
            throw IoError(e.what());
        }
        *width = img.width;
        *height = img.height;
        if (rgba_out) {
            if ((int64_t)img.rgba.size() > capacity_floats) throw CapacityError("rgba_out holds fewer than 4*width*height floats");
            std::memcpy(rgba_out, img.rgba.data(), sizeof(float) * img.rgba.size());
        }
    });
}

int aur_batch_submit(aur_ctx *exec, const aur_job *jobs, int n_jobs, const aur_batch_opts *opts, aur_job_result *results,
                     aur_stats *total) {
    return guarded([&] {
        NEED(exec);
        NEED(jobs);
        NEED(results);
        if (n_jobs <= 0) throw std::invalid_argument("n_jobs must be positive");
        Ctx *c = &exec->c;
        set_device(c);
        std::vector<Ctx *> cs(n_jobs);
        std::vector<int> fp(n_jobs), np(n_jobs);
        for (int k = 0; k < n_jobs; ++k) {
            if (!jobs[k].scene) throw std::invalid_argument("job " + std::to_string(k) + ": null scene");
            cs[k] = &jobs[k].scene->c;
            fp[k] = jobs[k].first_pass;
            np[k] = jobs[k].n_passes;
        }
        aur_batch_opts o;
        o.order = 0;
        o.use_graph = 1;
        o.iters_per_poll = 8;
        if (opts) o = *opts;
        if (o.order != 0 && o.order != 1) throw std::invalid_argument("order must be 0 or 1");
        std::memset(results, 0, sizeof(aur_job_result) * n_jobs);
        aur_stats st;
        std::memset(&st, 0, sizeof st);
        run_jobs(c, cs.data(), fp.data(), np.data(), n_jobs, o.order, o.use_graph, o.iters_per_poll, results, &st);
        if (total) *total = st;
    });
}

int aur_set_launch(aur_ctx *ctx, const aur_launch_opts *opts) {
    return guarded([&] {
        NEED(ctx);
        NEED(opts);
        LaunchCfg &l = ctx->c.launch;
        if (opts->trace_block > 0) {
            if (opts->trace_block != 64 && opts->trace_block != 128 && opts->trace_block != 256)
                throw std::invalid_argument("trace_block must be 64, 128 or 256");
        }
        if (opts->trace_waves_per_cu > 64) throw std::invalid_argument("trace_waves_per_cu must be 1..64");
        if (opts->pool_paths > 0 && opts->pool_paths < 64) throw std::invalid_argument("pool_paths must be >= 64");
        if (opts->pool_paths > 0 && (size_t)opts->pool_paths > kMaxPoolPaths)
            throw std::invalid_argument("pool_paths must be <= " + std::to_string(kMaxPoolPaths));
        if (opts->trace_block > 0) l.trace_block = opts->trace_block;
        if (opts->trace_waves_per_cu > 0) l.trace_waves_per_cu = opts->trace_waves_per_cu;
        if (opts->pool_paths > 0) l.pool_paths = (size_t)opts->pool_paths;
        if (opts->shade_list_tiles > 0) l.shade_list_tiles = opts->shade_list_tiles;
    });
}

int aur_get_launch(aur_ctx *ctx, aur_launch_opts *out) {
    return guarded([&] {
        NEED(ctx);
        NEED(out);
        const LaunchCfg &l = ctx->c.launch;
        out->trace_block = l.trace_block;
        out->trace_waves_per_cu = l.trace_waves_per_cu;
        out->pool_paths = (int)std::min<size_t>(l.pool_paths, 0x7fffffff);
        out->shade_list_tiles = l.shade_list_tiles;
    });
}

int aur_render_paths(aur_ctx *ctx, const int *pixel, const int *sample, int n_paths, const float *slots, float *rgba_out,
                     aur_stats *out) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (!c->built) throw StateError("render before aur_build_accel");
        if (n_paths < 0) throw std::invalid_argument("negative path count");
        if (n_paths == 0) { if (out) std::memset(out, 0, sizeof *out); return; }
        NEED(pixel); NEED(sample); NEED(rgba_out);
        set_device(c);
        const int wh = c->scene.opt.width * c->scene.opt.height;
        for (int i = 0; i < n_paths; ++i)
            if (pixel[i] < 0 || pixel[i] >= wh || sample[i] < 0) throw std::invalid_argument("pixel or sample index out of range");
        const size_t nslots = 2 + AUR_SLOTS_PER_VERTEX * (size_t)c->scene.opt.maxdepth;
        DevBuf dp, ds, dslots, dout;
        dp.reserve(sizeof(int) * n_paths);
        ds.reserve(sizeof(int) * n_paths);
        dout.reserve(sizeof(float4) * n_paths);
        HIP_CHECK(hipMemcpyAsync(dp.p, pixel, sizeof(int) * n_paths, hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(ds.p, sample, sizeof(int) * n_paths, hipMemcpyHostToDevice, c->stream));
        if (slots) {
            dslots.reserve(sizeof(float) * nslots * n_paths);
            HIP_CHECK(hipMemcpyAsync(dslots.p, slots, sizeof(float) * nslots * n_paths, hipMemcpyHostToDevice, c->stream));
        }
        PathSource src;
        src.mode = 1;
        src.first_pass = 0;
        src.pixel = dp.as<int>();
        src.sample = ds.as<int>();
        src.slots = slots ? dslots.as<float>() : nullptr;
        aur_stats st;
        run_paths(c, src, (size_t)n_paths, dout.as<float4>(), &st);
        HIP_CHECK(hipMemcpyAsync(rgba_out, dout.p, sizeof(float4) * n_paths, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        if (out) *out = st;
    });
}

int aur_set_material_params(aur_ctx *ctx, const float *roughness, const float *spec_weight) {
    return guarded([&] {
        NEED(ctx);
        ++ctx->c.version;  // kept path records no longer match
        Ctx *c = &ctx->c;
        HostScene &s = c->scene;
        if (!c->have_scene) throw StateError("no scene loaded");
        for (size_t i = 0; i < s.attr_mats.size(); ++i) {
            aur_material &m = s.attr_mats[i];
            if (roughness && (m.type == AUR_MAT_KELEMEN || m.type == AUR_MAT_METAL)) {
                float r = std::max(roughness[i], 1e-6f);
                m.roughness = r;
                if (m.type == AUR_MAT_METAL) m.roughnessB = r;
            }
            if (spec_weight && m.type == AUR_MAT_KELEMEN) {
                float w = std::max(spec_weight[i], 1e-6f);
                m.spec[0] = m.spec[1] = m.spec[2] = w;
            }
        }
        if (c->built) {
            set_device(c);
            for (size_t i = 0; i < s.attr_mats.size() && i < c->dmats.size(); ++i) {
                const aur_material &m = s.attr_mats[i];
                DMat &d = c->dmats[i];
                d.rough = m.roughness;
                d.roughB = m.roughnessB;
                for (int k = 0; k < 3; ++k) d.spec[k] = m.spec[k];
            }
            HIP_CHECK(hipMemcpyAsync(c->d_mats.p, c->dmats.data(), sizeof(DMat) * c->dmats.size(), hipMemcpyHostToDevice,
                                     c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
        }
    });
}

int aur_keep_path_records(aur_ctx *ctx, int first_pass, int n_passes) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (first_pass < 0 || n_passes < 0) throw std::invalid_argument("negative pass range");
        if (n_passes == 0) {
            c->path_rec.release();
            c->rec_first = c->rec_n = 0;
            c->rec_wh = 0;
            c->rec_version.clear();
            return;
        }
        if (!c->have_scene) throw StateError("no scene loaded");
        set_device(c);
        const size_t wh = (size_t)c->scene.opt.width * c->scene.opt.height;
        if ((long long)wh * n_passes > 0x7fffffffLL) throw CapacityError("too many paths to keep records of");
        c->path_rec.reserve(sizeof(float4) * wh * (size_t)n_passes);
        c->rec_first = first_pass;
        c->rec_n = n_passes;
        c->rec_wh = wh;
        c->rec_version.assign((size_t)n_passes, 0);
    });
}

int aur_adjoint(aur_ctx *ctx, int first_pass, int n_passes, const float *dloss_dimage, aur_param_grads *out,
                aur_stats *stats) {
    return guarded([&] {
        NEED(ctx);
        NEED(dloss_dimage);
        NEED(out);
        Ctx *c = &ctx->c;
        if (!c->built) throw StateError("adjoint before aur_build_accel");
        if (first_pass < 0 || n_passes <= 0) throw std::invalid_argument("bad pass range");
        if (out->albedo_tri && c->tri_albedo.empty())
            throw StateError("albedo gradients need a per-triangle albedo (aur_set_tri_albedo)");
        if (out->albedo_texture && c->ptex.empty())
            throw StateError("albedo texture gradients need an albedo texture (aur_set_albedo_texture)");
        for (const DMat &m : c->dmats) {
            if (m.texmask & ~TEXMASK_PTEX)
                throw StateError("the adjoint does not differentiate shader-driven (textured) materials");
            if (m.type == MAT_GLASS) throw StateError("the adjoint does not support glass (specular) materials");
        }
        if (c->sampler != AUR_SAMPLER_HASH) throw StateError("the adjoint replays hash-sampled paths only");
        if (c->scene.opt.lightsamples != 1) throw StateError("the adjoint replays scenes with lightsamples = 1 only");
        for (const DLight &l : c->dlights)
            if (l.tex >= 0) throw StateError("the adjoint does not support environment maps");
        set_device(c);
        const size_t wh = (size_t)c->scene.opt.width * c->scene.opt.height;
        const size_t na = 3 * (size_t)c->n_tris, nm = (size_t)c->n_mats;
        const size_t nx = c->ptex.size();
        DevBuf dl, ga, gr, gs, gx;
        dl.reserve(sizeof(float4) * wh);
        if (out->albedo_texture) {
            gx.reserve(sizeof(float) * nx);
            fill_zero(c, gx.p, sizeof(float) * nx);
        }
        ga.reserve(sizeof(float) * std::max<size_t>(na, 1));
        gr.reserve(sizeof(float) * std::max<size_t>(nm, 1));
        gs.reserve(sizeof(float) * std::max<size_t>(nm, 1));
        HIP_CHECK(hipMemcpyAsync(dl.p, dloss_dimage, sizeof(float4) * wh, hipMemcpyHostToDevice, c->stream));
        fill_zero(c, ga.p, sizeof(float) * std::max<size_t>(na, 1));
        fill_zero(c, gr.p, sizeof(float) * std::max<size_t>(nm, 1));
        fill_zero(c, gs.p, sizeof(float) * std::max<size_t>(nm, 1));
        adjoint(c, first_pass, n_passes, dl.as<float4>(), out->albedo_tri ? ga.as<float>() : nullptr, gr.as<float>(),
                gs.as<float>(), out->albedo_texture ? gx.as<float>() : nullptr, stats);
        if (out->albedo_texture)
            HIP_CHECK(hipMemcpyAsync(out->albedo_texture, gx.p, sizeof(float) * nx, hipMemcpyDeviceToHost, c->stream));
        if (out->albedo_tri)
            HIP_CHECK(hipMemcpyAsync(out->albedo_tri, ga.p, sizeof(float) * na, hipMemcpyDeviceToHost, c->stream));
        if (out->roughness) HIP_CHECK(hipMemcpyAsync(out->roughness, gr.p, sizeof(float) * nm, hipMemcpyDeviceToHost, c->stream));
        if (out->spec_weight) HIP_CHECK(hipMemcpyAsync(out->spec_weight, gs.p, sizeof(float) * nm, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int aur_set_profiling(aur_ctx *ctx, int enable) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (enable) {
            set_device(c);
            timer_collect(c);
            for (int k = 0; k < KT_COUNT; ++k) { c->timer.total_ms[k] = 0; c->timer.launches[k] = 0; }
        }
        c->timer.on = enable != 0;
    });
}

int aur_kernel_stats(aur_ctx *ctx, aur_kernel_stat *out, int max, int *n_out) {
    return guarded([&] {
        NEED(ctx);
        NEED(n_out);
        Ctx *c = &ctx->c;
        if (c->device >= 0) timer_collect(c);
        int n = 0;
        for (int k = 0; k < KT_COUNT && n < max; ++k, ++n) {
            NEED(out);
            std::memset(&out[n], 0, sizeof out[n]);
            std::strncpy(out[n].name, kKernelNames[k], sizeof out[n].name - 1);
            out[n].launches = c->timer.launches[k];
            out[n].total_ms = c->timer.total_ms[k];
        }
        *n_out = n;
    });
}

int aur_trace(aur_ctx *ctx, const aur_ray *rays, int n, aur_hit *hits, int mode) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (n < 0) throw std::invalid_argument("negative ray count");
        if (mode != AUR_TRACE_CLOSEST && mode != AUR_TRACE_ANY && mode != AUR_TRACE_EXTEND)
            throw std::invalid_argument("unknown trace mode");
        if (mode == AUR_TRACE_EXTEND && (size_t)n > kMaxPoolPaths)
            throw CapacityError("more rays than the extension queue can index");
        if (n == 0) return;
        NEED(rays); NEED(hits);
        if (!c->built) throw StateError("trace before aur_build_accel");
        set_device(c);
        DevBuf dr, dh;
        dr.reserve(sizeof(aur_ray) * n);
        dh.reserve(sizeof(aur_hit) * n);
        HIP_CHECK(hipMemcpyAsync(dr.p, rays, sizeof(aur_ray) * n, hipMemcpyHostToDevice, c->stream));
        trace_rays(c, dr.as<aur_ray>(), n, dh.as<aur_hit>(), mode);
        HIP_CHECK(hipMemcpyAsync(hits, dh.p, sizeof(aur_hit) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int aur_trace_noalloc(aur_ctx *ctx, const aur_ray *d_rays, int n, aur_hit *d_hits, int mode) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (n < 0) throw std::invalid_argument("negative ray count");
        if (n == 0) return;
        NEED(d_rays); NEED(d_hits);
        set_device(c);
        trace_rays_noalloc(c, d_rays, n, d_hits, mode);
    });
}

int aur_light_queries(aur_ctx *ctx, const aur_light_query *q, int n, aur_light_result *r) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (n < 0) throw std::invalid_argument("negative query count");
        if (n == 0) return;
        NEED(q); NEED(r);
        if (!c->built) throw StateError("light queries before aur_build_accel");
        for (int i = 0; i < n; ++i)
            if (q[i].light < 0 || q[i].light >= c->n_lights) throw std::invalid_argument("light index out of range");
        set_device(c);
        DevBuf dq, drs;
        dq.reserve(sizeof(aur_light_query) * n);
        drs.reserve(sizeof(aur_light_result) * n);
        HIP_CHECK(hipMemcpyAsync(dq.p, q, sizeof(aur_light_query) * n, hipMemcpyHostToDevice, c->stream));
        light_queries(c, dq.as<aur_light_query>(), n, drs.as<aur_light_result>());
        HIP_CHECK(hipMemcpyAsync(r, drs.p, sizeof(aur_light_result) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int aur_bsdf(aur_ctx *ctx, const aur_bsdf_query *q, int n, aur_bsdf_result *r) {
    return guarded([&] {
        NEED(ctx);
        Ctx *c = &ctx->c;
        if (n < 0) throw std::invalid_argument("negative query count");
        if (n == 0) return;
        NEED(q); NEED(r);
        for (int i = 0; i < n; ++i)
            if (q[i].kind < BSDF_LAMBERT || q[i].kind > BSDF_TRANSMIT) throw std::invalid_argument("unknown BSDF kind");
        set_device(c);
        DevBuf dq, drs;
        dq.reserve(sizeof(aur_bsdf_query) * n);
        drs.reserve(sizeof(aur_bsdf_result) * n);
        HIP_CHECK(hipMemcpyAsync(dq.p, q, sizeof(aur_bsdf_query) * n, hipMemcpyHostToDevice, c->stream));
        bsdf_queries(c, dq.as<aur_bsdf_query>(), n, drs.as<aur_bsdf_result>());
        HIP_CHECK(hipMemcpyAsync(r, drs.p, sizeof(aur_bsdf_result) * n, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int aur_kelemen_table(aur_ctx *ctx, float reflectance, int distribution, int n_samples, uint32_t seed, float *out) {
    return guarded([&] {
        NEED(ctx);
        NEED(out);
        if (n_samples <= 0) throw std::invalid_argument("n_samples must be positive");
        Ctx *c = &ctx->c;
        set_device(c);
        DevBuf d;
        d.reserve(sizeof(float) * AUR_KEL_COS * AUR_KEL_ROUGH);
        kelemen_table(c, reflectance, distribution, n_samples, seed, d.as<float>());
        HIP_CHECK(hipMemcpyAsync(out, d.p, sizeof(float) * AUR_KEL_COS * AUR_KEL_ROUGH, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    });
}

int aur_loop_subdivide(aur_ctx *ctx, int nv, int nf, const float *P, const int *F, int levels, float amplitude,
                       float frequency, int octaves, uint32_t seed, float *P_out, int *F_out, float *N_out, int nv_cap,
                       int nf_cap, int *nv_out, int *nf_out) {
    bool over = false;
    int rc = guarded([&] {
        NEED(ctx);
        NEED(nv_out);
        NEED(nf_out);
        if (nv < 0 || nf < 0) throw std::invalid_argument("negative mesh size");
        if (nv > 0) NEED(P);
        if (nf > 0) NEED(F);
        Ctx *c = &ctx->c;
        set_device(c);
        std::vector<float> Pin(P, P + 3 * (size_t)nv);
        std::vector<int> Fin(F, F + 3 * (size_t)nf);
        Displace d;
        d.on = amplitude != 0.f;
        d.amplitude = amplitude;
        d.frequency = frequency;
        d.octaves = octaves;
        d.seed = seed;
        LoopMesh m = loop_subdivide(c, Pin, Fin, levels, d);
        int onv = (int)(m.P.size() / 3), onf = (int)(m.F.size() / 3);
        *nv_out = onv;
        *nf_out = onf;
        if (onv > nv_cap || onf > nf_cap) { over = true; return; }
        if (P_out) std::memcpy(P_out, m.P.data(), sizeof(float) * m.P.size());
        if (F_out) std::memcpy(F_out, m.F.data(), sizeof(int) * m.F.size());
        if (N_out) std::memcpy(N_out, m.N.data(), sizeof(float) * m.N.size());
    });
    if (rc == AUR_OK && over) {
        g_last_error = "subdivided mesh exceeds the output capacity";
        return AUR_ECAPACITY;
    }
    return rc;
}

}  // extern "C"
