// shade_common.h -- scene view and material evaluation shared by the
// wavefront kernels and the adjoint kernel.
#pragma once

#include "context.h"
#include "dev_common.h"
#include "halton.h"

namespace aur {

// Division by a divisor fixed for a launch, without the ~25-instruction
// integer divide: Granlund & Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994), unsigned case.  With l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1 (which fits 32 bits for d < 2^31),
// n / d = (mulhi(m, n) + n) >> l for every 0 <= n < 2^31 (mulhi(m, n) <= n,
// so the sum does not wrap).
struct UDiv {
    uint32_t m;
    int l;
};
inline UDiv udiv_make(uint32_t d) {
    int l = 0;
    while ((1ull << l) < d) ++l;
    return {(uint32_t)((((1ull << l) - d) << 32) / d + 1), l};
}
__device__ __host__ inline int udiv(int n, const UDiv &d) {
    const uint32_t t = (uint32_t)(((uint64_t)(uint32_t)n * d.m) >> 32);
    return (int)((t + (uint32_t)n) >> d.l);
}

struct SceneView {
    const void *nodes;
    const float4 *tris;
    const float4 *shade;    // 3 per triangle: (n0, attr), (n1, 0), (n2, 0)
    const DMat *mats;
    const DLight *lights;
    const float *tables;
    const float *albedo;    // optional per-triangle diffuse override
    const float2 *uv;       // 3 per triangle (texture coordinates), or null
    const DAttrX *attrx;    // per attribute, for materials with shaders
    const DShader *shaders;
    const DTex *texs;
    const ushort4 *texels;  // FP16 RGBA
    const HaltonView *qmc;  // Halton sample source (device memory), or null
    const float *ptex;      // FP32 RGB albedo texture (TEXMASK_PTEX materials), or null
    int ptex_w, ptex_h;
    int n_lights;
    int has_env;
    float env[3];
    float fov;
    int width, height, maxdepth, mindepth;
    UDiv div_w, div_8w, div_wh;  // width, 8 * width, width * height
};

SceneView scene_view(Ctx *c);
// Whether a scene needs the shade kernels built with shader and env-map code.
bool shade_needs_tex(const Ctx *c);

// Path order within one pass: bands of 8 rows, each walked in 8-column tiles
// (the last tile of a band may be narrower, the last band shorter), so the 64
// lanes of a wave hold an 8x8 block of neighbouring pixels and their camera
// rays walk the same part of the hierarchy.  A bijection on [0, w*h).
__device__ __host__ inline int tiled_pixel(int local, int w, int h) {
    const int band = local / (8 * w);
    const int y0 = band * 8;
    const int hb = h - y0 < 8 ? h - y0 : 8;
    const int l = local - y0 * w;
    const int cx = l / (hb * 8);
    const int cw = w - cx * 8 < 8 ? w - cx * 8 : 8;
    const int idx = l - cx * hb * 8;
    return (y0 + idx / cw) * w + cx * 8 + idx % cw;
}

// The same order on the device, with the scene's divisors and whole 8x8
// tiles (all but the image's right and bottom edges) decoded by shifts.
__device__ inline int tiled_pixel(const SceneView &sv, int local) {
    const int w = sv.width;
    const int y0 = udiv(local, sv.div_8w) * 8;
    const int hb = sv.height - y0 < 8 ? sv.height - y0 : 8;
    const int l = local - y0 * w;
    if (hb == 8 && l < (w & ~7) * 8)
        return (y0 + ((l >> 3) & 7)) * w + ((l >> 6) << 3) + (l & 7);
    const int cx = l / (hb * 8);
    const int cw = w - cx * 8 < 8 ? w - cx * 8 : 8;
    const int idx = l - cx * hb * 8;
    return (y0 + idx / cw) * w + cx * 8 + idx % cw;
}

// ---------------------------------------------------------------------------
// textures and shaders: texture2D.cpp:48-66 (nearest read), texture2DShader.h,
// triplanarTextureShader.h, material.cpp:19-33 (normal map)
// ---------------------------------------------------------------------------
__device__ inline float half_bits_to_float(unsigned short v) { return (float)__builtin_bit_cast(_Float16, v); }
__device__ inline unsigned short float_to_half_bits(float x) { return __builtin_bit_cast(unsigned short, (_Float16)x); }

// OpenexrDisplay::appendValue (openexrDisplay.cpp:61-72) for the pass that is
// the n-th sample of the pixel: weight = 1/n, pixel = old (1 - weight) +
// value weight in float, colour and alpha alike, stored as FP16 (Imf::Rgba,
// rounded to nearest even).
__device__ inline ushort4 display_update(ushort4 d, float4 v, int n) {
#pragma clang fp contract(off)
    const float w = 1.f / (float)n;
    const float k = 1.f - w;
    const float r = half_bits_to_float(d.x) * k + v.x * w;
    const float g = half_bits_to_float(d.y) * k + v.y * w;
    const float b = half_bits_to_float(d.z) * k + v.z * w;
    const float a = half_bits_to_float(d.w) * k + v.w * w;
    return make_ushort4(float_to_half_bits(r), float_to_half_bits(g), float_to_half_bits(b), float_to_half_bits(a));
}

// Texture2D::read with filterSize 0.  The reference clamps an index past the
// edge to [0, size], which only an argument beyond 1 + 1/size reaches (no
// caller produces one); the clamp here keeps every read inside the texture.
__device__ inline float4 tex_read(const SceneView &sv, int t, float s, float tt) {
    const DTex T = sv.texs[t];
    int x = (int)floorf(s * (float)T.w);
    int y = (int)floorf(tt * (float)T.h);
    if (y == T.h) y -= 1;
    if (x == T.w) x -= 1;
    x = min(max(x, 0), T.w - 1);
    y = min(max(y, 0), T.h - 1);
    const ushort4 q = sv.texels[T.off + (long long)y * T.w + x];
    return make_float4(half_bits_to_float(q.x), half_bits_to_float(q.y), half_bits_to_float(q.z),
                       half_bits_to_float(q.w));
}

// Texel of the FP32 albedo texture a hit at uv (s, t) reads: texture2DShader's
// mapping (u - floor(u), 1 - (v - floor(v))) and Texture2D::read's nearest
// texel with tex_read's clamps.
__device__ inline int ptex_texel(const SceneView &sv, float s, float t) {
    const float a = s - floorf(s);
    const float b = 1.f - (t - floorf(t));
    int x = (int)floorf(a * (float)sv.ptex_w);
    int y = (int)floorf(b * (float)sv.ptex_h);
    if (y == sv.ptex_h) y -= 1;
    if (x == sv.ptex_w) x -= 1;
    x = min(max(x, 0), sv.ptex_w - 1);
    y = min(max(y, 0), sv.ptex_h - 1);
    return y * sv.ptex_w + x;
}

// transform.cpp:206-232 on a row-major 4x4
__device__ inline f3 xform_vec16(const float *m, f3 v) {
    return mk3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
               m[8] * v.x + m[9] * v.y + m[10] * v.z);
}
__device__ inline f3 xform_point16(const float *m, f3 p) {
    f3 r = mk3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
               m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
    const float wp = m[12] * p.x + m[13] * p.y + m[14] * p.z + m[15];
    if (wp == 1.f) return r;
    const float inv = 1.f / wp;
    return mk3(r.x * inv, r.y * inv, r.z * inv);
}
__device__ inline float f3_at(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// What a shader sees of a hit (ShadingGeometry): the hit point, the
// face-forwarded interpolated normal Ng, the interpolated uv and the
// attribute (for its camera-to-object transform).
struct ShadeGeo {
    f3 P, Ng;
    float s, t;
    int attr;
};

// Shader<Color>::evaluate / Shader<float>::evaluate (red channel)
__device__ inline f3 shader_eval(const SceneView &sv, int idx, const ShadeGeo &g) {
    const DShader sh = sv.shaders[idx];
    if (sh.type == SH_CONSTANT) return mk3(sh.value[0], sh.value[1], sh.value[2]);
    if (sh.type == SH_TEXTURE2D) {
        // texture2DShader.h:28-31: read(u - floor(u), 1 - (v - floor(v)))
        const float s = g.s - floorf(g.s);
        const float t = 1.f - (g.t - floorf(g.t));
        const float4 c = tex_read(sv, sh.tex, s, t);
        return mk3(c.x, c.y, c.z);
    }
    // triplanar (color): triplanarTextureShader.h:28-53 (object-space normal,
    // threshold 0.25); float: 76-99 (camera-space Ng, threshold 0.2 in double)
    const float *c2o = sv.attrx[g.attr].c2o;
    f3 wv;
    if (sh.kind == SH_COLOR) {
        const f3 N = norm3(xform_vec16(c2o, g.Ng));
        wv = mk3(fabsf(N.x) - 0.25f, fabsf(N.y) - 0.25f, fabsf(N.z) - 0.25f);
    } else {
        wv = mk3((float)((double)fabsf(g.Ng.x) - 0.2), (float)((double)fabsf(g.Ng.y) - 0.2),
                 (float)((double)fabsf(g.Ng.z) - 0.2));
    }
    // std::max(w, 0.f), then the weights over their float sum
    wv = mk3(wv.x < 0.f ? 0.f : wv.x, wv.y < 0.f ? 0.f : wv.y, wv.z < 0.f ? 0.f : wv.z);
    const float wsum = (0.f + wv.x + wv.y) + wv.z;
    wv = mk3(wv.x / wsum, wv.y / wsum, wv.z / wsum);
    f3 P = xform_point16(c2o, g.P);
    const float inv = 1.f / sh.scale;
    P = mk3(P.x * inv, P.y * inv, P.z * inv);
    f3 r = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float wi = f3_at(wv, i);
        if (wi > 0.f) {
            const float s = fabsf(fmodf(f3_at(P, (i + 1) % 3), 1.f));
            const float t = fabsf(fmodf(f3_at(P, (i + 2) % 3), 1.f));
            const float4 c = tex_read(sv, sh.tex, s, t);
            r = mk3(r.x + c.x * wi, r.y + c.y * wi, r.z + c.z * wi);
        }
    }
    return r;
}

// A material parameter: the shader's value when the slot is shader-driven.
// TEX = false compiles the shader code out (scenes without shaders).
template <bool TEX>
__device__ inline f3 mat_color(const SceneView &sv, const DMat &m, int slot, const float *k, const ShadeGeo &g) {
    if constexpr (TEX) {
        if (m.texmask & (1 << slot)) return shader_eval(sv, sv.attrx[g.attr].slot[slot], g);
    }
    return mk3(k[0], k[1], k[2]);
}
template <bool TEX>
__device__ inline float mat_float(const SceneView &sv, const DMat &m, int slot, float k, const ShadeGeo &g) {
    if constexpr (TEX) {
        if (m.texmask & (1 << slot)) return shader_eval(sv, sv.attrx[g.attr].slot[slot], g).x;
    }
    return k;
}

// Material::runNormalShader (material.cpp:19-33): the map colour, unscaled,
// as a tangent-space direction around Ng.
__device__ inline f3 normal_map(const SceneView &sv, const DMat &m, const ShadeGeo &g) {
    const f3 c = shader_eval(sv, sv.attrx[g.attr].slot[SLOT_NORMAL], g);
    const f3 ns = tangent_to_world(c, g.Ng);
    if (!(ns.x != 0.f || ns.y != 0.f || ns.z != 0.f)) return g.Ng;
    return norm3(ns);
}

// ---------------------------------------------------------------------------
// environment maps: infiniteAreaLight.cpp:35-106
// ---------------------------------------------------------------------------
#define AUR_PI_D 3.14159265358979323846

// eval(): the map at the camera-space direction d, times 2^exposure
__device__ inline f3 env_eval(const SceneView &sv, const DLight &l, f3 d) {
    if (l.tex < 0) return mk3(l.Le[0], l.Le[1], l.Le[2]);
    const f3 w = xform_vec16(l.c2w, d);
    const float v = 0.5f * (1.f - w.y);
    float u = (float)((double)atan2f(w.z, w.x) / (AUR_PI_D * 2.0));
    if (u < 0.f) u += 1.f;
    const float4 c = tex_read(sv, l.tex, u, v);
    return mk3((float)(c.x * l.escale), (float)(c.y * l.escale), (float)(c.z * l.escale));
}

// texture pdf value at (s, t): Texture2D::pdf
__device__ inline float env_tex_pdf(const SceneView &sv, const DLight &l, float s, float t) {
    return tex_read(sv, l.tex, s, t).w / sv.texs[l.tex].pdf_norm;
}

// generateSample / eval / pdf with the map; other lights go to dev_common.h
__device__ inline LightSample light_sample_sv(const SceneView &sv, const DLight &l, f3 orig, f3 nn, float x, float y) {
    if (l.type != LIGHT_ENV || l.tex < 0) return light_sample(l, orig, nn, x, y);
    LightSample s;
    const float vy = 1.f - 2.f * y;
    const float r = sqrtf(fmaxf(0.f, 1.f - vy * vy));
    const float phi = (float)(2.0 * AUR_PI_D * (double)x);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    const f3 dw = xform_vec16(l.w2c, mk3(r * cp, vy, r * sp));
    const float pdf = env_tex_pdf(sv, l, x, y);
    s.dir = norm3(dw);
    s.tmax = 1000000.f;
    s.pdf = (float)((double)pdf / (AUR_PI_D * 4.0));
    s.color = env_eval(sv, l, s.dir);
    return s;
}
__device__ inline float light_pdf_sv(const SceneView &sv, const DLight &l, f3 o, f3 d, float tmin, float *tmax) {
    if (l.type != LIGHT_ENV || l.tex < 0) return light_pdf(l, o, d, tmin, tmax);
    const f3 w = xform_vec16(l.c2w, d);
    const float v = 0.5f * (1.f - w.y);
    float u = (float)((double)atan2f(w.z, w.x) / (AUR_PI_D * 2.0));
    if (u < 0.f) u += 1.f;
    return (float)((double)env_tex_pdf(sv, l, u, v) / (AUR_PI_D * 2.0));  // hemisphere domain
}
__device__ inline f3 light_le_sv(const SceneView &sv, const DLight &l, f3 d) {
    if (l.type != LIGHT_ENV || l.tex < 0) return mk3(l.Le[0], l.Le[1], l.Le[2]);
    return env_eval(sv, l, d);
}
// integrator.cpp:287-297: every env light's eval() in the camera direction
__device__ inline f3 env_miss(const SceneView &sv, f3 d) {
    f3 L = mk3(0.f, 0.f, 0.f);
    for (int i = 0; i < sv.n_lights; ++i) {
        const DLight &l = sv.lights[i];
        if (l.type == LIGHT_ENV) L = add3(L, env_eval(sv, l, d));
    }
    return L;
}

// ---------------------------------------------------------------------------
// Halton QMC source (halton_sampler.h / halton_enum.h of Aurora/external/halton)
// ---------------------------------------------------------------------------
__device__ inline float halton_sample(const HaltonView &h, uint32_t dim, uint32_t index) {
    if (dim == 0) return __uint_as_float(0x3f800000u | (__brev(index) >> 9)) - 1.f;
    const DHaltonDim D = h.dims[dim];
    uint32_t sum = 0, p = D.top, r = index;
    for (uint32_t k = 0; k < D.n_blocks; ++k) {
        const uint32_t q = r / D.block;
        sum += (uint32_t)h.tables[D.off + (r - q * D.block)] * p;
        r = q;
        p /= D.block;
    }
    return (float)sum * D.scale;
}

// Halton_enum::get_index: the sample-th point of pixel (x, y)
__device__ inline uint32_t halton_index(const HaltonView &h, int pixel, uint32_t sample) {
    const uint32_t x = (uint32_t)(pixel % h.width), y = (uint32_t)(pixel / h.width);
    const uint32_t hx = h.p2 ? (__brev(x) >> (32 - h.p2)) : 0u;
    uint32_t hy = 0, r = y;
    for (uint32_t d = 0; d < h.p3; ++d) {
        hy = hy * 3 + r % 3;
        r /= 3;
    }
    const uint32_t off = (uint32_t)(((unsigned long long)hx * h.mx + (unsigned long long)hy * h.my) % h.inc);
    return off + sample * h.inc;
}

// The key of light sample `sub` of a pixel's pass (sub 0: the pass's own key)
__device__ __host__ inline uint32_t sub_path_key(uint32_t key, int sub) {
    return sub == 0 ? key : pcg_hash(key ^ ((uint32_t)sub * 0x68E31DA4u));
}

// The random numbers of one path: injected, Halton, or the counter-based hash.
// A pass with lightsamples > 1 traces several paths through the pixel's one
// camera sample (integrator.cpp:162-173): the camera slots 0 and 1 come from
// the pass's stream, the vertex slots from the path's own.  QMC = false
// compiles the Halton branch out (kernels that never see it).
template <bool QMC>
struct PathRngT {
    const float *slots;
    uint32_t key;             // vertex slots
    uint32_t key_cam;         // camera slots 0 and 1
    int pixel;                // the path's pixel (for the camera direction)
    const HaltonView *qmc;    // Halton source, or null
    uint32_t hidx, hidx_cam;  // Halton indices of the vertex and camera slots
    __device__ inline float u(int slot) const {
        if (slots) return slots[slot];
        if constexpr (QMC) {
            if (qmc && slot < AUR_HALTON_DIMS) {
                const float v = halton_sample(*qmc, (uint32_t)slot, slot < 2 ? hidx_cam : hidx);
                // camera.cpp:39-45: the first two dimensions scaled to the
                // image, here as offsets inside the pixel
                if (slot == 0) return v * qmc->sx - (float)(pixel % qmc->width);
                if (slot == 1) return v * qmc->sy - (float)(pixel / qmc->width);
                return v;
            }
        }
        return slot_uniform(slot < 2 ? key_cam : key, (uint32_t)slot);
    }
};
typedef PathRngT<true> PathRng;

// Light sample `sub` of pass `sample` of a pixel: injected slots when given,
// else the scene's Halton source when set (QMC; the camera from point
// `sample`, the vertices from point sample * ls + sub), else the hash of
// (seed, pixel, sample) and sub.
template <bool QMC = true>
__device__ inline PathRngT<QMC> make_path_rng(const SceneView &sv, const float *slots, uint32_t seed, int pixel,
                                              uint32_t sample, int sub = 0, int ls = 1) {
    PathRngT<QMC> r;
    r.slots = slots;
    r.key_cam = path_key(seed, (uint32_t)pixel, sample);
    r.key = sub_path_key(r.key_cam, sub);
    r.pixel = pixel;
    r.qmc = QMC ? sv.qmc : nullptr;
    r.hidx_cam = (QMC && sv.qmc) ? halton_index(*sv.qmc, pixel, sample) : 0u;
    r.hidx = (QMC && sv.qmc && ls > 1) ? halton_index(*sv.qmc, pixel, sample * (uint32_t)ls + (uint32_t)sub) : r.hidx_cam;
    return r;
}

// Material::getBrdf: matteMaterial.cpp:30-35, metalMaterial.cpp:32-62,
// kelemenMaterial.cpp:225-255.  Reports the Kelemen selection probability and
// its derivative in the (blurred) roughness when asked (adjoint).
struct BrdfChoice {
    Bsdf b;
    bool kelemen_choice;  // the lobe was chosen against the albedo table
    float A, dA;          // selection probability of the spec lobe and dA/droughness
    float drough;         // d(roughness used) / d(material roughness)
};

// The diffuse colour of a Lambert lobe: the albedo texture's texel, the
// per-triangle override, or the material's (shader-driven) colour.  PTEX
// compiles the albedo-texture lookup in.
template <bool TEX, bool PTEX>
__device__ inline f3 diffuse_color(const SceneView &sv, const DMat &m, int prim, const ShadeGeo &g) {
    if constexpr (PTEX) {
        if (m.texmask & TEXMASK_PTEX) {
            const float *p = sv.ptex + 3 * (size_t)ptex_texel(sv, g.s, g.t);
            return mk3(p[0], p[1], p[2]);
        }
    }
    if (sv.albedo) return mk3(sv.albedo[3 * prim], sv.albedo[3 * prim + 1], sv.albedo[3 * prim + 2]);
    return mat_color<TEX>(sv, m, SLOT_DIFFUSE, m.diffuse, g);
}

template <bool TEX, bool PTEX = TEX>
__device__ inline BrdfChoice get_brdf_ex(const SceneView &sv, const DMat &m, int prim, f3 vn, f3 nn, bool matte,
                                         float u_lobe, const ShadeGeo &g) {
    BrdfChoice ch;
    ch.kelemen_choice = false;
    ch.A = 1.f;
    ch.dA = 0.f;
    ch.drough = 1.f;
    Bsdf &b = ch.b;
    if (m.type == MAT_MATTE) {
        const f3 diffuse = diffuse_color<TEX, PTEX>(sv, m, prim, g);
        b.kind = BSDF_LAMBERT; b.dist = 0; b.color = diffuse; b.rough = 0.f;
    } else if (m.type == MAT_METAL) {
        const float mix = mat_float<TEX>(sv, m, SLOT_MIX, m.mix, g);
        const f3 refl = mat_color<TEX>(sv, m, SLOT_SPEC, m.spec, g);
        const f3 gain = mat_color<TEX>(sv, m, SLOT_GAIN, m.gain, g);
        float ct = dot3(vn, nn);
        f3 sc = mk3(schlick(ct, refl.x), schlick(ct, refl.y), schlick(ct, refl.z));
        float rough = (u_lobe > mix) ? mat_float<TEX>(sv, m, SLOT_ROUGH, m.rough, g) : mat_float<TEX>(sv, m, SLOT_ROUGHB, m.roughB, g);
        if (matte) { rough = rough * (1.f - AUR_CAUSTIC_BLUR) + AUR_CAUSTIC_BLUR; ch.drough = 1.f - AUR_CAUSTIC_BLUR; }
        b.kind = BSDF_CT; b.dist = m.dist; b.color = mul3v(sc, gain); b.rough = rough;
    } else if (m.type == MAT_GLASS) {
        // glassMaterial.cpp:33-58: Fresnel of half the cosine picks reflection
        // or transmission; b.rough carries the ior
        b.dist = 0; b.rough = m.transmit[3];
        bool refl = true;
        if (m.refl != 1.f) {
            float ct = dot3(vn, nn) * 0.5f;
            if (ct < 0.f) ct = -ct;
            refl = !(u_lobe > schlick(ct, m.refl));
        }
        if (refl) { b.kind = BSDF_MIRROR; b.color = mk3(m.spec[0], m.spec[1], m.spec[2]); }
        else { b.kind = BSDF_TRANSMIT; b.color = mk3(m.transmit[0], m.transmit[1], m.transmit[2]); }
    } else {
        float rough = mat_float<TEX>(sv, m, SLOT_ROUGH, m.rough, g);
        if (matte) { rough = rough * (1.f - AUR_CAUSTIC_BLUR) + AUR_CAUSTIC_BLUR; ch.drough = 1.f - AUR_CAUSTIC_BLUR; }
        b.kind = BSDF_CT; b.dist = m.dist; b.color = mat_color<TEX>(sv, m, SLOT_SPEC, m.spec, g); b.rough = rough;
        if (m.refl < 1.f) {
            float c = fminf(fmaxf(dot3(nn, vn), 0.f), 1.f);
            const float *tab = sv.tables + (size_t)m.table * (AUR_KEL_COS * AUR_KEL_ROUGH);
            float albedo = kelemen_albedo(tab, c, rough);
            ch.kelemen_choice = true;
            ch.A = albedo;
            ch.dA = kelemen_albedo_drough(tab, c, rough);
            if (u_lobe > albedo) {
                const f3 diffuse = diffuse_color<TEX, PTEX>(sv, m, prim, g);
                b.kind = BSDF_LAMBERT; b.dist = 0; b.color = diffuse; b.rough = 0.f;
            }
        }
    }
    return ch;
}

template <bool TEX>
__device__ inline Bsdf get_brdf(const SceneView &sv, const DMat &m, int prim, f3 vn, f3 nn, bool matte, float u_lobe,
                                const ShadeGeo &g) {
    return get_brdf_ex<TEX>(sv, m, prim, vn, nn, matte, u_lobe, g).b;
}

}  // namespace aur
