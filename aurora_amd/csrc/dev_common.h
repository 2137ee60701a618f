// dev_common.h -- device-side data layout and per-vertex math of the tracer.
//
// Everything here runs inside the HIP kernels (extend / shade / shadow and
// the parity kernels).  Functions restate the reference's per-vertex routines;
// each carries the reference file:line it follows.  Arithmetic is FP32; where
// the reference promotes to double (e.g. `M_PI * 2 * r2`) the float form is
// used, which differs by at most an ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Every expression in the device code rounds as written, like the reference
// compiled for x86 without FMA: a contracted a*b + c changes the low bits of
// dot products, which the microfacet distribution amplifies near its peak.
// The slab tests of the traversal call fmaf() explicitly where they want it.
#pragma clang fp contract(off)

namespace aur {

// ---------------------------------------------------------------------------
// constants (Aurora/constants.h:13-23)
// ---------------------------------------------------------------------------
#define AUR_CAUSTIC_BLUR 0.2f
#define AUR_FIREFLY 200.f
// The clamp loop of integrator.cpp:277-281 takes at most this many 0.2 steps:
// any finite float drops below 200 within 53 of them, so finite radiance is
// clamped exactly as the unbounded loop would, while an infinite channel (an
// overflowing light or texel) ends the loop instead of spinning forever.
#define AUR_FIREFLY_MAX_STEPS 64
#define AUR_ISBLACK 0.0001f
#define AUR_RAY_BIAS 0.0005f
#define AUR_PI_F 3.14159265358979323846f
#define AUR_SLOTS_PER_VERTEX 9
#define AUR_KEL_COS 64
// Monte-Carlo samples per albedo-table entry (SPECALBEDO_SAMPLES, constants.h:26)
#define AUR_KEL_SAMPLES 250000
#define AUR_KEL_ROUGH 32

enum { MAT_MATTE = 0, MAT_METAL = 1, MAT_KELEMEN = 2, MAT_LIGHT = 3, MAT_GLASS = 4 };
enum { DIST_BECKMANN = 0, DIST_GGX = 1 };
enum { LIGHT_AREA = 0, LIGHT_ENV = 1 };
// BSDF_MIRROR / BSDF_TRANSMIT are the reference's MirrorBrdf type (SpecMirror,
// SpecTransmit): a delta lobe, no direct lighting, pdf 1
enum { BSDF_LAMBERT = 0, BSDF_CT = 1, BSDF_MIRROR = 2, BSDF_TRANSMIT = 3 };

// ---------------------------------------------------------------------------
// device records
// ---------------------------------------------------------------------------
// Material parameters a shader may drive (AUR_SLOT_* of the C ABI).
enum { SLOT_DIFFUSE = 0, SLOT_SPEC = 1, SLOT_GAIN = 2, SLOT_ROUGH = 3, SLOT_ROUGHB = 4, SLOT_MIX = 5, SLOT_NORMAL = 6 };
// DMat::texmask bit of a material whose diffuse colour is the FP32 albedo
// texture (aur_set_albedo_texture) at the hit's uv
enum { TEXMASK_PTEX = 1 << 7 };
enum { SH_TEXTURE2D = 0, SH_TRIPLANAR = 1, SH_CONSTANT = 2 };
enum { SH_COLOR = 0, SH_FLOAT = 1 };

struct alignas(16) DMat {
    int type, dist, table;
    int texmask;          // bit k set: SLOT k is driven by the shader DAttrX::slot[k]; TEXMASK_PTEX

    float diffuse[4];
    float spec[4];
    float gain[4];
    float rough, roughB, mix, refl;
    float emission[4];
    float transmit[4];    // glass: transmitColor, ior in [3]
};

// Per-attribute shading extras, read only when DMat::texmask is non-zero.
struct alignas(16) DAttrX {
    int slot[8];          // shader index per SLOT_*, -1 when constant
    float c2o[16];        // camera-to-object matrix of the attribute's object
};

struct alignas(16) DShader {
    int type, kind, tex, pad;
    float scale, value[3];
};

// A texture in the texel pool: w x h FP16 RGBA texels from `off`.  For env
// maps the alpha channel holds the sampling pdf value (Texture2D::initPdf).
struct alignas(16) DTex {
    long long off;
    int w, h;
    float pdf_norm;       // mean pdf value (m_pdf_normalization)
    int pad[3];
};

struct alignas(16) DLight {
    int type, tex, pad1, pad2;  // tex: env map texture, -1 for a constant env light
    float m[12];          // object-to-camera rows 0..2
    float sx, sy, intensity, pad3;
    float Le[4];          // eval(): area color*2^exp (float), env color*2^exp (double-rounded)
    float N[4];           // camera-space normal
    float pcam[4][4];     // corners p1..p4
    float tri[2][12];     // (p, e1, e2) of the pdf() triangles, padded to 4
    float c2w[12];        // env: camera-to-world rows (InfiniteAreaLight::cameraToWorld)
    float w2c[12];        // env: world-to-camera rows
    double escale;        // env: pow(2, exposure) in double (infiniteAreaLight.cpp:76)
    double pad4;
};

// ---------------------------------------------------------------------------
// small vector type with explicit evaluation order
// ---------------------------------------------------------------------------
struct f3 { float x, y, z; };
__device__ __host__ inline f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ inline f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline f3 mul3(f3 a, float f) { return mk3(a.x * f, a.y * f, a.z * f); }
__device__ inline f3 mul3v(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
// math_utils.h:31-33
__device__ inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// math_utils.h:49-53
__device__ inline f3 cross3(f3 a, f3 b) {
    return mk3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
// vector.h:69-79
__device__ inline float len3(f3 a) {
    float l = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return l == 0.f ? 0.0001f : l;
}
// math_utils.h:39-41 (division multiplies by the reciprocal, vector.h:116-120)
__device__ inline f3 norm3(f3 a) { float inv = 1.f / len3(a); return mul3(a, inv); }

// math_utils.h:114-129
__device__ inline void coordinate_system(f3 n, f3 *t, f3 *b) {
    if (!n.x && n.y && !n.z) *t = cross3(n, mk3(1.f, 0.f, 0.f));
    else *t = cross3(n, mk3(0.f, 1.f, 0.f));
    *b = cross3(n, *t);
}
// math_utils.h:81-90
__device__ inline f3 tangent_to_world(f3 dir, f3 n) {
    f3 t, b;
    coordinate_system(n, &t, &b);
    f3 r = mul3(norm3(t), dir.x);
    r = add3(r, mul3(norm3(b), dir.y));
    r = add3(r, mul3(n, dir.z));
    return r;
}
// math_utils.h:68-76
__device__ inline f3 sample_hemisphere_uniform(float r1, float r2) {
    float z = r1;
    float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    float phi = (2.f * AUR_PI_F) * r2;
    float sphi, cphi;
    sincosf(phi, &sphi, &cphi);
    return mk3(r * cphi, r * sphi, z);
}
// math_utils.h:222-224
__device__ inline float schlick(float vdoth, float refl) {
    float a = 1.f - vdoth;
    float a2 = a * a;
    return refl + (a2 * a2 * a) * (1.f - refl);
}
// color.h lum() (double constants, float result) / isBlack()
__device__ inline float lum3(f3 c) { return (float)(0.2126 * (double)c.x + 0.7152 * (double)c.y + 0.0722 * (double)c.z); }
__device__ inline bool black3(f3 c) { return c.x < AUR_ISBLACK && c.y < AUR_ISBLACK && c.z < AUR_ISBLACK; }
// color.h operator/(float): multiplies by 1/f
__device__ inline f3 div3(f3 a, float f) { float d = 1.f / f; return mul3(a, d); }

// ---------------------------------------------------------------------------
// counter-based random numbers: one uniform per (seed, pixel, sample, slot)
// ---------------------------------------------------------------------------
__device__ __host__ inline uint32_t pcg_hash(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __host__ inline uint32_t path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return pcg_hash(pcg_hash(pcg_hash(seed) ^ pixel) ^ (sample * 0x9E3779B9u));
}
__device__ __host__ inline float slot_uniform(uint32_t key, uint32_t slot) {
    return (float)(pcg_hash(key + slot * 0x85EBCA6Bu) >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// Moller-Trumbore, src/geometry/renderableTriangle.cpp:33-66.  Contraction is
// off so the result is bit-identical to an IEEE evaluation of the same
// expression tree on any other machine.
// ---------------------------------------------------------------------------
__device__ inline bool intersect_tri(f3 o, f3 d, f3 p, f3 e1, f3 e2, float tmin, float tmax,
                                     float *t_out, float *b1_out, float *b2_out) {
#pragma clang fp contract(off)
    f3 s1 = mk3((d.y * e2.z) - (d.z * e2.y), (d.z * e2.x) - (d.x * e2.z), (d.x * e2.y) - (d.y * e2.x));
    float divisor = s1.x * e1.x + s1.y * e1.y + s1.z * e1.z;
    if (divisor == 0.f) return false;
    float inv = 1.f / divisor;
    f3 dv = mk3(o.x - p.x, o.y - p.y, o.z - p.z);
    float b1 = (dv.x * s1.x + dv.y * s1.y + dv.z * s1.z) * inv;
    if (b1 < 0.f || b1 > 1.f) return false;
    f3 s2 = mk3((dv.y * e1.z) - (dv.z * e1.y), (dv.z * e1.x) - (dv.x * e1.z), (dv.x * e1.y) - (dv.y * e1.x));
    float b2 = (d.x * s2.x + d.y * s2.y + d.z * s2.z) * inv;
    if (b2 < 0.f || (b1 + b2) > 1.f) return false;
    float t = (e2.x * s2.x + e2.y * s2.y + e2.z * s2.z) * inv;
    if (t < tmin || t > tmax) return false;
    *t_out = t; *b1_out = b1; *b2_out = b2;
    return true;
}

// ---------------------------------------------------------------------------
// BSDFs: lambert.h:44-80, cookTorrance.cpp:40-144 and the GGX analogue
// ---------------------------------------------------------------------------
struct Bsdf { int kind; int dist; f3 color; float rough; };

// tan^2(acos(c)) without the two transcendentals: (1 - c)(1 + c) / c^2.
// 1 - c is exact for c in [0.5, 1] (Sterbenz), so the peak of the
// distribution (c -> 1) keeps full relative precision, and the result stays
// closer to the host's tanf(acosf(c)) than the device's own tanf(acosf(c))
// does (largest relative error of D 9.5e-6 against 4.3e-5 on 20,000 queries).
// c > 1 (a rounded dot product) gives a negative value, which the callers
// treat as the NaN that acos(c > 1) produces in the reference.
__device__ inline float tan2_of_cos(float c) { return ((1.f - c) * (1.f + c)) / (c * c); }

// cookTorrance.cpp:40-45: Beckmann D(h) * cos(theta_h); GGX analogue.
// tan2 < 0 stands for the reference's NaN (acos of a cosine above one).
__device__ inline float ct_d(int dist, float rough, float cosa, float tan2) {
    const float r2 = fmaxf(0.001f, rough * rough);
    const float c3 = cosa * cosa * cosa;
    if (tan2 < 0.f) return __int_as_float(0x7fc00000);
    if (dist == DIST_GGX) {
        float t = r2 + tan2;
        return r2 / (AUR_PI_F * c3 * t * t);
    }
    const float e = expf(-tan2 / r2);
    return e / (r2 * AUR_PI_F * c3);
}
// cookTorrance.cpp:136 (no floor)
__device__ inline float ct_d_eval(int dist, float rough, float ndoth) {
    const float r2 = rough * rough;
    const float tan2 = tan2_of_cos(ndoth);
    if (tan2 < 0.f) return __int_as_float(0x7fc00000);
    const float c3 = ndoth * ndoth * ndoth;
    if (dist == DIST_GGX) {
        float t = r2 + tan2;
        return (1.f / (AUR_PI_F * c3 * t * t)) * r2;
    }
    return (1.f / (r2 * AUR_PI_F * c3)) * expf(-tan2 / r2);
}

struct BsdfSample { f3 dir; float pdf; f3 color; float tmax; };

__device__ inline BsdfSample bsdf_sample(const Bsdf &b, f3 vn, f3 nn, float r1, float r2) {
    BsdfSample s;
    if (b.kind == BSDF_LAMBERT) {
        // lambert.h:44-59
        f3 dir = sample_hemisphere_uniform(r1, r2);
        s.dir = norm3(tangent_to_world(dir, nn));
        s.pdf = 0.15915494309189535f;          // 1/pi * 0.5
        s.color = div3(b.color, AUR_PI_F);     // albedo * 1.f / M_PI
        s.tmax = 1000000.f;
        return s;
    }
    // cookTorrance.cpp:47-95.  z keeps the reference's cos(atan(x)) rounding:
    // r = sqrt(1 - z^2) below turns one ulp of z near 1 into ~1e-5 of direction.
    float ta;
    if (b.dist == DIST_GGX) ta = b.rough * sqrtf(r1 / (1.f - r1));
    else ta = b.rough * sqrtf(-logf(1.f - r1));
    float z = cosf(atanf(ta));
    float r = sqrtf(fmaxf(0.f, 1.f - z * z));
    float phi = (2.f * AUR_PI_F) * r2;
    float sphi, cphi;
    sincosf(phi, &sphi, &cphi);
    f3 dir = norm3(mk3(r * cphi, r * sphi, z));
    f3 hn = norm3(tangent_to_world(dir, nn));
    float vdoth = dot3(vn, hn);
    f3 ln = norm3(sub3(mul3(mul3(hn, 2.f), vdoth), vn));
    float ndoth = dot3(nn, hn);
    f3 res = mk3(0.f, 0.f, 0.f);
    float pdf = 1.f;
    if (vdoth > 0.0001f) {
        const float tan2 = (ndoth > 0.999f) ? 0.f : tan2_of_cos(ndoth);
        float D = ct_d(b.dist, b.rough, ndoth, tan2);
        D = fminf(D, 1000.f);
        pdf = D / (4.f * vdoth);
        res = div3(mul3(b.color, D), 4.f * vdoth * vdoth);
        if (pdf > 100000.f) { pdf *= 0.01f; res = mul3(res, 0.01f); }
    }
    s.dir = ln; s.pdf = pdf; s.color = res; s.tmax = 10000000.f;
    return s;
}

// math_utils.h:91-95: the tangent frame of coordinateSystem, not normalised
__device__ inline f3 world_to_tangent(f3 dir, f3 n) {
    f3 t, b;
    coordinate_system(n, &t, &b);
    return mk3(dot3(dir, t), dot3(dir, b), dot3(dir, n));
}

// specMirror.cpp:30-35 and specTransmit.cpp:30-58 (b.rough holds the ior).
// The reflected / refracted tangent-space vector goes back through
// tangentToWorld, whose frame is normalised while worldToTangent's is not.
__device__ inline BsdfSample delta_sample(const Bsdf &b, f3 vn, f3 nn) {
    BsdfSample s;
    s.color = b.color;
    s.pdf = 1.f;
    const f3 vt = world_to_tangent(vn, nn);
    if (b.kind == BSDF_MIRROR) {
        s.dir = norm3(tangent_to_world(mk3(-vt.x, -vt.y, vt.z), nn));
        s.tmax = 100000000.f;
        return s;
    }
    const float costheta = dot3(nn, vn);
    float ei = 1.f, et = b.rough;
    if (costheta < 0.f) { ei = b.rough; et = 1.f; }
    const float sini2 = fmaxf(0.f, 1.f - costheta * costheta);
    const float eta = ei / et;
    const float sint2 = eta * eta * sini2;
    if (sint2 >= 1.f) {
        s.pdf = 0.f;
        s.color = mk3(0.f, 0.f, 0.f);
        s.dir = mk3(0.f, 0.f, 0.f);
        s.tmax = 0.f;
        return s;
    }
    float cost = sqrtf(fmaxf(0.f, 1.f - sint2));
    if (costheta > 0.f) cost = -cost;
    s.dir = norm3(tangent_to_world(mk3(eta * -vt.x, eta * -vt.y, cost), nn));
    s.tmax = 10000000.f;
    return s;
}

// lambert.h:76-79; cookTorrance.cpp:97-117
__device__ inline float bsdf_pdf(const Bsdf &b, f3 ln, f3 vn, f3 nn) {
    if (b.kind == BSDF_LAMBERT) return 0.15915494309189535f;
    f3 hn = norm3(add3(ln, vn));
    float vdoth = dot3(vn, hn);
    float ndoth = dot3(nn, hn);
    float D = ct_d(b.dist, b.rough, ndoth, tan2_of_cos(ndoth));
    if (D != D) D = 0.f;
    return D / (4.f * vdoth);
}

// lambert.h:65-74; cookTorrance.cpp:125-144
__device__ inline f3 bsdf_eval(const Bsdf &b, f3 ln, f3 vn, f3 nn) {
    if (b.kind == BSDF_LAMBERT) {
        if (dot3(ln, nn) > 0.f) return div3(b.color, AUR_PI_F);
        return mk3(0.f, 0.f, 0.f);
    }
    f3 hn = norm3(add3(ln, vn));
    const float vdoth = dot3(vn, hn);
    const float vdotn = dot3(vn, nn);
    const float ndoth = dot3(hn, nn);
    const float ldotn = dot3(ln, nn);
    const float D = ct_d_eval(b.dist, b.rough, ndoth);
    if (vdoth <= 0.001f || ldotn <= 0.001f || vdotn <= 0.001f || D != D) return mk3(0.f, 0.f, 0.f);
    return div3(mul3(b.color, D), 4.f * vdoth * vdoth);
}

// ---------------------------------------------------------------------------
// Kelemen albedo lookup, kelemenMaterial.cpp:180-223 (rB = max(rA+1, 31)
// always selects the last row; indices are clamped where the reference would
// read past the table).
// ---------------------------------------------------------------------------
__device__ inline int ref_floor(float x) { return (int)x - (x < 0.f && (float)(int)x != x); }

__device__ inline float kelemen_albedo(const float *tab, float costheta, float rough) {
    int i = ref_floor(costheta * (AUR_KEL_COS - 1));
    int ra = ref_floor(rough * (AUR_KEL_ROUGH - 1));
    i = min(max(i, 0), AUR_KEL_COS - 1);
    ra = min(max(ra, 0), AUR_KEL_ROUGH - 1);
    int rb = AUR_KEL_ROUGH - 1;
    int i1 = min(i + 1, AUR_KEL_COS - 1);
    float va = tab[i * AUR_KEL_ROUGH + ra];
    float vb = tab[i1 * AUR_KEL_ROUGH + ra];
    float mix = costheta * (AUR_KEL_COS - 1) - i;
    float res_a = (1.f - mix) * va + mix * vb;
    va = tab[i * AUR_KEL_ROUGH + rb];
    vb = tab[i1 * AUR_KEL_ROUGH + rb];
    float res_b = (1.f - mix) * va + mix * vb;
    float rmix = rough * (AUR_KEL_ROUGH - 1) - ra;
    return (1.f - rmix) * res_a + rmix * res_b;
}

// d(kelemen_albedo)/d(roughness): the lookup is linear in roughness inside a
// table cell, so the derivative is (ROUGH-1) * (row rB - row rA) at this cosine.
__device__ inline float kelemen_albedo_drough(const float *tab, float costheta, float rough) {
    int i = ref_floor(costheta * (AUR_KEL_COS - 1));
    int ra = ref_floor(rough * (AUR_KEL_ROUGH - 1));
    i = min(max(i, 0), AUR_KEL_COS - 1);
    ra = min(max(ra, 0), AUR_KEL_ROUGH - 1);
    int rb = AUR_KEL_ROUGH - 1;
    int i1 = min(i + 1, AUR_KEL_COS - 1);
    float mix = costheta * (AUR_KEL_COS - 1) - i;
    float res_a = (1.f - mix) * tab[i * AUR_KEL_ROUGH + ra] + mix * tab[i1 * AUR_KEL_ROUGH + ra];
    float res_b = (1.f - mix) * tab[i * AUR_KEL_ROUGH + rb] + mix * tab[i1 * AUR_KEL_ROUGH + rb];
    return (AUR_KEL_ROUGH - 1) * (res_b - res_a);
}

// d log D / d roughness of the distribution at cos(theta_h) = ndoth (the
// evalSampleWorld form, cookTorrance.cpp:136, without floors):
//   Beckmann  D ~ exp(-tan^2/m^2) / (m^2 cos^3):  2 tan^2 / m^3 - 2 / m
//   GGX       D ~ a^2 / (cos^3 (a^2 + tan^2)^2):  2 / a - 4 a / (a^2 + tan^2)
__device__ inline float ct_dlogd_drough(int dist, float rough, float ndoth) {
    float c2 = fmaxf(ndoth * ndoth, 1e-12f);
    float t2 = fmaxf(1.f - c2, 0.f) / c2;
    if (dist == DIST_GGX) return 2.f / rough - 4.f * rough / (rough * rough + t2);
    return 2.f * t2 / (rough * rough * rough) - 2.f / rough;
}

// ---------------------------------------------------------------------------
// lights: squareLight.h:57-122, infiniteAreaLight.cpp:35-106
// ---------------------------------------------------------------------------
__device__ inline f3 xform_point(const float *m, f3 p) {
    // transform.cpp:206-214 (affine: wp == 1)
    return mk3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3],
               m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
               m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}

struct LightSample { f3 dir; float tmax; float pdf; f3 color; };

__device__ inline LightSample light_sample(const DLight &l, f3 orig, f3 nn, float x, float y) {
    LightSample s;
    if (l.type == LIGHT_AREA) {
        f3 lp = mk3((x - 0.5f) * l.sx * 2.f, (y - 0.5f) * l.sy * 2.f, 0.f);
        lp = xform_point(l.m, lp);
        f3 dir = sub3(lp, orig);
        float len = len3(dir);
        s.dir = norm3(dir);
        s.tmax = len - AUR_RAY_BIAS * 10.f;
        float d = l.N[0] * (-s.dir.x) + l.N[1] * (-s.dir.y) + l.N[2] * (-s.dir.z);
        if (signbit(d)) d = 0.f;
        s.color = mk3(l.Le[0], l.Le[1], l.Le[2]);
        s.pdf = (d == 0.f) ? 0.f : (len * len) / (d * l.sx * l.sy * 4.f);
    } else {
        f3 dt = sample_hemisphere_uniform(x, y);
        s.dir = norm3(tangent_to_world(dt, nn));
        s.tmax = 1000000.f;
        s.pdf = 0.15915494309189535f;  // 2 / (4 pi)
        s.color = mk3(l.Le[0], l.Le[1], l.Le[2]);
    }
    return s;
}

// squareLight.h:88-97: a hit on the quad shortens tmax; returns the pdf
__device__ inline float light_pdf(const DLight &l, f3 o, f3 d, float tmin, float *tmax) {
    if (l.type != LIGHT_AREA) return 0.15915494309189535f;  // 1 / (2 pi), hemisphere domain
    float t, b1, b2;
    bool hit = intersect_tri(o, d, mk3(l.tri[0][0], l.tri[0][1], l.tri[0][2]), mk3(l.tri[0][4], l.tri[0][5], l.tri[0][6]),
                             mk3(l.tri[0][8], l.tri[0][9], l.tri[0][10]), tmin, *tmax, &t, &b1, &b2);
    if (!hit)
        hit = intersect_tri(o, d, mk3(l.tri[1][0], l.tri[1][1], l.tri[1][2]), mk3(l.tri[1][4], l.tri[1][5], l.tri[1][6]),
                            mk3(l.tri[1][8], l.tri[1][9], l.tri[1][10]), tmin, *tmax, &t, &b1, &b2);
    if (!hit) return 0.f;
    f3 dn = norm3(d);
    float c = l.N[0] * dn.x + l.N[1] * dn.y + l.N[2] * dn.z;
    c = c * (float)((0.f < c) - (c < 0.f));
    float pdf = (t * t) / (c * l.sx * l.sy * 4.f);
    *tmax = t - AUR_RAY_BIAS * 10.f;
    return pdf;
}

// squareLight.h:99-110
__device__ inline bool light_visible(const DLight &l, f3 orig, f3 nn) {
    if (l.type != LIGHT_AREA) return true;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        f3 v = sub3(mk3(l.pcam[i][0], l.pcam[i][1], l.pcam[i][2]), orig);
        if (dot3(v, nn) > 0.f) return true;
    }
    return false;
}

// integrateDirect's light choice (integrator.cpp:100-130): the lights whose
// visible() test passes from (orig, nn) are counted and the
// min(floor(u * n), n - 1)-th of them, in light order, is returned (-1 when
// none is visible).  One pass over the lights keeps a bit mask of the first 64.
__device__ inline int pick_visible_light(const DLight *lights, int n_lights, f3 orig, f3 nn, float u, int *nvis_out) {
    unsigned long long mask = 0ull;
    int nvis = 0;
    const int nm = n_lights < 64 ? n_lights : 64;
    for (int li = 0; li < nm; ++li)
        if (light_visible(lights[li], orig, nn)) { mask |= 1ull << li; ++nvis; }
    for (int li = 64; li < n_lights; ++li) nvis += light_visible(lights[li], orig, nn) ? 1 : 0;
    *nvis_out = nvis;
    if (!nvis) return -1;
    int pick = min((int)(u * nvis), nvis - 1);
    const int in_mask = __popcll(mask);
    if (pick < in_mask) {
        for (int k = 0; k < pick; ++k) mask &= mask - 1ull;
        return __ffsll((long long)mask) - 1;
    }
    for (int li = 64, k = in_mask; li < n_lights; ++li)
        if (light_visible(lights[li], orig, nn)) {
            if (k == pick) return li;
            ++k;
        }
    return -1;
}

// integrator.cpp:28-31
__device__ inline float power_heuristic(float fpdf, float gpdf) {
    float f = fpdf, g = gpdf;
    return (f * f) / (f * f + g * g);
}

}  // namespace aur
