// shade_common.h -- scene view and material evaluation shared by the
// wavefront kernels and the adjoint kernel.
#pragma once

#include "context.h"
#include "dev_common.h"

namespace aur {

struct SceneView {
    const void *nodes;
    const float4 *tris;
    const float4 *shade;    // 3 per triangle: (n0, attr), (n1, 0), (n2, 0)
    const DMat *mats;
    const DLight *lights;
    const float *tables;
    const float *albedo;    // optional per-triangle diffuse override
    int n_lights;
    int has_env;
    float env[3];
    float fov;
    int width, height, maxdepth, mindepth;
};

SceneView scene_view(Ctx *c);

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

struct PathRng {
    const float *slots;
    uint32_t key;
    __device__ inline float u(int slot) const { return slots ? slots[slot] : slot_uniform(key, (uint32_t)slot); }
};

// Material::getBrdf: matteMaterial.cpp:30-35, metalMaterial.cpp:32-62,
// kelemenMaterial.cpp:225-255.  Reports the Kelemen selection probability and
// its derivative in the (blurred) roughness when asked (adjoint).
struct BrdfChoice {
    Bsdf b;
    bool kelemen_choice;  // the lobe was chosen against the albedo table
    float A, dA;          // selection probability of the spec lobe and dA/droughness
    float drough;         // d(roughness used) / d(material roughness)
};

__device__ inline BrdfChoice get_brdf_ex(const SceneView &sv, const DMat &m, int prim, f3 vn, f3 nn, bool matte,
                                         float u_lobe) {
    BrdfChoice ch;
    ch.kelemen_choice = false;
    ch.A = 1.f;
    ch.dA = 0.f;
    ch.drough = 1.f;
    Bsdf &b = ch.b;
    f3 diffuse = mk3(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
    if (sv.albedo) diffuse = mk3(sv.albedo[3 * prim], sv.albedo[3 * prim + 1], sv.albedo[3 * prim + 2]);
    if (m.type == MAT_MATTE) {
        b.kind = BSDF_LAMBERT; b.dist = 0; b.color = diffuse; b.rough = 0.f;
    } else if (m.type == MAT_METAL) {
        float ct = dot3(vn, nn);
        f3 sc = mk3(schlick(ct, m.spec[0]), schlick(ct, m.spec[1]), schlick(ct, m.spec[2]));
        float rough = (u_lobe > m.mix) ? m.rough : m.roughB;
        if (matte) { rough = rough * (1.f - AUR_CAUSTIC_BLUR) + AUR_CAUSTIC_BLUR; ch.drough = 1.f - AUR_CAUSTIC_BLUR; }
        b.kind = BSDF_CT; b.dist = m.dist; b.color = mul3v(sc, mk3(m.gain[0], m.gain[1], m.gain[2])); b.rough = rough;
    } else {
        float rough = m.rough;
        if (matte) { rough = rough * (1.f - AUR_CAUSTIC_BLUR) + AUR_CAUSTIC_BLUR; ch.drough = 1.f - AUR_CAUSTIC_BLUR; }
        b.kind = BSDF_CT; b.dist = m.dist; b.color = mk3(m.spec[0], m.spec[1], m.spec[2]); b.rough = rough;
        if (m.refl < 1.f) {
            float c = fminf(fmaxf(dot3(nn, vn), 0.f), 1.f);
            const float *tab = sv.tables + (size_t)m.table * (AUR_KEL_COS * AUR_KEL_ROUGH);
            float albedo = kelemen_albedo(tab, c, rough);
            ch.kelemen_choice = true;
            ch.A = albedo;
            ch.dA = kelemen_albedo_drough(tab, c, rough);
            if (u_lobe > albedo) { b.kind = BSDF_LAMBERT; b.dist = 0; b.color = diffuse; b.rough = 0.f; }
        }
    }
    return ch;
}

__device__ inline Bsdf get_brdf(const SceneView &sv, const DMat &m, int prim, f3 vn, f3 nn, bool matte, float u_lobe) {
    return get_brdf_ex(sv, m, prim, vn, nn, matte, u_lobe).b;
}

}  // namespace aur
