/*
 * aurora_oracle.h -- CPU restatement of Aurora's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the pytest suite,
 * the smoke check and the CPU-baseline leg of bench.py call it to produce
 * reference values.  The product (aurora_amd/libaurora_hip.so) never links,
 * loads or calls it.
 *
 * Every function restates a reference routine; the citation (file:line, paths
 * relative to the reference repository root, `src/` = `Aurora/src/`) is on the
 * definition in aurora_oracle.c.  Arithmetic follows the reference's own
 * evaluation order, including the places where it silently promotes to double
 * (e.g. `1. - u - v`, `M_PI * 2 * r2`, `pdf / (d * sx * sy * 4.)`).
 *
 * Random numbers: the reference draws `rand()/RAND_MAX` at fixed call sites.
 * Here every draw is a numbered "slot" of a path: 2 camera slots, then 9 per
 * vertex (lobe, light pick, light sample x2, MIS BSDF sample x2, continuation
 * x2, roulette).  A slot value comes either from an injected table or from a
 * counter-based hash of (seed, pixel, sample, slot) -- the same hash the GPU
 * renderer uses, so both renderers consume identical sample streams.
 */
#ifndef AURORA_ORACLE_H
#define AURORA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MAT_MATTE = 0, ORC_MAT_METAL = 1, ORC_MAT_KELEMEN = 2, ORC_MAT_LIGHT = 3, ORC_MAT_GLASS = 4 };
enum { ORC_DIST_BECKMANN = 0, ORC_DIST_GGX = 1 };
enum { ORC_LIGHT_AREA = 0, ORC_LIGHT_ENV = 1 };
enum { ORC_BSDF_LAMBERT = 0, ORC_BSDF_CT = 1, ORC_BSDF_MIRROR = 2, ORC_BSDF_TRANSMIT = 3 };

enum { ORC_SLOT_DIFFUSE = 0, ORC_SLOT_SPEC = 1, ORC_SLOT_GAIN = 2, ORC_SLOT_ROUGH = 3, ORC_SLOT_ROUGHB = 4,
       ORC_SLOT_MIX = 5, ORC_SLOT_NORMAL = 6, ORC_N_SLOTS = 7 };
enum { ORC_SHADER_TEXTURE2D = 0, ORC_SHADER_TRIPLANAR = 1, ORC_SHADER_CONSTANT = 2 };
enum { ORC_SHADER_COLOR = 0, ORC_SHADER_FLOAT = 1 };

#define ORC_SLOTS_PER_VERTEX 9
#define ORC_KELEMEN_COS 64
#define ORC_KELEMEN_ROUGH 32

typedef struct {
    int type;            /* ORC_MAT_* */
    int distribution;    /* ORC_DIST_* for Cook-Torrance lobes */
    float diffuse[3];    /* matte color / kelemen diffusecolor */
    float spec[3];       /* kelemen speccolor / metal reflectance (rgb F0) */
    float gain[3];       /* metal gain */
    float roughness;     /* kelemen roughness / metal roughnessA */
    float roughnessB;    /* metal roughnessB */
    float mix;           /* metal mix */
    float reflectance;   /* kelemen scalar reflectance */
    float emission[3];   /* light triangles: color * 2^exposure */
    const float *table;  /* kelemen albedo table [64*32], index cos*32 + rough */
    int shader[ORC_N_SLOTS]; /* shader driving each ORC_SLOT_*, -1 = the constant above */
    float c2o[16];       /* camera-to-object matrix of the attribute's object (triplanar) */
    float transmit[3];   /* glass transmitColor (spec = specColor) */
    float ior;           /* glass */
} orc_material;

typedef struct {
    int type;            /* ORC_LIGHT_* */
    float m[16];         /* area: object-to-camera matrix, row major */
    float sx, sy;        /* area: half extents (scaleX, scaleY) */
    float color[3];
    float exposure;
    /* derived by orc_light_init */
    float intensity;     /* powf(2, exposure) */
    float N[3];          /* camera-space normal */
    float pcam[4][3];    /* camera-space corners p1..p4 */
    float tri[2][9];     /* (p, e1, e2) of the two pdf() triangles */
    int envmap;          /* env: texture index, -1 for the constant colour */
    float c2w[16], w2c[16]; /* env: camera-to-world and world-to-camera */
} orc_light;

typedef struct {
    int type;            /* ORC_SHADER_* */
    int kind;            /* ORC_SHADER_COLOR / ORC_SHADER_FLOAT */
    int texture;
    float scale;         /* triplanar */
    float value[3];      /* constant */
} orc_shader;

/* Texture2D: rgba is width*height*4 FP16-exact values (Imf::Rgba).  For an
 * env map orc_texture_init_pdf rewrites alpha as initPdf does and sets
 * pdf_norm. */
typedef struct {
    int width, height;
    float *rgba;
    float pdf_norm;
} orc_texture;

typedef struct {
    int n_verts, n_tris;
    const float *verts;    /* 3*n_verts, camera space */
    const int *tris;       /* 3*n_tris vertex indices */
    const float *normals;  /* 9*n_tris per-corner normals, camera space */
    const int *attr;       /* n_tris attribute (material) index */
    const float *tri_albedo; /* optional 3*n_tris diffuse albedo override (inverse-rendering scenes) */
    int n_attrs;
    const orc_material *mats;
    int n_lights;
    const orc_light *lights;
    float fov;
    int width, height;
    int maxdepth, mindepth;
    const float *uvs;      /* 6*n_tris per-corner uvs, or NULL (all zero) */
    int n_shaders;
    const orc_shader *shaders;
    int n_textures;
    const orc_texture *textures;
    int sampler;           /* 0: hash of (seed, pixel, sample, slot); 1: Halton (oracle_halton.c) */
    int lightsamples;      /* paths per camera sample after pass 0 */
    /* optional FP32 RGB albedo texture (row 0 at the top, 3 floats a texel)
     * replacing the diffuse colour of the attributes flagged in ptex_mat */
    const float *ptex;
    int ptex_w, ptex_h;
    const int *ptex_mat;   /* n_attrs flags, or NULL */
} orc_scene;

typedef struct {
    float o[3], d[3];
    float tmin, tmax;
} orc_ray;

typedef struct {
    float t, u, v;
    int prim;             /* -1 on miss */
} orc_hit;

typedef struct {
    int64_t paths, ext, hit, shadow;
} orc_counters;

typedef struct orc_bvh orc_bvh;

/* ---- geometry ---- */
int orc_intersect_tri(const float *p1, const float *e1, const float *e2,
                      const orc_ray *r, float *t, float *b1, float *b2);
void orc_trace_brute(const orc_scene *s, const orc_ray *rays, int n, int any_hit, orc_hit *out);
orc_bvh *orc_bvh_build(const orc_scene *s, int leaf_size);
void orc_bvh_free(orc_bvh *b);
void orc_trace_bvh(const orc_scene *s, const orc_bvh *b, const orc_ray *rays, int n, int any_hit, orc_hit *out);

/* ---- lights ---- */
void orc_light_init(orc_light *l);
void orc_texture_init_pdf(orc_texture *t);

/* ---- textures and shaders (single-routine checks) ---- */
float orc_half_round(float x);
void orc_texture_read(const orc_texture *t, float s, float tt, float *rgb_out);
void orc_shader_eval(const orc_scene *sc, int shader, const float *P, const float *Ng, float s, float t, int attr,
                     float *out);
void orc_env_eval(const orc_scene *sc, const orc_light *l, const float *dir, float *rgb_out);

/* ---- BSDF (query interface used by the parity tests) ---- */
/* params: [0..2] color (albedo or spec), [3] roughness, [4] distribution */
void orc_bsdf_eval(int kind, const float *params, const float *V, const float *N,
                   const float *L, float *f_out, float *pdf_out);
void orc_bsdf_sample(int kind, const float *params, const float *V, const float *N,
                     float u1, float u2, float *dir_out, float *pdf_out, float *color_out);

/* ---- Kelemen albedo table ---- */
void orc_kelemen_table(float reflectance, int distribution, int n_samples, uint32_t seed, float *out);
float orc_kelemen_albedo(const float *table, float costheta, float roughness);

/* ---- random numbers ---- */
uint32_t orc_path_key(uint32_t seed, uint32_t pixel, uint32_t sample);
float orc_slot_uniform(uint32_t key, uint32_t slot);

/* ---- integrator ---- */
/* One path for pixel (x, y).  slots==NULL -> hash(seed, pixel, sample). */
void orc_path(const orc_scene *s, const orc_bvh *b, int x, int y, uint32_t seed, uint32_t sample,
              const float *slots, float *rgba_out, orc_counters *cnt);
/* Passes [first_pass, first_pass+n_passes): accum[4*W*H] += per-pass rgba. */
void orc_render(const orc_scene *s, const orc_bvh *b, uint32_t seed, int first_pass, int n_passes,
                float *accum, orc_counters *cnt, int n_threads);
/* Paths listed by (pixel, sample); out[4*n]. */
void orc_render_paths(const orc_scene *s, const orc_bvh *b, uint32_t seed, const int *pixel,
                      const int *sample, int n, const float *slots, float *out, orc_counters *cnt,
                      int n_threads);

/* ---- Loop subdivision (indexed restatement) ---- */
/* UV (optional): 6 per input face, the uvs of its corners; UV_out then
 * receives 6 per output face (not pushed to the limit, loopSubdivMesh.cpp:477-482).
 * push_limit = 0 (a check of the checker) returns the last level's positions
 * instead of the limit surface's. */
int orc_loop_subdivide(int nv, int nf, const float *P, const int *F, int levels,
                       float *P_out, int *F_out, float *N_out, int nv_cap, int nf_cap,
                       int *nv_out, int *nf_out, const float *UV, float *UV_out, int push_limit);

/* ---- value-noise displacement + area-weighted normals ---- */
void orc_displace_radial(int nv, float *P, float amplitude, float frequency, int octaves, uint32_t seed);
void orc_vertex_normals(int nv, int nf, const float *P, const int *F, float *N_out);

/* ---- single routines for closed-form tests ---- */
void orc_camera_ray(const orc_scene *s, int x, int y, float j0, float j1, orc_ray *r);
void orc_light_sample(const orc_light *l, const float *P, const float *N, float x, float y, orc_ray *ray,
                      float *pdf, float *color);
float orc_light_pdf(const orc_light *l, orc_ray *ray);
/* the same for lights that read the scene's textures (env maps) */
void orc_light_sample_in(const orc_scene *sc, const orc_light *l, const float *P, const float *N, float x, float y,
                         orc_ray *ray, float *pdf, float *color);
float orc_light_pdf_in(const orc_scene *sc, const orc_light *l, orc_ray *ray);
int orc_light_visible(const orc_light *l, const float *P, const float *N);

/* ---- Halton QMC (oracle_halton.c) ---- */
float orc_halton_sample(unsigned dim, unsigned index);
void orc_halton_enum(unsigned w, unsigned h, unsigned *p2, unsigned *p3, unsigned *mx, unsigned *my, unsigned *inc);
unsigned orc_halton_index(unsigned w, unsigned h, unsigned i, unsigned x, unsigned y);
float orc_halton_slot(unsigned w, unsigned h, unsigned i, unsigned x, unsigned y, unsigned slot);

/* ---- adjoint (path-replay) ---- */
void orc_adjoint(const orc_scene *s, const orc_bvh *b, uint32_t seed, int first_pass, int n_passes,
                 const float *dl_dimage, const float *albedo_tri, float *g_albedo_tri,
                 float *g_rough_mat, float *g_spec_mat, float *g_ptex, int n_threads);

#ifdef __cplusplus
}
#endif
#endif
