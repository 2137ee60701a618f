/*
 * aurora_hip.h -- C ABI of the MI355X path tracer (libaurora_hip.so).
 *
 * The library replaces Aurora's CPU render loop: the TBB pixel loop
 * (`Renderer::renderImage` / `IntegrateParallel`, src/frontend/renderer.cpp:147-271)
 * and everything it calls per pixel -- `Integrator::integrateCameraSample`
 * (src/frontend/integrator.cpp:152-300), the BRDFs, the lights, and Embree's
 * `intersect` / `intersectBinary` (src/geometry/embreeAccelerator.cpp:41-100).
 * It keeps the reference's scene input (the PyAurora JSON scene plus OBJ meshes
 * and config.json).
 *
 * Conventions:
 *   - every call returns 0 on success or a negative AUR_E* code; the message of
 *     the last failure on the calling thread is returned by aur_last_error();
 *     no C++ exception or abort crosses this boundary;
 *   - host buffers belong to the caller; device memory belongs to the library;
 *   - a context is used by one host thread at a time (it is not locked).
 */
#ifndef AURORA_HIP_H
#define AURORA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AUR_OK 0
#define AUR_EINVAL (-1)   /* bad argument */
#define AUR_EIO (-2)      /* file missing or unreadable */
#define AUR_EPARSE (-3)   /* malformed scene / OBJ / config */
#define AUR_EHIP (-4)     /* HIP runtime failure */
#define AUR_ESTATE (-5)   /* call out of order (e.g. render before build) */
#define AUR_ECAPACITY (-6) /* a size limit of the library was exceeded */

typedef struct aur_ctx aur_ctx;

/* A ray: origin, [tmin, tmax] interval, direction (32 bytes).  The fields of
 * `Ray` (src/core/ray.h:17-34). */
typedef struct {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} aur_ray;

/* A hit: distance, barycentrics (u weights vertex 1, v weights vertex 2, as
 * embreeAccelerator.cpp:113-129), global triangle id or -1 on a miss. */
typedef struct {
    float t, u, v;
    int prim;
} aur_hit;

/* Material kinds (the classes of src/shading/materials/). */
#define AUR_MAT_MATTE 0
#define AUR_MAT_METAL 1
#define AUR_MAT_KELEMEN 2
#define AUR_MAT_LIGHT 3 /* ConstantMaterial on light triangles (renderer.cpp:131-133) */
#define AUR_MAT_GLASS 4 /* GlassMaterial: SpecMirror / SpecTransmit (glassMaterial.cpp) */

#define AUR_DIST_BECKMANN 0
#define AUR_DIST_GGX 1

#define AUR_LIGHT_AREA 0 /* SquareLight (src/lights/squareLight.*) */
#define AUR_LIGHT_ENV 1  /* InfiniteAreaLight (src/geometry/infiniteAreaLight.*) */

/* Material parameters that a shader may drive (ShadingEngine indices held by
 * the materials, jsonParser.cpp:312-365); AUR_SLOT_NORMAL is the normal map
 * (Material::runNormalShader, material.cpp:19-33). */
#define AUR_SLOT_DIFFUSE 0 /* matte color / kelemen diffusecolor */
#define AUR_SLOT_SPEC 1    /* kelemen speccolor / metal reflectance */
#define AUR_SLOT_GAIN 2    /* metal gain */
#define AUR_SLOT_ROUGH 3   /* kelemen roughness / metal roughnessA */
#define AUR_SLOT_ROUGHB 4  /* metal roughnessB */
#define AUR_SLOT_MIX 5     /* metal mix */
#define AUR_SLOT_NORMAL 6  /* tangent-space normal map */
#define AUR_N_SLOTS 7

typedef struct {
    int type;          /* AUR_MAT_* */
    int distribution;  /* AUR_DIST_* (Cook-Torrance lobes) */
    float diffuse[3];  /* matte color / kelemen diffusecolor */
    float spec[3];     /* kelemen speccolor / metal reflectance */
    float gain[3];     /* metal gain */
    float roughness;   /* kelemen roughness / metal roughnessA */
    float roughnessB;  /* metal roughnessB */
    float mix;         /* metal mix */
    float reflectance; /* kelemen reflectance (selects the albedo table) */
    float emission[3]; /* light triangles: color * 2^exposure */
    int shader[AUR_N_SLOTS]; /* aur_shader index driving each AUR_SLOT_*, or -1 for the constant above */
    float c2o[16];     /* camera-to-object matrix of the attribute's object, row major (triplanar shaders) */
    float transmit[3]; /* glass transmitColor (specColor is spec, reflectance is reflectance) */
    float ior;         /* glass index of refraction */
} aur_material;

typedef struct {
    int type;        /* AUR_LIGHT_* */
    float m[16];     /* area light: object-to-camera matrix, row major */
    float sx, sy;    /* area light: scaleX, scaleY */
    float color[3];
    float exposure;
    int envmap;      /* env light: aur_texture index of its map, or -1 for the constant color */
} aur_light;

/* Shaders (src/shading/shaders/): a texture read by uv (Texture2DShader), a
 * texture projected along the three object axes (TriplanarTextureShader), or
 * a constant (ConstantShader, only needed for a constant normal map).  kind
 * selects the colour or the float (red channel) variant. */
#define AUR_SHADER_TEXTURE2D 0
#define AUR_SHADER_TRIPLANAR 1
#define AUR_SHADER_CONSTANT 2
#define AUR_SHADER_COLOR 0
#define AUR_SHADER_FLOAT 1

typedef struct {
    int type;       /* AUR_SHADER_* */
    int kind;       /* AUR_SHADER_COLOR or AUR_SHADER_FLOAT */
    int texture;    /* aur_texture index (texture shaders) */
    float scale;    /* triplanar: object-space units per texture repeat */
    float value[3]; /* constant shader */
} aur_shader;

/* A texture (Texture2D, texture2D.cpp:23-46): width x height RGBA texels,
 * row 0 first, at `offset` texels into the scene's texel pool.  Values are
 * the FP16 values the reference's Imf::Rgba buffer holds. */
typedef struct {
    int width, height;
    int64_t offset;
} aur_texture;

/* Flattened, camera-space scene: what `EmbreeMesh::preRender`
 * (src/geometry/embreeMesh.h:30-58) hands to Embree plus the per-attribute
 * materials, the lights and the render options. */
typedef struct {
    int n_verts, n_tris;
    const float *verts;   /* 3 * n_verts */
    const int *tris;      /* 3 * n_tris */
    const float *normals; /* 9 * n_tris: per-corner normals */
    const int *attr;      /* n_tris: attribute index (objects, then lights) */
    int n_attrs;
    const aur_material *mats; /* n_attrs */
    int n_lights;
    const aur_light *lights;
    float fov;
    int width, height;
    int pixelsamples, lightsamples, mindepth, maxdepth;
    /* shading inputs (all may be empty / NULL) */
    const float *uvs;     /* 6 * n_tris per-corner texture coordinates, or NULL for all zero */
    int n_shaders;
    const aur_shader *shaders;
    int n_textures;
    const aur_texture *textures;
    int64_t n_texels;
    const float *texels;  /* 4 * n_texels */
    float cam_to_world[16], world_to_cam[16]; /* camera transform and its inverse (env maps) */
} aur_scene_desc;

typedef struct {
    int n_verts, n_tris, n_attrs, n_lights, n_objects;
    int width, height, pixelsamples, lightsamples, mindepth, maxdepth;
    float fov;
    int bvh_width, bvh_nodes, bvh_depth, bvh_max_leaf;
    int n_shaders, n_textures;
    int64_t n_texels;
} aur_scene_info;

typedef struct {
    int width;      /* BVH branching factor: 4 or 8 (0 = default 4) */
    int max_leaf;   /* triangles per leaf, 1..16 (0 = default 2) */
} aur_accel_opts;

/* Counters, counted as the reference counts rays (integrator.cpp:44,71,88,162,260). */
typedef struct {
    int64_t paths;       /* camera paths started */
    int64_t ext_rays;    /* closest-hit rays traced (camera + extension) */
    int64_t hit_rays;    /* closest-hit rays that hit */
    int64_t shadow_rays; /* any-hit rays traced */
    double ms;           /* wall time of the call on the device stream */
} aur_stats;

/* One BSDF query: kind (0 Lambert, 1 Cook-Torrance, 2 SpecMirror,
 * 3 SpecTransmit with roughness = ior), params, the frame and two uniforms for
 * getSample. */
typedef struct {
    int kind;
    int distribution;
    float color[3];
    float roughness;
    float V[3], N[3], L[3];
    float u1, u2;
} aur_bsdf_query;

/* One light query against light `light` of the built scene (the lights in
 * scene order, as aur_scene_export returns them): the calls the integrator
 * makes on a Light (light.h:23-81; squareLight.h:57-122,
 * infiniteAreaLight.cpp:35-106). */
typedef struct {
    int light;
    float P[3], N[3];  /* shading point and normal: generateSample, visible */
    float u1, u2;      /* generateSample's two uniforms */
    float D[3];        /* a unit direction from P: eval and pdf */
    float tmax;        /* pdf: the ray's maxt before the call (a quad hit shortens it) */
} aur_light_query;

typedef struct {
    float sample_dir[3];   /* generateSample: unit direction to the sampled point */
    float sample_tmax;     /* and the shadow ray's maxt */
    float sample_pdf;      /* solid-angle pdf (0 where a quad faces away) */
    float sample_color[3]; /* radiance of the sample */
    float Le[3];           /* eval(D) */
    float pdf;             /* pdf of the ray (P, D, [RAY_BIAS, tmax]) */
    float pdf_tmax;        /* the ray's maxt after pdf() */
    int visible;           /* visible(P, N) */
} aur_light_result;

typedef struct {
    float f[3];           /* evalSampleWorld(L, V, N) */
    float pdf;            /* pdf(L, V, N) */
    float sample_dir[3];  /* getSample(V, N) direction */
    float sample_pdf;
    float sample_color[3];
} aur_bsdf_result;

/* ---- lifetime / errors ---- */
int aur_create(aur_ctx **out, int hip_device);
void aur_destroy(aur_ctx *ctx);
const char *aur_last_error(void);
/* hip_stream: a hipStream_t, or NULL for the library's own stream.  Renders
 * of at least 2 M paths also run on three streams the context creates for
 * itself (a second path pipeline and the side streams that trace shadow rays
 * beside extension rays); they start after work already on the context's
 * stream and the context's stream waits for them before the call returns. */
int aur_set_stream(aur_ctx *ctx, void *hip_stream);
int aur_set_seed(aur_ctx *ctx, uint32_t seed);
/* Sample source of the paths' random numbers.  AUR_SAMPLER_HASH (default):
 * a counter-based hash of (seed, pixel, pass, slot).  AUR_SAMPLER_HALTON: the
 * reference's optional QMC source (Aurora/external/halton, enabled there by
 * USE_HALTON, constants.h:31-41): Faure-permuted Halton points enumerated per
 * pixel (Halton_enum); the pass-th point of a pixel gives the camera position
 * (camera.cpp:39-45) and, dimension by dimension, the path's other slots
 * (2..255; later slots use the hash).  The seed does not affect it. */
#define AUR_SAMPLER_HASH 0
#define AUR_SAMPLER_HALTON 1
int aur_set_sampler(aur_ctx *ctx, int sampler);

/* ---- scene ---- */
/* Replaces JsonParser::parseScene (jsonParser.cpp:30-59) + ObjTriangleMesh
 * (objTriangleMesh.cpp:26-92).  config_json_path may be NULL. */
int aur_scene_load(aur_ctx *ctx, const char *scene_path, const char *config_json_path);
/* Replaces the flattened scene wholesale (already camera space). */
int aur_scene_upload(aur_ctx *ctx, const aur_scene_desc *desc);
/* Per-triangle diffuse albedo override (3 * n_tris floats) or NULL to clear. */
int aur_set_tri_albedo(aur_ctx *ctx, const float *albedo_rgb);
/* Loop subdivision, flattening and the wide-BVH build, all on the GPU.
 * Replaces EmbreeAccelerator's constructor (embreeAccelerator.cpp:19-35). */
int aur_build_accel(aur_ctx *ctx, const aur_accel_opts *opts);
int aur_scene_info_get(aur_ctx *ctx, aur_scene_info *out);
/* The built acceleration structure, for inspection (Embree keeps its
 * bvh4.spatialsplit internal, embreeAccelerator.cpp:19-35; no counterpart).
 * nodes (bvh_nodes * 32 * bvh_width bytes, may be NULL): each node is six
 * rows of bvh_width floats (lo x, hi x, lo y, hi y, lo z, hi z of every child
 * box), bvh_width int child codes (>= 0 inner node index, 0x7fffffff empty,
 * < 0 leaf: ~code = first << 4 | (count - 1), a range of triangle records)
 * and bvh_width ints of which the first is the index of the node's first
 * inner child (inner children take the first slots and consecutive indices)
 * and the rest are 0; node 0 is the root.  tri_records (12 * n_tris
 * floats, may be NULL): per triangle in BVH order v0 (3 floats) and the
 * global triangle id (int bits), e1 and 0, e2 and 0. */
int aur_bvh_export(aur_ctx *ctx, void *nodes, float *tri_records);
/* Copies the flattened scene to host buffers (any pointer may be NULL). */
int aur_scene_export(aur_ctx *ctx, float *verts, int *tris, float *normals, int *attr,
                     aur_material *mats, aur_light *lights);
/* The shading inputs of the flattened scene (any pointer may be NULL):
 * 6 * n_tris uvs, the shaders, the textures, 4 * n_texels texel values and
 * the camera transform with its inverse (32 floats). */
int aur_scene_export_shading(aur_ctx *ctx, float *uvs, aur_shader *shaders, aur_texture *textures,
                             float *texels, float *cam_to_world_and_inverse);

/* ---- rendering ---- */
/* Renders passes [first_pass, first_pass + n_passes) into the device
 * accumulation buffer (sum of per-pass RGBA).  rgba_sum_host (4*W*H floats)
 * receives a copy of the sum when non-NULL.  first_pass == 0 clears it. */
int aur_render(aur_ctx *ctx, int first_pass, int n_passes, float *rgba_sum_host, aur_stats *out);
/* Paths given by (pixel index, sample index); rgba_out receives 4 floats per
 * path.  slots (n_paths * (2 + 9*maxdepth) floats) injects every random number
 * of every path; NULL uses the counter-based hash of (seed, pixel, sample). */
int aur_render_paths(aur_ctx *ctx, const int *pixel, const int *sample, int n_paths,
                     const float *slots, float *rgba_out, aur_stats *out);

/* Copies the accumulation buffer (sum of per-pass RGBA, 4*W*H floats) of the
 * last aur_render or aur_batch_submit on this context.  The reference's
 * framebuffer is OpenexrDisplay's running mean (openexrDisplay.cpp:61-72);
 * divide by the pass count for it. */
int aur_framebuffer(aur_ctx *ctx, float *rgba_sum_host);
/* Copies that running mean as the reference keeps it (OpenexrDisplay::
 * appendValue, openexrDisplay.cpp:61-72): updated with every resolved pass in
 * pass order, n-th pass weight 1/n, pixel = old (1 - 1/n) + value / n in
 * float, stored as FP16 RGBA (4*W*H halves, rounded to nearest even).  Passes
 * are counted from pass 0. */
int aur_display(aur_ctx *ctx, uint16_t *rgba_half_host);

/* ---- image files (OpenEXR, scanline, single part) ---- */
/* Replaces OpenexrDisplay::draw (openexrDisplay.cpp:132-169): writes the mean
 * image (framebuffer sum / n_passes; with n_passes = 0 the FP16 running mean
 * of aur_display, the reference's own pixel values) as FP16 RGBA, ZIP
 * compressed, with a
 * "renderer" string attribute plus n_meta key/value strings (e.g. the
 * reference's renderTime and RaySpeed, renderer.cpp:292-297), to
 * <path>_tmp and renames it over path. */
int aur_write_image(aur_ctx *ctx, const char *path, int n_passes, const char *const *meta_keys,
                    const char *const *meta_values, int n_meta);
/* Replaces Renderer::renderImage, outputStats and postRenderCleanup
 * (renderer.cpp:200-304): renders the scene's pixelsamples passes from pass 0
 * and writes the display (aur_display's FP16 running mean, as
 * aur_write_image with n_passes = 0) to path -- NULL: the scene's options
 * "filename" -- after pass 0, again whenever more than draw_delay_s seconds
 * (<= 0: the reference's DRAW_DELAY, 1 s) have passed since the last write,
 * and a final time with the "renderTime" ("H h M min S sec.") and "RaySpeed"
 * ("NK rays/sec.") attributes.  Passes are traced in chunks of about a quarter
 * of the delay, so a timed write lands at the end of the first chunk that ends
 * past the delay; a file is encoded on a host thread while the next chunk
 * traces.  RaySpeed counts camera, extension and shadow rays per second of
 * device time (the reference averages per-pass rates over whole seconds, at
 * least one).  out receives the summed counters with ms = the call's wall
 * time; n_draws (may be NULL) the number of files written. */
int aur_render_image(aur_ctx *ctx, const char *path, double draw_delay_s, aur_stats *out, int *n_draws);
/* Generic writer: rgba is width*height*4 floats, row 0 first; compression
 * 0 (NONE), 2 (ZIPS) or 3 (ZIP); half != 0 stores FP16 channels. */
int aur_exr_write(const char *path, int width, int height, const float *rgba, int compression, int half,
                  const char *const *meta_keys, const char *const *meta_values, int n_meta);
/* Reader (the texture loader's Imf::RgbaInputFile, texture2D.cpp:24-46):
 * channels R, G, B, A as HALF or FLOAT, NONE / ZIPS / ZIP.  Writes width and
 * height; when rgba_out is non-NULL also width*height*4 floats (missing
 * channels read 0, alpha 1), failing with AUR_ECAPACITY if capacity_floats is
 * too small. */
int aur_exr_read(const char *path, int *width, int *height, float *rgba_out, int64_t capacity_floats);

/* ---- batched persistent launch (no counterpart in the reference, which
 * renders one scene per process: Renderer::renderImage, renderer.cpp:201-271) */
typedef struct {
    aur_ctx *scene;   /* a context with a built scene, on exec's device */
    int first_pass;   /* passes [first_pass, first_pass + n_passes) are added to */
    int n_passes;     /* the job's framebuffer; first_pass == 0 clears it first */
} aur_job;
typedef struct {
    int64_t paths, ext_rays, hit_rays, shadow_rays;
    double latency_ms; /* from the start of the launch to the job's last finished path */
    int order;         /* the job's position in the launch order */
} aur_job_result;
typedef struct {
    int order;          /* 0: shortest estimated job first, 1: as given */
    int use_graph;      /* nonzero: replay a captured hipGraph of two bounces */
    int iters_per_poll; /* bounces between completion checks by the host (0: 8) */
} aur_batch_opts;
/* Renders every job through one pool of path slots on exec's stream (exec
 * may be one of the jobs' contexts); launches of at least 2 M paths split the
 * jobs over two such pools, the second on one of exec's own streams.  All jobs need the same BVH width.
 * opts may be NULL (shortest first, graph replay).  results has n_jobs
 * entries, in the order of jobs; total (may be NULL) gets the sums and the
 * device time of the whole launch. */
int aur_batch_submit(aur_ctx *exec, const aur_job *jobs, int n_jobs, const aur_batch_opts *opts,
                     aur_job_result *results, aur_stats *total);

/* Launch shape of the tracer kernels on this context (the autotuner's knobs). */
typedef struct {
    int trace_block;        /* traversal workgroup size: 64, 128 or 256 */
    int trace_waves_per_cu; /* persistent traversal grid, in waves per CU (1..64) */
    int pool_paths;         /* path slots in flight (64 .. 357913941): the aur_batch_submit pool, and
                               the largest chunk of whole passes aur_render traces at once */
    int shade_list_tiles;   /* aur_render's shading: a bounce after the camera bounce whose queue
                               holds at least this many 1024-entry tiles first finishes its misses
                               and shades full waves of the hits; shorter queues, and the camera
                               bounce, are shaded in place (default: four tiles per shade
                               workgroup, 16 per CU; the image is the same either way) */
} aur_launch_opts;
/* Fields <= 0 keep their current value. */
int aur_set_launch(aur_ctx *ctx, const aur_launch_opts *opts);
int aur_get_launch(aur_ctx *ctx, aur_launch_opts *out);

/* ---- inverse rendering (no counterpart in the reference, which is not
 * differentiable) ---- */
/* Per-material parameters (n_attrs floats each; NULL leaves a field as is):
 * roughness sets Kelemen roughness and metal roughnessA/B; spec_weight sets a
 * Kelemen material's speccolor to (s, s, s).  Values are clamped to >= 1e-6. */
int aur_set_material_params(aur_ctx *ctx, const float *roughness, const float *spec_weight);
/* An FP32 RGB albedo texture (width * height texels, row 0 at the top, 3
 * floats each) that replaces the diffuse colour, and any diffuse shader, of
 * the materials listed in attrs (matte or Kelemen).  A hit reads the texel
 * texture2DShader.h:28-31 would read at its interpolated uv (nearest texel,
 * Texture2D::read's clamps), without the FP16 rounding of Imf::Rgba, so the
 * texels can be optimised.  Needs a built scene with per-corner uvs; rgb =
 * NULL removes it.  Calls that keep the size and the materials only upload
 * the new texels. */
int aur_set_albedo_texture(aur_ctx *ctx, int width, int height, const float *rgb, const int *attrs, int n_attrs);
typedef struct {
    float *albedo_tri;   /* 3 * n_tris, or NULL (needs aur_set_tri_albedo) */
    float *roughness;    /* n_attrs, or NULL */
    float *spec_weight;  /* n_attrs, or NULL */
    float *albedo_texture; /* 3 * width * height of the albedo texture, or NULL (needs aur_set_albedo_texture) */
} aur_param_grads;
/* Path-replay adjoint of passes [first_pass, first_pass + n_passes) rendered
 * with the current seed: given dL/dI for the mean image I (4 * W * H floats,
 * host, alpha ignored) writes dL/dtheta into the non-NULL fields of out.
 * Roughness gradients are produced for Kelemen materials and for metal
 * materials (whose roughnessA and roughnessB aur_set_material_params sets to
 * the one value), spec-weight gradients for Kelemen materials.  The adjoint
 * replays every path twice (its radiance, then its gradient); when
 * aur_render kept the records of all these passes with the current inputs
 * (aur_keep_path_records), the first replay is skipped.  stats counts the
 * rays the call traced. */
int aur_adjoint(aur_ctx *ctx, int first_pass, int n_passes, const float *dloss_dimage,
                aur_param_grads *out, aur_stats *stats);
/* No reference counterpart (the reference is not differentiable).  From now
 * on, aur_render keeps the per-path record the adjoint's first replay would
 * compute -- the path's radiance before the firefly clamp, its clamp steps
 * and ray count, 16 bytes per path in device memory -- for every path of
 * passes [first_pass, first_pass + n_passes) it traces (scenes with
 * lightsamples = 1).  A record stays valid until the seed, sampler, scene,
 * acceleration structure, material parameters or albedo overrides change;
 * aur_adjoint uses records only when every pass of its range has a valid
 * one.  n_passes = 0 frees them.  Calling it again re-arms (and invalidates)
 * the records. */
int aur_keep_path_records(aur_ctx *ctx, int first_pass, int n_passes);

/* ---- instrumentation ---- */
/* Kernel timing, measured with HIP events recorded on the library stream
 * around each launch.  The reference only logs a running "K rays/sec"
 * (renderer.cpp:265-268); this is its device-side counterpart. */
typedef struct {
    char name[32];
    int64_t launches;
    double total_ms;
} aur_kernel_stat;
/* enable != 0 turns timing on and resets the totals.  While timing is on,
 * the tracer's kernels run one after another on the context's stream
 * (normally a bounce's shadow and finish kernels overlap the next bounce's
 * extension kernel on a second stream, and large renders run two path
 * pipelines at once), so each kernel's time is its own. */
int aur_set_profiling(aur_ctx *ctx, int enable);
int aur_kernel_stats(aur_ctx *ctx, aur_kernel_stat *out, int max, int *n_out);

/* ---- parity queries ---- */
/* Batch ray queries against the built BVH: AccelerationStructure::intersect
 * (closest hit) and intersectBinary (any hit), accelerationStructure.h:17-27,
 * embreeAccelerator.cpp:41-100.  AUR_TRACE_CLOSEST and AUR_TRACE_ANY run one
 * ray per lane to completion; AUR_TRACE_EXTEND sends the rays through the
 * renderer's own closest-hit kernel (the extension queue, the persistent
 * launch shape and the per-lane ray refill of a render), so its hits are
 * those a render would see. */
#define AUR_TRACE_CLOSEST 0
#define AUR_TRACE_ANY 1
#define AUR_TRACE_EXTEND 2
int aur_trace(aur_ctx *ctx, const aur_ray *rays, int n, aur_hit *hits, int mode);
/* No counterpart in the reference.  aur_trace (AUR_TRACE_CLOSEST or
 * AUR_TRACE_ANY) for callers that cannot allocate device memory -- code under
 * stream capture, tight serving loops: d_rays and d_hits are device buffers of
 * n entries, the call only enqueues work on the context's stream (results are
 * ready when that stream has reached it) and allocates nothing.  The traversal
 * stack's overflow lives in a fixed 16 MiB per GPU shared by every call on
 * that GPU, so calls that overlap on one GPU race; a BVH that needs more than
 * 128 overflow entries per ray fails with AUR_ECAPACITY. */
int aur_trace_noalloc(aur_ctx *ctx, const aur_ray *d_rays, int n, aur_hit *d_hits, int mode);
int aur_bsdf(aur_ctx *ctx, const aur_bsdf_query *q, int n, aur_bsdf_result *r);
/* Light::generateSample / eval / pdf / visible per query (light.h:23-81). */
int aur_light_queries(aur_ctx *ctx, const aur_light_query *q, int n, aur_light_result *r);
/* Kelemen albedo table (64 cos x 32 roughness), kelemenMaterial.cpp:113-178. */
int aur_kelemen_table(aur_ctx *ctx, float reflectance, int distribution, int n_samples,
                      uint32_t seed, float *out);
/* Loop subdivision of an indexed object-space mesh on the GPU
 * (LoopSubdivMesh::LoopSubdivMesh, loopSubdivMesh.cpp:167-562).  Corners are
 * welded by exact position.  amplitude > 0 adds the radial value-noise
 * displacement with area-weighted normals.  Outputs: limit positions
 * (3 * nv_out), faces (3 * nf_out), vertex normals (3 * nv_out); returns
 * AUR_ECAPACITY if the result exceeds nv_cap / nf_cap (the counts are still
 * written). */
int aur_loop_subdivide(aur_ctx *ctx, int nv, int nf, const float *P, const int *F, int levels,
                       float amplitude, float frequency, int octaves, uint32_t seed,
                       float *P_out, int *F_out, float *N_out, int nv_cap, int nf_cap,
                       int *nv_out, int *nf_out);

#ifdef __cplusplus
}
#endif
#endif
