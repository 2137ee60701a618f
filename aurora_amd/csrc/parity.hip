// parity.hip -- batched BSDF and light queries and the Kelemen albedo-table kernel.
#include "shade_common.h"

namespace aur {

// One query per lane: evalSampleWorld + pdf for (L, V, N), and getSample for
// (V, N, u1, u2).  lambert.h:44-80, cookTorrance.cpp:47-144.
__global__ void k_bsdf(const aur_bsdf_query *__restrict__ q, int n, aur_bsdf_result *__restrict__ r) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    aur_bsdf_query a = q[i];
    Bsdf b;
    b.kind = a.kind;
    b.dist = a.distribution;
    b.color = mk3(a.color[0], a.color[1], a.color[2]);
    b.rough = a.roughness;
    f3 V = mk3(a.V[0], a.V[1], a.V[2]), N = mk3(a.N[0], a.N[1], a.N[2]), L = mk3(a.L[0], a.L[1], a.L[2]);
    // SpecMirror / SpecTransmit: eval and pdf are zero, the sample is the delta direction
    const bool delta = b.kind >= BSDF_MIRROR;
    f3 f = delta ? mk3(0.f, 0.f, 0.f) : bsdf_eval(b, L, V, N);
    float pdf = delta ? 0.f : bsdf_pdf(b, L, V, N);
    BsdfSample s = delta ? delta_sample(b, V, N) : bsdf_sample(b, V, N, a.u1, a.u2);
    aur_bsdf_result o;
    o.f[0] = f.x; o.f[1] = f.y; o.f[2] = f.z;
    o.pdf = pdf;
    o.sample_dir[0] = s.dir.x; o.sample_dir[1] = s.dir.y; o.sample_dir[2] = s.dir.z;
    o.sample_pdf = s.pdf;
    o.sample_color[0] = s.color.x; o.sample_color[1] = s.color.y; o.sample_color[2] = s.color.z;
    r[i] = o;
}

// One query per lane on the scene's light q.light: generateSample at (P, N,
// u1, u2), eval and pdf along D (squareLight.h:57-122 and the environment
// light of infiniteAreaLight.cpp:35-106, maps included), visible(P, N).
__global__ void k_light(SceneView sv, const aur_light_query *__restrict__ q, int n, aur_light_result *__restrict__ r) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const aur_light_query a = q[i];
    const DLight &l = sv.lights[a.light];
    const f3 P = mk3(a.P[0], a.P[1], a.P[2]), N = mk3(a.N[0], a.N[1], a.N[2]), D = mk3(a.D[0], a.D[1], a.D[2]);
    const LightSample s = light_sample_sv(sv, l, P, N, a.u1, a.u2);
    const f3 le = light_le_sv(sv, l, D);
    float tmax = a.tmax;
    const float pdf = light_pdf_sv(sv, l, P, D, AUR_RAY_BIAS, &tmax);
    aur_light_result o;
    o.sample_dir[0] = s.dir.x; o.sample_dir[1] = s.dir.y; o.sample_dir[2] = s.dir.z;
    o.sample_tmax = s.tmax;
    o.sample_pdf = s.pdf;
    o.sample_color[0] = s.color.x; o.sample_color[1] = s.color.y; o.sample_color[2] = s.color.z;
    o.Le[0] = le.x; o.Le[1] = le.y; o.Le[2] = le.z;
    o.pdf = pdf;
    o.pdf_tmax = tmax;
    o.visible = light_visible(l, P, N) ? 1 : 0;
    r[i] = o;
}

void light_queries(Ctx *c, const aur_light_query *d_q, int n, aur_light_result *d_r) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_light, dim3((n + 255) / 256), dim3(256), 0, c->stream, scene_view(c), d_q, n, d_r);
    HIP_CHECK(hipGetLastError());
}

void bsdf_queries(Ctx *c, const aur_bsdf_query *d_q, int n, aur_bsdf_result *d_r) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_bsdf, dim3((n + 255) / 256), dim3(256), 0, c->stream, d_q, n, d_r);
    HIP_CHECK(hipGetLastError());
}

// Kelemen albedo table, one workgroup per (cos theta, roughness) entry:
// the estimator of kelemenMaterial.cpp:113-178 with a counter-based sample
// stream; partial sums in double, reduced in LDS.
// half_store: the entry is stored as FP16, as the reference writes its table
// through the FP16 display and reads it back as Imf::HALF
// (kelemenMaterial.cpp:53-76, 170-177)
__global__ void __launch_bounds__(256) k_kelemen(float reflectance, int dist, int n_samples, uint32_t seed, int half_store,
                                                 float *__restrict__ out) {
    __shared__ double s_sum[256];
    const int e = blockIdx.x;
    const int i = e / AUR_KEL_ROUGH, ri = e % AUR_KEL_ROUGH;
    const float rough = fmaxf(0.00001f, (float)ri / (AUR_KEL_ROUGH - 1.f));
    const uint32_t key = path_key(seed, (uint32_t)e, 0u);
    const float zv = fmaxf(0.0001f, (float)i / (AUR_KEL_COS - 1));
    const float rv = sqrtf(fmaxf(0.f, 1.f - zv * zv));
    double acc = 0.0;
    for (int j = threadIdx.x; j < n_samples; j += blockDim.x) {
        float r1 = slot_uniform(key, 3u * (uint32_t)j);
        float r2 = slot_uniform(key, 3u * (uint32_t)j + 1u);
        float r3 = slot_uniform(key, 3u * (uint32_t)j + 2u);
        float z;
        if (dist == DIST_GGX) z = cosf(atanf(rough * sqrtf(r1 / (1.f - r1))));
        else z = cosf(atanf(rough * sqrtf(-logf(1.f - r1))));
        float rr = sqrtf(fmaxf(0.f, 1.f - z * z));
        float phi = (2.f * AUR_PI_F) * r2;
        f3 hn = norm3(mk3(rr * cosf(phi), rr * sinf(phi), z));
        float phv = r3 * 2.f * AUR_PI_F;
        f3 vn = norm3(mk3(rv * cosf(phv), rv * sinf(phv), zv));
        if (dot3(hn, vn) < 0.f) hn = neg3(hn);
        float vdoth = dot3(vn, hn);
        if (vdoth > 0.000001f) {
            f3 ln = norm3(sub3(mul3(mul3(hn, 2.f), vdoth), vn));
            float F = schlick(dot3(hn, ln), reflectance);
            acc += (double)(F * fmaxf(0.f, ln.z) / vdoth);
        }
    }
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float v = (float)(s_sum[0] / n_samples);
        out[e] = half_store ? (float)(_Float16)v : v;  // to nearest even
    }
}

void kelemen_table(Ctx *c, float reflectance, int dist, int n_samples, uint32_t seed, float *d_out, bool half_store) {
    hipLaunchKernelGGL(k_kelemen, dim3(AUR_KEL_COS * AUR_KEL_ROUGH), dim3(256), 0, c->stream, reflectance, dist, n_samples,
                       seed, half_store ? 1 : 0, d_out);
    HIP_CHECK(hipGetLastError());
}

}  // namespace aur
