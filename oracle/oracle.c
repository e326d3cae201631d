/*
 * oracle.c — CPU restatement of Bulbit's unidirectional path tracer.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the reference
 * file:line it restates; paths are relative to the Bulbit source tree.
 *
 * Argument evaluation order.  Two reference call sites draw random numbers
 * inside one argument list, an order C++ leaves unspecified:
 *   camera->SampleRay(&ray, pixel, sampler->Next2D(), sampler->Next2D())
 *       src/integrator/integrator.cpp:190
 *   bsdf.Sample_f(&s, wo, sampler.Next1D(), sampler.Next2D())
 *       src/integrator/path_integrator.cpp:115
 * g++ 11 (the reference's Linux compiler) evaluates them right to left, so the
 * lens sample u1 is drawn before the filter sample u0, and u12 before u0.
 * oracle/probe_argorder.cpp pins this; the draws below are sequenced
 * explicitly in that order.
 *
 * std::complex<float> arithmetic as g++ 11 compiles it (pinned by
 * oracle/probe_complex.cpp): multiplication is the float formula
 * (ac - bd, ad + bc); division is the naive formula evaluated in double and
 * rounded to float; std::abs is hypotf.
 */
#include "oracle.h"
#include "orc_math.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ================================================================ hashing */

/* Transcendentals on the sampling path.  The reference calls glibc's float
 * functions, and so does the oracle by default.  orc_set_cr_libm(1) (tests
 * only) evaluates them in double and rounds, as the GPU path does, so that a
 * comparison can separate what the libm difference causes from the rest. */
static int orc_cr_libm = 0;
static inline float o_sinf(float x) { return orc_cr_libm ? (float)sin((double)x) : sinf(x); }
static inline float o_cosf(float x) { return orc_cr_libm ? (float)cos((double)x) : cosf(x); }
static inline float o_acosf(float x) { return orc_cr_libm ? (float)acos((double)x) : acosf(x); }
static inline float o_atan2f(float y, float x) { return orc_cr_libm ? (float)atan2((double)y, (double)x) : atan2f(y, x); }
static inline float o_powf(float x, float y) { return orc_cr_libm ? (float)pow((double)x, (double)y) : powf(x, y); }

/* include/bulbit/hash.h:17-70 (MurmurHash64A) */
uint64_t orc_murmur64a(const void* key_, int32_t len, uint64_t seed)
{
    const unsigned char* key = (const unsigned char*)key_;
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t h = seed ^ ((uint64_t)len * m);
    const unsigned char* end = key + 8 * (len / 8);
    while (key != end) {
        uint64_t k;
        memcpy(&k, key, 8);
        key += 8;
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    switch (len & 7) {
    case 7: h ^= (uint64_t)key[6] << 48; /* fallthrough */
    case 6: h ^= (uint64_t)key[5] << 40; /* fallthrough */
    case 5: h ^= (uint64_t)key[4] << 32; /* fallthrough */
    case 4: h ^= (uint64_t)key[3] << 24; /* fallthrough */
    case 3: h ^= (uint64_t)key[2] << 16; /* fallthrough */
    case 2: h ^= (uint64_t)key[1] << 8;  /* fallthrough */
    case 1:
        h ^= (uint64_t)key[0];
        h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

/* hash.h:111-120: Hash(args...) hashes the packed bytes of its arguments;
 * IndependentSampler passes (Point2i pixel, int32 seed, int32 sample_index)
 * (src/sampler/independent_sampler.cpp:14-17) = 16 bytes. */
uint64_t orc_pixel_sample_hash(int32_t px, int32_t py, int32_t seed, int32_t sample_index)
{
    int32_t buf[4] = { px, py, seed, sample_index };
    return orc_murmur64a(buf, 16, 0);
}

/* ================================================================ PCG32 */

typedef struct {
    uint64_t state, inc;
    int32_t draws;
} pcg32;

/* include/bulbit/random.h:39-46 */
static inline uint32_t pcg_next(pcg32* g)
{
    uint64_t old = g->state;
    g->state = old * 0x5851f42d4c957f2dULL + g->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
}

/* random.h:27-34 */
static inline void pcg_seed(pcg32* g, uint64_t initstate, uint64_t initseq)
{
    g->state = 0u;
    g->inc = (initseq << 1u) | 1u;
    pcg_next(g);
    g->state += initstate;
    pcg_next(g);
    g->draws = 0;
}

/* random.h:87-90: fmin(1 - epsilon, float(u32) * 2^-32) */
static inline float pcg_float(pcg32* g)
{
    g->draws++;
    return fminf(1.0f - FLT_EPSILON, (float)pcg_next(g) * 0x1p-32f);
}

/* IndependentSampler::Next1D / Next2D (independent_sampler.cpp:19-27);
 * Next2D's braced initializer is sequenced: x first. */
static inline float next1d(pcg32* g) { return pcg_float(g); }
static inline v2 next2d(pcg32* g)
{
    float x = pcg_float(g);
    float y = pcg_float(g);
    return V2(x, y);
}

static inline void sampler_start(pcg32* g, int32_t px, int32_t py, int32_t seed, int32_t sample_index)
{
    pcg_seed(g, orc_pixel_sample_hash(px, py, seed, sample_index), 1);
}

void orc_pcg32_outputs(uint64_t initstate, uint64_t initseq, int32_t n, uint32_t* out)
{
    pcg32 g;
    pcg_seed(&g, initstate, initseq);
    for (int32_t i = 0; i < n; ++i) out[i] = pcg_next(&g);
}

void orc_sampler_floats(int32_t px, int32_t py, int32_t seed, int32_t sample_index, int32_t n, float* out)
{
    pcg32 g;
    sampler_start(&g, px, py, seed, sample_index);
    for (int32_t i = 0; i < n; ++i) out[i] = next1d(&g);
}

/* ============================================================== sampling */

/* include/bulbit/sampling.h:114-134 */
static v2 sample_concentric(v2 u)
{
    v2 uo = V2(u.x * 2.0f - 1.0f, u.y * 2.0f - 1.0f);
    if (uo.x == 0.0f && uo.y == 0.0f) return V2(0.0f, 0.0f);
    float theta, r;
    if (fabsf(uo.x) > fabsf(uo.y)) {
        r = uo.x;
        theta = (orc_pi / 4.0f) * (uo.y / uo.x);
    } else {
        r = uo.y;
        theta = (orc_pi / 2.0f) - (orc_pi / 4.0f) * (uo.x / uo.y);
    }
    return V2(o_cosf(theta) * r, o_sinf(theta) * r);
}

/* sampling.h:64-74 */
static v3 sample_uniform_sphere(v2 u)
{
    float z = 1.0f - 2.0f * u.x;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = orc_two_pi * u.y;
    return V3(r * o_cosf(phi), r * o_sinf(phi), z);
}

/* sampling.h:151-169 (Malley's method) */
static v3 sample_cosine_hemisphere(v2 u)
{
    v2 d = sample_concentric(u);
    float z = safe_sqrt(1.0f - sqrf(d.x) - sqrf(d.y));
    return V3(d.x, d.y, z);
}

/* sampling.h:42-48 with nf = ng = 1 */
static inline float power_heuristic(float pdf_f, float pdf_g)
{
    float f = 1.0f * pdf_f;
    float g = 1.0f * pdf_g;
    return (f * f) / (f * f + g * g);
}

void orc_sample_concentric(const float u[2], float out[2])
{
    v2 r = sample_concentric(V2(u[0], u[1]));
    out[0] = r.x;
    out[1] = r.y;
}

void orc_sample_uniform_sphere(const float u[2], float out[3])
{
    v3 r = sample_uniform_sphere(V2(u[0], u[1]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

/* ============================================================ quaternions */

typedef struct { float x, y, z, w; } quat;

/* include/bulbit/quaternion.h:142-170 (Rotate / RotateInv) */
static v3 quat_rotate(quat q, v3 v)
{
    float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
    float w2 = q.w * q.w - 0.5f;
    float dot2 = (q.x * vx + q.y * vy + q.z * vz);
    return V3((vx * w2 + (q.y * vz - q.z * vy) * q.w + q.x * dot2),
              (vy * w2 + (q.z * vx - q.x * vz) * q.w + q.y * dot2),
              (vz * w2 + (q.x * vy - q.y * vx) * q.w + q.z * dot2));
}

static v3 quat_rotate_inv(quat q, v3 v)
{
    float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
    float w2 = q.w * q.w - 0.5f;
    float dot2 = (q.x * vx + q.y * vy + q.z * vz);
    return V3((vx * w2 - (q.y * vz - q.z * vy) * q.w + q.x * dot2),
              (vy * w2 - (q.z * vx - q.x * vz) * q.w + q.y * dot2),
              (vz * w2 - (q.x * vy - q.y * vx) * q.w + q.z * dot2));
}

/* quaternion.h:181-205 */
static v3 quat_basis_x(quat q)
{
    float x2 = q.x * 2.0f, w2 = q.w * 2.0f;
    return V3((q.w * w2) - 1.0f + q.x * x2, (q.z * w2) + q.y * x2, (-q.y * w2) + q.z * x2);
}
static v3 quat_basis_y(quat q)
{
    float y2 = q.y * 2.0f, w2 = q.w * 2.0f;
    return V3((-q.z * w2) + q.x * y2, (q.w * w2) - 1.0f + q.y * y2, (q.x * w2) + q.z * y2);
}
static v3 quat_basis_z(quat q)
{
    float z2 = q.z * 2.0f, w2 = q.w * 2.0f;
    return V3((q.y * w2) + q.x * z2, (-q.x * w2) + q.y * z2, (q.w * w2) - 1.0f + q.z * z2);
}

static quat quat_scale(float x, float y, float z, float w, float s)
{
    quat q = { x * s, y * s, z * s, w * s };
    return q;
}

/* quaternion.h:34-62: Quat(const Mat3&) with columns ex, ey, ez */
static quat quat_from_mat3(v3 ex, v3 ey, v3 ez)
{
    if (ez.z < 0.0f) {
        if (ex.x > ey.y) {
            float t = 1.0f + ex.x - ey.y - ez.z;
            return quat_scale(t, ex.y + ey.x, ez.x + ex.z, ey.z - ez.y, 0.5f / sqrtf(t));
        } else {
            float t = 1.0f - ex.x + ey.y - ez.z;
            return quat_scale(ex.y + ey.x, t, ey.z + ez.y, ez.x - ex.z, 0.5f / sqrtf(t));
        }
    } else {
        if (ex.x < -ey.y) {
            float t = 1.0f - ex.x - ey.y + ez.z;
            return quat_scale(ez.x + ex.z, ey.z + ez.y, t, ex.y - ey.x, 0.5f / sqrtf(t));
        } else {
            float t = 1.0f + ex.x + ey.y + ez.z;
            return quat_scale(ey.z - ez.y, ez.x - ex.z, ex.y - ey.x, t, 0.5f / sqrtf(t));
        }
    }
}

/* transform.h:112-122 (LookAt) and quaternion.h:64-75 (Quat(front, up)) */
void orc_look_at(const float position[3], const float target[3], const float up_[3], float q_out[4])
{
    v3 w = vsub(V3(target[0], target[1], target[2]), V3(position[0], position[1], position[2]));
    vnormalize_m(&w);
    v3 up = V3(up_[0], up_[1], up_[2]);
    v3 ez = w;
    v3 ex = vcross(up, ez);
    vnormalize_m(&ex);
    v3 ey = vcross(ez, ex);
    quat q = quat_from_mat3(ex, ey, ez);
    q_out[0] = q.x;
    q_out[1] = q.y;
    q_out[2] = q.z;
    q_out[3] = q.w;
}

/* ================================================================= camera */

/* Distribution1D over n <= ORC_FILTER_N values (sampling.h:343-379) */
#define ORC_FILTER_N 32
typedef struct {
    float func[ORC_FILTER_N], cdf[ORC_FILTER_N + 1], func_integral;
    int32_t n;
} dist1d_t;

static void dist1d_init(dist1d_t* d, const float* f, int32_t n)
{
    d->n = n;
    for (int32_t i = 0; i < n; ++i) d->func[i] = f[i];
    d->cdf[0] = 0.0f;
    for (int32_t i = 1; i < n + 1; ++i) d->cdf[i] = d->cdf[i - 1] + d->func[i - 1] / (float)n;
    d->func_integral = d->cdf[n];
    if (d->func_integral == 0.0f) {
        for (int32_t i = 1; i < n + 1; ++i) d->cdf[i] = (float)i / (float)n;
    } else {
        for (int32_t i = 1; i < n + 1; ++i) d->cdf[i] /= d->func_integral;
    }
}

/* Distribution1D::SampleContinuous (sampling.h:381-398) with FindInterval
 * (math.h:117-135) */
static float dist1d_sample(const dist1d_t* d, float u, int32_t* off)
{
    int32_t size = d->n + 1;
    int32_t first = 0, len = size;
    while (len > 0) {
        int32_t half = len >> 1, middle = first + half;
        if (d->cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    int32_t offset = first - 1;
    if (offset < 0) offset = 0;
    if (offset > size - 2) offset = size - 2;
    if (off) *off = offset;
    float du = (u - d->cdf[offset]) / (d->cdf[offset + 1] - d->cdf[offset]);
    return ((float)offset + du) / (float)d->n;
}

/* Distribution2D (sampling.h:426-478) of the Gaussian filter table */
typedef struct {
    dist1d_t cond[ORC_FILTER_N];
    dist1d_t marginal;
} dist2d_t;

typedef struct {
    v3 origin, lower_left, horizontal, vertical, u, v, w;
    float lens_radius, filter_extent, filter_sigma;
    int32_t filter_type;
    int32_t width, height;
    dist2d_t gauss;
} camera_t;

/* Gaussian (sampling.h:9-12) */
static float gaussian(float x, float mu, float sigma)
{
    float xm = x - mu;
    return 1.0f / sqrtf(orc_two_pi * sigma * sigma) * expf(-(xm * xm) / (2.0f * sigma * sigma));
}

/* GaussianFilter::GaussianFilter (gaussian_filter.cpp:8-27): 32 x 32 table
 * of the separable Gaussian over [-extent/2, extent/2]^2 */
static void gaussian_filter_init(dist2d_t* g, float sigma, float extent)
{
    const int32_t n = ORC_FILTER_N;
    float values[ORC_FILTER_N * ORC_FILTER_N];
    float half_extent = extent / 2.0f;
    for (int32_t y = 0; y < n; ++y)
        for (int32_t x = 0; x < n; ++x) {
            float px = (((float)x + 0.5f) / (float)n) * extent - half_extent;
            float py = (((float)y + 0.5f) / (float)n) * extent - half_extent;
            values[y * n + x] = gaussian(px, 0.0f, sigma) * gaussian(py, 0.0f, sigma);
        }
    float marginal[ORC_FILTER_N];
    for (int32_t v = 0; v < n; ++v) {
        dist1d_init(&g->cond[v], &values[v * n], n);
        marginal[v] = g->cond[v].func_integral;
    }
    dist1d_init(&g->marginal, marginal, n);
}

/* Filter::Sample of the three filters: box_filter.cpp:17-21,
 * tent_filter.cpp:22-30, gaussian_filter.cpp:40-43 */
static v2 filter_sample(const camera_t* c, v2 u)
{
    float e = c->filter_extent;
    if (c->filter_type == BB_FILTER_TENT) {
        float h = e / 2.0f;
        float x = u.x < 0.5f ? h * (sqrtf(2.0f * u.x) - 1.0f) : h * (1.0f - sqrtf(1.0f - 2.0f * (u.x - 0.5f)));
        float y = u.y < 0.5f ? h * (sqrtf(2.0f * u.y) - 1.0f) : h * (1.0f - sqrtf(1.0f - 2.0f * (u.y - 0.5f)));
        return V2(x, y);
    }
    if (c->filter_type == BB_FILTER_GAUSSIAN) {
        int32_t v = 0;
        float d1 = dist1d_sample(&c->gauss.marginal, u.y, &v);
        float d0 = dist1d_sample(&c->gauss.cond[v], u.x, NULL);
        float h = e / 2.0f;
        return V2(d0 * e - h, d1 * e - h);
    }
    return V2((u.x * 2.0f - 1.0f) * e / 2.0f, (u.y * 2.0f - 1.0f) * e / 2.0f);
}

/* src/camera/perspective_camera.cpp:8-38 */
static void camera_init(camera_t* c, const BbCameraDesc* d)
{
    quat q = { d->orientation[0], d->orientation[1], d->orientation[2], d->orientation[3] };
    float theta = d->vfov_deg * orc_pi / 180.0f; /* DegToRad, math.h:36-39 */
    float h = tanf(theta / 2.0f);
    float aspect_ratio = (float)d->width / (float)d->height;
    float viewport_height = 2.0f * h;
    float viewport_width = aspect_ratio * viewport_height;
    c->u = vneg(quat_basis_x(q));
    c->v = quat_basis_y(q);
    c->w = quat_basis_z(q);
    c->origin = V3(d->position[0], d->position[1], d->position[2]);
    c->horizontal = vmul(c->u, d->focus_distance * viewport_width);
    c->vertical = vmul(c->v, d->focus_distance * viewport_height);
    c->lower_left = vadd(vsub(vsub(c->origin, vdiv(c->horizontal, 2.0f)), vdiv(c->vertical, 2.0f)),
                         vmul(c->w, d->focus_distance));
    c->lens_radius = d->aperture_radius;
    c->filter_extent = d->filter_extent;
    c->filter_type = d->filter_type;
    c->filter_sigma = d->filter_sigma;
    if (c->filter_type == BB_FILTER_GAUSSIAN) gaussian_filter_init(&c->gauss, d->filter_sigma, d->filter_extent);
    c->width = d->width;
    c->height = d->height;
}

/* perspective_camera.cpp:40-52 with the camera's Filter::Sample */
static void camera_sample_ray(const camera_t* c, int32_t px, int32_t py, v2 u0, v2 u1, v3* o, v3* dir)
{
    v2 off = filter_sample(c, u0);
    off = V2(off.x + 0.5f, off.y + 0.5f);
    v3 pixel_center = vadd(vadd(c->lower_left, vdiv(vmul(c->horizontal, (float)px + off.x), (float)c->width)),
                           vdiv(vmul(c->vertical, (float)py + off.y), (float)c->height));
    v2 ap = sample_concentric(u1);
    ap = V2(ap.x * c->lens_radius, ap.y * c->lens_radius);
    v3 offset = vadd(vmul(c->u, ap.x), vmul(c->v, ap.y));
    *o = vadd(c->origin, offset);
    *dir = vnormalize(vsub(pixel_center, *o));
}

void orc_filter_sample(const BbCameraDesc* cam, const float u[2], float out[2])
{
    camera_t c;
    camera_init(&c, cam);
    v2 r = filter_sample(&c, V2(u[0], u[1]));
    out[0] = r.x;
    out[1] = r.y;
}

void orc_camera_ray(const BbCameraDesc* cam, int32_t px, int32_t py, const float u0[2], const float u1[2], BbRay* ray)
{
    camera_t c;
    camera_init(&c, cam);
    v3 o, d;
    camera_sample_ray(&c, px, py, V2(u0[0], u0[1]), V2(u1[0], u1[1]), &o, &d);
    ray->o[0] = o.x;
    ray->o[1] = o.y;
    ray->o[2] = o.z;
    ray->d[0] = d.x;
    ray->d[1] = d.y;
    ray->d[2] = d.z;
    ray->t_min = 1e-4f;
    ray->t_max = INFINITY;
}

/* ================================================================= scene */

#define BX_REFLECTION 1
#define BX_TRANSMISSION 2
#define BX_DIFFUSE 4
#define BX_GLOSSY 8
#define BX_SPECULAR 16

typedef struct {
    v3 z;                     /* spot frame z = -Normalize(direction) */
    float cos_min, cos_max;   /* spot */
    float world_radius;       /* infinite */
} light_extra;

struct OrcScene {
    int32_t n_vertices, n_triangles, n_spheres, n_prims, n_materials, n_lights;
    float* pos;
    float* nrm;
    float* tan;
    int32_t* idx;
    BbSphere* spheres;
    BbPrimitive* prims;
    BbMaterial* mats;
    BbLight* lights;
    light_extra* lx;
    float* uvs;        /* 2 per vertex, NULL: every texture coordinate is (0, 0) */
    int32_t n_textures;
    BbTexture* texs;   /* texels point at owned copies */
    /* BVH */
    BbBvhNode* nodes;
    int32_t n_nodes;
    int32_t* ordered; /* slot -> primitive */
    /* integrator set-up */
    int32_t* prim_light; /* primitive -> area light index, -1 */
    int32_t n_inf;
    int32_t* inf_lights;
    float* func; /* n_lights */
    float* cdf;  /* n_lights + 1 */
    float func_integral;
    v3 world_min, world_max;
    double build_seconds;
};

static inline v3 vtx(const OrcScene* s, const float* a, int32_t i) { (void)s; return V3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

/* ============================================================= triangles */

/* src/geometry/triangle.cpp:40-52 */
static void triangle_aabb(const OrcScene* s, int32_t tri, v3* mn, v3* mx)
{
    const int32_t* v = s->idx + 3 * tri;
    v3 p0 = vtx(s, s->pos, v[0]), p1 = vtx(s, s->pos, v[1]), p2 = vtx(s, s->pos, v[2]);
    float off = FLT_EPSILON * 10.0f;
    *mn = vsub(vmin(vmin(p0, p1), p2), V3(off, off, off));
    *mx = vadd(vmax(vmax(p0, p1), p2), V3(off, off, off));
}

/* triangle.cpp:55-97: Möller–Trumbore with the direction normalized by the
 * member Normalize */
static inline int mt_intersect(v3 p0, v3 p1, v3 p2, v3 o, v3 d, float t_min, float t_max, float* t_out, float* u_out,
                               float* v_out)
{
    v3 e1 = vsub(p1, p0);
    v3 e2 = vsub(p2, p0);
    v3 dn = d;
    float l = vnormalize_m(&dn);
    v3 pvec = vcross(dn, e2);
    float det = vdot(e1, pvec);
    if (fabsf(det) < FLT_EPSILON) return 0;
    float inv_det = 1.0f / det;
    v3 tvec = vsub(o, p0);
    float u = vdot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return 0;
    v3 qvec = vcross(tvec, e1);
    float v = vdot(dn, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = vdot(e2, qvec) * inv_det / l;
    if (t < t_min || t > t_max) return 0;
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return 1;
}

int orc_triangle_intersect(const float a[3], const float b[3], const float c[3], const BbRay* ray, float* t, float* u,
                           float* v)
{
    return mt_intersect(V3(a[0], a[1], a[2]), V3(b[0], b[1], b[2]), V3(c[0], c[1], c[2]),
                        V3(ray->o[0], ray->o[1], ray->o[2]), V3(ray->d[0], ray->d[1], ray->d[2]), ray->t_min,
                        ray->t_max, t, u, v);
}

/* ================================================================ spheres */

static inline quat sphere_q(const BbSphere* sp)
{
    quat q = { sp->q[0], sp->q[1], sp->q[2], sp->q[3] };
    return q;
}

/* src/geometry/sphere.cpp:30-33 */
static void sphere_aabb(const BbSphere* sp, v3* mn, v3* mx)
{
    v3 p = V3(sp->p[0], sp->p[1], sp->p[2]);
    v3 r = V3(sp->radius, sp->radius, sp->radius);
    *mn = vsub(p, r);
    *mx = vadd(p, r);
}

/* sphere.cpp:35-61 (+ MulT(Transform, Ray), include/bulbit/ray.h:40-43;
 * MulT(Transform, Vec3), transform.h:142-146).  Returns the root. */
static inline int sphere_intersect(const BbSphere* sp, v3 o, v3 d, float t_min, float t_max, float* t_out)
{
    quat q = sphere_q(sp);
    v3 p = V3(sp->p[0], sp->p[1], sp->p[2]);
    v3 s = V3(sp->s[0], sp->s[1], sp->s[2]);
    v3 ro = vdivv(quat_rotate_inv(q, vsub(o, p)), s);
    v3 rd = quat_rotate_inv(q, vdivv(d, s));
    float a = vlen2(rd);
    float half_b = vdot(ro, rd);
    float c = vlen2(ro) - sp->radius * sp->radius;
    float discriminant = half_b * half_b - a * c;
    if (discriminant < 0.0f) return 0;
    float sqrt_d = sqrtf(discriminant);
    float root = (-half_b - sqrt_d) / a;
    if (root < t_min || t_max < root) {
        root = (-half_b + sqrt_d) / a;
        if (root < t_min || t_max < root) return 0;
    }
    *t_out = root;
    return 1;
}

/* ============================================================ hit record */

typedef struct {
    int32_t prim;
    float t, u, v;
    v3 ro, rd; /* the ray the accepted candidate was found on (a restarted one after an alpha rejection) */
    v3 point, normal, ns, ts;
    v2 uv;
    int front_face;
} isect_t;

/* include/bulbit/shape.h:32-43 */
static void set_face_normal(isect_t* is, v3 wi, v3 outward, v3 shading_n, v3 shading_t)
{
    int front = vdot(wi, outward) < 0.0f;
    float sign = front ? 1.0f : -1.0f;
    is->front_face = front;
    is->normal = vmul(outward, sign);
    is->ns = vmul(shading_n, sign);
    is->ts = vmul(shading_t, sign);
}

/* Sphere's ComputeTexCoord (src/geometry/sphere.cpp:9-16) */
static v2 sphere_texcoord(v3 n)
{
    float theta = o_acosf(clampf_(n.y, -1.0f, 1.0f));
    float r = o_atan2f(n.z, n.x);
    float phi = r < 0.0f ? r + orc_two_pi : r;
    return V2(phi * orc_inv_two_pi, 1.0f - theta * orc_inv_pi);
}

/* Triangle::GetTexCoord (triangle.cpp:31-38) */
static v2 triangle_texcoord(const OrcScene* s, const int32_t* v, float u, float vv, float w)
{
    if (!s->uvs) return V2(0.0f, 0.0f);
    const float* a = s->uvs + 2 * v[0];
    const float* b = s->uvs + 2 * v[1];
    const float* c = s->uvs + 2 * v[2];
    return V2(w * a[0] + u * b[0] + vv * c[0], w * a[1] + u * b[1] + vv * c[1]);
}

/* The shading part of Triangle::Intersect (triangle.cpp:99-107) and
 * Sphere::Intersect (sphere.cpp:63-82), evaluated once for the final closest
 * candidate on the ray it was found on; the reference evaluates it for every
 * accepted candidate with the same inputs, so the last evaluation is
 * identical. */
static void finish_isect(const OrcScene* s, v3 o, v3 d, isect_t* is)
{
    const BbPrimitive* pr = &s->prims[is->prim];
    if (pr->shape_kind == BB_SHAPE_TRIANGLE) {
        const int32_t* v = s->idx + 3 * pr->shape_index;
        v3 p0 = vtx(s, s->pos, v[0]), p1 = vtx(s, s->pos, v[1]), p2 = vtx(s, s->pos, v[2]);
        v3 e1 = vsub(p1, p0), e2 = vsub(p2, p0);
        float u = is->u, vv = is->v;
        float w = 1.0f - u - vv;
        is->point = vadd(o, vmul(d, is->t));
        is->uv = triangle_texcoord(s, v, u, vv, w);
        v3 normal = vnormalize(vcross(e1, e2));
        /* Triangle::GetNormal / GetTangent (triangle.cpp:13-29) */
        v3 n0 = vtx(s, s->nrm, v[0]), n1 = vtx(s, s->nrm, v[1]), n2 = vtx(s, s->nrm, v[2]);
        v3 t0 = vtx(s, s->tan, v[0]), t1 = vtx(s, s->tan, v[1]), t2 = vtx(s, s->tan, v[2]);
        v3 ns = vnormalize(vadd(vadd(vmul(n0, w), vmul(n1, u)), vmul(n2, vv)));
        v3 ts = vnormalize(vadd(vadd(vmul(t0, w), vmul(t1, u)), vmul(t2, vv)));
        set_face_normal(is, d, normal, ns, ts);
    } else {
        const BbSphere* sp = &s->spheres[pr->shape_index];
        quat q = sphere_q(sp);
        v3 p = V3(sp->p[0], sp->p[1], sp->p[2]);
        v3 sc = V3(sp->s[0], sp->s[1], sp->s[2]);
        v3 ro = vdivv(quat_rotate_inv(q, vsub(o, p)), sc);
        v3 rd = quat_rotate_inv(q, vdivv(d, sc));
        v3 point = vadd(ro, vmul(rd, is->t));
        v3 normal = vnormalize(vdiv(point, sp->radius));
        is->point = vadd(quat_rotate(q, vmulv(sc, point)), p);
        is->uv = sphere_texcoord(normal);
        v3 n = quat_rotate(q, normal);
        v3 t;
        if (n.y < 1.0f - 1e-4f)
            t = vnormalize(vcross(V3(0.0f, 1.0f, 0.0f), n));
        else
            t = vnormalize(vcross(n, V3(1.0f, 0.0f, 0.0f)));
        set_face_normal(is, d, n, n, t);
    }
}

/* ================================================================== BVH */

typedef struct {
    int32_t index;
    v3 bmin, bmax;
} bvh_prim;

typedef struct {
    const OrcScene* s;
    bvh_prim* prims;
    BbBvhNode* nodes;
    int32_t n_nodes;
    int32_t ordered_offset;
    int32_t* ordered;
    int error;
} bvh_builder;

/* bounding_box.h:376-381 */
static inline float box_area(v3 mn, v3 mx)
{
    v3 w = vsub(mx, mn);
    return 2.0f * ((w.x * w.y) + (w.y * w.z) + (w.z * w.x));
}

static inline float prim_center(const bvh_prim* p, int axis)
{
    return (vcomp(p->bmin, axis) + vcomp(p->bmax, axis)) * 0.5f;
}

/* bvh.cpp:162-168 and 215-221: the SAH bucket of a primitive */
static inline int bucket_of(const bvh_prim* p, int axis, float cmin, float extent)
{
    int b = (int)(12.0f * (prim_center(p, axis) - cmin) / extent);
    if (b == 12) b = 11;
    return b;
}

static void make_leaf(bvh_builder* B, int32_t node, int32_t start, int32_t n, v3 mn, v3 mx)
{
    int32_t offset = B->ordered_offset;
    B->ordered_offset += n;
    for (int32_t i = 0; i < n; ++i) B->ordered[offset + i] = B->prims[start + i].index;
    if (n > 65535) B->error = 1;
    BbBvhNode* nd = &B->nodes[node];
    nd->bmin[0] = mn.x; nd->bmin[1] = mn.y; nd->bmin[2] = mn.z;
    nd->bmax[0] = mx.x; nd->bmax[1] = mx.y; nd->bmax[2] = mx.z;
    nd->offset = offset;
    nd->count = (uint16_t)n;
    nd->axis = 0;
    nd->pad = 0;
}

/* bvh.cpp:56-271, BuildRecursive, with the pre-order numbering of FlattenBVH
 * (bvh.cpp:273-297): a node's index is taken before its children are built,
 * child1 = index + 1, child2 stored in offset.  Sequential, so leaf offsets
 * follow depth-first order (the reference's parallel build above 65,536
 * primitives only permutes leaf offsets, never leaf contents or topology). */
static int32_t bvh_build(bvh_builder* B, int32_t start, int32_t end)
{
    int32_t node = B->n_nodes++;
    int32_t n = end - start;
    bvh_prim* P = B->prims + start;

    v3 mn = V3(FLT_MAX, FLT_MAX, FLT_MAX), mx = V3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int32_t i = 0; i < n; ++i) {
        mn = vmin(mn, P[i].bmin);
        mx = vmax(mx, P[i].bmax);
    }
    if (box_area(mn, mx) == 0.0f || n == 1) {
        make_leaf(B, node, start, n, mn, mx);
        return node;
    }

    v3 cmn = V3(FLT_MAX, FLT_MAX, FLT_MAX), cmx = V3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int32_t i = 0; i < n; ++i) {
        v3 c = vmul(vadd(P[i].bmin, P[i].bmax), 0.5f);
        cmn = vmin(cmn, c);
        cmx = vmax(cmx, c);
    }
    v3 ext = vsub(cmx, cmn);
    int axis = 0;
    if (ext.y > ext.x) axis = 1;
    if (ext.z > vcomp(ext, axis)) axis = 2;
    float extent = vcomp(ext, axis);
    if (extent == 0.0f) {
        make_leaf(B, node, start, n, mn, mx);
        return node;
    }

    int32_t mid;
    if (n <= 2) {
        /* std::nth_element(begin, begin + 1, end) on two elements is an
         * insertion sort: swap when the second centre is strictly smaller */
        mid = n / 2;
        if (prim_center(&P[1], axis) < prim_center(&P[0], axis)) {
            bvh_prim t = P[0];
            P[0] = P[1];
            P[1] = t;
        }
    } else {
        int count[12] = { 0 };
        v3 bmn[12], bmx[12];
        for (int b = 0; b < 12; ++b) {
            bmn[b] = V3(FLT_MAX, FLT_MAX, FLT_MAX);
            bmx[b] = V3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        }
        float cmin = vcomp(cmn, axis);
        for (int32_t i = 0; i < n; ++i) {
            int b = bucket_of(&P[i], axis, cmin, extent);
            count[b]++;
            bmn[b] = vmin(bmn[b], P[i].bmin);
            bmx[b] = vmax(bmx[b], P[i].bmax);
        }
        v3 lmn = V3(FLT_MAX, FLT_MAX, FLT_MAX), lmx = V3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        v3 rmn = lmn, rmx = lmx;
        int left_count[11], right_count[11];
        float left_area[11], right_area[11];
        int left_sum = 0, right_sum = 0;
        for (int i = 0; i < 11; ++i) {
            left_sum += count[i];
            left_count[i] = left_sum;
            lmn = vmin(lmn, bmn[i]);
            lmx = vmax(lmx, bmx[i]);
            left_area[i] = box_area(lmn, lmx);
            right_sum += count[11 - i];
            right_count[10 - i] = right_sum;
            rmn = vmin(rmn, bmn[11 - i]);
            rmx = vmax(rmx, bmx[11 - i]);
            right_area[10 - i] = box_area(rmn, rmx);
        }
        int best = -1;
        float min_cost = INFINITY;
        for (int i = 0; i < 11; ++i) {
            float cost = (float)left_count[i] * left_area[i] + (float)right_count[i] * right_area[i];
            if (cost < min_cost) {
                min_cost = cost;
                best = i;
            }
        }
        min_cost = 0.5f + min_cost / box_area(mn, mx);
        if (min_cost < (float)n) {
            /* std::partition, libstdc++'s bidirectional-iterator form */
            int32_t first = 0, last = n;
            for (;;) {
                for (;;) {
                    if (first == last) goto done;
                    if (bucket_of(&P[first], axis, cmin, extent) <= best)
                        ++first;
                    else
                        break;
                }
                --last;
                for (;;) {
                    if (first == last) goto done;
                    if (!(bucket_of(&P[last], axis, cmin, extent) <= best))
                        --last;
                    else
                        break;
                }
                bvh_prim t = P[first];
                P[first] = P[last];
                P[last] = t;
                ++first;
            }
        done:
            mid = first;
        } else {
            make_leaf(B, node, start, n, mn, mx);
            return node;
        }
    }

    int32_t c1 = bvh_build(B, start, start + mid);
    int32_t c2 = bvh_build(B, start + mid, end);
    BbBvhNode* nd = &B->nodes[node];
    /* BuildNode::InitInternal (bvh.h:90-97): box = Union(child1, child2) */
    for (int k = 0; k < 3; ++k) {
        nd->bmin[k] = smin(B->nodes[c1].bmin[k], B->nodes[c2].bmin[k]);
        nd->bmax[k] = smax(B->nodes[c1].bmax[k], B->nodes[c2].bmax[k]);
    }
    nd->offset = c2;
    nd->count = 0;
    nd->axis = (uint8_t)axis;
    nd->pad = 0;
    return node;
}

/* bounding_box.h:443-484, the ordered slab test (x, y tighten; z does not) */
static inline int box_test(const BbBvhNode* nd, v3 o, float t_min, float t_max, v3 inv, const int neg[3])
{
    float bx0 = neg[0] ? nd->bmax[0] : nd->bmin[0], bx1 = neg[0] ? nd->bmin[0] : nd->bmax[0];
    float t_min_x = (bx0 - o.x) * inv.x;
    float t_max_x = (bx1 - o.x) * inv.x;
    if (t_min > t_max_x || t_max < t_min_x) return 0;
    if (t_min_x > t_min) t_min = t_min_x;
    if (t_max_x < t_max) t_max = t_max_x;
    float by0 = neg[1] ? nd->bmax[1] : nd->bmin[1], by1 = neg[1] ? nd->bmin[1] : nd->bmax[1];
    float t_min_y = (by0 - o.y) * inv.y;
    float t_max_y = (by1 - o.y) * inv.y;
    if (t_min > t_max_y || t_max < t_min_y) return 0;
    if (t_min_y > t_min) t_min = t_min_y;
    if (t_max_y < t_max) t_max = t_max_y;
    float bz0 = neg[2] ? nd->bmax[2] : nd->bmin[2], bz1 = neg[2] ? nd->bmin[2] : nd->bmax[2];
    float t_min_z = (bz0 - o.z) * inv.z;
    float t_max_z = (bz1 - o.z) * inv.z;
    if (t_min > t_max_z || t_max < t_min_z) return 0;
    return 1;
}

/* ------------------------------------------------------------ textures */

/* FilterTexCoord (include/bulbit/texture.h:26-46) */
static int32_t filter_coord(int32_t i, int32_t n, int32_t filter)
{
    if (filter == BB_TEXCOORD_CLAMP) return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
    return i >= 0 ? i % n : n - (-i % n) - 1;
}

/* Texture<T>::Evaluate: ConstantTexture, ImageTexture (bilinear,
 * textures.h:60-88), CheckerTexture (textures.h:109-130) */
static void texture_eval(const OrcScene* s, int32_t idx, v2 uv, float out[3])
{
    const BbTexture* t = &s->texs[idx];
    int c = t->channels;
    if (t->type == BB_TEX_CONSTANT) {
        for (int k = 0; k < 3; ++k) out[k] = c == 1 ? t->value[0] : t->value[k];
        return;
    }
    if (t->type == BB_TEX_CHECKER) {
        int32_t sx = (int32_t)(uv.x * t->resolution[0]);
        int32_t sy = (int32_t)(uv.y * t->resolution[1]);
        int32_t cs = sx + sy;
        texture_eval(s, (cs % 2) ? t->checker[0] - 1 : t->checker[1] - 1, uv, out);
        return;
    }
    float w = uv.x * (float)t->width - 0.5f;
    float h = uv.y * (float)t->height - 0.5f;
    int32_t i0 = (int32_t)floorf(w), i1 = i0 + 1;
    int32_t j0 = (int32_t)floorf(h), j1 = j0 + 1;
    i0 = filter_coord(i0, t->width, t->filter);
    j0 = filter_coord(j0, t->height, t->filter);
    i1 = filter_coord(i1, t->width, t->filter);
    j1 = filter_coord(j1, t->height, t->filter);
    float fu = w - floorf(w);
    float fv = h - floorf(h);
    const float* v00 = t->texels + (size_t)(i0 + j0 * t->width) * c;
    const float* v10 = t->texels + (size_t)(i1 + j0 * t->width) * c;
    const float* v01 = t->texels + (size_t)(i0 + j1 * t->width) * c;
    const float* v11 = t->texels + (size_t)(i1 + j1 * t->width) * c;
    for (int k = 0; k < c; ++k)
        out[k] = (1 - fu) * (1 - fv) * v00[k] + (1 - fu) * (fv) * v01[k] + (fu) * (1 - fv) * v10[k] + (fu) * (fv) * v11[k];
    for (int k = c; k < 3; ++k) out[k] = out[0];
}

static float texture1(const OrcScene* s, int32_t idx, v2 uv)
{
    float v[3];
    texture_eval(s, idx, uv, v);
    return v[0];
}

static spec texture3(const OrcScene* s, int32_t idx, v2 uv)
{
    float v[3];
    texture_eval(s, idx, uv, v);
    return S3(v[0], v[1], v[2]);
}

void orc_texture_eval(const OrcScene* s, int32_t idx, const float uv[2], float out[3])
{
    texture_eval(s, idx, V2(uv[0], uv[1]), out);
}

/* HashFloat(ray, point) (hash.h:111-126): MurmurHash64A over Ray{o, d} and the point */
static float hash_float_ray_point(v3 o, v3 d, v3 p)
{
    float buf[9] = { o.x, o.y, o.z, d.x, d.y, d.z, p.x, p.y, p.z };
    return (float)(uint32_t)orc_murmur64a(buf, 36, 0) * 0x1p-32f;
}

/* Shape::Intersect of a primitive's shape */
static inline int shape_intersect(const OrcScene* s, int32_t prim, v3 o, v3 d, float t_min, float t_max, float* t,
                                  float* u, float* v)
{
    const BbPrimitive* pr = &s->prims[prim];
    if (pr->shape_kind == BB_SHAPE_TRIANGLE) {
        const int32_t* vi = s->idx + 3 * pr->shape_index;
        return mt_intersect(vtx(s, s->pos, vi[0]), vtx(s, s->pos, vi[1]), vtx(s, s->pos, vi[2]), o, d, t_min, t_max,
                            t, u, v);
    }
    *u = 0.0f;
    *v = 0.0f;
    return sphere_intersect(&s->spheres[pr->shape_index], o, d, t_min, t_max, t);
}

/* Primitive::Intersect (include/bulbit/primitive.h:45-77): the shape test,
 * then the material's alpha test; a rejected hit re-intersects the same
 * primitive from the hit point along Normalize(d) over
 * [Ray::epsilon, t_max - t].  *o and *d become the ray of the accepted hit. */
static int prim_intersect(const OrcScene* s, int32_t prim, v3* o, v3* d, float t_min, float t_max, float* t, float* u,
                          float* v)
{
    if (!shape_intersect(s, prim, *o, *d, t_min, t_max, t, u, v)) return 0;
    const BbPrimitive* pr = &s->prims[prim];
    if (pr->material < 0) return 1;
    const BbMaterial* m = &s->mats[pr->material];
    if (m->tex_alpha == 0 || m->type == BB_MAT_DIELECTRIC) return 1;
    for (;;) {
        isect_t is;
        memset(&is, 0, sizeof(is));
        is.prim = prim;
        is.t = *t;
        is.u = *u;
        is.v = *v;
        finish_isect(s, *o, *d, &is);
        float alpha = texture1(s, m->tex_alpha - 1, is.uv);
        if (!(alpha < 1.0f)) return 1;
        float p = alpha <= 0.0f ? 1.0f : hash_float_ray_point(*o, *d, is.point);
        if (!(p > alpha)) return 1;
        float t_max2 = t_max - *t;
        *d = vnormalize(*d);
        *o = is.point;
        t_min = 1e-4f;
        t_max = t_max2;
        if (!shape_intersect(s, prim, *o, *d, t_min, t_max, t, u, v)) return 0;
    }
}

#define ORC_STACK 1024

/* include/bulbit/bvh.h:101-155 (RayCast) with the callbacks of
 * BVH::Intersect (bvh.cpp:299-332) and BVH::IntersectAny (bvh.cpp:334-361).
 * Returns 1 on hit; closest-hit fills prim/t/u/v. */
static int bvh_raycast(const OrcScene* s, v3 o, v3 d, float t_min, float t_max, int any_hit, isect_t* out,
                       BbTraceCounters* cnt)
{
    v3 inv = V3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int neg[3] = { inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f };
    int32_t stack[ORC_STACK];
    int sp = 0;
    stack[sp++] = 0;
    int hit = 0;
    while (sp > 0) {
        int32_t index = stack[--sp];
        const BbBvhNode* nd = &s->nodes[index];
        if (cnt) cnt->nodes++;
        if (!box_test(nd, o, t_min, t_max, inv, neg)) continue;
        if (nd->count > 0) {
            for (int32_t i = 0; i < nd->count; ++i) {
                int32_t prim = s->ordered[nd->offset + i];
                float t, u, v;
                v3 po = o, pd = d;
                if (cnt) cnt->prims++;
                if (prim_intersect(s, prim, &po, &pd, t_min, t_max, &t, &u, &v)) {
                    if (cnt) cnt->accepts++;
                    hit = 1;
                    if (any_hit) return 1;
                    out->prim = prim;
                    out->t = t;
                    out->u = u;
                    out->v = v;
                    out->ro = po;
                    out->rd = pd;
                    if (t <= t_min) return 1;
                    t_max = t;
                }
            }
        } else {
            if (sp + 2 > ORC_STACK) {
                fprintf(stderr, "oracle: traversal stack overflow\n");
                abort();
            }
            int32_t child1 = index + 1, child2 = nd->offset;
            if (neg[nd->axis]) {
                stack[sp++] = child1;
                stack[sp++] = child2;
            } else {
                stack[sp++] = child2;
                stack[sp++] = child1;
            }
        }
    }
    return hit;
}

/* ============================================================ complex math */

typedef struct { float re, im; } cpx;
static inline cpx C2(float re, float im) { cpx c = { re, im }; return c; }
static inline cpx cmul(cpx a, cpx b) { return C2(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cpx cmulf(cpx a, float f) { return C2(a.re * f, a.im * f); }
/* std::complex<float> division as g++ 11 emits it: naive formula in double */
static inline cpx cdiv(cpx a, cpx b)
{
    double ar = a.re, ai = a.im, br = b.re, bi = b.im;
    double den = br * br + bi * bi;
    return C2((float)((ar * br + ai * bi) / den), (float)((ai * br - ar * bi) / den));
}
static inline float cabs_(cpx z) { return hypotf(z.re, z.im); }
static inline float cnorm(cpx z) { return z.re * z.re + z.im * z.im; }

/* include/bulbit/scattering.h:81-100 (SqrtComplex) */
static cpx csqrt_(cpx z)
{
    float n = cabs_(z);
    float t1 = sqrtf(0.5f * (n + fabsf(z.re)));
    float t2 = 0.5f * z.im / t1;
    if (n == 0.0f) return C2(0.0f, 0.0f);
    if (z.re >= 0.0f) return C2(t1, t2);
    return C2(fabsf(t2), copysignf(t1, z.im));
}

/* scattering.h:102-121 (FresnelComplex) */
static float fresnel_complex(float cos_theta_i, cpx eta)
{
    cos_theta_i = clampf_(cos_theta_i, 0.0f, 1.0f);
    float sin2_theta_i = 1.0f - sqrf(cos_theta_i);
    cpx sin2_theta_t = cdiv(C2(sin2_theta_i, 0.0f), cmul(eta, eta));
    cpx cos_theta_t = csqrt_(C2(1.0f - sin2_theta_t.re, 0.0f - sin2_theta_t.im));
    cpx ec = cmulf(eta, cos_theta_i);
    cpx r_parl = cdiv(C2(ec.re - cos_theta_t.re, ec.im - cos_theta_t.im), C2(ec.re + cos_theta_t.re, ec.im + cos_theta_t.im));
    cpx et = cmul(eta, cos_theta_t);
    cpx r_perp = cdiv(C2(cos_theta_i - et.re, -et.im), C2(cos_theta_i + et.re, et.im));
    return (cnorm(r_parl) + cnorm(r_perp)) / 2.0f;
}

/* scattering.h:123-132 */
static spec fresnel_complex_spec(float cos_theta_i, spec eta, spec k)
{
    return S3(fresnel_complex(cos_theta_i, C2(eta.r, k.r)), fresnel_complex(cos_theta_i, C2(eta.g, k.g)),
              fresnel_complex(cos_theta_i, C2(eta.b, k.b)));
}

/* scattering.h:51-77 */
static float fresnel_dielectric(float cos_theta_i, float eta)
{
    cos_theta_i = clampf_(cos_theta_i, -1.0f, 1.0f);
    if (cos_theta_i < 0.0f) {
        eta = 1.0f / eta;
        cos_theta_i = -cos_theta_i;
    }
    float sin2_theta_i = 1.0f - sqrf(cos_theta_i);
    float sin2_theta_t = sin2_theta_i / sqrf(eta);
    if (sin2_theta_t >= 1.0f) return 1.0f;
    float cos_theta_t = sqrtf(smax(0.0f, 1.0f - sin2_theta_t));
    float r_parl = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    float r_perp = (cos_theta_i - eta * cos_theta_t) / (cos_theta_i + eta * cos_theta_t);
    return (sqrf(r_parl) + sqrf(r_perp)) / 2.0f;
}

float orc_fresnel_dielectric(float c, float eta) { return fresnel_dielectric(c, eta); }
float orc_fresnel_complex(float c, float eta, float k) { return fresnel_complex(c, C2(eta, k)); }

/* scattering.h:9-39 */
static int refract(v3* wt, v3 wi, v3 n, float eta, float* eta_p)
{
    float cos_theta_i = vdot(n, wi);
    if (cos_theta_i < 0.0f) {
        eta = 1.0f / eta;
        cos_theta_i = -cos_theta_i;
        n = vneg(n);
    }
    float sin2_theta_i = smax(0.0f, 1.0f - sqrf(cos_theta_i));
    float sin2_theta_t = sin2_theta_i / sqrf(eta);
    if (sin2_theta_t >= 1.0f) return 0;
    float cos_theta_t = sqrtf(1.0f - sin2_theta_t);
    float inv_eta = 1.0f / eta;
    *wt = vadd(vmul(vneg(wi), inv_eta), vmul(n, cos_theta_i * inv_eta - cos_theta_t));
    if (eta_p) *eta_p = eta;
    return 1;
}

/* math.h:109-113: Reflect(v, n) = -v + 2 * Dot(v, n) * n */
static inline v3 reflect(v3 v, v3 n) { return vadd(vneg(v), vmul(n, 2.0f * vdot(v, n))); }

/* ======================================================= microfacet (GGX) */

typedef struct { float ax, ay; } ggx;

/* include/bulbit/microfacet.h:27-30 */
static inline int ggx_smooth(ggx m) { return smax(m.ax, m.ay) < 1e-3f; }

/* microfacet.h:37-54 */
static float ggx_D(ggx m, v3 wm)
{
    float tan2 = tan2_theta(wm);
    if (isinf(tan2)) return 0.0f;
    float cos4 = sqrf(cos2_theta(wm));
    if (cos4 < 1e-16f) return 0.0f;
    float e = tan2 * (sqrf(cos_phi(wm) / m.ax) + sqrf(sin_phi(wm) / m.ay));
    return 1.0f / (orc_pi * m.ax * m.ay * cos4 * sqrf(1.0f + e));
}

/* microfacet.h:76-86 */
static float ggx_lambda(ggx m, v3 w)
{
    float tan2 = tan2_theta(w);
    if (isinf(tan2)) return 0.0f;
    float alpha2 = sqrf(cos_phi(w) * m.ax) + sqrf(sin_phi(w) * m.ay);
    return (sqrtf(1.0f + alpha2 * tan2) - 1.0f) / 2.0f;
}

static inline float ggx_G1(ggx m, v3 w) { return 1.0f / (1.0f + ggx_lambda(m, w)); }
static inline float ggx_G(ggx m, v3 wo, v3 wi) { return 1.0f / (1.0f + ggx_lambda(m, wo) + ggx_lambda(m, wi)); }
/* microfacet.h:56-64: D(w, wm) = G1(w) / |cos w| * D(wm) * |w.wm| */
static inline float ggx_pdf(ggx m, v3 w, v3 wm)
{
    return ggx_G1(m, w) / abs_cos_theta(w) * ggx_D(m, wm) * vabsdot(w, wm);
}

/* sampling.h:203-232 (Dupuy & Benyoub spherical-cap VNDF sampling) with the
 * explicit fused multiply-add of sampling.h:210 */
static v3 ggx_sample_wm(ggx m, v3 wo, v2 u)
{
    v3 wo_std = vnormalize(V3(wo.x * m.ax, wo.y * m.ay, wo.z));
    float phi = orc_two_pi * u.x;
    float z = fmaf((1.0f - u.y), (1.0f + wo_std.z), -wo_std.z);
    float sin_t = sqrtf(clampf_(1.0f - z * z, 0.0f, 1.0f));
    float x = sin_t * o_cosf(phi);
    float y = sin_t * o_sinf(phi);
    v3 h = vadd(V3(x, y, z), wo_std);
    return vnormalize(V3(h.x * m.ax, h.y * m.ay, h.z));
}

/* ===================================================== reflectance tables */

/* PermutationElement (hash.h:128-159): Kensler's permutation of [0, l) */
static int32_t permutation_element(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (int32_t)((i + p) % l);
}

int32_t orc_permutation_element(uint32_t i, uint32_t l, uint32_t p) { return permutation_element(i, l, p); }

/* PrepareReflectanceTextures (microfacet.cpp:239-270) with texture_size 32
 * and 1024 samples (integrator.cpp:149-151) */
#define ORC_TABLE_N 32
#define ORC_TABLE_SAMPLES 1024

typedef struct {
    float u0[ORC_TABLE_SAMPLES];
    v2 u12[ORC_TABLE_SAMPLES];
    float E[ORC_TABLE_N * ORC_TABLE_N]; /* TrowbridgeReitz E(cos_theta = i, alpha = j) at [j * 32 + i] */
    float E_avg[ORC_TABLE_N];
} refl_tables_t;

static refl_tables_t g_tables;
static pthread_once_t g_tables_once = PTHREAD_ONCE_INIT;

/* TrowbridgeReitzDistribution::rho (microfacet.h:109-131) */
static float ggx_rho(ggx m, v3 wo, const v2* u, int32_t n)
{
    float r = 0.0f;
    for (int32_t i = 0; i < n; ++i) {
        v3 wm = ggx_sample_wm(m, wo, u[i]);
        v3 wi = reflect(wo, wm);
        if (!same_hemisphere(wo, wi)) continue;
        float pdf = ggx_pdf(m, wo, wm) / (4.0f * vabsdot(wo, wm));
        float f = ggx_D(m, wm) * ggx_G(m, wo, wi) / (4.0f * abs_cos_theta(wo) * abs_cos_theta(wi));
        r += f * abs_cos_theta(wi) / pdf;
    }
    return r / (float)n;
}

static void tables_build(void)
{
    refl_tables_t* t = &g_tables;
    const int32_t samples = ORC_TABLE_SAMPLES, xs = 32, ys = 32;
    pcg32 rng;
    pcg_seed(&rng, 3, 1); /* RNG rng(3) */
    for (int32_t i = 0; i < samples; ++i) {
        int32_t stratum = permutation_element((uint32_t)i, (uint32_t)samples, 1u);
        t->u0[i] = ((float)stratum + pcg_float(&rng)) / (float)samples;
        stratum = permutation_element((uint32_t)i, (uint32_t)samples, 2u);
        int32_t x = stratum % xs, y = stratum / xs;
        float ux = ((float)x + pcg_float(&rng)) / (float)xs;
        float uy = ((float)y + pcg_float(&rng)) / (float)ys;
        t->u12[i] = V2(ux, uy);
    }
    /* TrowbridgeReitzDistribution::PrepareReflectanceTexture (microfacet.cpp:31-75) */
    const int32_t n = ORC_TABLE_N;
    const float d = 1.0f / (float)n;
    for (int32_t j = 0; j < n; ++j) {
        float a = d / 2.0f + d * (float)j;
        float r_sum = 0.0f;
        for (int32_t i = 0; i < n; ++i) {
            float cos_theta = d / 2.0f + d * (float)i;
            float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
            v3 wo = V3(sin_theta, 0.0f, cos_theta);
            ggx m = { a, a };
            float r = ggx_rho(m, wo, t->u12, samples);
            r_sum += r * cos_theta * d;
            t->E[j * n + i] = r;
        }
        t->E_avg[j] = 2.0f * r_sum;
    }
}

static const refl_tables_t* tables(void)
{
    pthread_once(&g_tables_once, tables_build);
    return &g_tables;
}

void orc_reflectance_tables(float* E, float* E_avg, float* u0, float* u12)
{
    const refl_tables_t* t = tables();
    if (E) memcpy(E, t->E, sizeof(t->E));
    if (E_avg) memcpy(E_avg, t->E_avg, sizeof(t->E_avg));
    if (u0) memcpy(u0, t->u0, sizeof(t->u0));
    if (u12)
        for (int32_t i = 0; i < ORC_TABLE_SAMPLES; ++i) {
            u12[2 * i] = t->u12[i].x;
            u12[2 * i + 1] = t->u12[i].y;
        }
}

/* ImageTexture<Float>::Evaluate, bilinear with TexCoordFilter::clamp
 * (textures.h:60-88, texture.h:26-46) */
static float tex_bilinear(const float* img, int32_t w, int32_t h, float u, float v)
{
    float fw = u * (float)w - 0.5f;
    float fh = v * (float)h - 0.5f;
    int32_t i0 = (int32_t)floorf(fw), i1 = i0 + 1;
    int32_t j0 = (int32_t)floorf(fh), j1 = j0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > w - 1 ? w - 1 : i0);
    j0 = j0 < 0 ? 0 : (j0 > h - 1 ? h - 1 : j0);
    i1 = i1 < 0 ? 0 : (i1 > w - 1 ? w - 1 : i1);
    j1 = j1 < 0 ? 0 : (j1 > h - 1 ? h - 1 : j1);
    float fu = fw - floorf(fw);
    float fv = fh - floorf(fh);
    float v00 = img[i0 + j0 * w], v10 = img[i1 + j0 * w];
    float v01 = img[i0 + j1 * w], v11 = img[i1 + j1 * w];
    return (1.0f - fu) * (1.0f - fv) * v00 + (1.0f - fu) * fv * v01 + fu * (1.0f - fv) * v10 + fu * fv * v11;
}

/* TrowbridgeReitzDistribution::E / E_avg (microfacet.h:133-145), GetMeanAlpha (:157-160) */
static float ggx_E(ggx m, v3 wo)
{
    return tex_bilinear(tables()->E, ORC_TABLE_N, ORC_TABLE_N, wo.z, sqrtf(m.ax * m.ay));
}
static float ggx_E_avg(ggx m)
{
    return tex_bilinear(tables()->E_avg, ORC_TABLE_N, 1, sqrtf(m.ax * m.ay), 0.0f);
}

float orc_ggx_E(float ax, float ay, const float wo[3])
{
    ggx m = { ax, ay };
    return ggx_E(m, V3(wo[0], wo[1], wo[2]));
}

float orc_ggx_E_avg(float ax, float ay)
{
    ggx m = { ax, ay };
    return ggx_E_avg(m);
}

/* ================================================================= BxDFs */

enum { K_LAMBERT = 0, K_DIELECTRIC = 1, K_CONDUCTOR = 2, K_MR = 3, K_CONDUCTOR_MS = 4, K_DIELECTRIC_MS = 5, K_EON = 6, K_MIRROR = 7 };

typedef struct {
    int kind;
    spec r;      /* Lambert reflectance; dielectric Sqrt(reflectance); conductor lobe reflectance */
    float eta;   /* dielectric (already eta_p) */
    spec ceta, ck;
    spec color;
    float metallic;
    ggx mf;
} bxdf_t;

typedef struct {
    spec f;
    v3 wi;
    float pdf;
    int flags;
    float eta;
    int stochastic;
} bsample;

static int bxdf_flags(const bxdf_t* b)
{
    switch (b->kind) {
    case K_LAMBERT:
    case K_EON: return !sblack(b->r) ? (BX_DIFFUSE | BX_REFLECTION) : 0;
    case K_MIRROR: return !sblack(b->r) ? (BX_SPECULAR | BX_REFLECTION) : 0; /* bxdfs.h:55-58 */
    case K_DIELECTRIC:
    case K_DIELECTRIC_MS: {
        int f = (b->eta == 1.0f) ? BX_TRANSMISSION : (BX_REFLECTION | BX_TRANSMISSION);
        return f | (ggx_smooth(b->mf) ? BX_SPECULAR : BX_GLOSSY);
    }
    case K_CONDUCTOR:
    case K_CONDUCTOR_MS: return ggx_smooth(b->mf) ? (BX_SPECULAR | BX_REFLECTION) : (BX_GLOSSY | BX_REFLECTION);
    default: return BX_DIFFUSE | BX_GLOSSY | BX_REFLECTION;
    }
}

/* ---- Lambertian: src/bxdf/lambertian_bxdf.cpp:8-53 ---- */
static spec lambert_f(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi)) return S1(0.0f);
    return smulf(b->r, orc_inv_pi);
}
static float lambert_pdf(v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi)) return 0.0f;
    return abs_cos_theta(wi) * orc_inv_pi;
}
static int lambert_sample(const bxdf_t* b, v3 wo, v2 u12, bsample* s)
{
    v3 wi = sample_cosine_hemisphere(u12);
    float pdf = wi.z * orc_inv_pi;
    if (wo.z < 0.0f) wi.z = -wi.z;
    s->f = smulf(b->r, orc_inv_pi);
    s->wi = wi;
    s->pdf = pdf;
    s->flags = BX_DIFFUSE | BX_REFLECTION;
    s->eta = 1.0f;
    s->stochastic = 0;
    return 1;
}

/* ---- EON: src/bxdf/eon_bxdf.cpp (Portsmouth et al., JCGT 14(1), 2025),
 * exact variant; b->r = rho, b->mf.ax = roughness ---- */
static const float fon1 = 0.5f - 2.0f / (3.0f * ORC_PI);
static const float fon2 = 2.0f / 3.0f - 28.0f / (15.0f * ORC_PI);

static float eon_E(float mu, float r)
{
    float AF = 1.0f / (1.0f + fon1 * r);
    float BF = r * AF;
    float Si = sqrtf(1.0f - (mu * mu));
    float G = Si * (o_acosf(mu) - Si * mu) + (2.0f / 3.0f) * ((Si / mu) * (1.0f - (Si * Si * Si)) - Si);
    return AF + (BF * orc_inv_pi) * G;
}

static void eon_ltc(float mu, float r, float* a, float* b, float* c, float* d)
{
    *a = 1.0f + r * (0.303392f + (-0.518982f + 0.111709f * mu) * mu + (-0.276266f + 0.335918f * mu) * r);
    *b = r * (-1.16407f + 1.15859f * mu + (0.150815f - 0.150105f * mu) * r) / (mu * mu * mu - 1.43545f);
    *c = 1.0f + (0.20013f + (-0.506373f + 0.261777f * mu) * mu) * r;
    *d = ((0.540852f + (-1.01625f + 0.475392f * mu) * mu) * r) / (-1.0743f + mu * (0.0725628f + mu));
}

/* orthonormal_basis_ltc: columns X, Y = (-X.y, X.x, 0), Z = (0, 0, 1) */
static void eon_basis(v3 w, v3* X, v3* Y)
{
    float len2 = w.x * w.x + w.y * w.y;
    *X = len2 > 0.0f ? V3(w.x / sqrtf(len2), w.y / sqrtf(len2), 0.0f / sqrtf(len2)) : V3(1.0f, 0.0f, 0.0f);
    *Y = V3(-X->y, X->x, 0.0f);
}

static float eon_cltc_pdf(v3 wo, v3 wl, float r)
{
    v3 X, Y;
    eon_basis(wo, &X, &Y);
    v3 wi = V3(X.x * wl.x + X.y * wl.y + X.z * wl.z, Y.x * wl.x + Y.y * wl.y + Y.z * wl.z,
               0.0f * wl.x + 0.0f * wl.y + 1.0f * wl.z);
    float a, b, c, d;
    eon_ltc(wo.z, r, &a, &b, &c, &d);
    float detM = c * (a - b * d);
    v3 wh = V3(c * (wi.x - b * wi.z), (a - b * d) * wi.y, -c * (d * wi.x - a * wi.z));
    float len2 = vdot(wh, wh);
    float vz = 1.0f / sqrtf(d * d + 1.0f);
    float s = 0.5f * (1.0f + vz);
    return detM * detM / (len2 * len2) * smax(wh.z, 0.0f) / (ORC_PI * s);
}

static v3 eon_cltc_sample(v3 wo, float r, float u1, float u2, float* pdf)
{
    float a, b, c, d;
    eon_ltc(wo.z, r, &a, &b, &c, &d);
    float R = sqrtf(u1);
    float phi = orc_two_pi * u2;
    float x = R * o_cosf(phi);
    float y = R * o_sinf(phi);
    float vz = 1.0f / sqrtf(d * d + 1.0f);
    float s = 0.5f * (1.0f + vz);
    x = -(sqrtf(1.0f - y * y) * (1.0f - s) + x * s);
    v3 wh = V3(x, y, sqrtf(smax(1.0f - (x * x + y * y), 0.0f)));
    float pdf_wh = wh.z / (ORC_PI * s);
    v3 wi = V3(a * wh.x + b * wh.z, c * wh.y, d * wh.x + wh.z);
    float len = vlen(wi);
    float detM = c * (a - b * d);
    *pdf = pdf_wh * len * len * len / detM;
    v3 X, Y;
    eon_basis(wo, &X, &Y);
    return vnormalize(V3(X.x * wi.x + Y.x * wi.y + 0.0f * wi.z, X.y * wi.x + Y.y * wi.y + 0.0f * wi.z,
                         X.z * wi.x + Y.z * wi.y + 1.0f * wi.z));
}

static spec eon_f(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi)) return S1(0.0f);
    float r = b->mf.ax;
    float mu_i = wi.z, mu_o = wo.z;
    float s = vdot(wi, wo) - mu_i * mu_o;
    float sovertF = s > 0.0f ? s / smax(mu_i, mu_o) : s;
    float AF = 1.0f / (1.0f + fon1 * r);
    spec f_ss = smulf(smulf(smulf(b->r, orc_inv_pi), AF), 1.0f + r * sovertF);
    float EFo = eon_E(mu_o, r), EFi = eon_E(mu_i, r);
    float avgEF = AF * (1.0f + fon2 * r);
    spec rho_ms = sdiv(smulf(smul(b->r, b->r), avgEF), ssub(S1(1.0f), smulf(b->r, 1.0f - avgEF)));
    const float eps = 1.0e-7f;
    spec f_ms = sdivf(smulf(smulf(smulf(rho_ms, orc_inv_pi), smax(eps, 1.0f - EFo)), smax(eps, 1.0f - EFi)),
                      smax(eps, 1.0f - avgEF));
    return sadd(f_ss, f_ms);
}

static float eon_P_u(float mu, float r)
{
    return o_powf(r, 0.1f) * (0.162925f + mu * (-0.372058f + (0.538233f - 0.290822f * mu) * mu));
}

static float eon_pdf(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi)) return 0.0f;
    float P_u = eon_P_u(wo.z, b->mf.ax);
    float P_c = 1.0f - P_u;
    return P_u * orc_inv_two_pi + P_c * eon_cltc_pdf(wo, wi, b->mf.ax);
}

static int eon_sample(const bxdf_t* b, v3 wo, float u0, v2 u12, bsample* out)
{
    float r = b->mf.ax;
    float P_u = eon_P_u(wo.z, r);
    float P_c = 1.0f - P_u;
    v3 wi;
    float pdf_c;
    if (u0 <= P_u) {
        /* SampleUniformHemisphere (sampling.h:50-57) */
        float z = u12.x;
        float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float phi = orc_two_pi * u12.y;
        wi = V3(rr * o_cosf(phi), rr * o_sinf(phi), z);
        pdf_c = eon_cltc_pdf(wo, wi, r);
    } else {
        wi = eon_cltc_sample(wo, r, u12.x, u12.y, &pdf_c);
    }
    out->pdf = P_u * orc_inv_two_pi + P_c * pdf_c;
    out->f = eon_f(b, wo, wi);
    out->wi = wi;
    out->flags = BX_DIFFUSE | BX_REFLECTION;
    out->eta = 1.0f;
    out->stochastic = 0;
    return 1;
}

/* ---- Conductor: src/bxdf/conductor_bxdf.cpp:7-112 ---- */
static spec conductor_f(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b->mf)) return S1(0.0f);
    float cos_o = abs_cos_theta(wo), cos_i = abs_cos_theta(wi);
    if (cos_i == 0.0f || cos_o == 0.0f) return S1(0.0f);
    v3 wm = vadd(wo, wi);
    if (vlen2(wm) == 0.0f) return S1(0.0f);
    vnormalize_m(&wm);
    spec F = fresnel_complex_spec(vabsdot(wo, wm), b->ceta, b->ck);
    spec r = smulf(smulf(smul(b->r, F), ggx_D(b->mf, wm)), ggx_G(b->mf, wo, wi));
    return sdivf(r, 4.0f * cos_i * cos_o);
}
static float conductor_pdf(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b->mf)) return 0.0f;
    v3 wm = vadd(wo, wi);
    if (vlen2(wm) == 0.0f) return 0.0f;
    vnormalize_m(&wm);
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    return ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
}
static int conductor_sample(const bxdf_t* b, v3 wo, v2 u12, bsample* s)
{
    if (ggx_smooth(b->mf)) {
        v3 wi = V3(-wo.x, -wo.y, wo.z);
        spec f = sdivf(fresnel_complex_spec(abs_cos_theta(wi), b->ceta, b->ck), abs_cos_theta(wi));
        s->f = smul(b->r, f);
        s->wi = wi;
        s->pdf = 1.0f;
        s->flags = BX_SPECULAR | BX_REFLECTION;
        s->eta = 1.0f;
        s->stochastic = 0;
        return 1;
    }
    if (wo.z == 0.0f) return 0;
    v3 wm = ggx_sample_wm(b->mf, wo, u12);
    v3 wi = reflect(wo, wm);
    if (!same_hemisphere(wo, wi)) return 0;
    float cos_o = abs_cos_theta(wo), cos_i = abs_cos_theta(wi);
    if (cos_i == 0.0f || cos_o == 0.0f) return 0;
    float pdf = ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
    spec F = fresnel_complex_spec(vabsdot(wo, wm), b->ceta, b->ck);
    spec f = sdivf(smulf(smulf(F, ggx_D(b->mf, wm)), ggx_G(b->mf, wo, wi)), 4.0f * cos_i * cos_o);
    s->f = smul(b->r, f);
    s->wi = wi;
    s->pdf = pdf;
    s->flags = BX_GLOSSY | BX_REFLECTION;
    s->eta = 1.0f;
    s->stochastic = 0;
    return 1;
}

/* ---- Conductor with multiple-scattering compensation:
 *      src/bxdf/conductor_multi_scattering_bxdf.cpp:7-148 ---- */

/* FresnelConductorAverage (scattering.h:174-185) */
static spec fresnel_conductor_average(spec eta, spec k)
{
    spec F0 = fresnel_complex_spec(1.0f, eta, k);
    spec F90 = fresnel_complex_spec(0.0f, eta, k);
    return sdivf(sadd(smulf(F0, 20.0f), F90), 21.0f);
}

/* f_ms of the compensated conductor: fresnel_ms (1 - E(wo)) (1 - E(wi)) / (pi (1 - E_avg)) */
static spec conductor_ms_lobe(const bxdf_t* b, v3 wo, v3 wi)
{
    float E_avg = ggx_E_avg(b->mf);
    spec fa = fresnel_conductor_average(b->ceta, b->ck);
    spec fresnel_ms = sdiv(smulf(smul(fa, fa), E_avg), ssub(S1(1.0f), smulf(fa, 1.0f - E_avg)));
    return sdivf(smulf(fresnel_ms, (1.0f - ggx_E(b->mf, wo)) * (1.0f - ggx_E(b->mf, wi))), orc_pi * (1.0f - E_avg));
}

static spec conductor_ms_f(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b->mf)) return S1(0.0f);
    float cos_o = abs_cos_theta(wo), cos_i = abs_cos_theta(wi);
    if (cos_i == 0.0f || cos_o == 0.0f) return S1(0.0f);
    v3 wm = vadd(wo, wi);
    if (vlen2(wm) == 0.0f) return S1(0.0f);
    vnormalize_m(&wm);
    spec F = fresnel_complex_spec(vabsdot(wo, wm), b->ceta, b->ck);
    spec f_ss = sdivf(smulf(smulf(F, ggx_D(b->mf, wm)), ggx_G(b->mf, wo, wi)), 4.0f * cos_i * cos_o);
    return smul(b->r, sadd(f_ss, conductor_ms_lobe(b, wo, wi)));
}

static float conductor_ms_pdf(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b->mf)) return 0.0f;
    v3 wm = vadd(wo, wi);
    if (vlen2(wm) == 0.0f) return 0.0f;
    vnormalize_m(&wm);
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    float pdf_ss = ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
    float pdf_ms = wi.z * orc_inv_pi; /* CosineHemispherePDF(wi.z) */
    float e = ggx_E(b->mf, wo);
    return pdf_ms * (1.0f - e) + pdf_ss * e; /* Lerp(pdf_ms, pdf_ss, E(wo)), math.h:87-90 */
}

static int conductor_ms_sample(const bxdf_t* b, v3 wo, float u0, v2 u12, bsample* s)
{
    if (ggx_smooth(b->mf)) return conductor_sample(b, wo, u12, s);
    if (wo.z == 0.0f) return 0;
    v3 wm, wi;
    int flags;
    float e = ggx_E(b->mf, wo);
    if (u0 < e) {
        wm = ggx_sample_wm(b->mf, wo, u12);
        wi = reflect(wo, wm);
        flags = BX_GLOSSY | BX_REFLECTION;
    } else {
        wi = sample_cosine_hemisphere(u12);
        wm = vnormalize(vadd(wo, wi));
        flags = BX_DIFFUSE | BX_REFLECTION;
    }
    if (!same_hemisphere(wo, wi)) return 0;
    float cos_o = abs_cos_theta(wo), cos_i = abs_cos_theta(wi);
    if (cos_i == 0.0f || cos_o == 0.0f) return 0;
    spec F = fresnel_complex_spec(vabsdot(wo, wm), b->ceta, b->ck);
    spec f_ss = sdivf(smulf(smulf(F, ggx_D(b->mf, wm)), ggx_G(b->mf, wo, wi)), 4.0f * cos_i * cos_o);
    spec f_ms = conductor_ms_lobe(b, wo, wi);
    float pdf_ss = ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
    float pdf_ms = wi.z * orc_inv_pi;
    s->f = smul(b->r, sadd(f_ss, f_ms));
    s->wi = wi;
    s->pdf = pdf_ms * (1.0f - e) + pdf_ss * e;
    s->flags = flags;
    s->eta = 1.0f;
    s->stochastic = 0;
    return 1;
}

/* ---- Dielectric: src/bxdf/dielectric_bxdf.cpp:9-254 ---- */
static spec dielectric_f(const bxdf_t* b, v3 wo, v3 wi)
{
    if (b->eta == 1.0f || ggx_smooth(b->mf)) return S1(0.0f);
    float cos_o = wo.z, cos_i = wi.z;
    float eta_p = 1.0f;
    int refl = cos_i * cos_o > 0.0f;
    if (!refl) eta_p = cos_o > 0.0f ? b->eta : (1.0f / b->eta);
    v3 wm = vadd(vmul(wi, eta_p), wo);
    if (cos_i == 0.0f || cos_o == 0.0f || vlen2(wm) == 0.0f) return S1(0.0f);
    vnormalize_m(&wm);
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    if (vdot(wm, wi) * cos_i < 0.0f || vdot(wm, wo) * cos_o < 0.0f) return S1(0.0f);
    float F = fresnel_dielectric(vdot(wo, wm), b->eta);
    if (refl) {
        return S1(F * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) / fabsf(4.0f * cos_i * cos_o));
    }
    float denom = sqrf(vdot(wi, wm) + vdot(wo, wm) / eta_p) * cos_i * cos_o;
    float ft = (1.0f - F) * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) * fabsf(vdot(wi, wm) * vdot(wo, wm) / denom);
    ft /= sqrf(eta_p);
    return smulf(b->r, ft);
}
static float dielectric_pdf(const bxdf_t* b, v3 wo, v3 wi)
{
    if (b->eta == 1.0f || ggx_smooth(b->mf)) return 0.0f;
    float cos_o = wo.z, cos_i = wi.z;
    float eta_p = 1.0f;
    int refl = cos_i * cos_o > 0.0f;
    if (!refl) eta_p = cos_o > 0.0f ? b->eta : (1.0f / b->eta);
    v3 wm = vadd(vmul(wi, eta_p), wo);
    if (cos_i == 0.0f || cos_o == 0.0f || vlen2(wm) == 0.0f) return 0.0f;
    vnormalize_m(&wm);
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    if (vdot(wm, wi) * cos_i < 0.0f || vdot(wm, wo) * cos_o < 0.0f) return 0.0f;
    float R = fresnel_dielectric(vdot(wo, wm), b->eta);
    float T = 1.0f - R;
    float pr = R, pt = T;
    if (pr == 0.0f && pt == 0.0f) return 0.0f;
    if (refl) return ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm)) * pr / (pr + pt);
    float dwm_dwi = vabsdot(wi, wm) / sqrf(vdot(wi, wm) + vdot(wo, wm) / eta_p);
    return ggx_pdf(b->mf, wo, wm) * dwm_dwi * pt / (pr + pt);
}
/* to_light: TransportDirection::ToLight (radiance, the path tracer's
 * direction); 0 = ToCamera (the reflectance tables) */
static int dielectric_sample(const bxdf_t* b, v3 wo, float u0, v2 u12, int to_light, bsample* s)
{
    s->stochastic = 0;
    if (b->eta == 1.0f || ggx_smooth(b->mf)) {
        float R = fresnel_dielectric(wo.z, b->eta);
        float T = 1.0f - R;
        float pr = R, pt = T;
        if (pr == 0.0f && pt == 0.0f) return 0;
        if (u0 < pr / (pr + pt)) {
            v3 wi = V3(-wo.x, -wo.y, wo.z);
            s->f = S1(R / abs_cos_theta(wi));
            s->wi = wi;
            s->pdf = pr / (pr + pt);
            s->flags = BX_SPECULAR | BX_REFLECTION;
            s->eta = 1.0f;
        } else {
            v3 wi;
            float eta_p;
            if (!refract(&wi, wo, V3(0.0f, 0.0f, 1.0f), b->eta, &eta_p)) return 0;
            float ft = T / abs_cos_theta(wi);
            if (to_light) ft /= sqrf(eta_p);
            s->f = smulf(b->r, ft);
            s->wi = wi;
            s->pdf = pt / (pr + pt);
            s->flags = BX_SPECULAR | BX_TRANSMISSION;
            s->eta = eta_p;
        }
        return 1;
    }
    v3 wm = ggx_sample_wm(b->mf, wo, u12);
    float R = fresnel_dielectric(vdot(wo, wm), b->eta);
    float T = 1.0f - R;
    float pr = R, pt = T;
    if (pr == 0.0f && pt == 0.0f) return 0;
    if (u0 < pr / (pr + pt)) {
        v3 wi = reflect(wo, wm);
        if (!same_hemisphere(wo, wi)) return 0;
        float pdf = ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm)) * pr / (pr + pt);
        s->f = S1(R * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) / (4.0f * wi.z * wo.z));
        s->wi = wi;
        s->pdf = pdf;
        s->flags = BX_GLOSSY | BX_REFLECTION;
        s->eta = 1.0f;
    } else {
        float eta_p = 0.0f;
        v3 wi = V3(0.0f, 0.0f, 0.0f);
        int tir = !refract(&wi, wo, wm, b->eta, &eta_p);
        if (same_hemisphere(wo, wi) || wi.z == 0.0f || tir) return 0;
        float denom = sqrf(vdot(wi, wm) + vdot(wo, wm) / eta_p);
        float dwm_dwi = vabsdot(wi, wm) / denom;
        float pdf = ggx_pdf(b->mf, wo, wm) * dwm_dwi * pt / (pr + pt);
        float ft = T * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) *
                   fabsf(vdot(wi, wm) * vdot(wo, wm) / (wi.z * wo.z * denom));
        if (to_light) ft /= sqrf(eta_p);
        s->f = smulf(b->r, ft);
        s->wi = wi;
        s->pdf = pdf;
        s->flags = BX_GLOSSY | BX_TRANSMISSION;
        s->eta = eta_p;
    }
    return 1;
}

/* ---- Dielectric with multiple-scattering compensation:
 *      src/bxdf/dielectric_multi_scattering_bxdf.cpp:9-438, include/bulbit/bxdfs.h:145-247 ---- */

/* DielectricMultiScatteringBxDF::PrepareReflectanceTexture (microfacet.cpp:173-236):
 * E and E_inv at (f0 = x, cos_theta = y, alpha = z), [(z * 32 + y) * 32 + x]
 * (image3d.h:54-62), E_avg and E_inv_avg at (f0 = x, alpha = y), [y * 32 + x]
 * (image.h:55-63); E_inv is the table of the inverted interface (1 / ior) */
typedef struct {
    float E[ORC_TABLE_N * ORC_TABLE_N * ORC_TABLE_N];
    float E_inv[ORC_TABLE_N * ORC_TABLE_N * ORC_TABLE_N];
    float E_avg[ORC_TABLE_N * ORC_TABLE_N];
    float E_inv_avg[ORC_TABLE_N * ORC_TABLE_N];
} dms_tables_t;

static dms_tables_t g_dms;
static pthread_once_t g_dms_once = PTHREAD_ONCE_INIT;

/* F0toIOR / MapF0toIOR / MapIORtoF0 (scattering.h:188-208) */
static float f0_to_ior(float f0)
{
    float sqrt_f0 = sqrtf(f0);
    return (1.0f + sqrt_f0) / (1.0f - sqrt_f0);
}
static float map_f0_to_ior(float f) { return f0_to_ior(sqrf(sqrf(f))); }
static float map_ior_to_f0(float ior) { return sqrtf((ior - 1.0f) / (ior + 1.0f)); }

/* BxDF::rho(wo, uc, u, ToCamera) (bxdf.cpp:8-21) of DielectricBxDF(ior, GGX(a, a), 1),
 * then Spectrum::Average (spectrum.h:59-71) of its three equal channels */
static float dielectric_rho(float ior, float a, v3 wo, const refl_tables_t* t)
{
    bxdf_t b;
    memset(&b, 0, sizeof(b));
    b.kind = K_DIELECTRIC;
    b.eta = ior;
    b.mf.ax = a;
    b.mf.ay = a;
    b.r = S1(1.0f); /* Sqrt(Spectrum(1)) */
    float r = 0.0f;
    for (int32_t i = 0; i < ORC_TABLE_SAMPLES; ++i) {
        bsample smp;
        memset(&smp, 0, sizeof(smp));
        if (dielectric_sample(&b, wo, t->u0[i], t->u12[i], 0, &smp)) r += smp.f.r * abs_cos_theta(smp.wi) / smp.pdf;
    }
    r = r / (float)ORC_TABLE_SAMPLES;
    return (r + r + r) / 3.0f;
}

typedef struct {
    int32_t first, step;
} dms_job;

static void* dms_worker(void* arg)
{
    const dms_job* job = (const dms_job*)arg;
    const refl_tables_t* t = tables();
    dms_tables_t* o = &g_dms;
    const int32_t n = ORC_TABLE_N;
    const float d = 1.0f / (float)n;
    for (int32_t k = job->first; k < n; k += job->step) {
        float f = d / 2.0f + d * (float)k;
        float ior = map_f0_to_ior(f);
        for (int32_t j = 0; j < n; ++j) {
            float a = d / 2.0f + d * (float)j;
            float r_sum = 0.0f, r_inv_sum = 0.0f;
            for (int32_t i = 0; i < n; ++i) {
                float cos_theta = d / 2.0f + d * (float)i;
                float sin_theta = sqrtf(1.0f - sqrf(cos_theta));
                v3 wo = V3(sin_theta, 0.0f, cos_theta);
                float r = dielectric_rho(ior, a, wo, t);
                float r_inv = dielectric_rho(1.0f / ior, a, wo, t);
                r_sum += r * cos_theta * d;
                r_inv_sum += r_inv * cos_theta * d;
                o->E[(j * n + i) * n + k] = r;
                o->E_inv[(j * n + i) * n + k] = r_inv;
            }
            o->E_avg[j * n + k] = 2.0f * r_sum;
            o->E_inv_avg[j * n + k] = 2.0f * r_inv_sum;
        }
    }
    return NULL;
}

/* the 32 f0 slices are independent (the reference's ParallelFor over k) */
static void dms_build(void)
{
    enum { MAX_THREADS = 16 };
    long cpus = sysconf(_SC_NPROCESSORS_ONLN);
    int32_t nt = cpus < 1 ? 1 : (cpus > MAX_THREADS ? MAX_THREADS : (int32_t)cpus);
    pthread_t th[MAX_THREADS];
    dms_job jobs[MAX_THREADS];
    (void)tables();
    int32_t started = 0;
    for (int32_t i = 0; i < nt; ++i) {
        jobs[i].first = i;
        jobs[i].step = nt;
        if (pthread_create(&th[i], NULL, dms_worker, &jobs[i]) != 0) break;
        ++started;
    }
    if (started < nt) { /* run the slices of threads that did not start here */
        for (int32_t i = started; i < nt; ++i) dms_worker(&jobs[i]);
    }
    for (int32_t i = 0; i < started; ++i) pthread_join(th[i], NULL);
}

static const dms_tables_t* dms_tables(void);

/* the tables are built first, in the default mode, so that they stay equal
 * to the library's whatever the mode */
void orc_set_cr_libm(int32_t on)
{
    (void)tables();
    (void)dms_tables();
    orc_cr_libm = on;
}

static const dms_tables_t* dms_tables(void)
{
    pthread_once(&g_dms_once, dms_build);
    return &g_dms;
}

void orc_dielectric_tables(float* E, float* E_inv, float* E_avg, float* E_inv_avg)
{
    const dms_tables_t* t = dms_tables();
    if (E) memcpy(E, t->E, sizeof(t->E));
    if (E_inv) memcpy(E_inv, t->E_inv, sizeof(t->E_inv));
    if (E_avg) memcpy(E_avg, t->E_avg, sizeof(t->E_avg));
    if (E_inv_avg) memcpy(E_inv_avg, t->E_inv_avg, sizeof(t->E_inv_avg));
}

/* ImageTexture3D<Float>::Evaluate, trilinear with TexCoordFilter::clamp
 * (textures3d.h:36-88, texture3d.h:16-43), on an n^3 image */
static float tex_trilinear(const float* img, int32_t n, float u, float v, float w)
{
    float fw = u * (float)n - 0.5f;
    float fh = v * (float)n - 0.5f;
    float fd = w * (float)n - 0.5f;
    int32_t i0 = (int32_t)floorf(fw), i1 = i0 + 1;
    int32_t j0 = (int32_t)floorf(fh), j1 = j0 + 1;
    int32_t k0 = (int32_t)floorf(fd), k1 = k0 + 1;
#define ORC_CLAMPI(x) ((x) < 0 ? 0 : ((x) > n - 1 ? n - 1 : (x)))
    i0 = ORC_CLAMPI(i0);
    j0 = ORC_CLAMPI(j0);
    k0 = ORC_CLAMPI(k0);
    i1 = ORC_CLAMPI(i1);
    j1 = ORC_CLAMPI(j1);
    k1 = ORC_CLAMPI(k1);
#undef ORC_CLAMPI
    float fu = fw - floorf(fw);
    float fv = fh - floorf(fh);
    float fz = fd - floorf(fd);
#define ORC_AT(i, j, k) img[((k) * n + (j)) * n + (i)]
    float v000 = ORC_AT(i0, j0, k0), v001 = ORC_AT(i0, j0, k1);
    float v010 = ORC_AT(i0, j1, k0), v011 = ORC_AT(i0, j1, k1);
    float v100 = ORC_AT(i1, j0, k0), v101 = ORC_AT(i1, j0, k1);
    float v110 = ORC_AT(i1, j1, k0), v111 = ORC_AT(i1, j1, k1);
#undef ORC_AT
    float v00 = (1.0f - fu) * v000 + fu * v100;
    float v01 = (1.0f - fu) * v001 + fu * v101;
    float v10 = (1.0f - fu) * v010 + fu * v110;
    float v11 = (1.0f - fu) * v011 + fu * v111;
    float v0 = (1.0f - fv) * v00 + fv * v10;
    float v1 = (1.0f - fv) * v01 + fv * v11;
    return (1.0f - fz) * v0 + fz * v1;
}

/* DielectricMultiScatteringBxDF::E / E_avg (bxdfs.h:180-212) */
static float dms_E(const bxdf_t* b, v3 w, float eta)
{
    const dms_tables_t* t = dms_tables();
    float a = sqrtf(b->mf.ax * b->mf.ay);
    if (eta >= 1.0f) return tex_trilinear(t->E, ORC_TABLE_N, map_ior_to_f0(eta), fabsf(w.z), a);
    return tex_trilinear(t->E_inv, ORC_TABLE_N, map_ior_to_f0(1.0f / eta), fabsf(w.z), a);
}
static float dms_E_avg(const bxdf_t* b, float eta)
{
    const dms_tables_t* t = dms_tables();
    float a = sqrtf(b->mf.ax * b->mf.ay);
    if (eta >= 1.0f) return tex_bilinear(t->E_avg, ORC_TABLE_N, ORC_TABLE_N, map_ior_to_f0(eta), a);
    return tex_bilinear(t->E_inv_avg, ORC_TABLE_N, ORC_TABLE_N, map_ior_to_f0(1.0f / eta), a);
}

/* FresnelDielectricAverage (scattering.h:154-172), eta_p = eta_i / eta_t */
static float fresnel_dielectric_average(float eta_p)
{
    if (eta_p < 1.0f) {
        const float eta2 = eta_p * eta_p;
        const float eta3 = eta_p * eta2;
        return 0.997118f + 0.1014f * eta_p - 0.965241f * eta2 - 0.130607f * eta3;
    }
    return (eta_p - 1.0f) / (4.08567f + 1.00071f * eta_p);
}

/* ComputeScatteringRatio (bxdfs.h:222-242) */
static float dms_ratio(const bxdf_t* b, float eta_o)
{
    float eta_i = b->eta >= 1.0f ? b->eta : 1.0f / b->eta;
    float eta_t = 1.0f / eta_i;
    float a = (1.0f - fresnel_dielectric_average(eta_i)) / (1.0f - dms_E_avg(b, eta_t));
    float bb = sqrf(eta_i) * (1.0f - fresnel_dielectric_average(eta_t)) / (1.0f - dms_E_avg(b, eta_i));
    float x = 1.0f / (a / bb + 1.0f);
    if (eta_o >= 1.0f) return 1.0f - x * (1.0f - fresnel_dielectric_average(eta_o));
    return 1.0f - (1.0f - x) * (1.0f - fresnel_dielectric_average(eta_o));
}

/* the multiple-scattering lobes: ratio (1 - E(wi)) (1 - E(wo)) / max(1e-4, pi (1 - E_avg))
 * for reflection, (1 - ratio) ... with eta_i for transmission */
static float dms_fr_ms(const bxdf_t* b, v3 wo, v3 wi, float ratio, float eta_o)
{
    return ratio * (1.0f - dms_E(b, wi, eta_o)) * (1.0f - dms_E(b, wo, eta_o)) /
           fmaxf(1e-4f, orc_pi * (1.0f - dms_E_avg(b, eta_o)));
}
static float dms_ft_ms(const bxdf_t* b, v3 wo, v3 wi, float ratio, float eta_o, float eta_i)
{
    return (1.0f - ratio) * (1.0f - dms_E(b, wi, eta_i)) * (1.0f - dms_E(b, wo, eta_o)) /
           fmaxf(1e-4f, orc_pi * (1.0f - dms_E_avg(b, eta_i)));
}

/* f (dielectric_multi_scattering_bxdf.cpp:9-89), TransportDirection::ToLight */
static spec dielectric_ms_f(const bxdf_t* b, v3 wo, v3 wi)
{
    if (b->eta == 1.0f || ggx_smooth(b->mf)) return S1(0.0f);
    float cos_o = wo.z, cos_i = wi.z;
    float eta_p = 1.0f;
    int refl = cos_i * cos_o > 0.0f;
    if (!refl) eta_p = cos_o > 0.0f ? b->eta : (1.0f / b->eta);
    float eta_o = cos_o > 0.0f ? b->eta : 1.0f / b->eta;
    float eta_i = cos_o > 0.0f ? 1.0f / b->eta : b->eta;
    v3 wm = vadd(vmul(wi, eta_p), wo);
    if (cos_i == 0.0f || cos_o == 0.0f || vlen2(wm) == 0.0f) return S1(0.0f);
    vnormalize_m(&wm);
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    if (vdot(wm, wi) * cos_i < 0.0f || vdot(wm, wo) * cos_o < 0.0f) return S1(0.0f);
    float ratio = dms_ratio(b, eta_o);
    float F = fresnel_dielectric(vdot(wo, wm), b->eta);
    if (refl) {
        float fr_ss = F * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) / (4.0f * cos_i * cos_o);
        float fr_ms = dms_fr_ms(b, wo, wi, ratio, eta_o);
        return S1(fr_ss + fr_ms);
    }
    float denom = sqrf(vdot(wi, wm) + vdot(wo, wm) / eta_p);
    float ft_ss = (1.0f - F) * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) *
                  fabsf(vdot(wi, wm) * vdot(wo, wm) / (cos_i * cos_o * denom));
    float ft_ms = dms_ft_ms(b, wo, wi, ratio, eta_o, eta_i);
    ft_ss /= sqrf(eta_p);
    return smulf(b->r, ft_ss + ft_ms);
}

/* PDF (dielectric_multi_scattering_bxdf.cpp:91-187), all sampling flags */
static float dielectric_ms_pdf(const bxdf_t* b, v3 wo, v3 wi)
{
    if (b->eta == 1.0f || ggx_smooth(b->mf)) return 0.0f;
    float cos_o = wo.z, cos_i = wi.z;
    float eta_p = 1.0f;
    int refl = cos_i * cos_o > 0.0f;
    if (!refl) eta_p = cos_o > 0.0f ? b->eta : (1.0f / b->eta);
    float eta_o = cos_o > 0.0f ? b->eta : 1.0f / b->eta;
    v3 wm = vadd(vmul(wi, eta_p), wo);
    if (cos_i == 0.0f || cos_o == 0.0f || vlen2(wm) == 0.0f) return 0.0f;
    vnormalize_m(&wm);
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    if (vdot(wm, wi) * cos_i < 0.0f || vdot(wm, wo) * cos_o < 0.0f) return 0.0f;
    float ratio = dms_ratio(b, eta_o);
    float E_o = clampf_(dms_E(b, wo, eta_o), 1e-4f, 1.0f - 1e-4f);
    float R = fresnel_dielectric(vdot(wo, wm), b->eta);
    float T = 1.0f - R;
    float pr_ss = R, pt_ss = T, pr_ms = ratio, pt_ms = 1.0f - ratio;
    if (pr_ss == 0.0f && pt_ss == 0.0f && pr_ms == 0.0f && pt_ms == 0.0f) return 0.0f;
    float p_ss_sum = pr_ss + pt_ss, p_ms_sum = pr_ms + pt_ms;
    pr_ss /= p_ss_sum;
    pt_ss /= p_ss_sum;
    pr_ms /= p_ms_sum;
    pt_ms /= p_ms_sum;
    float pdf_ss, pdf_ms;
    if (refl) {
        pdf_ss = pr_ss * ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
        pdf_ms = pr_ms * (abs_cos_theta(wi) * orc_inv_pi);
    } else {
        float dwm_dwi = vabsdot(wi, wm) / sqrf(vdot(wi, wm) + vdot(wo, wm) / eta_p);
        pdf_ss = pt_ss * ggx_pdf(b->mf, wo, wm) * dwm_dwi;
        pdf_ms = pt_ms * (abs_cos_theta(wi) * orc_inv_pi);
    }
    return pdf_ms * (1.0f - E_o) + pdf_ss * E_o; /* Lerp(pdf_ms, pdf_ss, E_o) */
}

/* Sample_f (dielectric_multi_scattering_bxdf.cpp:189-433), rough branch,
 * TransportDirection::ToLight, all sampling flags */
static int dielectric_ms_sample(const bxdf_t* b, v3 wo, float u0, v2 u12, bsample* s)
{
    s->stochastic = 0;
    if (b->eta == 1.0f || ggx_smooth(b->mf)) return dielectric_sample(b, wo, u0, u12, 1, s);
    float eta_o = wo.z > 0.0f ? b->eta : 1.0f / b->eta;
    float eta_i = wo.z > 0.0f ? 1.0f / b->eta : b->eta;
    float ratio = dms_ratio(b, eta_o);
    float E_o = clampf_(dms_E(b, wo, eta_o), 1e-4f, 1.0f - 1e-4f);
    float pr_ms = ratio, pt_ms = 1.0f - ratio;
    if (u0 < E_o) {
        /* single scattering at the rough interface */
        v3 wm = ggx_sample_wm(b->mf, wo, u12);
        vnormalize_m(&wm);
        if (dot_z(wm) < 0.0f) wm = vneg(wm);
        float R = fresnel_dielectric(vdot(wo, wm), b->eta);
        float T = 1.0f - R;
        float pr_ss = R, pt_ss = T;
        if (pr_ss == 0.0f && pt_ss == 0.0f) return 0;
        float p_ss_sum = pr_ss + pt_ss, p_ms_sum = pr_ms + pt_ms;
        pr_ss /= p_ss_sum;
        pt_ss /= p_ss_sum;
        pr_ms /= p_ms_sum;
        pt_ms /= p_ms_sum;
        u0 /= E_o;
        if (u0 < pr_ss) {
            v3 wi = reflect(wo, wm);
            if (!same_hemisphere(wo, wi)) return 0;
            float pdf_ss = pr_ss * ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
            float pdf_ms = pr_ms * (abs_cos_theta(wi) * orc_inv_pi);
            float fr_ss = R * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) / (4.0f * wi.z * wo.z);
            float fr_ms = dms_fr_ms(b, wo, wi, ratio, eta_o);
            s->f = S1(fr_ss + fr_ms);
            s->wi = wi;
            s->pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
            s->flags = BX_GLOSSY | BX_REFLECTION;
            s->eta = 1.0f;
            return 1;
        }
        v3 wi = V3(0.0f, 0.0f, 0.0f);
        int tir = !refract(&wi, wo, wm, b->eta, NULL);
        if (same_hemisphere(wo, wi) || wi.z == 0.0f || tir) return 0;
        float eta_p = wo.z > 0.0f ? b->eta : (1.0f / b->eta);
        float denom = sqrf(vdot(wi, wm) + vdot(wo, wm) / eta_p);
        float dwm_dwi = vabsdot(wi, wm) / denom;
        float pdf_ss = pt_ss * ggx_pdf(b->mf, wo, wm) * dwm_dwi;
        float pdf_ms = pt_ms * (abs_cos_theta(wi) * orc_inv_pi);
        float ft_ss = T * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) *
                      fabsf(vdot(wi, wm) * vdot(wo, wm) / (wi.z * wo.z * denom));
        float ft_ms = dms_ft_ms(b, wo, wi, ratio, eta_o, eta_i);
        ft_ss /= sqrf(eta_p);
        s->f = smulf(b->r, ft_ss + ft_ms);
        s->wi = wi;
        s->pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
        s->flags = BX_GLOSSY | BX_TRANSMISSION;
        s->eta = eta_p;
        return 1;
    }
    /* the multiple-scattering lobe: cosine-distributed reflection or transmission */
    if (pr_ms == 0.0f && pt_ms == 0.0f) return 0;
    float p_ms_sum = pr_ms + pt_ms;
    pr_ms /= p_ms_sum;
    pt_ms /= p_ms_sum;
    u0 = (u0 - E_o) / (1.0f - E_o);
    if (u0 < pr_ms) {
        v3 wi = sample_cosine_hemisphere(u12);
        if (wo.z < 0.0f) wi = vneg(wi);
        v3 wm = vnormalize(vadd(wi, wo));
        float R = fresnel_dielectric(vdot(wo, wm), b->eta);
        float pr_ss = R, pt_ss = 1.0f - R;
        float p_ss_sum = pr_ss + pt_ss;
        pr_ss /= p_ss_sum;
        pt_ss /= p_ss_sum;
        float pdf_ss = pr_ss * ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
        float pdf_ms = pr_ms * (abs_cos_theta(wi) * orc_inv_pi);
        float fr_ss = R * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) / (4.0f * wi.z * wo.z);
        float fr_ms = dms_fr_ms(b, wo, wi, ratio, eta_o);
        s->f = S1(fr_ss + fr_ms);
        s->wi = wi;
        s->pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
        s->flags = BX_DIFFUSE | BX_REFLECTION;
        s->eta = 1.0f;
        return 1;
    }
    v3 wi = sample_cosine_hemisphere(u12);
    if (wo.z > 0.0f) wi = vneg(wi);
    v3 wm = vnormalize(vadd(vmul(wi, eta_o), wo));
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    float R = fresnel_dielectric(vdot(wo, wm), b->eta);
    float T = 1.0f - R;
    float pr_ss = R, pt_ss = 1.0f - R;
    float p_ss_sum = pr_ss + pt_ss;
    pr_ss /= p_ss_sum;
    pt_ss /= p_ss_sum;
    float denom = sqrf(vdot(wi, wm) + vdot(wo, wm) / eta_o);
    float dwm_dwi = vabsdot(wi, wm) / denom;
    float pdf_ss = pt_ss * ggx_pdf(b->mf, wo, wm) * dwm_dwi;
    float pdf_ms = pt_ms * (abs_cos_theta(wi) * orc_inv_pi);
    float ft_ss = T * ggx_D(b->mf, wm) * ggx_G(b->mf, wo, wi) *
                  fabsf(vdot(wi, wm) * vdot(wo, wm) / (wi.z * wo.z * denom));
    float ft_ms = dms_ft_ms(b, wo, wi, ratio, eta_o, eta_i);
    ft_ss /= sqrf(eta_o);
    s->f = smulf(b->r, ft_ss + ft_ms);
    s->wi = wi;
    s->pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
    s->flags = BX_DIFFUSE | BX_TRANSMISSION;
    s->eta = eta_o;
    return 1;
}

/* ---- MetallicRoughness: src/bxdf/metallic_roughness_bxdf.cpp:8-154,
 *      include/bulbit/bxdfs.h:448-489 ---- */
static spec mr_F0(const bxdf_t* b)
{
    /* Lerp(0.04, color, metallic) = (1 - m) * f0 + m * color (spectrum.h:228-232) */
    return sadd(smulf(S1(0.04f), 1.0f - b->metallic), smulf(b->color, b->metallic));
}
static spec mr_schlick(spec f0, float c)
{
    float p = o_powf(1.0f - c, 5.0f);
    return sadd(f0, smulf(ssub(S1(1.0f), f0), p));
}
static spec mr_f(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi)) return S1(0.0f);
    float cos_o = abs_cos_theta(wo), cos_i = abs_cos_theta(wi);
    if (cos_i == 0.0f || cos_o == 0.0f) return S1(0.0f);
    v3 wm = vadd(wo, wi);
    if (vlen2(wm) == 0.0f) return S1(0.0f);
    vnormalize_m(&wm);
    spec F = mr_schlick(mr_F0(b), vdot(wi, wm));
    spec f_s = sdivf(smulf(smulf(F, ggx_D(b->mf, wm)), ggx_G(b->mf, wo, wi)), 4.0f * cos_i * cos_o);
    spec f_d = smul(smulf(ssub(S1(1.0f), F), 1.0f - b->metallic), smulf(b->color, orc_inv_pi));
    return sadd(f_d, f_s);
}
static float mr_pdf(const bxdf_t* b, v3 wo, v3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b->mf)) return 0.0f;
    v3 wm = vadd(wo, wi);
    if (vlen2(wm) == 0.0f) return 0.0f;
    vnormalize_m(&wm);
    if (dot_z(wm) < 0.0f) wm = vneg(wm);
    float R = fresnel_dielectric(vdot(wo, wm), 1.5f);
    float T = 1.0f - R;
    float pr = R, pt = T * (1.0f - b->metallic);
    if (pr == 0.0f && pt == 0.0f) return 0.0f;
    float p_sum = pr + pt;
    pr /= p_sum;
    pt /= p_sum;
    float pdf_r = ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
    float pdf_d = abs_cos_theta(wi) * orc_inv_pi;
    return pr * pdf_r + pt * pdf_d;
}
static int mr_sample(const bxdf_t* b, v3 wo, float u0, v2 u12, bsample* s)
{
    if (wo.z == 0.0f) return 0;
    v3 wm = ggx_sample_wm(b->mf, wo, u12);
    v3 wi = reflect(wo, wm);
    float R = fresnel_dielectric(vdot(wo, wm), 1.5f);
    float T = 1.0f - R;
    float pr = R, pt = T * (1.0f - b->metallic);
    if (pr == 0.0f && pt == 0.0f) return 0;
    float p_sum = pr + pt;
    pr /= p_sum;
    pt /= p_sum;
    int flag;
    if (u0 < pr) {
        if (!same_hemisphere(wo, wi)) return 0;
        flag = BX_GLOSSY | BX_REFLECTION;
    } else {
        wi = sample_cosine_hemisphere(u12);
        wm = vnormalize(vadd(wi, wo));
        flag = BX_DIFFUSE | BX_REFLECTION;
    }
    float cos_o = abs_cos_theta(wo), cos_i = abs_cos_theta(wi);
    if (cos_i == 0.0f || cos_o == 0.0f) return 0;
    spec F = mr_schlick(mr_F0(b), vdot(wi, wm));
    spec f_r = sdivf(smulf(smulf(F, ggx_D(b->mf, wm)), ggx_G(b->mf, wo, wi)), 4.0f * cos_i * cos_o);
    spec f_d = smul(smulf(ssub(S1(1.0f), F), 1.0f - b->metallic), smulf(b->color, orc_inv_pi));
    float pdf_r = ggx_pdf(b->mf, wo, wm) / (4.0f * vabsdot(wo, wm));
    float pdf_d = cos_i * orc_inv_pi;
    s->f = sadd(f_r, f_d);
    s->wi = wi;
    s->pdf = pr * pdf_r + pt * pdf_d;
    s->flags = flag;
    s->eta = 1.0f;
    s->stochastic = 0;
    return 1;
}

/* ============================================================ BSDF layer */

typedef struct {
    frame3 frame;
    bxdf_t bx;
} bsdf_t;

/* Intersection::GetBSDF + Material::GetBSDF (src/geometry/intersection.cpp:32-52,
 * diffuse_material.cpp:21-36, dielectric_material.cpp:28-52,
 * conductor_material.cpp:52-99, metallic_roughness_material.cpp:28-39).
 * BSDF(n, t, bxdf) builds Frame::FromXZ(t, n) (bsdf.h:64-68). */
static int make_bsdf_c(const BbMaterial* m, v3 ns, v3 ts, int front_face, bsdf_t* out)
{
    bxdf_t* b = &out->bx;
    memset(b, 0, sizeof(*b));
    out->frame = frame_from_xz(ts, ns);
    switch (m->type) {
    case BB_MAT_DIFFUSE:
        /* EONBxDF(rho, r) when the roughness is > 0, else Lambertian */
        b->kind = m->u_roughness > 0.0f ? K_EON : K_LAMBERT;
        b->r = S3(m->color[0], m->color[1], m->color[2]);
        b->mf.ax = b->mf.ay = m->u_roughness;
        return 1;
    case BB_MAT_DIELECTRIC:
        /* DielectricMultiScatteringBxDF when energy_compensation (dielectric_material.cpp:36-42);
         * effectively smooth, or with eta 1, it is the plain BxDF's specular branch
         * (dielectric_multi_scattering_bxdf.cpp:11-14, 90-93, 196-244) */
        b->eta = front_face ? m->eta[0] : 1.0f / m->eta[0];
        b->r = ssqrt(S3(m->color[0], m->color[1], m->color[2]));
        b->mf.ax = sqrf(m->u_roughness);
        b->mf.ay = sqrf(m->v_roughness);
        b->kind = (m->energy_compensation && !ggx_smooth(b->mf) && b->eta != 1.0f) ? K_DIELECTRIC_MS : K_DIELECTRIC;
        return 1;
    case BB_MAT_CONDUCTOR:
    case BB_MAT_CONDUCTOR_R:
        /* ConductorMultiScatteringBxDF when energy_compensation (conductor_material.cpp:63-95) */
        b->kind = m->energy_compensation ? K_CONDUCTOR_MS : K_CONDUCTOR;
        b->mf.ax = sqrf(m->u_roughness);
        b->mf.ay = sqrf(m->v_roughness);
        b->r = S3(m->color[0], m->color[1], m->color[2]);
        if (m->type == BB_MAT_CONDUCTOR) {
            b->ceta = S3(m->eta[0], m->eta[1], m->eta[2]);
            b->ck = S3(m->k[0], m->k[1], m->k[2]);
        } else {
            /* ConductorBxDF(R, ...) (bxdfs.h:262-272): eta = 1,
             * k = 2 * Sqrt(r) / Sqrt(Max(1 - r, 0)), r = Clamp(R, 0, 1 - 1e-4) */
            spec R = S3(m->k[0], m->k[1], m->k[2]);
            float hi = 1.0f - 1e-4f;
            spec r = S3(clampf_(R.r, 0.0f, hi), clampf_(R.g, 0.0f, hi), clampf_(R.b, 0.0f, hi));
            spec one_minus = ssub(S1(1.0f), r);
            spec mx = S3(smax(0.0f, one_minus.r), smax(0.0f, one_minus.g), smax(0.0f, one_minus.b));
            b->ceta = S1(1.0f);
            b->ck = sdiv(smulf(ssqrt(r), 2.0f), ssqrt(mx));
        }
        return 1;
    case BB_MAT_MIRROR: /* mirror_material.cpp:18-23: SpecularReflectionBxDF(r) */
        b->kind = K_MIRROR;
        b->r = S3(m->color[0], m->color[1], m->color[2]);
        return 1;
    case BB_MAT_METALLIC_ROUGHNESS:
        b->kind = K_MR;
        b->color = S3(m->color[0], m->color[1], m->color[2]);
        b->metallic = m->metallic;
        b->mf.ax = fmaxf(sqrf(m->u_roughness), 1e-3f);
        b->mf.ay = fmaxf(sqrf(m->v_roughness), 1e-3f);
        return 1;
    }
    return 0;
}

/* make_bsdf_c with the material's textures evaluated at uv: each texture
 * slot replaces the constant the material would otherwise pass (the
 * *Material::GetBSDF functions evaluate every parameter texture at isect.uv) */
static int make_bsdf(const OrcScene* s, const BbMaterial* m0, v2 uv, v3 ns, v3 ts, int front_face, bsdf_t* out)
{
    BbMaterial m = *m0;
    if (s) {
        if (m.tex_color) {
            spec c = texture3(s, m.tex_color - 1, uv);
            m.color[0] = c.r;
            m.color[1] = c.g;
            m.color[2] = c.b;
        }
        if (m.tex_u_roughness) m.u_roughness = texture1(s, m.tex_u_roughness - 1, uv);
        if (m.tex_v_roughness && m.type != BB_MAT_DIFFUSE) m.v_roughness = texture1(s, m.tex_v_roughness - 1, uv);
        if (m.tex_metallic && m.type == BB_MAT_METALLIC_ROUGHNESS) m.metallic = texture1(s, m.tex_metallic - 1, uv);
    }
    return make_bsdf_c(&m, ns, ts, front_face, out);
}

/* NormalMapping (src/geometry/intersection.cpp:9-30) */
static void normal_mapping(const OrcScene* s, int32_t idx, v3 wo, isect_t* is)
{
    spec sp = texture3(s, idx, is->uv);
    v3 sn = V3(sp.r * 2 - 1, sp.g * 2 - 1, sp.b * 2 - 1);
    vnormalize_m(&sn);
    frame3 tbn = frame_from_xz(is->ts, is->ns);
    v3 n = frame_from_local(&tbn, sn);
    if (vdot(n, wo) < 0.0f) n = reflect(n, is->normal);
    v3 t = vsub(is->ts, vmul(n, vdot(is->ts, n)));
    is->ns = n;
    is->ts = t;
}

/* BSDF::f (bsdf.h:75-86) */
static spec bsdf_f(const bsdf_t* s, v3 wo_w, v3 wi_w)
{
    v3 wi = frame_to_local(&s->frame, wi_w);
    v3 wo = frame_to_local(&s->frame, wo_w);
    if (wo.z == 0.0f) return S1(0.0f);
    switch (s->bx.kind) {
    case K_LAMBERT: return lambert_f(&s->bx, wo, wi);
    case K_EON: return eon_f(&s->bx, wo, wi);
    case K_DIELECTRIC: return dielectric_f(&s->bx, wo, wi);
    case K_DIELECTRIC_MS: return dielectric_ms_f(&s->bx, wo, wi);
    case K_CONDUCTOR: return conductor_f(&s->bx, wo, wi);
    case K_CONDUCTOR_MS: return conductor_ms_f(&s->bx, wo, wi);
    case K_MIRROR: return S1(0.0f); /* specular_reflection_bxdf.cpp:7-13 */
    default: return mr_f(&s->bx, wo, wi);
    }
}

/* BSDF::PDF (bsdf.h:88-99) */
static float bsdf_pdf(const bsdf_t* s, v3 wo_w, v3 wi_w)
{
    v3 wo = frame_to_local(&s->frame, wo_w);
    v3 wi = frame_to_local(&s->frame, wi_w);
    if (wo.z == 0.0f) return 0.0f;
    switch (s->bx.kind) {
    case K_LAMBERT: return lambert_pdf(wo, wi);
    case K_EON: return eon_pdf(&s->bx, wo, wi);
    case K_DIELECTRIC: return dielectric_pdf(&s->bx, wo, wi);
    case K_DIELECTRIC_MS: return dielectric_ms_pdf(&s->bx, wo, wi);
    case K_CONDUCTOR: return conductor_pdf(&s->bx, wo, wi);
    case K_CONDUCTOR_MS: return conductor_ms_pdf(&s->bx, wo, wi);
    case K_MIRROR: return 0.0f;
    default: return mr_pdf(&s->bx, wo, wi);
    }
}

/* BSDF::Sample_f (bsdf.h:101-125) */
static int bsdf_sample(const bsdf_t* s, v3 wo_w, float u0, v2 u12, bsample* out)
{
    v3 wo = frame_to_local(&s->frame, wo_w);
    if (wo.z == 0.0f || !(bxdf_flags(&s->bx) & (BX_REFLECTION | BX_TRANSMISSION))) return 0;
    int ok;
    switch (s->bx.kind) {
    case K_LAMBERT: ok = lambert_sample(&s->bx, wo, u12, out); break;
    case K_EON: ok = eon_sample(&s->bx, wo, u0, u12, out); break;
    case K_DIELECTRIC: ok = dielectric_sample(&s->bx, wo, u0, u12, 1, out); break;
    case K_DIELECTRIC_MS: ok = dielectric_ms_sample(&s->bx, wo, u0, u12, out); break;
    case K_CONDUCTOR: ok = conductor_sample(&s->bx, wo, u12, out); break;
    case K_CONDUCTOR_MS: ok = conductor_ms_sample(&s->bx, wo, u0, u12, out); break;
    case K_MIRROR: { /* specular_reflection_bxdf.cpp:24-38 */
        v3 wi = V3(-wo.x, -wo.y, wo.z);
        out->f = sdivf(s->bx.r, fabsf(wi.z));
        out->wi = wi;
        out->pdf = 1.0f;
        out->flags = BX_SPECULAR | BX_REFLECTION;
        out->eta = 1.0f;
        out->stochastic = 0;
        ok = 1;
        break;
    }
    default: ok = mr_sample(&s->bx, wo, u0, u12, out); break;
    }
    if (!ok) return 0;
    if (sblack(out->f) || out->pdf == 0.0f || out->wi.z == 0.0f) return 0;
    out->wi = frame_from_local(&s->frame, out->wi);
    return 1;
}

static int material_supported(const BbMaterial* m, char* err, int32_t err_size)
{
    switch (m->type) {
    case BB_MAT_DIFFUSE: return 1;
    case BB_MAT_CONDUCTOR:
    case BB_MAT_CONDUCTOR_R: return 1;
    case BB_MAT_DIELECTRIC: return 1;
    case BB_MAT_METALLIC_ROUGHNESS: return 1;
    case BB_MAT_MIRROR: return 1;
    }
    snprintf(err, err_size, "unknown material type %d", m->type);
    return 0;
}

int32_t orc_bsdf_eval(const BbMaterial* mat, const float n[3], const float t[3], int32_t front_face, const float wo[3],
                      const float wi[3], float f_out[3], float* pdf_out)
{
    bsdf_t b;
    make_bsdf(NULL, mat, V2(0.0f, 0.0f), V3(n[0], n[1], n[2]), V3(t[0], t[1], t[2]), front_face, &b);
    spec f = bsdf_f(&b, V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]));
    f_out[0] = f.r;
    f_out[1] = f.g;
    f_out[2] = f.b;
    *pdf_out = bsdf_pdf(&b, V3(wo[0], wo[1], wo[2]), V3(wi[0], wi[1], wi[2]));
    return bxdf_flags(&b.bx);
}

int32_t orc_bsdf_sample(const BbMaterial* mat, const float n[3], const float t[3], int32_t front_face,
                        const float wo[3], float u0, const float u12[2], float f_out[3], float wi_out[3],
                        float* pdf_out, int32_t* flags_out, float* eta_out)
{
    bsdf_t b;
    make_bsdf(NULL, mat, V2(0.0f, 0.0f), V3(n[0], n[1], n[2]), V3(t[0], t[1], t[2]), front_face, &b);
    bsample s;
    memset(&s, 0, sizeof(s));
    int ok = bsdf_sample(&b, V3(wo[0], wo[1], wo[2]), u0, V2(u12[0], u12[1]), &s);
    f_out[0] = s.f.r;
    f_out[1] = s.f.g;
    f_out[2] = s.f.b;
    wi_out[0] = s.wi.x;
    wi_out[1] = s.wi.y;
    wi_out[2] = s.wi.z;
    *pdf_out = s.pdf;
    *flags_out = s.flags;
    *eta_out = s.eta;
    return ok;
}

/* ================================================================ lights */

/* Triangle::Area (triangle.cpp:255-264) */
static float triangle_area(const OrcScene* s, int32_t tri)
{
    const int32_t* v = s->idx + 3 * tri;
    v3 p0 = vtx(s, s->pos, v[0]), p1 = vtx(s, s->pos, v[1]), p2 = vtx(s, s->pos, v[2]);
    return 0.5f * vlen(vcross(vsub(p1, p0), vsub(p2, p0)));
}

/* Light::Phi (spot_light.cpp:111-114, uniform_infinite_light.cpp:90-93,
 * diffuse_area_light.cpp:144-148) */
static spec light_phi(const OrcScene* s, int32_t li)
{
    const BbLight* L = &s->lights[li];
    const light_extra* x = &s->lx[li];
    spec c = S3(L->spectrum[0], L->spectrum[1], L->spectrum[2]);
    switch (L->type) {
    case BB_LIGHT_SPOT:
        return smulf(smulf(c, orc_two_pi), ((1.0f - x->cos_min) + (x->cos_min - x->cos_max) / 2.0f));
    case BB_LIGHT_UNIFORM_INFINITE:
        return smulf(smulf(smulf(smulf(c, L->scale), orc_four_pi), orc_pi), sqrf(x->world_radius));
    case BB_LIGHT_POINT: /* point_light.cpp:93-96 */
        return smulf(c, orc_four_pi);
    case BB_LIGHT_DIRECTIONAL: /* directional_light.cpp:92-95 */
        return smulf(smulf(c, orc_pi), sqrf(x->world_radius));
    default: {
        int32_t tri = s->prims[L->primitive].shape_index;
        return smulf(smulf(smulf(c, triangle_area(s, tri)), orc_pi), (float)(L->two_sided ? 2 : 1));
    }
    }
}

typedef struct {
    v3 wi;
    float visibility;
    spec Li;
    float pdf;
} light_sample;

/* Light::Sample_Li (spot_light.cpp:46-68, uniform_infinite_light.cpp:30-43,
 * diffuse_area_light.cpp:42-72 with Triangle::Sample, triangle.cpp:159-229) */
static int light_sample_li(const OrcScene* s, int32_t li, v3 ref, v2 u, light_sample* ls)
{
    const BbLight* L = &s->lights[li];
    const light_extra* x = &s->lx[li];
    spec c = S3(L->spectrum[0], L->spectrum[1], L->spectrum[2]);
    if (L->type == BB_LIGHT_SPOT) {
        v3 p = V3(L->position[0], L->position[1], L->position[2]);
        v3 wi = vsub(p, ref);
        float distance = vnormalize_m(&wi);
        float cos_t = -vdot(wi, x->z);
        spec l = smulf(c, smooth_step(x->cos_max, x->cos_min, cos_t));
        if (sblack(l)) return 0;
        ls->wi = wi;
        ls->visibility = distance;
        ls->pdf = 1.0f;
        ls->Li = sdivf(l, distance * distance);
        return 1;
    }
    if (L->type == BB_LIGHT_DIRECTIONAL) { /* directional_light.cpp:29-38 */
        ls->wi = x->z;
        ls->visibility = 2.0f * x->world_radius;
        ls->pdf = 1.0f;
        ls->Li = c;
        return 1;
    }
    if (L->type == BB_LIGHT_POINT) { /* point_light.cpp:30-42 */
        v3 wi = vsub(V3(L->position[0], L->position[1], L->position[2]), ref);
        float distance = vnormalize_m(&wi);
        ls->wi = wi;
        ls->visibility = distance;
        ls->pdf = 1.0f;
        ls->Li = sdivf(c, distance * distance);
        return 1;
    }
    if (L->type == BB_LIGHT_UNIFORM_INFINITE) {
        v3 wi = sample_uniform_sphere(u);
        ls->wi = wi;
        ls->pdf = orc_inv_four_pi;
        ls->visibility = 2.0f * x->world_radius;
        ls->Li = smulf(c, L->scale);
        return 1;
    }
    /* diffuse area light on a triangle */
    int32_t tri = s->prims[L->primitive].shape_index;
    const int32_t* v = s->idx + 3 * tri;
    v3 p0 = vtx(s, s->pos, v[0]), p1 = vtx(s, s->pos, v[1]), p2 = vtx(s, s->pos, v[2]);
    v3 e1 = vsub(p1, p0), e2 = vsub(p2, p0);
    float uu = u.x, vv = u.y;
    if (uu + vv > 1.0f) {
        uu = 1.0f - uu;
        vv = 1.0f - vv;
    }
    v3 normal = vcross(e1, e2);
    v3 point = vadd(vadd(p0, vmul(e1, uu)), vmul(e2, vv));
    float area = vnormalize_m(&normal) * 0.5f;
    float pdf = 1.0f / area;
    v3 d = vsub(point, ref);
    float dist2 = vdot(d, d);
    float cosine = vdot(d, normal) / sqrtf(dist2);
    if (cosine < 0.0f) cosine = -cosine;
    pdf *= dist2 / cosine;
    v3 wi = vsub(point, ref);
    ls->visibility = vnormalize_m(&wi) - 1e-4f;
    ls->wi = wi;
    ls->pdf = pdf;
    int front = vdot(normal, wi) < 0.0f;
    ls->Li = (front || L->two_sided) ? c : S1(0.0f);
    return 1;
}

/* Distribution1D::SampleDiscrete (sampling.h:383-397) + FindInterval
 * (math.h:117-135) */
static int32_t light_select(const OrcScene* s, float u, float* pmf)
{
    int32_t size = s->n_lights + 1;
    int32_t first = 0, len = size;
    while (len > 0) {
        int32_t half = len >> 1, middle = first + half;
        if (s->cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    int32_t offset = first - 1;
    if (offset < 0) offset = 0;
    if (offset > size - 2) offset = size - 2;
    *pmf = s->func[offset] / (s->func_integral * (float)s->n_lights);
    return offset;
}

/* PowerLightSampler::EvaluatePMF -> Distribution1D::DiscretePDF (sampling.h:404-407) */
static inline float light_pmf(const OrcScene* s, int32_t li)
{
    return s->func[li] / (s->func_integral * (float)s->n_lights);
}

void orc_scene_light_pmf(const OrcScene* s, float* out)
{
    for (int32_t i = 0; i < s->n_lights; ++i) out[i] = light_pmf(s, i);
}

/* Triangle::PDF(isect, ray) (triangle.cpp:247-253) */
static float triangle_pdf_solid_angle(const OrcScene* s, int32_t tri, const isect_t* is, v3 d)
{
    float distance_squared = is->t * is->t * vlen2(d);
    float cosine = fabsf(vdot(d, is->normal) / vlen(d));
    return distance_squared / (cosine * triangle_area(s, tri));
}

void orc_light_sample(const OrcScene* s, int32_t li, const float ref[3], int32_t n, const float* u, float* wi,
                      float* Li, float* pdf, float* visibility, int32_t* ok)
{
    for (int32_t i = 0; i < n; ++i) {
        light_sample ls;
        memset(&ls, 0, sizeof(ls));
        ok[i] = light_sample_li(s, li, V3(ref[0], ref[1], ref[2]), V2(u[2 * i], u[2 * i + 1]), &ls);
        wi[3 * i] = ls.wi.x;
        wi[3 * i + 1] = ls.wi.y;
        wi[3 * i + 2] = ls.wi.z;
        Li[3 * i] = ls.Li.r;
        Li[3 * i + 1] = ls.Li.g;
        Li[3 * i + 2] = ls.Li.b;
        pdf[i] = ls.pdf;
        visibility[i] = ls.visibility;
    }
}

void orc_light_hit_pdf(const OrcScene* s, int32_t li, const float o[3], int32_t n, const float* d, float* pdf)
{
    for (int32_t i = 0; i < n; ++i) {
        v3 ro = V3(o[0], o[1], o[2]), rd = V3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        isect_t is;
        pdf[i] = 0.0f;
        if (!bvh_raycast(s, ro, rd, 1e-4f, INFINITY, 0, &is, NULL)) continue;
        finish_isect(s, is.ro, is.rd, &is);
        if (s->prim_light[is.prim] != li) continue;
        /* path_li's emitter-hit density before the light PMF */
        pdf[i] = triangle_pdf_solid_angle(s, s->prims[is.prim].shape_index, &is, rd);
    }
}

/* ================================================================ scenes */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void* dup_bytes(const void* p, size_t n)
{
    if (!n) return NULL;
    void* r = malloc(n);
    if (r && p) memcpy(r, p, n);
    return r;
}

void orc_scene_destroy(OrcScene* s)
{
    if (!s) return;
    free(s->pos);
    free(s->nrm);
    free(s->tan);
    free(s->idx);
    free(s->spheres);
    free(s->prims);
    free(s->mats);
    free(s->lights);
    free(s->uvs);
    if (s->texs)
        for (int32_t i = 0; i < s->n_textures; ++i) free((void*)s->texs[i].texels);
    free(s->texs);
    free(s->lx);
    free(s->nodes);
    free(s->ordered);
    free(s->prim_light);
    free(s->inf_lights);
    free(s->func);
    free(s->cdf);
    free(s);
}

#define FAIL(...)                                                                                                      \
    do {                                                                                                               \
        if (err) snprintf(err, err_size, __VA_ARGS__);                                                                 \
        orc_scene_destroy(s);                                                                                          \
        return NULL;                                                                                                   \
    } while (0)

/* Scene -> BVH::BVH (bvh.cpp:8-49) -> Integrator::Integrator
 * (integrator.cpp:107-147) -> PowerLightSampler::Init
 * (power_light_sampler.cpp:8-23) */
OrcScene* orc_scene_create(const BbSceneDesc* d, char* err, int32_t err_size)
{
    OrcScene* s = (OrcScene*)calloc(1, sizeof(OrcScene));
    if (!s || !d) FAIL("null scene");
    if (d->n_primitives <= 0) FAIL("scene has no primitives");
    s->n_vertices = d->n_vertices;
    s->n_triangles = d->n_triangles;
    s->n_spheres = d->n_spheres;
    s->n_prims = d->n_primitives;
    s->n_materials = d->n_materials;
    s->n_lights = d->n_lights;
    s->pos = (float*)dup_bytes(d->positions, sizeof(float) * 3 * (size_t)d->n_vertices);
    s->nrm = (float*)dup_bytes(d->normals, sizeof(float) * 3 * (size_t)d->n_vertices);
    s->tan = (float*)dup_bytes(d->tangents, sizeof(float) * 3 * (size_t)d->n_vertices);
    s->idx = (int32_t*)dup_bytes(d->indices, sizeof(int32_t) * 3 * (size_t)d->n_triangles);
    s->spheres = (BbSphere*)dup_bytes(d->spheres, sizeof(BbSphere) * (size_t)d->n_spheres);
    s->prims = (BbPrimitive*)dup_bytes(d->primitives, sizeof(BbPrimitive) * (size_t)d->n_primitives);
    s->mats = (BbMaterial*)dup_bytes(d->materials, sizeof(BbMaterial) * (size_t)d->n_materials);
    s->lights = (BbLight*)dup_bytes(d->lights, sizeof(BbLight) * (size_t)d->n_lights);
    s->uvs = d->uvs ? (float*)dup_bytes(d->uvs, sizeof(float) * 2 * (size_t)d->n_vertices) : NULL;
    s->n_textures = d->n_textures;
    s->texs = (BbTexture*)calloc((size_t)(d->n_textures + 1), sizeof(BbTexture));
    for (int32_t i = 0; i < d->n_textures; ++i) {
        s->texs[i] = d->textures[i];
        if (d->textures[i].type == BB_TEX_IMAGE)
            s->texs[i].texels = (float*)dup_bytes(d->textures[i].texels, sizeof(float) * (size_t)d->textures[i].width *
                                                                                 (size_t)d->textures[i].height *
                                                                                 (size_t)d->textures[i].channels);
        else
            s->texs[i].texels = NULL;
    }
    if (d->n_triangles > 0 && (!d->positions || !d->normals || !d->tangents || !d->indices))
        FAIL("triangles need positions, normals, tangents and indices");
    for (int32_t i = 0; i < 3 * d->n_triangles; ++i)
        if (d->indices[i] < 0 || d->indices[i] >= d->n_vertices) FAIL("vertex index out of range");
    for (int32_t i = 0; i < d->n_primitives; ++i) {
        const BbPrimitive* p = &d->primitives[i];
        if (p->shape_kind == BB_SHAPE_TRIANGLE) {
            if (p->shape_index < 0 || p->shape_index >= d->n_triangles) FAIL("primitive %d: bad triangle", i);
        } else if (p->shape_kind == BB_SHAPE_SPHERE) {
            if (p->shape_index < 0 || p->shape_index >= d->n_spheres) FAIL("primitive %d: bad sphere", i);
        } else {
            FAIL("primitive %d: bad shape kind", i);
        }
        if (p->material >= d->n_materials) FAIL("primitive %d: bad material", i);
    }
    for (int32_t i = 0; i < d->n_materials; ++i)
        if (!material_supported(&d->materials[i], err, err_size)) {
            orc_scene_destroy(s);
            return NULL;
        }

    double t0 = now_s();
    /* BVH */
    bvh_builder B;
    memset(&B, 0, sizeof(B));
    B.s = s;
    B.prims = (bvh_prim*)malloc(sizeof(bvh_prim) * (size_t)s->n_prims);
    B.nodes = (BbBvhNode*)calloc((size_t)(2 * s->n_prims), sizeof(BbBvhNode));
    B.ordered = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->n_prims);
    for (int32_t i = 0; i < s->n_prims; ++i) {
        B.prims[i].index = i;
        const BbPrimitive* p = &s->prims[i];
        if (p->shape_kind == BB_SHAPE_TRIANGLE)
            triangle_aabb(s, p->shape_index, &B.prims[i].bmin, &B.prims[i].bmax);
        else
            sphere_aabb(&s->spheres[p->shape_index], &B.prims[i].bmin, &B.prims[i].bmax);
    }
    bvh_build(&B, 0, s->n_prims);
    free(B.prims);
    s->nodes = B.nodes;
    s->n_nodes = B.n_nodes;
    s->ordered = B.ordered;
    if (B.error) FAIL("a BVH leaf holds more than 65535 primitives");
    s->build_seconds = now_s() - t0;
    s->world_min = V3(s->nodes[0].bmin[0], s->nodes[0].bmin[1], s->nodes[0].bmin[2]);
    s->world_max = V3(s->nodes[0].bmax[0], s->nodes[0].bmax[1], s->nodes[0].bmax[2]);

    /* lights */
    s->lx = (light_extra*)calloc((size_t)(s->n_lights + 1), sizeof(light_extra));
    s->prim_light = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->n_prims);
    for (int32_t i = 0; i < s->n_prims; ++i) s->prim_light[i] = -1;
    s->inf_lights = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->n_lights + 1));
    s->n_inf = 0;
    for (int32_t i = 0; i < s->n_lights; ++i) {
        const BbLight* L = &s->lights[i];
        light_extra* x = &s->lx[i];
        if (L->type == BB_LIGHT_SPOT) {
            /* spot_light.cpp:7-24 */
            x->z = vneg(vnormalize(V3(L->direction[0], L->direction[1], L->direction[2])));
            x->cos_min = cosf(L->angle_falloff_start * orc_pi / 180.0f);
            x->cos_max = cosf(L->angle_max * orc_pi / 180.0f);
        } else if (L->type == BB_LIGHT_UNIFORM_INFINITE) {
            /* uniform_infinite_light.cpp:14-17, bounding_box.h:524-528 */
            v3 c = vmul(vadd(s->world_min, s->world_max), 0.5f);
            int inside = !(s->world_min.x > c.x || s->world_max.x < c.x) && !(s->world_min.y > c.y || s->world_max.y < c.y) &&
                         !(s->world_min.z > c.z || s->world_max.z < c.z);
            x->world_radius = inside ? vlen(vsub(s->world_max, c)) : 0.0f;
            s->inf_lights[s->n_inf++] = i;
        } else if (L->type == BB_LIGHT_DIFFUSE_AREA) {
            if (L->primitive < 0 || L->primitive >= s->n_prims) FAIL("light %d: bad primitive", i);
            if (s->prims[L->primitive].shape_kind != BB_SHAPE_TRIANGLE) FAIL("light %d: area lights must be triangles", i);
            s->prim_light[L->primitive] = i;
        } else if (L->type == BB_LIGHT_DIRECTIONAL) {
            /* directional_light.cpp:7-17: w = Normalize(direction), the world's bounding sphere */
            x->z = vnormalize(V3(L->direction[0], L->direction[1], L->direction[2]));
            v3 c = vmul(vadd(s->world_min, s->world_max), 0.5f);
            int inside = !(s->world_min.x > c.x || s->world_max.x < c.x) && !(s->world_min.y > c.y || s->world_max.y < c.y) &&
                         !(s->world_min.z > c.z || s->world_max.z < c.z);
            x->world_radius = inside ? vlen(vsub(s->world_max, c)) : 0.0f;
        } else if (L->type != BB_LIGHT_POINT) {
            FAIL("light %d: unknown type", i);
        }
    }
    /* PowerLightSampler::Init + Distribution1D (sampling.h:345-375) */
    int32_t n = s->n_lights;
    s->func = (float*)calloc((size_t)(n + 1), sizeof(float));
    s->cdf = (float*)calloc((size_t)(n + 2), sizeof(float));
    for (int32_t i = 0; i < n; ++i) s->func[i] = slum(light_phi(s, i));
    s->cdf[0] = 0.0f;
    for (int32_t i = 1; i < n + 1; ++i) s->cdf[i] = s->cdf[i - 1] + s->func[i - 1] / (float)n;
    s->func_integral = s->cdf[n];
    if (s->func_integral == 0.0f) {
        for (int32_t i = 1; i < n + 1; ++i) s->cdf[i] = (float)i / (float)n;
    } else {
        for (int32_t i = 1; i < n + 1; ++i) s->cdf[i] /= s->func_integral;
    }
    return s;
}

int32_t orc_scene_node_count(const OrcScene* s) { return s->n_nodes; }
double orc_scene_build_seconds(const OrcScene* s) { return s->build_seconds; }

void orc_scene_bvh(const OrcScene* s, BbBvhNode* nodes_out, int32_t* ordered_out)
{
    if (nodes_out) memcpy(nodes_out, s->nodes, sizeof(BbBvhNode) * (size_t)s->n_nodes);
    if (ordered_out) memcpy(ordered_out, s->ordered, sizeof(int32_t) * (size_t)s->n_prims);
}

void orc_scene_world_bounds(const OrcScene* s, float bmin[3], float bmax[3])
{
    bmin[0] = s->world_min.x;
    bmin[1] = s->world_min.y;
    bmin[2] = s->world_min.z;
    bmax[0] = s->world_max.x;
    bmax[1] = s->world_max.y;
    bmax[2] = s->world_max.z;
}

/* ============================================================ integrator */

/* PathIntegrator::SampleDirectLight (path_integrator.cpp:152-194) */
static spec sample_direct_light(const OrcScene* s, v3 wo, const isect_t* is, const bsdf_t* bsdf, pcg32* g, spec beta)
{
    float u0 = next1d(g);
    v2 u12 = next2d(g);
    if (s->n_lights == 0) return S1(0.0f);
    float pmf;
    int32_t li = light_select(s, u0, &pmf);
    light_sample ls;
    if (!light_sample_li(s, li, is->point, u12, &ls)) return S1(0.0f);
    float bsdf_pdf_v = bsdf_pdf(bsdf, wo, ls.wi);
    if (sblack(ls.Li) || bsdf_pdf_v == 0.0f) return S1(0.0f);
    if (bvh_raycast(s, is->point, ls.wi, 1e-4f, ls.visibility, 1, NULL, NULL)) return S1(0.0f);
    float light_pdf = pmf * ls.pdf;
    spec f_cos = smulf(bsdf_f(bsdf, wo, ls.wi), vabsdot(is->ns, ls.wi));
    if (s->lights[li].type == BB_LIGHT_SPOT || s->lights[li].type == BB_LIGHT_POINT ||
        s->lights[li].type == BB_LIGHT_DIRECTIONAL) /* delta lights */
        return sdivf(smul(smul(beta, ls.Li), f_cos), light_pdf);
    float mis = power_heuristic(light_pdf, bsdf_pdf_v);
    return sdivf(smul(smul(smulf(beta, mis), ls.Li), f_cos), light_pdf);
}

/* PathIntegrator::Li (path_integrator.cpp:26-150) */
static spec path_li(const OrcScene* s, const BbRenderParams* P, v3 ro, v3 rd, pcg32* g)
{
    int32_t bounce = 0;
    spec L = S1(0.0f), beta = S1(1.0f);
    int specular_bounce = 0;
    float eta_scale = 1.0f;
    float prev_bsdf_pdf = 0.0f;
    for (;;) {
        isect_t is;
        if (!bvh_raycast(s, ro, rd, 1e-4f, INFINITY, 0, &is, NULL)) {
            if (bounce == 0 || specular_bounce) {
                for (int32_t k = 0; k < s->n_inf; ++k) {
                    const BbLight* Lt = &s->lights[s->inf_lights[k]];
                    L = sadd(L, smul(beta, smulf(S3(Lt->spectrum[0], Lt->spectrum[1], Lt->spectrum[2]), Lt->scale)));
                }
            } else {
                for (int32_t k = 0; k < s->n_inf; ++k) {
                    int32_t li = s->inf_lights[k];
                    const BbLight* Lt = &s->lights[li];
                    float light_pdf = orc_inv_four_pi * light_pmf(s, li);
                    float mis = power_heuristic(prev_bsdf_pdf, light_pdf);
                    L = sadd(L, smul(smulf(beta, mis), smulf(S3(Lt->spectrum[0], Lt->spectrum[1], Lt->spectrum[2]), Lt->scale)));
                }
            }
            break;
        }
        finish_isect(s, is.ro, is.rd, &is);
        v3 wo = vnormalize(vneg(rd));

        int32_t al = s->prim_light[is.prim];
        if (al >= 0) {
            const BbLight* Lt = &s->lights[al];
            spec Le = (is.front_face || Lt->two_sided) ? S3(Lt->spectrum[0], Lt->spectrum[1], Lt->spectrum[2]) : S1(0.0f);
            if (!sblack(Le)) {
                if (bounce == 0 || specular_bounce) {
                    L = sadd(L, smul(beta, Le));
                } else {
                    float light_pdf =
                        triangle_pdf_solid_angle(s, s->prims[is.prim].shape_index, &is, rd) * light_pmf(s, al);
                    float mis = power_heuristic(prev_bsdf_pdf, light_pdf);
                    L = sadd(L, smul(smulf(beta, mis), Le));
                }
            }
        }

        if (bounce++ >= P->max_bounces) break;

        int32_t mi = s->prims[is.prim].material;
        if (mi < 0) {
            /* null material: continue straight, path_integrator.cpp:95-100 */
            ro = is.point;
            rd = vneg(wo);
            --bounce;
            continue;
        }
        /* Intersection::GetBSDF (intersection.cpp:32-52): normal map, then the material */
        if (s->mats[mi].tex_normal) normal_mapping(s, s->mats[mi].tex_normal - 1, wo, &is);
        bsdf_t bsdf;
        make_bsdf(s, &s->mats[mi], is.uv, is.ns, is.ts, is.front_face, &bsdf);
        int flags = bxdf_flags(&bsdf.bx);
        if (flags & (BX_DIFFUSE | BX_GLOSSY)) L = sadd(L, sample_direct_light(s, wo, &is, &bsdf, g, beta));

        /* Sample_f(wo, Next1D(), Next2D()): u12 drawn first (g++ order) */
        v2 u12 = next2d(g);
        float u0 = next1d(g);
        bsample bs;
        if (!bsdf_sample(&bsdf, wo, u0, u12, &bs)) break;
        specular_bounce = (bs.flags & BX_SPECULAR) != 0;
        if (bs.flags & BX_TRANSMISSION) eta_scale *= sqrf(bs.eta);
        prev_bsdf_pdf = bs.stochastic ? bsdf_pdf(&bsdf, wo, bs.wi) : bs.pdf;
        beta = smul(beta, sdivf(smulf(bs.f, vabsdot(is.ns, bs.wi)), bs.pdf));
        ro = is.point;
        rd = bs.wi;
        if (bounce > P->rr_min_bounces) {
            float p = smaxc(beta) * eta_scale;
            if (p < 1.0f) {
                if (next1d(g) > p)
                    break;
                else
                    beta = sdivf(beta, p);
            }
        }
    }
    return L;
}

/* One (pixel, sample) of UniDirectionalRayIntegrator::Render
 * (integrator.cpp:187-196) */
static spec render_sample(const OrcScene* s, const camera_t* cam, const BbRenderParams* P, int32_t px, int32_t py,
                          int32_t sample, int32_t* draws)
{
    pcg32 g;
    sampler_start(&g, px, py, P->seed, sample);
    /* SampleRay(&ray, pixel, Next2D(), Next2D()): u1 (lens) drawn first (g++ order) */
    v2 u1 = next2d(&g);
    v2 u0 = next2d(&g);
    v3 o, d;
    camera_sample_ray(cam, px, py, u0, u1, &o, &d);
    spec L = path_li(s, P, o, d, &g);
    if (draws) *draws = g.draws;
    return L;
}

void orc_sample_li(const OrcScene* s, const BbCameraDesc* cd, const BbRenderParams* P, int32_t px, int32_t py,
                   int32_t sample, float L_out[3], int32_t* draws)
{
    camera_t cam;
    camera_init(&cam, cd);
    spec L = render_sample(s, &cam, P, px, py, sample, draws);
    L_out[0] = L.r;
    L_out[1] = L.g;
    L_out[2] = L.b;
}

/* ================================================================ threads */

typedef struct {
    const OrcScene* s;
    camera_t cam;
    const BbRenderParams* P;
    int32_t x0, y0, x1, y1, tiles_x, tiles;
    float* sum;
    int32_t* count;
    int32_t next;
    pthread_mutex_t lock;
} render_job;

static void* render_worker(void* arg)
{
    render_job* J = (render_job*)arg;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int32_t tile = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (tile >= J->tiles) break;
        int32_t tx = tile % J->tiles_x, ty = tile / J->tiles_x;
        int32_t xa = J->x0 + 16 * tx, ya = J->y0 + 16 * ty;
        int32_t xb = xa + 16 < J->x1 ? xa + 16 : J->x1;
        int32_t yb = ya + 16 < J->y1 ? ya + 16 : J->y1;
        for (int32_t py = ya; py < yb; ++py)
            for (int32_t px = xa; px < xb; ++px) {
                size_t pi = (size_t)py * (size_t)J->cam.width + (size_t)px;
                for (int32_t k = 0; k < J->P->spp; ++k) {
                    spec L = render_sample(J->s, &J->cam, J->P, px, py, J->P->sample_begin + k, NULL);
                    /* Film::AddSample (film.cpp:26-37); NaN/Inf dropped (integrator.cpp:193-196) */
                    if (!snullish(L)) {
                        J->sum[3 * pi + 0] += L.r;
                        J->sum[3 * pi + 1] += L.g;
                        J->sum[3 * pi + 2] += L.b;
                        J->count[pi] += 1;
                    }
                }
            }
    }
    return NULL;
}

double orc_render(const OrcScene* s, const BbCameraDesc* cd, const BbRenderParams* P, int32_t x0, int32_t y0, int32_t x1,
                  int32_t y1, float* sum, int32_t* count, int32_t threads)
{
    render_job J;
    memset(&J, 0, sizeof(J));
    J.s = s;
    camera_init(&J.cam, cd);
    J.P = P;
    J.x0 = x0;
    J.y0 = y0;
    J.x1 = x1;
    J.y1 = y1;
    J.tiles_x = (x1 - x0 + 15) / 16;
    J.tiles = J.tiles_x * ((y1 - y0 + 15) / 16);
    J.sum = sum;
    J.count = count;
    pthread_mutex_init(&J.lock, NULL);
    if (threads < 1) threads = 1;
    double t0 = now_s();
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int32_t i = 0; i < threads; ++i) pthread_create(&th[i], NULL, render_worker, &J);
    for (int32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&J.lock);
    return now_s() - t0;
}

typedef struct {
    const OrcScene* s;
    const BbRay* rays;
    BbHit* hits;
    int32_t n, any_hit, next;
    BbTraceCounters total;
    pthread_mutex_t lock;
} trace_job;

static void* trace_worker(void* arg)
{
    trace_job* J = (trace_job*)arg;
    BbTraceCounters c;
    memset(&c, 0, sizeof(c));
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int32_t start = J->next;
        J->next += 4096;
        pthread_mutex_unlock(&J->lock);
        if (start >= J->n) break;
        int32_t end = start + 4096 < J->n ? start + 4096 : J->n;
        for (int32_t i = start; i < end; ++i) {
            const BbRay* r = &J->rays[i];
            isect_t is;
            memset(&is, 0, sizeof(is));
            c.rays++;
            int hit = bvh_raycast(J->s, V3(r->o[0], r->o[1], r->o[2]), V3(r->d[0], r->d[1], r->d[2]), r->t_min,
                                  r->t_max, J->any_hit, &is, &c);
            BbHit* h = &J->hits[i];
            if (J->any_hit) {
                h->prim = hit ? 1 : -1;
                h->t = h->u = h->v = 0.0f;
            } else if (hit) {
                h->prim = is.prim;
                h->t = is.t;
                h->u = is.u;
                h->v = is.v;
            } else {
                h->prim = -1;
                h->t = h->u = h->v = 0.0f;
            }
        }
    }
    pthread_mutex_lock(&J->lock);
    J->total.rays += c.rays;
    J->total.nodes += c.nodes;
    J->total.prims += c.prims;
    J->total.accepts += c.accepts;
    pthread_mutex_unlock(&J->lock);
    return NULL;
}

void orc_trace(const OrcScene* s, const BbRay* rays, int32_t n, int32_t any_hit, BbHit* hits, int32_t threads,
               BbTraceCounters* counters)
{
    trace_job J;
    memset(&J, 0, sizeof(J));
    J.s = s;
    J.rays = rays;
    J.hits = hits;
    J.n = n;
    J.any_hit = any_hit;
    pthread_mutex_init(&J.lock, NULL);
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int32_t i = 0; i < threads; ++i) pthread_create(&th[i], NULL, trace_worker, &J);
    for (int32_t i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&J.lock);
    if (counters) *counters = J.total;
}

void orc_trace_brute(const OrcScene* s, const BbRay* rays, int32_t n, BbHit* hits)
{
    for (int32_t i = 0; i < n; ++i) {
        const BbRay* r = &rays[i];
        v3 o = V3(r->o[0], r->o[1], r->o[2]), d = V3(r->d[0], r->d[1], r->d[2]);
        float t_max = r->t_max;
        hits[i].prim = -1;
        hits[i].t = hits[i].u = hits[i].v = 0.0f;
        for (int32_t p = 0; p < s->n_prims; ++p) {
            float t, u, v;
            v3 po = o, pd = d;
            if (prim_intersect(s, p, &po, &pd, r->t_min, t_max, &t, &u, &v)) {
                hits[i].prim = p;
                hits[i].t = t;
                hits[i].u = u;
                hits[i].v = v;
                t_max = t;
            }
        }
    }
}
