/*
 * bbgpu.h — C ABI of the MI355X path-tracer library (libbbgpu.so).
 *
 * The library replaces the hot path of Bulbit's unidirectional path tracer:
 *   UniDirectionalRayIntegrator::Render  src/integrator/integrator.cpp:162-210
 *   PathIntegrator::Li / SampleDirectLight src/integrator/path_integrator.cpp:26-194
 *   BVH::BVH / Intersect / IntersectAny   src/accel/bvh.cpp:8-361, include/bulbit/bvh.h:101-155
 * behind plain C entry points: opaque handles, plain pointers and sizes, int
 * status codes (0 = ok), no C++ or torch types.  The C++ drop-in classes in
 * include/bulbit/ are the intended callers; tests and benchmarks may call this
 * ABI directly.
 *
 * All floating point data is IEEE float32.  Device pointers are HIP device
 * allocations; a stream argument of NULL means the default stream.
 */
#ifndef BBGPU_H
#define BBGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BBGPU_ABI_VERSION 20

/* status codes */
enum {
    BBGPU_OK = 0,
    BBGPU_ERR_INVALID = 1,     /* bad argument: null pointer, bad size, bad index */
    BBGPU_ERR_UNSUPPORTED = 2, /* a feature outside the implemented subset */
    BBGPU_ERR_DEVICE = 3,      /* a HIP runtime call failed */
    BBGPU_ERR_NOMEM = 4,
};

/* ---------------------------------------------------------------- scene -- */

enum { BB_SHAPE_TRIANGLE = 0, BB_SHAPE_SPHERE = 1 };

/* One entry of Scene::GetPrimitives() (include/bulbit/scene.h:40,116-121,
 * include/bulbit/primitive.h:14-31): a shape plus an optional material.
 * Hit identities reported by the library are indices into this array. */
typedef struct BbPrimitive {
    int32_t shape_kind;  /* BB_SHAPE_* */
    int32_t shape_index; /* triangle index into BbSceneDesc.indices, or sphere index */
    int32_t material;    /* index into BbSceneDesc.materials, -1 = none (pass-through) */
    int32_t reserved;
} BbPrimitive;

/* Sphere(const Transform&, Float radius)  include/bulbit/shapes.h:12-31,
 * src/geometry/sphere.cpp:18-28.  Transform = position, quaternion (x,y,z,w),
 * per-axis scale (include/bulbit/transform.h:9-13). */
typedef struct BbSphere {
    float p[3];
    float q[4];
    float s[3];
    float radius;
} BbSphere;

enum {
    BB_MAT_DIFFUSE = 0,            /* DiffuseMaterial     src/material/diffuse_material.cpp:21-36 */
    BB_MAT_DIELECTRIC = 1,         /* DielectricMaterial  src/material/dielectric_material.cpp:28-52 */
    BB_MAT_CONDUCTOR = 2,          /* ConductorMaterial(eta, k, ...)  conductor_material.cpp:52-99 */
    BB_MAT_CONDUCTOR_R = 3,        /* ConductorMaterial(R, ...)       conductor_material.cpp:33-50 */
    BB_MAT_METALLIC_ROUGHNESS = 4, /* MetallicRoughnessMaterial  metallic_roughness_material.cpp:28-39 */
    BB_MAT_MIRROR = 5,             /* MirrorMaterial(R)   src/material/mirror_material.cpp:18-23; R in color */
};

/* Materials.  Field meaning per type:
 *   DIFFUSE:      color = reflectance, u_roughness = roughness (only 0 is supported: Lambertian)
 *   DIELECTRIC:   eta[0] = eta, color = reflectance, u/v_roughness
 *   CONDUCTOR:    eta, k, color = lobe reflectance (1,1,1 when absent), u/v_roughness
 *   CONDUCTOR_R:  k = physical reflectance R, color = lobe reflectance, u/v_roughness
 *   METALLIC_ROUGHNESS: color = basecolor, metallic, u/v_roughness
 * energy_compensation selects the *MultiScattering BxDF variants.
 * Texture slots hold 1 + an index into BbSceneDesc.textures, or 0 for the
 * constant above (so a zeroed struct is a constant material):
 *   tex_color       the reflectance / basecolor texture (3 channels)
 *   tex_u_roughness, tex_v_roughness, tex_metallic   (1 channel; not on DIFFUSE)
 *   tex_normal      normal map, Intersection::GetBSDF's NormalMapping
 *                   (src/geometry/intersection.cpp:9-30, 3 channels)
 *   tex_alpha       alpha, Primitive::Intersect's stochastic alpha test
 *                   (include/bulbit/primitive.h:45-77, 1 channel; triangles
 *                   only, DielectricMaterial has none) */
typedef struct BbMaterial {
    int32_t type;
    int32_t energy_compensation;
    float color[3];
    float eta[3];
    float k[3];
    float u_roughness;
    float v_roughness;
    float metallic;
    int32_t tex_color;
    int32_t tex_u_roughness;
    int32_t tex_v_roughness;
    int32_t tex_metallic;
    int32_t tex_normal;
    int32_t tex_alpha;
} BbMaterial;

enum {
    BB_TEX_CONSTANT = 0, /* ConstantTexture  include/bulbit/textures.h:11-33 */
    BB_TEX_IMAGE = 1,    /* ImageTexture (bilinear)  textures.h:35-107 */
    BB_TEX_CHECKER = 2,  /* CheckerTexture  textures.h:109-142 */
};
enum { BB_TEXCOORD_REPEAT = 0, BB_TEXCOORD_CLAMP = 1 }; /* TexCoordFilter, include/bulbit/texture.h:20-46 */

/* A Float (channels 1) or Spectrum (channels 3) texture.
 *   CONSTANT: value (value[0] for one channel)
 *   IMAGE:    width x height texels of `channels` floats, texel (x, y) at
 *             x + y * width (Image<T>::operator[]), linear values; filter
 *   CHECKER:  checker = 1 + the indices of textures a and b (constant or
 *             image, with this texture's channel count), resolution */
typedef struct BbTexture {
    int32_t type;
    int32_t channels;
    int32_t width;
    int32_t height;
    int32_t filter;
    int32_t checker[2];
    float resolution[2];
    float value[3];
    const float* texels;
} BbTexture;

enum {
    BB_LIGHT_SPOT = 0,              /* SpotLight            src/light/spot_light.cpp:7-24 */
    BB_LIGHT_UNIFORM_INFINITE = 1,  /* UniformInfiniteLight src/light/uniform_infinite_light.cpp:7-12 */
    BB_LIGHT_DIFFUSE_AREA = 2,      /* DiffuseAreaLight     src/light/diffuse_area_light.cpp:9-15 */
    BB_LIGHT_POINT = 3,             /* PointLight           src/light/point_light.cpp:7-13 */
    BB_LIGHT_DIRECTIONAL = 4,       /* DirectionalLight     src/light/directional_light.cpp:7-17 */
};

/* Constructor arguments of the light types, in Scene::GetLights() order
 * (the order fixes the power-sampler CDF, power_light_sampler.cpp:8-23). */
typedef struct BbLight {
    int32_t type;
    int32_t primitive;          /* DIFFUSE_AREA: index into primitives (a triangle) */
    int32_t two_sided;          /* DIFFUSE_AREA */
    int32_t reserved;
    float position[3];          /* SPOT, POINT */
    float direction[3];         /* SPOT, DIRECTIONAL: the constructor's direction argument */
    float spectrum[3];          /* SPOT, POINT, DIRECTIONAL intensity / INFINITE l / AREA emission */
    float scale;                /* INFINITE scale */
    float angle_max;            /* SPOT, degrees */
    float angle_falloff_start;  /* SPOT, degrees */
} BbLight;

/* Host-side scene description.  Vertex data is in world space (the Mesh
 * constructor's transform already applied, src/geometry/mesh.cpp:8-39).
 * normals and tangents are required when triangles are present; uvs may be
 * NULL (every texture coordinate is then (0, 0)); they feed texture lookups. */
typedef struct BbSceneDesc {
    int32_t n_vertices;
    const float* positions;  /* 3 * n_vertices */
    const float* normals;    /* 3 * n_vertices */
    const float* tangents;   /* 3 * n_vertices */
    const float* uvs;        /* 2 * n_vertices, may be NULL */
    int32_t n_triangles;
    const int32_t* indices;  /* 3 * n_triangles */
    int32_t n_spheres;
    const BbSphere* spheres;
    int32_t n_primitives;
    const BbPrimitive* primitives;
    int32_t n_materials;
    const BbMaterial* materials;
    int32_t n_lights;
    const BbLight* lights;
    int32_t n_textures;
    const BbTexture* textures;
} BbSceneDesc;

/* LinearBVHNode, include/bulbit/bvh.h:46-58: 32 bytes, same field meaning. */
typedef struct BbBvhNode {
    float bmin[3];
    float bmax[3];
    int32_t offset;   /* leaf: first slot in the ordered primitive list; internal: child2 index */
    uint16_t count;   /* leaf: primitive count; internal: 0 */
    uint8_t axis;     /* internal: split axis; leaf: bit 0 set when the leaf holds a sphere */
    uint8_t pad;
} BbBvhNode;

typedef struct BbScene BbScene; /* opaque */

typedef struct BbSceneInfo {
    int32_t n_nodes;
    int32_t n_primitives;
    int32_t n_lights;
    int32_t n_infinite_lights;
    float world_min[3];       /* root box, BVH::GetAABB  src/accel/bvh.cpp:363-366 */
    float world_max[3];
    double build_seconds;     /* host BVH build time */
    size_t device_bytes;      /* device memory held by the scene */
} BbSceneInfo;

/* ----------------------------------------------------------- traversal -- */

/* One ray: origin, t_min, direction, t_max (8 floats, 32 bytes). */
typedef struct BbRay {
    float o[3];
    float t_min;
    float d[3];
    float t_max;
} BbRay;

/* Closest hit: prim = index into BbSceneDesc.primitives (-1 = miss), t, and
 * the barycentrics (u, v) of the triangle test (0 for spheres).  Any-hit
 * queries fill prim with 1 (occluded) or -1 only. */
typedef struct BbHit {
    int32_t prim;
    float t;
    float u;
    float v;
} BbHit;

/* Traversal counters, summed over a batch: node boxes tested, primitive tests,
 * accepted closer candidates.  They equal the work of the reference's
 * RayCast (include/bulbit/bvh.h:101-155) on the same rays. */
typedef struct BbTraceCounters {
    uint64_t rays;
    uint64_t nodes;
    uint64_t prims;
    uint64_t accepts;
} BbTraceCounters;

/* ------------------------------------------------------------- render -- */

/* The camera's reconstruction filter (include/bulbit/filters.h) */
enum {
    BB_FILTER_BOX = 0,      /* BoxFilter(extent = 1)       src/reconstruction/box_filter.cpp:17-21 */
    BB_FILTER_TENT = 1,     /* TentFilter(extent = 2)      src/reconstruction/tent_filter.cpp:22-30 */
    BB_FILTER_GAUSSIAN = 2, /* GaussianFilter(sigma = 0.5, extent = 3)  gaussian_filter.cpp:8-43 */
};

/* PerspectiveCamera(const Transform&, vfov, aperture_radius, focus_distance,
 * resolution, medium, filter): src/camera/perspective_camera.cpp:8-52.
 * The transform's scale does not enter the camera.  A zero-filled tail
 * (filter_type 0) is the box filter. */
typedef struct BbCameraDesc {
    float position[3];
    float orientation[4]; /* quaternion x, y, z, w */
    float vfov_deg;
    float aperture_radius;
    float focus_distance;
    int32_t width;
    int32_t height;
    float filter_extent;  /* the filter's extent */
    int32_t filter_type;  /* BB_FILTER_* */
    float filter_sigma;   /* GAUSSIAN: sigma */
} BbCameraDesc;

/* PathIntegrator(accel, lights, sampler, max_bounces, rr_min_bounces,
 * regularize_bsdf) + IndependentSampler(spp, seed):
 * include/bulbit/integrators.h:92-114, src/sampler/independent_sampler.cpp:8-12.
 * The remaining fields tune the GPU pipeline; 0 selects the default. */
typedef struct BbRenderParams {
    int32_t spp;
    int32_t seed;
    int32_t max_bounces;
    int32_t rr_min_bounces;
    int32_t regularize_bsdf;  /* only 0 is supported */
    int32_t wavefront_paths;  /* paths in flight per pass (0 = auto: as many whole samples per pixel as an
                                 eighth of the device memory holds in path state, at least 2^22 paths) */
    int32_t trace_lds_stack;  /* traversal stack entries per lane kept in LDS (0 = auto: 12) */
    int32_t sample_begin;     /* render samples [sample_begin, sample_begin+spp) */
    int32_t trace_refill;     /* a closest-hit traversal wave takes new rays once this many of its lanes are idle (0 = 32) */
    int32_t shadow_refill;    /* the same for shadow rays (0 = trace_refill's default) */
    int32_t graph;            /* 1: launch the render as a HIP graph, captured on first use and kept with the
                                 scene for calls that differ only in sample_begin (speed only; not with stats
                                 or the radiance cache) */
    int32_t trace_block;      /* threads per block of the traversal kernels: 64, 128 or 256 (0 = 128) */
    int32_t trace_stack;      /* BB_STACK_LDS (0): stacks in LDS, deeper entries in a global overflow area;
                                 BB_STACK_SCRATCH: each lane's stack in a private scratch array of 64 entries
                                 (no LDS; scenes whose traversal needs more are refused) */
    int32_t light_sampler;    /* BB_LIGHTS_POWER (0): PowerLightSampler's CDF search, as the reference;
                                 BB_LIGHTS_ALIAS: an alias table over the same powers (opt-in: same
                                 probabilities and estimator, a different draw-to-light map, so images
                                 agree in expectation, not sample for sample) */
} BbRenderParams;

enum { BB_LIGHTS_POWER = 0, BB_LIGHTS_ALIAS = 1 };
enum { BB_STACK_LDS = 0, BB_STACK_SCRATCH = 1 };

/* Per-render statistics, filled when a pointer is passed to bbgpu_render.
 * Kernel times come from HIP events recorded around each launch on the
 * stream it runs on.  On scenes of 8 primitives or more (not in
 * counting frames) each depth's shadow rays run on a second stream beside
 * the next depth's closest hits, so the two sums overlap in time. */
typedef struct BbRenderStats {
    double ms_total;          /* render wall time on the stream (events) */
    double ms_trace_closest;  /* summed over all closest-hit launches */
    double ms_trace_shadow;   /* summed over all any-hit launches, each from
                                 the end of its depth's shading to its own end */
    double ms_shade;          /* raygen + shade + film kernels */
    int64_t launches_closest;
    int64_t launches_shadow;
    int64_t rays_closest;     /* closest-hit rays traced */
    int64_t rays_shadow;      /* shadow rays traced */
    int64_t paths;            /* camera samples */
    BbTraceCounters closest;  /* filled only with count_traversal = 1 */
    BbTraceCounters shadow;
} BbRenderStats;

/* ------------------------------------------------------------ entry points */

/* ABI version, BBGPU_ABI_VERSION. */
int bbgpu_abi_version(void);

/* Copies the last error message of the calling thread into buf. */
int bbgpu_last_error(char* buf, size_t buf_size);

/* Number of visible HIP devices (0 when none); selects nothing. */
int bbgpu_device_count(int32_t* count);

/* The reflectance tables that PrepareReflectanceTextures computes
 * (src/bxdf/microfacet.cpp:31-75, 173-270) and the library uploads with a
 * scene that needs them, as computed on the host.  Host-only.  Writes the
 * table's floats to out when capacity is large enough and always sets
 * *length_out.  Layouts, index [..] from the slowest axis:
 *   BB_TABLE_GGX_E             [alpha 32][cos_theta 32]
 *   BB_TABLE_GGX_E_AVG         [alpha 32]
 *   BB_TABLE_DIELECTRIC_E      [alpha 32][cos_theta 32][f0 32]  (eta >= 1)
 *   BB_TABLE_DIELECTRIC_E_INV  same, for the inverted interface
 *   BB_TABLE_DIELECTRIC_E_AVG  [alpha 32][f0 32], and _E_INV_AVG
 * (f0 = MapIORtoF0, scattering.h:188-208).  The reference keeps them in
 * texture objects and has no accessor; this entry point is for inspection. */
enum {
    BB_TABLE_GGX_E = 0,
    BB_TABLE_GGX_E_AVG = 1,
    BB_TABLE_DIELECTRIC_E = 2,
    BB_TABLE_DIELECTRIC_E_INV = 3,
    BB_TABLE_DIELECTRIC_E_AVG = 4,
    BB_TABLE_DIELECTRIC_E_INV_AVG = 5
};
int bbgpu_reflectance_table(int32_t table, float* out, int32_t capacity, int32_t* length_out);

/* Walker's alias table of n non-negative weights, as the render's
 * BB_LIGHTS_ALIAS mode builds it over the light powers (no counterpart in
 * the reference, which searches a CDF): draw i uniformly, keep it with
 * probability prob_out[i], else take alias_out[i].  All-zero weights select
 * uniformly.  Host-only: makes no HIP call. */
int bbgpu_alias_table(const float* weights, int32_t n, float* prob_out, int32_t* alias_out);

/* Builds the SAH BVH on the host exactly as BVH::BVH
 * (src/accel/bvh.cpp:8-49, 56-271, 273-297) and returns the flattened nodes and
 * the primitive order.  Host-only: makes no HIP call.  nodes_out must hold
 * 2 * n_primitives - 1 entries, ordered_prims_out n_primitives entries. */
int bbgpu_build_bvh(const BbSceneDesc* desc, BbBvhNode* nodes_out, int32_t* ordered_prims_out,
                    int32_t* n_nodes_out);

/* Replaces BVH construction + Integrator::Integrator's light set-up
 * (src/accel/bvh.cpp:8-49, src/integrator/integrator.cpp:107-152): builds the
 * BVH, preprocesses lights, builds the power-sampler CDF
 * (src/light_sampler/power_light_sampler.cpp:8-23) and uploads everything to
 * the current HIP device.  The library owns the device memory. */
int bbgpu_scene_create(const BbSceneDesc* desc, BbScene** scene_out);
int bbgpu_scene_destroy(BbScene* scene);
int bbgpu_scene_info(const BbScene* scene, BbSceneInfo* info_out);

/* Replaces BVH::Intersect / BVH::IntersectAny over a ray batch
 * (src/accel/bvh.cpp:299-361).  rays_dev and hits_dev are device pointers for
 * the fast path; buffers in host memory are also accepted and staged through
 * device memory (PCIe copies included).  counters (host pointer, may be NULL)
 * requests traversal counts.  The call returns when the batch is traced. */
int bbgpu_trace(BbScene* scene, const BbRay* rays_dev, int32_t n_rays, int32_t any_hit, BbHit* hits_dev,
                BbTraceCounters* counters, void* stream);

/* bbgpu_trace without any device allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and counters as bbgpu_trace, but rays_dev and
 * hits_dev must be device memory (host buffers are rejected).  The kernels'
 * scratch (error flag, work-queue counter, traversal counters and the
 * stacks' overflow area) is one static 16 MiB device array per GPU, held
 * until the process ends; a scene whose stack overflow needs more than that
 * for one block is rejected.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other threads) race. */
int bbgpu_trace_noalloc(BbScene* scene, const BbRay* rays_dev, int32_t n_rays, int32_t any_hit, BbHit* hits_dev,
                        BbTraceCounters* counters, void* stream);

/* Replaces UniDirectionalRayIntegrator::Render with PathIntegrator::Li
 * (src/integrator/integrator.cpp:162-210, path_integrator.cpp:26-194):
 * accumulates samples [sample_begin, sample_begin + spp) of every pixel into
 * film_sum_dev (3 floats per pixel, row-major, added to the existing contents
 * in sample order) and film_count_dev (1 int32 per pixel), as Film::AddSample
 * does (src/reconstruction/film.cpp:26-37).  NaN/Inf samples are dropped.
 * The film buffers are device pointers for the fast path; a film in host
 * memory is also accepted (copied in, accumulated, copied back).
 * stats (host pointer, may be NULL) requests timing; count_traversal = 1 also
 * fills the traversal counters (slower kernels; for measurement only).  The
 * call returns when the render is complete. */
int bbgpu_render(BbScene* scene, const BbCameraDesc* camera, const BbRenderParams* params, float* film_sum_dev,
                 int32_t* film_count_dev, void* stream, BbRenderStats* stats, int32_t count_traversal);

/* ---------------------------------------------------------------- batch -- */

/* One render of a batch: bbgpu_render's arguments (device films only). */
typedef struct BbRenderJob {
    BbScene* scene;
    BbCameraDesc camera;
    BbRenderParams params;
    float* film_sum_dev;
    int32_t* film_count_dev;
} BbRenderJob;

typedef struct BbBatchStats {
    double ms_total;  /* host wall time from the first launch to the last stream's end */
    int64_t samples;  /* camera samples of all jobs */
    int32_t streams;
    double* job_ms;   /* in: NULL or room for n_jobs values; out: each job's completion time
                         from the start of its stream (HIP events) */
} BbBatchStats;

/* Renders independent jobs (no counterpart in Bulbit: many scenes or camera
 * jobs of uneven size): each scene's jobs run in order on one of `streams`
 * HIP streams (0: 16, at most 16; default priority), scenes are assigned
 * largest first to the least loaded stream, and small jobs overlap on the
 * GPU.  Every job's film equals what bbgpu_render would accumulate.  Returns
 * when all jobs are done; stats may be NULL. */
int bbgpu_render_batch(BbRenderJob* jobs, int32_t n_jobs, int32_t streams, BbBatchStats* stats);

/* Picks the traversal refill thresholds, block size (64, 128, 256), stack
 * placement (8, 16 or 24 LDS entries, or the scratch stack) and
 * wavefront_paths for a job by timing short renders of probe_spp samples
 * (coordinate descent, one knob at a time); the knobs change speed, never
 * results.  tuned_out = *params with them set. */
int bbgpu_autotune(BbScene* scene, const BbCameraDesc* camera, const BbRenderParams* params, int32_t probe_spp,
                   BbRenderParams* tuned_out);

/* ------------------------------------------------------- radiance cache -- */

/* A small MLP radiance cache trained online on path-tail radiance (no
 * counterpart in Bulbit; the configuration of the cached path tracer):
 * input = 16 floats per row (position in [0,1]^3, direction, normal,
 * roughness, albedo, 3 unused), frequency-encoded to 74 values; hidden_layers
 * hidden layers of width hidden_width with ReLU; 3 linear outputs (RGB).
 * bf16 weights and activations, fp32 accumulation and master weights, Adam
 * (beta1 0.9, beta2 0.99), relative L2 loss (y - t)^2 / (y^2 + 0.01). */
typedef struct BbNrc BbNrc; /* opaque */

typedef struct BbNrcDesc {
    int32_t hidden_width;   /* 64 (with 4 layers) or 128 (with 5 layers) */
    int32_t hidden_layers;  /* 4 or 5, as above */
    float learning_rate;    /* Adam step size (0: 1e-3) */
    uint32_t seed;          /* initial weights (Xavier uniform from PCG32) */
} BbNrcDesc;

/* number of fp32 parameters of the network desc describes (the weight
 * matrices, k-major, input padded to 80 rows and output to 16 columns);
 * NULL: the 64 x 4 network; -1: a shape that is not built */
int bbgpu_nrc_param_count(const BbNrcDesc* desc);
int bbgpu_nrc_create(const BbNrcDesc* desc, BbNrc** nrc_out);
int bbgpu_nrc_destroy(BbNrc* nrc);
int bbgpu_nrc_get_weights(BbNrc* nrc, float* host_out, int32_t count);
int bbgpu_nrc_set_weights(BbNrc* nrc, const float* host_in, int32_t count);
/* predicted radiance (3 floats per row) of n rows of 16 raw floats; device pointers */
int bbgpu_nrc_query(BbNrc* nrc, const float* inputs_dev, int32_t n, float* out_dev, void* stream);
/* one training step on n rows (targets: 3 floats per row, device pointers):
 * the loss gradient over the batch, then (apply != 0) an Adam update.
 * loss_out and grad_out (host, may be NULL; grad_out holds
 * bbgpu_nrc_param_count(desc) floats) make the call wait for the step. */
int bbgpu_nrc_train(BbNrc* nrc, const float* inputs_dev, const float* targets_dev, int32_t n, int32_t apply,
                    float* loss_out, float* grad_out, void* stream);
/* 1 in *diverged_out when the network has diverged (a training step met a
 * non-finite loss or produced a non-finite weight, or a cached render met a
 * non-finite prediction); new weights (bbgpu_nrc_set_weights) clear it.
 * Waits for the device. */
int bbgpu_nrc_diverged(BbNrc* nrc, int32_t* diverged_out);

/* The cached path tracer: rendering paths end at their query vertex (the
 * first rough (diffuse or glossy) vertex, from vertex 1 on unless
 * query_primary, or vertex query_bounce at the latest) with throughput x
 * cache(vertex); every train_stride-th path is a
 * training path that continues train_extra_bounces vertices further without
 * the cache, and the reflected radiance it measured at each vertex trains
 * the cache after the pass: train_steps Adam steps of up to train_batch
 * rows, taking the pass's rows batch by batch and starting over when they
 * run out (fewer rows than one batch: every step trains on all of them). */
typedef struct BbCacheParams {
    int32_t query_bounce;         /* 2 */
    int32_t train_stride;         /* 16 */
    int32_t train_extra_bounces;  /* 4 */
    int32_t train_batch;          /* rows per optimizer step (0: 65536) */
    int32_t train_steps;          /* optimizer steps per pass (0: inference only) */
    int32_t query_primary;        /* 1: the primary hit may be the query vertex; 0: vertex 1 at the earliest */
    int32_t query_first;          /* the first vertex that may be a query vertex when query_primary is 0
                                     (0: vertex 1; 2: after a 2-bounce prefix) */
    int32_t query_direct;         /* 1: the query vertex takes its direct light (NEE) as the plain tracer does
                                     and the cache stands in for the indirect rest only; 0: the cache stands in
                                     for everything the vertex reflects */
} BbCacheParams;

typedef struct BbCacheStats {
    double ms_total;   /* the whole call on the stream */
    double ms_query;   /* the network's inference (query kernel) */
    double ms_train;   /* the network's optimizer steps (training kernels) */
    int64_t queries;
    int64_t train_rows; /* rows trained on, counted once per step that used them */
    float loss;        /* relative L2 loss of the last step */
    int32_t diverged;  /* 1: the network is diverged (non-finite loss, weight or prediction); the passes from the
                          detection on were rendered by the plain tracer, and so are later calls until new
                          weights are set (bbgpu_nrc_set_weights) */
    int32_t passes_redone; /* passes whose cached samples were dropped and rendered again without the cache */
    double ms_rows;    /* the predictions applied to the paths and the training rows built and counted */
} BbCacheStats;

/* bbgpu_render with the radiance cache (same film and parameter contract;
 * stats, a host pointer, may be NULL).  Row counts stay on the device: the
 * call synchronizes its stream once, at the end.  Divergence: a non-finite
 * loss or weight after a training step, or a non-finite prediction, marks
 * the network diverged.  A pass whose predictions were not finite is dropped
 * from the film and rendered again by the plain tracer; later passes, and
 * later calls until new weights are set, are plain renders (the fallback). */
int bbgpu_render_cached(BbScene* scene, BbNrc* nrc, const BbCameraDesc* camera, const BbRenderParams* params,
                        const BbCacheParams* cache_params, float* film_sum_dev, int32_t* film_count_dev, void* stream,
                        BbCacheStats* stats);

/* Transform::LookAt (include/bulbit/transform.h:112-122): position and
 * orientation quaternion (x, y, z, w) of a camera at `position` looking at
 * `target`.  Host-only helper for callers building a BbCameraDesc. */
int bbgpu_look_at(const float position[3], const float target[3], const float up[3], float orientation_out[4]);

#ifdef __cplusplus
}
#endif

#endif /* BBGPU_H */
