/*
 * oracle.h — CPU restatement of Bulbit's unidirectional path tracer.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the GPU library
 * (libbbgpu.so) and the CPU baseline of the benchmark.  Only tests/, the smoke
 * check and the benchmark's CPU-baseline leg load it (liboracle.so, ctypes).
 * The product never calls it and has no CPU fallback through it.
 *
 * It consumes the same scene description as the library (include/bbgpu.h),
 * restates the reference's arithmetic op for op (file:line citations at each
 * function in oracle.c), and renders with a pthread tile pool.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include "../include/bbgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrcScene OrcScene;

/* --- primitives of the arithmetic, exposed for known-answer tests --- */
uint64_t orc_murmur64a(const void* key, int32_t len, uint64_t seed);
/* IndependentSampler seed: Hash(pixel, seed, sample_index) (hash.h:111-120) */
uint64_t orc_pixel_sample_hash(int32_t px, int32_t py, int32_t seed, int32_t sample_index);
/* PCG32: seed with (initstate, initseq) and draw n outputs */
void orc_pcg32_outputs(uint64_t initstate, uint64_t initseq, int32_t n, uint32_t* out);
/* the first n NextFloat() values of a (pixel, sample) stream */
void orc_sampler_floats(int32_t px, int32_t py, int32_t seed, int32_t sample_index, int32_t n, float* out);
/* uniform and cosine sampling helpers (sampling.h) */
void orc_sample_concentric(const float u[2], float out[2]);
void orc_sample_uniform_sphere(const float u[2], float out[3]);
/* Möller–Trumbore for one triangle: returns 1 on hit with t, u, v */
int orc_triangle_intersect(const float p0[3], const float p1[3], const float p2[3], const BbRay* ray, float* t,
                           float* u, float* v);
float orc_fresnel_dielectric(float cos_theta_i, float eta);
float orc_fresnel_complex(float cos_theta_i, float eta, float k);
/* Transform::LookAt (transform.h:112-122) */
void orc_look_at(const float position[3], const float target[3], const float up[3], float q_out[4]);
/* PermutationElement (hash.h:128-159) */
int32_t orc_permutation_element(uint32_t i, uint32_t l, uint32_t p);
/* the reflectance tables of PrepareReflectanceTextures (microfacet.cpp:239-270):
 * E[32 * 32] (alpha-major), E_avg[32], the 1024 u0 and 1024 (u1, u2) samples;
 * any pointer may be NULL */
void orc_reflectance_tables(float* E, float* E_avg, float* u0, float* u12);
/* the compensated dielectric's tables (DielectricMultiScatteringBxDF::
 * PrepareReflectanceTexture, microfacet.cpp:173-236): E and E_inv[32^3] at
 * [(alpha * 32 + cos_theta) * 32 + f0], E_avg and E_inv_avg[32^2] at
 * [alpha * 32 + f0]; any pointer may be NULL */
void orc_dielectric_tables(float* E, float* E_inv, float* E_avg, float* E_inv_avg);
/* TrowbridgeReitzDistribution::E(wo) and E_avg() of a GGX with (ax, ay) */
float orc_ggx_E(float ax, float ay, const float wo[3]);
float orc_ggx_E_avg(float ax, float ay);
/* the camera's Filter::Sample(u) (box, tent or Gaussian, BbCameraDesc.filter_type) */
void orc_filter_sample(const BbCameraDesc* cam, const float u[2], float out[2]);
/* PerspectiveCamera::SampleRay with explicit filter sample u0 and lens sample u1 */
void orc_camera_ray(const BbCameraDesc* cam, int32_t px, int32_t py, const float u0[2], const float u1[2],
                    BbRay* ray_out);

/* --- BSDF evaluation on a standalone material, for invariant tests ---
 * frame: shading normal n and tangent t (world); front_face as at a hit.
 * Returns the BxDF flags (bxdf.h:7-22); f_out/pdf_out for (wo, wi). */
int32_t orc_bsdf_eval(const BbMaterial* mat, const float n[3], const float t[3], int32_t front_face,
                      const float wo[3], const float wi[3], float f_out[3], float* pdf_out);
/* BSDF::Sample_f; returns 1 on success with f, wi, pdf, flags, eta, is_stochastic */
int32_t orc_bsdf_sample(const BbMaterial* mat, const float n[3], const float t[3], int32_t front_face,
                        const float wo[3], float u0, const float u12[2], float f_out[3], float wi_out[3],
                        float* pdf_out, int32_t* flags_out, float* eta_out);

/* --- scenes --- */
/* Builds the BVH (bvh.cpp restated, sequential) and the light set-up.
 * Returns NULL and writes a message to err on invalid input. */
OrcScene* orc_scene_create(const BbSceneDesc* desc, char* err, int32_t err_size);
void orc_scene_destroy(OrcScene* scene);
int32_t orc_scene_node_count(const OrcScene* scene);
/* copies the flattened nodes (n_nodes) and the primitive order (n_primitives) */
void orc_scene_bvh(const OrcScene* scene, BbBvhNode* nodes_out, int32_t* ordered_prims_out);
void orc_scene_world_bounds(const OrcScene* scene, float bmin[3], float bmax[3]);
/* power-sampler table: pmf per light (Scene light order) */
void orc_scene_light_pmf(const OrcScene* scene, float* pmf_out);
double orc_scene_build_seconds(const OrcScene* scene);

/* BVH::Intersect / IntersectAny over a batch of rays on `threads` threads.
 * counters (may be NULL) receives node/prim/accept counts. */
void orc_trace(const OrcScene* scene, const BbRay* rays, int32_t n, int32_t any_hit, BbHit* hits, int32_t threads,
               BbTraceCounters* counters);
/* Brute force closest hit over every primitive in scene order (ties: later wins). */
void orc_trace_brute(const OrcScene* scene, const BbRay* rays, int32_t n, BbHit* hits);

/* Render the pixel rectangle [x0,x1) x [y0,y1) of the camera's image with
 * samples [sample_begin, sample_begin + spp): film_sum / film_count cover the
 * full image (width * height) and are accumulated into.  Returns seconds. */
double orc_render(const OrcScene* scene, const BbCameraDesc* cam, const BbRenderParams* params, int32_t x0, int32_t y0,
                  int32_t x1, int32_t y1, float* film_sum, int32_t* film_count, int32_t threads);

/* Light::Sample_Li of light li (Scene light order) at the point ref for n
 * sample pairs u[2n] (spot_light.cpp:46-68, diffuse_area_light.cpp:42-72 with
 * Triangle::Sample, triangle.cpp:159-229, ...): wi[3n], Li[3n], pdf[n],
 * visibility[n] and ok[n] (the boolean the reference returns). */
void orc_light_sample(const OrcScene* scene, int32_t li, const float ref[3], int32_t n, const float* u, float* wi,
                      float* Li, float* pdf, float* visibility, int32_t* ok);
/* For n rays from o along d[3n]: Triangle::PDF(isect, ray) (triangle.cpp:
 * 247-253), the solid-angle density PathIntegrator::Li weighs an emitter hit
 * with (path_integrator.cpp:67-84) before the light PMF, when the ray's
 * closest hit is area light li's triangle; 0 otherwise. */
void orc_light_hit_pdf(const OrcScene* scene, int32_t li, const float o[3], int32_t n, const float* d, float* pdf);

/* One path sample, returning L (NaN/Inf kept) and the number of draws used. */
void orc_sample_li(const OrcScene* scene, const BbCameraDesc* cam, const BbRenderParams* params, int32_t px, int32_t py,
                   int32_t sample_index, float L_out[3], int32_t* draws_out);

#ifdef __cplusplus
}
#endif

#endif /* ORACLE_H */
