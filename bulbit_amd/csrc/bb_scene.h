// bb_scene.h — device-resident scene layout and the host-side scene object.
//
// HBM layout (all arrays 16-byte aligned, indexed as noted):
//   inner   [n_inner]  64 B per internal node of the reference's BVH
//                      (LinearBVHNode, include/bulbit/bvh.h:46-58), in
//                      pre-order: the boxes of its two children and their
//                      references,
//                        {c1.min.xyz, c1.max.x} {c1.max.yz, c2.min.xy}
//                        {c2.min.z, c2.max.xyz} {ref1, ref2, split axis, 0}
//                      ref >= 0: inner record; ref < 0: leaf whose first
//                      slot is ~ref (its last slot carries kSlotLast)
//   tris    [n_slots]  3 x float4 = 48 B per BVH slot (leaf order):
//                      {p0.xyz, prim}, {e1.xyz, sphere index},
//                      {e2.xyz, kind | kSlotLast on a leaf's last slot}
//                      with e1 = p1 - p0 and e2 = p2 - p0 computed on the host
//                      in float32, bit-identical to Triangle::Intersect's
//                      (src/geometry/triangle.cpp:61-62)
//   slot_shade[n_slots] int4 {v0, v1, v2, material}  (vertex indices, material)
//   slot_light[n_slots] int  area light index or -1
//   normals, tangents [n_vertices] float4 (xyz used)
//   spheres [n_spheres] BbSphere
//   materials [n_materials] GpuMaterial, lights [n_lights] GpuLight,
//   light_cdf [n_lights + 1], light_func [n_lights], light_alias [n_lights]
#pragma once

#include "../../include/bbgpu.h"
#include "bb_math.h"

#include <string>
#include <vector>

namespace bb
{

enum : int32_t
{
    KIND_TRIANGLE = 0,
    KIND_SPHERE = 1,
    KIND_MASK = 0xff,
    kSlotLast = 0x100, // on the last slot of a leaf
    kSlotAlpha = 0x200 // the slot's material has an alpha texture
};

// BxDF flags, include/bulbit/bxdf.h:7-22
enum : int32_t
{
    BX_REFLECTION = 1,
    BX_TRANSMISSION = 2,
    BX_DIFFUSE = 4,
    BX_GLOSSY = 8,
    BX_SPECULAR = 16
};

enum : int32_t
{
    MK_LAMBERT = 0,
    MK_DIELECTRIC = 1,
    MK_CONDUCTOR = 2,
    MK_METALLIC_ROUGHNESS = 3,
    MK_CONDUCTOR_MS = 4, // ConductorMultiScatteringBxDF (energy_compensation)
    MK_DIELECTRIC_MS = 5, // DielectricMultiScatteringBxDF (energy_compensation, rough, eta != 1)
    MK_EON = 6,           // EONBxDF: DiffuseMaterial with roughness > 0 (roughness kept in ax and ay)
    MK_MIRROR = 7         // SpecularReflectionBxDF: MirrorMaterial
};

// Reflectance tables of PrepareReflectanceTextures (microfacet.cpp:239-270):
// 32 x 32 GGX albedo E(cos_theta, alpha) stored [alpha][cos_theta] and its
// cosine-weighted average E_avg(alpha); computed once per process.
constexpr int32_t kTableSize = 32;
constexpr int32_t kTableSamples = 1024;

struct ReflectanceTables
{
    float ggx_E[kTableSize * kTableSize];
    float ggx_E_avg[kTableSize];
};

const ReflectanceTables& reflectance_tables();

// DielectricMultiScatteringBxDF's tables (microfacet.cpp:173-236): E and E_inv
// at (f0 = x, cos_theta = y, alpha = z), [(z * 32 + y) * 32 + x]; E_avg and
// E_inv_avg at (f0 = x, alpha = y), [y * 32 + x].  On the device they are one
// array in this order (kDielectricTableFloats floats).
struct DielectricTables
{
    float E[kTableSize * kTableSize * kTableSize];
    float E_inv[kTableSize * kTableSize * kTableSize];
    float E_avg[kTableSize * kTableSize];
    float E_inv_avg[kTableSize * kTableSize];
};
constexpr int32_t kDielectricTableFloats = 2 * kTableSize * kTableSize * kTableSize + 2 * kTableSize * kTableSize;

const DielectricTables& dielectric_tables();

// Material with its BxDF parameters precomputed on the host
// (the Material::GetBSDF arithmetic that does not depend on the hit).
struct GpuMaterial
{
    int32_t kind;
    float eta;   // dielectric eta (front face); the back face uses 1 / eta
    float ax, ay;
    spec r;      // Lambert reflectance | dielectric Sqrt(reflectance) | conductor lobe reflectance
    spec ceta;   // conductor eta
    spec ck;     // conductor k
    spec color;  // metallic-roughness base colour
    float metallic;
    int32_t flags_front; // BxDF flags (front and back face give the same flags)
    int32_t type;        // BB_MAT_* of the description
    int32_t energy_compensation;
    // texture slots (index into SceneView::textures, -1: the constants above)
    int32_t tex_color, tex_u, tex_v, tex_metallic, tex_normal, tex_alpha;
    int32_t textured;    // any of tex_color, tex_u, tex_v, tex_metallic
    int32_t pad;
};

// A texture (BbTexture): constant value, bilinear image, or checker of two
// non-checker textures a and b
struct GpuTexture
{
    int32_t type, channels, width, height;
    int32_t filter, a, b, pad;
    float res_x, res_y;
    float value[3];
    int32_t texel_offset; // first float in SceneView::texels
};

struct GpuLight
{
    int32_t type;
    int32_t two_sided;
    float cos_min, cos_max; // spot
    f3 p;                   // spot position | triangle p0
    float world_radius;     // infinite
    f3 e1, e2;              // triangle edges
    f3 z;                   // spot frame z
    f3 n;                   // triangle unit normal (Triangle::Sample)
    spec L;                 // spot intensity | infinite scale * l | area emission
    float area;             // triangle area (Triangle::Sample, Triangle::Area)
    float pmf;              // PowerLightSampler::EvaluatePMF
    float pad;
};

struct SceneView
{
    const float4* inner; // 4 float4 per internal node
    const float4* tris;  // 3 float4 per slot
    const int4* slot_shade;
    const int32_t* slot_light;
    const float4* normals;
    const float4* tangents;
    const BbSphere* spheres;
    const GpuMaterial* materials;
    const GpuLight* lights;
    const float* light_cdf;
    const float* light_func;
    const float2* light_alias; // alias table over the lights: {probability of keeping, alias index bits}
    const int32_t* inf_lights;
    const float* ggx_E;     // ReflectanceTables::ggx_E
    const float* ggx_E_avg; // ReflectanceTables::ggx_E_avg
    const float* dielectric; // DielectricTables (kDielectricTableFloats), or null when unused
    const GpuTexture* textures;
    const float* texels;
    const float2* uvs;       // per vertex, null when the scene has none (uv = (0, 0))
    int32_t has_alpha;       // some triangle's material has an alpha texture
    int32_t has_textures;    // some material has a textured parameter or a normal map
    int32_t kind_mask;       // bit MK_* set for every BxDF kind the scene's materials use
    int32_t n_lights;
    int32_t n_inf;
    float func_integral;
    int32_t stack_depth;
    int32_t has_null_material;
    float root_box[6];  // root node box: min.xyz, max.xyz
    int32_t root_ref;   // the root as a reference (inner record 0, or a leaf)
    // Two-level records (8 float4 each) for the traversal kernels that keep
    // no node counters: record of internal node N holds the entries of N's
    // children, an inner child replaced by its own two children (so up to 4
    // entries: node N's grandchildren, or a child that is a leaf):
    //   {b0.min.xyz, b0.max.x} {b0.max.yz, b1.min.xy} {b1.min.z, b1.max.xyz}
    //   {b2.min.xyz, b2.max.x} {b2.max.yz, b3.min.xy} {b3.min.z, b3.max.xyz}
    //   {ref0, ref1, ref2, ref3} {meta, 0, 0, 0}
    // entries 0, 1 come from child1 (N + 1), entries 2, 3 from child2; an
    // entry is a two-level record (>= 0) or a leaf (~first slot).  meta:
    // bits 0-1 N's split axis, 2-3 child1's, 4-5 child2's, bit 6: entry 1
    // exists (child1 is inner), bit 7: entry 3 exists (child2 is inner).
    const float4* quads;
    int32_t quad_root;        // the root as a two-level reference (record 0, or a leaf)
    int32_t quad_stack_depth; // most stack entries a two-level traversal holds (+ 1)
};

// Host-side build products
struct HostBvh
{
    std::vector<BbBvhNode> nodes;
    std::vector<int32_t> ordered; // slot -> primitive
    int32_t max_depth = 0;
};

// Walker's alias table for weights w[0..n) (Vose's construction, in double):
// draw i uniformly, keep it with probability prob[i], else take alias[i].
// All-zero weights select uniformly.
void build_alias_table(const float* w, int32_t n, float* prob, int32_t* alias);

// Builds the reference's SAH BVH (src/accel/bvh.cpp:56-297) on the host.
bool build_bvh(const BbSceneDesc& d, HostBvh& out, std::string& err);

} // namespace bb

struct BbScene
{
    int device = 0;
    int32_t n_prims = 0, n_slots = 0, n_nodes = 0, n_lights = 0, n_inf = 0;
    int32_t max_depth = 0;
    float world_min[3], world_max[3];
    double build_seconds = 0.0;
    size_t device_bytes = 0;
    bb::SceneView view{};
    std::vector<void*> allocations;
};
