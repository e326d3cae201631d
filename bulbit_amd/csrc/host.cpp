// host.cpp — host side of libbbgpu.so: the SAH BVH build, scene packing and
// upload, error reporting and host-only helpers of the C ABI.
#include "bb_scene.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <future>
#include <memory>

namespace bb
{

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// ------------------------------------------------------------------- BVH

namespace
{

struct BPrim
{
    int32_t index;
    f3 bmin, bmax;
};

struct BNode
{
    f3 bmin, bmax;
    int32_t axis, offset, count;
    int32_t c1, c2;
};

// bounding_box.h:376-381
inline float box_area(f3 mn, f3 mx)
{
    f3 w = mx - mn;
    return 2.0f * ((w.x * w.y) + (w.y * w.z) + (w.z * w.x));
}

inline f3 empty_min() { return mk3(FLT_MAX); }
inline f3 empty_max() { return mk3(-FLT_MAX); }

inline float center(const BPrim& p, int axis) { return (comp(p.bmin, axis) + comp(p.bmax, axis)) * 0.5f; }

// bvh.cpp:162-168 and 215-221
inline int bucket_of(const BPrim& p, int axis, float cmin, float extent)
{
    int b = int(12.0f * (center(p, axis) - cmin) / extent);
    if (b == 12) b = 11;
    return b;
}

struct Builder
{
    std::vector<BPrim> prims;
    std::vector<BNode> pool;
    std::atomic<int32_t> next{ 0 };
    std::atomic<int32_t> too_big{ 0 };

    int32_t alloc() { return next.fetch_add(1, std::memory_order_relaxed); }

    int32_t leaf(int32_t start, int32_t n, f3 mn, f3 mx)
    {
        int32_t id = alloc();
        BNode& nd = pool[id];
        nd.bmin = mn;
        nd.bmax = mx;
        nd.axis = 0;
        // Leaves cover consecutive spans in depth-first order, so the
        // reference's running offset (bvh.cpp:78-85) equals the span start.
        nd.offset = start;
        nd.count = n;
        nd.c1 = nd.c2 = -1;
        if (n > 65535) too_big = 1;
        return id;
    }

    // BVH::BuildRecursive, src/accel/bvh.cpp:56-271
    int32_t build(int32_t start, int32_t end, int depth)
    {
        int32_t n = end - start;
        BPrim* P = prims.data() + start;
        f3 mn = empty_min(), mx = empty_max();
        for (int32_t i = 0; i < n; ++i) {
            mn = vmin(mn, P[i].bmin);
            mx = vmax(mx, P[i].bmax);
        }
        if (box_area(mn, mx) == 0.0f || n == 1) return leaf(start, n, mn, mx);

        f3 cmn = empty_min(), cmx = empty_max();
        for (int32_t i = 0; i < n; ++i) {
            f3 c = (P[i].bmin + P[i].bmax) * 0.5f;
            cmn = vmin(cmn, c);
            cmx = vmax(cmx, c);
        }
        f3 ext = cmx - cmn;
        int axis = 0;
        if (ext.y > ext.x) axis = 1;
        if (ext.z > comp(ext, axis)) axis = 2;
        const float extent = comp(ext, axis);
        if (extent == 0.0f) return leaf(start, n, mn, mx);

        int32_t mid;
        if (n <= 2) {
            mid = n / 2;
            std::nth_element(P, P + mid, P + n,
                             [axis](const BPrim& a, const BPrim& b) { return center(a, axis) < center(b, axis); });
        } else {
            int count[12] = { 0 };
            f3 bmn[12], bmx[12];
            for (int b = 0; b < 12; ++b) {
                bmn[b] = empty_min();
                bmx[b] = empty_max();
            }
            const float cmin = comp(cmn, axis);
            for (int32_t i = 0; i < n; ++i) {
                int b = bucket_of(P[i], axis, cmin, extent);
                count[b]++;
                bmn[b] = vmin(bmn[b], P[i].bmin);
                bmx[b] = vmax(bmx[b], P[i].bmax);
            }
            f3 lmn = empty_min(), lmx = empty_max(), rmn = empty_min(), rmx = empty_max();
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
                float cost = float(left_count[i]) * left_area[i] + float(right_count[i]) * right_area[i];
                if (cost < min_cost) {
                    min_cost = cost;
                    best = i;
                }
            }
            min_cost = 0.5f + min_cost / box_area(mn, mx);
            if (!(min_cost < float(n))) return leaf(start, n, mn, mx);
            // std::partition over a random-access range: libstdc++'s
            // bidirectional form, the same call the reference makes
            BPrim* it = std::partition(P, P + n, [=](const BPrim& p) { return bucket_of(p, axis, cmin, extent) <= best; });
            mid = int32_t(it - P);
        }

        int32_t c1, c2;
        if (n > 64 * 1024 && depth < 16) {
            auto f = std::async(std::launch::async, [this, start, mid, depth] { return build(start, start + mid, depth + 1); });
            c2 = build(start + mid, end, depth + 1);
            c1 = f.get();
        } else {
            c1 = build(start, start + mid, depth + 1);
            c2 = build(start + mid, end, depth + 1);
        }
        int32_t id = alloc();
        BNode& nd = pool[id];
        // BuildNode::InitInternal (include/bulbit/bvh.h:90-97)
        nd.bmin = vmin(pool[c1].bmin, pool[c2].bmin);
        nd.bmax = vmax(pool[c1].bmax, pool[c2].bmax);
        nd.axis = axis;
        nd.offset = 0;
        nd.count = 0;
        nd.c1 = c1;
        nd.c2 = c2;
        return id;
    }
};

// BVH::FlattenBVH, src/accel/bvh.cpp:273-297: depth-first pre-order
void flatten(const Builder& B, int32_t root, HostBvh& out)
{
    struct Item
    {
        int32_t node, parent, depth;
    };
    std::vector<Item> stack;
    stack.push_back({ root, -1, 0 });
    out.nodes.clear();
    out.nodes.reserve(B.next.load());
    out.max_depth = 0;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const BNode& bn = B.pool[it.node];
        int32_t idx = int32_t(out.nodes.size());
        if (it.parent >= 0) out.nodes[it.parent].offset = idx; // this item is child2 of parent
        BbBvhNode ln{};
        ln.bmin[0] = bn.bmin.x;
        ln.bmin[1] = bn.bmin.y;
        ln.bmin[2] = bn.bmin.z;
        ln.bmax[0] = bn.bmax.x;
        ln.bmax[1] = bn.bmax.y;
        ln.bmax[2] = bn.bmax.z;
        out.max_depth = std::max(out.max_depth, it.depth);
        if (bn.count > 0) {
            ln.offset = bn.offset;
            ln.count = uint16_t(bn.count);
            ln.axis = 0;
            out.nodes.push_back(ln);
        } else {
            ln.count = 0;
            ln.axis = uint8_t(bn.axis);
            out.nodes.push_back(ln);
            // child1 is visited next (index + 1); child2 after child1's subtree
            stack.push_back({ bn.c2, idx, it.depth + 1 });
            stack.push_back({ bn.c1, -1, it.depth + 1 });
        }
    }
}

inline f3 load3(const float* a, int64_t i) { return mk3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

inline float i2f(int32_t v)
{
    float f;
    memcpy(&f, &v, 4);
    return f;
}

inline int32_t f2i(float f)
{
    int32_t v;
    memcpy(&v, &f, 4);
    return v;
}

} // namespace

bool validate_desc(const BbSceneDesc& d, std::string& err)
{
    char buf[256];
    if (d.n_primitives <= 0 || !d.primitives) {
        err = "scene has no primitives";
        return false;
    }
    if (d.n_triangles > 0 && (!d.positions || !d.normals || !d.tangents || !d.indices)) {
        err = "triangles need positions, normals, tangents and indices";
        return false;
    }
    if (d.n_spheres > 0 && !d.spheres) {
        err = "spheres pointer is null";
        return false;
    }
    if (d.n_lights > 0 && !d.lights) {
        err = "lights pointer is null";
        return false;
    }
    if (d.n_materials > 0 && !d.materials) {
        err = "materials pointer is null";
        return false;
    }
    for (int64_t i = 0; i < 3 * int64_t(d.n_triangles); ++i)
        if (d.indices[i] < 0 || d.indices[i] >= d.n_vertices) {
            err = "vertex index out of range";
            return false;
        }
    for (int32_t i = 0; i < d.n_primitives; ++i) {
        const BbPrimitive& p = d.primitives[i];
        bool ok = (p.shape_kind == BB_SHAPE_TRIANGLE && p.shape_index >= 0 && p.shape_index < d.n_triangles) ||
                  (p.shape_kind == BB_SHAPE_SPHERE && p.shape_index >= 0 && p.shape_index < d.n_spheres);
        if (!ok || p.material >= d.n_materials) {
            snprintf(buf, sizeof(buf), "primitive %d: bad shape or material index", i);
            err = buf;
            return false;
        }
    }
    return true;
}

bool build_bvh(const BbSceneDesc& d, HostBvh& out, std::string& err)
{
    if (!validate_desc(d, err)) return false;
    const int32_t n = d.n_primitives;
    Builder B;
    B.prims.resize(n);
    B.pool.resize(size_t(2) * size_t(n));
    for (int32_t i = 0; i < n; ++i) {
        const BbPrimitive& p = d.primitives[i];
        BPrim& bp = B.prims[i];
        bp.index = i;
        if (p.shape_kind == BB_SHAPE_TRIANGLE) {
            // Triangle::GetAABB, src/geometry/triangle.cpp:40-52
            const int32_t* v = d.indices + 3 * int64_t(p.shape_index);
            f3 p0 = load3(d.positions, v[0]), p1 = load3(d.positions, v[1]), p2 = load3(d.positions, v[2]);
            f3 off = mk3(kEps * 10.0f);
            bp.bmin = vmin(vmin(p0, p1), p2) - off;
            bp.bmax = vmax(vmax(p0, p1), p2) + off;
        } else {
            // Sphere::GetAABB, src/geometry/sphere.cpp:30-33
            const BbSphere& s = d.spheres[p.shape_index];
            f3 c = mk3(s.p[0], s.p[1], s.p[2]);
            bp.bmin = c - mk3(s.radius);
            bp.bmax = c + mk3(s.radius);
        }
    }
    int32_t root = B.build(0, n, 0);
    if (B.too_big) {
        err = "a BVH leaf holds more than 65535 primitives";
        return false;
    }
    flatten(B, root, out);
    out.ordered.resize(n);
    for (int32_t i = 0; i < n; ++i) out.ordered[i] = B.prims[i].index;
    return true;
}

// ------------------------------------------------------------ scene set-up

namespace
{

bool material_constants(const BbMaterial& m, GpuMaterial& g, std::string& err)
{
    memset(&g, 0, sizeof(g));
    g.type = m.type;
    g.energy_compensation = m.energy_compensation;
    switch (m.type) {
    case BB_MAT_DIFFUSE:
        // DiffuseMaterial::GetBSDF: EONBxDF when roughness > 0, else Lambertian
        // (diffuse_material.cpp:21-36); the EON roughness is kept in ax and ay
        g.kind = m.u_roughness > 0.0f ? MK_EON : MK_LAMBERT;
        g.ax = g.ay = m.u_roughness > 0.0f ? m.u_roughness : 0.0f;
        g.r = mks(m.color[0], m.color[1], m.color[2]);
        g.flags_front = !is_black(g.r) ? (BX_DIFFUSE | BX_REFLECTION) : 0;
        return true;
    case BB_MAT_DIELECTRIC: {
        // DielectricMaterial::GetBSDF (dielectric_material.cpp:28-52), DielectricBxDF (bxdfs.h:80-90)
        g.kind = MK_DIELECTRIC;
        g.eta = m.eta[0];
        g.ax = sqr(m.u_roughness);
        g.ay = sqr(m.v_roughness);
        g.r = sqrt_s(mks(m.color[0], m.color[1], m.color[2]));
        bool smooth = smax(g.ax, g.ay) < 1e-3f;
        // DielectricMultiScatteringBxDF (bxdfs.h:145-247) when compensated; smooth
        // or with eta 1 it takes the plain BxDF's specular branch
        // (dielectric_multi_scattering_bxdf.cpp:11-14, 90-93, 196-244)
        if (m.energy_compensation && !smooth && g.eta != 1.0f) g.kind = MK_DIELECTRIC_MS;
        int f = (g.eta == 1.0f) ? BX_TRANSMISSION : (BX_REFLECTION | BX_TRANSMISSION);
        g.flags_front = f | (smooth ? BX_SPECULAR : BX_GLOSSY);
        return true;
    }
    case BB_MAT_CONDUCTOR:
    case BB_MAT_CONDUCTOR_R: {
        // ConductorMaterial::GetBSDF (conductor_material.cpp:52-99): ConductorBxDF (bxdfs.h:253-277),
        // or ConductorMultiScatteringBxDF (bxdfs.h:303-355) with energy compensation
        g.kind = m.energy_compensation ? MK_CONDUCTOR_MS : MK_CONDUCTOR;
        g.ax = sqr(m.u_roughness);
        g.ay = sqr(m.v_roughness);
        g.r = mks(m.color[0], m.color[1], m.color[2]);
        if (m.type == BB_MAT_CONDUCTOR) {
            g.ceta = mks(m.eta[0], m.eta[1], m.eta[2]);
            g.ck = mks(m.k[0], m.k[1], m.k[2]);
        } else {
            const float hi = 1.0f - 1e-4f;
            spec r = mks(clampf_(m.k[0], 0.0f, hi), clampf_(m.k[1], 0.0f, hi), clampf_(m.k[2], 0.0f, hi));
            spec om = mks(1.0f) - r;
            spec mx = mks(smax(0.0f, om.r), smax(0.0f, om.g), smax(0.0f, om.b));
            g.ceta = mks(1.0f);
            g.ck = (sqrt_s(r) * 2.0f) / sqrt_s(mx);
        }
        bool smooth = smax(g.ax, g.ay) < 1e-3f;
        g.flags_front = smooth ? (BX_SPECULAR | BX_REFLECTION) : (BX_GLOSSY | BX_REFLECTION);
        return true;
    }
    case BB_MAT_MIRROR:
        // MirrorMaterial::GetBSDF (mirror_material.cpp:18-23): SpecularReflectionBxDF (bxdfs.h:48-73)
        g.kind = MK_MIRROR;
        g.r = mks(m.color[0], m.color[1], m.color[2]);
        g.flags_front = !is_black(g.r) ? (BX_SPECULAR | BX_REFLECTION) : 0;
        return true;
    case BB_MAT_METALLIC_ROUGHNESS:
        // MetallicRoughnessMaterial::GetBSDF (metallic_roughness_material.cpp:28-39)
        g.kind = MK_METALLIC_ROUGHNESS;
        g.color = mks(m.color[0], m.color[1], m.color[2]);
        g.metallic = m.metallic;
        g.ax = fmaxf(sqr(m.u_roughness), 1e-3f);
        g.ay = fmaxf(sqr(m.v_roughness), 1e-3f);
        g.flags_front = BX_DIFFUSE | BX_GLOSSY | BX_REFLECTION;
        return true;
    }
    err = "unknown material type";
    return false;
}

template <typename T>
bool upload(BbScene* s, const T* host, size_t count, const T** dev_out, std::string& err)
{
    size_t bytes = std::max<size_t>(sizeof(T) * count, 16);
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        err = "hipMalloc failed";
        return false;
    }
    s->allocations.push_back(p);
    s->device_bytes += bytes;
    if (count && hipMemcpy(p, host, sizeof(T) * count, hipMemcpyHostToDevice) != hipSuccess) {
        err = "hipMemcpy failed";
        return false;
    }
    *dev_out = reinterpret_cast<const T*>(p);
    return true;
}

// a texture slot of a material: 1 + index (0: none) -> index or -1, with
// the channel count the slot needs
bool texture_slot(const BbSceneDesc& d, int32_t slot, int32_t channels, const char* what, int32_t& out,
                  std::string& err)
{
    out = -1;
    if (slot == 0) return true;
    if (slot < 0 || slot > d.n_textures || !d.textures) {
        err = std::string("material ") + what + " texture index out of range";
        return false;
    }
    if (d.textures[slot - 1].channels != channels) {
        err = std::string("material ") + what + " texture must have " + std::to_string(channels) + " channel(s)";
        return false;
    }
    out = slot - 1;
    return true;
}

// the constants of material_constants plus the texture slots that the
// reference's material evaluates (the *Material::GetBSDF functions,
// Intersection::GetBSDF for the normal map, Primitive::Intersect for alpha)
bool material_to_gpu(const BbSceneDesc& d, const BbMaterial& m, GpuMaterial& g, std::string& err)
{
    if (!material_constants(m, g, err)) return false;
    const bool mr = m.type == BB_MAT_METALLIC_ROUGHNESS;
    if (!mr && m.tex_metallic != 0) {
        err = "a metallic texture belongs to a metallic-roughness material";
        return false;
    }
    bool ok = texture_slot(d, m.tex_color, 3, "colour", g.tex_color, err) &&
              texture_slot(d, m.tex_u_roughness, 1, "u roughness", g.tex_u, err) &&
              texture_slot(d, m.tex_v_roughness, 1, "v roughness", g.tex_v, err) &&
              texture_slot(d, m.tex_metallic, 1, "metallic", g.tex_metallic, err) &&
              texture_slot(d, m.tex_normal, 3, "normal", g.tex_normal, err) &&
              texture_slot(d, m.tex_alpha, 1, "alpha", g.tex_alpha, err);
    if (!ok) return false;
    // DielectricMaterial has no alpha texture (GetAlphaTexture returns null)
    if (m.type == BB_MAT_DIELECTRIC) g.tex_alpha = -1;
    // DiffuseMaterial's v roughness does not exist
    if (m.type == BB_MAT_DIFFUSE) g.tex_v = -1;
    g.textured = (g.tex_color >= 0 || g.tex_u >= 0 || g.tex_v >= 0 || g.tex_metallic >= 0) ? 1 : 0;
    return true;
}

// BbTexture -> GpuTexture, texels appended to `texels`
bool textures_to_gpu(const BbSceneDesc& d, std::vector<GpuTexture>& out, std::vector<float>& texels,
                     std::string& err)
{
    out.assign(std::max(1, d.n_textures), GpuTexture{});
    if (d.n_textures < 0 || (d.n_textures > 0 && !d.textures)) {
        err = "texture array missing";
        return false;
    }
    for (int32_t i = 0; i < d.n_textures; ++i) {
        const BbTexture& t = d.textures[i];
        GpuTexture& g = out[i];
        if (t.channels != 1 && t.channels != 3) {
            err = "texture channels must be 1 or 3";
            return false;
        }
        g.type = t.type;
        g.channels = t.channels;
        g.a = g.b = -1;
        for (int k = 0; k < 3; ++k) g.value[k] = t.channels == 1 ? t.value[0] : t.value[k];
        if (t.type == BB_TEX_CONSTANT) continue;
        if (t.type == BB_TEX_IMAGE) {
            if (t.width <= 0 || t.height <= 0 || !t.texels ||
                int64_t(t.width) * t.height * t.channels > (int64_t(1) << 31) - int64_t(texels.size())) {
                err = "image texture needs a positive size and texels";
                return false;
            }
            if (t.filter != BB_TEXCOORD_REPEAT && t.filter != BB_TEXCOORD_CLAMP) {
                err = "unknown texture coordinate filter";
                return false;
            }
            g.width = t.width;
            g.height = t.height;
            g.filter = t.filter;
            g.texel_offset = int32_t(texels.size());
            texels.insert(texels.end(), t.texels, t.texels + size_t(t.width) * t.height * t.channels);
            continue;
        }
        if (t.type == BB_TEX_CHECKER) {
            for (int k = 0; k < 2; ++k) {
                const int32_t c = t.checker[k] - 1;
                if (c < 0 || c >= d.n_textures || d.textures[c].type == BB_TEX_CHECKER ||
                    d.textures[c].channels != t.channels) {
                    err = "checker textures take two constant or image textures of their channel count";
                    return false;
                }
            }
            g.a = t.checker[0] - 1;
            g.b = t.checker[1] - 1;
            g.res_x = t.resolution[0];
            g.res_y = t.resolution[1];
            continue;
        }
        err = "unknown texture type";
        return false;
    }
    return true;
}

} // namespace

bool create_scene(const BbSceneDesc& d, BbScene* s, std::string& err)
{
    auto t0 = std::chrono::steady_clock::now();
    HostBvh bvh;
    if (!build_bvh(d, bvh, err)) return false;
    s->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    s->n_prims = d.n_primitives;
    s->n_slots = d.n_primitives;
    s->n_nodes = int32_t(bvh.nodes.size());
    s->max_depth = bvh.max_depth;
    for (int k = 0; k < 3; ++k) {
        s->world_min[k] = bvh.nodes[0].bmin[k];
        s->world_max[k] = bvh.nodes[0].bmax[k];
    }

    std::vector<GpuMaterial> mats(std::max(1, d.n_materials));
    for (int32_t i = 0; i < d.n_materials; ++i)
        if (!material_to_gpu(d, d.materials[i], mats[i], err)) return false;
    std::vector<GpuTexture> texs;
    std::vector<float> texels;
    if (!textures_to_gpu(d, texs, texels, err)) return false;
    if (d.n_primitives >= (1 << 30)) { // hit records keep bit 30 for alpha restarts
        err = "at most 2^30 - 1 primitives";
        return false;
    }

    // lights: Integrator::Integrator (integrator.cpp:107-147)
    std::vector<int32_t> prim_light(d.n_primitives, -1);
    std::vector<GpuLight> lights(std::max(1, d.n_lights));
    std::vector<int32_t> inf;
    f3 wmin = mk3(s->world_min[0], s->world_min[1], s->world_min[2]);
    f3 wmax = mk3(s->world_max[0], s->world_max[1], s->world_max[2]);
    std::vector<float> func(std::max(1, d.n_lights)), cdf(d.n_lights + 1);
    for (int32_t i = 0; i < d.n_lights; ++i) {
        const BbLight& L = d.lights[i];
        GpuLight& g = lights[i];
        memset(&g, 0, sizeof(g));
        g.type = L.type;
        g.two_sided = L.two_sided;
        spec c = mks(L.spectrum[0], L.spectrum[1], L.spectrum[2]);
        spec phi;
        if (L.type == BB_LIGHT_SPOT) {
            // SpotLight ctor (spot_light.cpp:7-24), Phi (:111-114)
            g.p = mk3(L.position[0], L.position[1], L.position[2]);
            g.z = -normalize(mk3(L.direction[0], L.direction[1], L.direction[2]));
            g.cos_min = std::cos(L.angle_falloff_start * kPi / 180.0f);
            g.cos_max = std::cos(L.angle_max * kPi / 180.0f);
            g.L = c;
            phi = (c * kTwoPi) * ((1.0f - g.cos_min) + (g.cos_min - g.cos_max) / 2.0f);
        } else if (L.type == BB_LIGHT_UNIFORM_INFINITE) {
            // UniformInfiniteLight::Preprocess (uniform_infinite_light.cpp:14-17) with
            // AABB::ComputeBoundingSphere (bounding_box.h:524-528); Phi (:90-93)
            f3 ctr = (wmin + wmax) * 0.5f;
            bool inside = !(wmin.x > ctr.x || wmax.x < ctr.x) && !(wmin.y > ctr.y || wmax.y < ctr.y) &&
                          !(wmin.z > ctr.z || wmax.z < ctr.z);
            g.world_radius = inside ? length(wmax - ctr) : 0.0f;
            g.L = c * L.scale;
            phi = (((c * L.scale) * kFourPi) * kPi) * sqr(g.world_radius);
            inf.push_back(i);
        } else if (L.type == BB_LIGHT_DIFFUSE_AREA) {
            if (L.primitive < 0 || L.primitive >= d.n_primitives ||
                d.primitives[L.primitive].shape_kind != BB_SHAPE_TRIANGLE) {
                err = "area lights must reference a triangle primitive";
                return false;
            }
            const int32_t* v = d.indices + 3 * int64_t(d.primitives[L.primitive].shape_index);
            f3 p0 = load3(d.positions, v[0]), p1 = load3(d.positions, v[1]), p2 = load3(d.positions, v[2]);
            g.p = p0;
            g.e1 = p1 - p0;
            g.e2 = p2 - p0;
            // Triangle::Sample (triangle.cpp:159-185): normal.Normalize() * 0.5f
            f3 n = cross(g.e1, g.e2);
            g.area = normalize_m(n) * 0.5f;
            g.n = n;
            g.L = c;
            prim_light[L.primitive] = i;
            // DiffuseAreaLight::Phi (diffuse_area_light.cpp:144-148), Triangle::Area (triangle.cpp:255-264)
            float area = 0.5f * length(cross(g.e1, g.e2));
            phi = ((c * area) * kPi) * float(L.two_sided ? 2 : 1);
        } else if (L.type == BB_LIGHT_DIRECTIONAL) {
            // DirectionalLight ctor and Preprocess (directional_light.cpp:7-17), Phi (:92-95)
            g.z = normalize(mk3(L.direction[0], L.direction[1], L.direction[2]));
            f3 ctr = (wmin + wmax) * 0.5f;
            bool inside = !(wmin.x > ctr.x || wmax.x < ctr.x) && !(wmin.y > ctr.y || wmax.y < ctr.y) &&
                          !(wmin.z > ctr.z || wmax.z < ctr.z);
            g.world_radius = inside ? length(wmax - ctr) : 0.0f;
            g.L = c;
            phi = (c * kPi) * sqr(g.world_radius);
        } else if (L.type == BB_LIGHT_POINT) {
            // PointLight ctor (point_light.cpp:7-13), Phi (:93-96)
            g.p = mk3(L.position[0], L.position[1], L.position[2]);
            g.L = c;
            phi = c * kFourPi;
        } else {
            err = "unknown light type";
            return false;
        }
        func[i] = luminance(phi);
    }
    // Distribution1D (sampling.h:345-375)
    const int32_t nl = d.n_lights;
    float func_integral = 0.0f;
    if (nl > 0) {
        cdf[0] = 0.0f;
        for (int32_t i = 1; i < nl + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / float(nl);
        func_integral = cdf[nl];
        if (func_integral == 0.0f) {
            for (int32_t i = 1; i < nl + 1; ++i) cdf[i] = float(i) / float(nl);
        } else {
            for (int32_t i = 1; i < nl + 1; ++i) cdf[i] /= func_integral;
        }
        // Distribution1D::DiscretePDF (sampling.h:404-407)
        for (int32_t i = 0; i < nl; ++i) lights[i].pmf = func[i] / (func_integral * float(nl));
    }
    s->n_lights = nl;
    s->n_inf = int32_t(inf.size());
    // the opt-in alias-table selection over the same powers
    std::vector<float2> alias_tab(std::max(1, nl), make_float2(1.0f, 0.0f));
    if (nl > 0) {
        std::vector<float> ap(nl);
        std::vector<int32_t> aa(nl);
        build_alias_table(func.data(), nl, ap.data(), aa.data());
        for (int32_t i = 0; i < nl; ++i) alias_tab[i] = make_float2(ap[i], i2f(aa[i]));
    }

    // BVH slot data
    const int32_t ns = s->n_slots;
    std::vector<float4> tris(size_t(3) * ns);
    std::vector<int4> slot_shade(ns);
    std::vector<int32_t> slot_light(ns);
    std::vector<BbBvhNode> nodes = bvh.nodes;
    bool has_null = false;
    for (int32_t j = 0; j < ns; ++j) {
        int32_t prim = bvh.ordered[j];
        const BbPrimitive& p = d.primitives[prim];
        int32_t prim_bits = prim;
        float prim_f;
        memcpy(&prim_f, &prim_bits, 4);
        slot_light[j] = prim_light[prim];
        if (p.material < 0) has_null = true;
        if (p.shape_kind == BB_SHAPE_TRIANGLE) {
            const int32_t* v = d.indices + 3 * int64_t(p.shape_index);
            f3 p0 = load3(d.positions, v[0]), p1 = load3(d.positions, v[1]), p2 = load3(d.positions, v[2]);
            f3 e1 = p1 - p0, e2 = p2 - p0;
            tris[3 * size_t(j) + 0] = make_float4(p0.x, p0.y, p0.z, prim_f);
            tris[3 * size_t(j) + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
            const bool alpha = p.material >= 0 && mats[p.material].tex_alpha >= 0;
            tris[3 * size_t(j) + 2] = make_float4(e2.x, e2.y, e2.z, i2f(KIND_TRIANGLE | (alpha ? kSlotAlpha : 0)));
            slot_shade[j] = make_int4(v[0], v[1], v[2], p.material);
        } else {
            if (p.material >= 0 && mats[p.material].tex_alpha >= 0) {
                err = "alpha textures on spheres are not supported";
                return false;
            }
            int32_t kind = KIND_SPHERE, sidx = p.shape_index;
            float kind_f, sidx_f;
            memcpy(&kind_f, &kind, 4);
            memcpy(&sidx_f, &sidx, 4);
            tris[3 * size_t(j) + 0] = make_float4(0.0f, 0.0f, 0.0f, prim_f);
            tris[3 * size_t(j) + 1] = make_float4(0.0f, 0.0f, 0.0f, sidx_f);
            tris[3 * size_t(j) + 2] = make_float4(0.0f, 0.0f, 0.0f, kind_f);
            slot_shade[j] = make_int4(0, 0, 0, p.material);
        }
    }
    // flag the last slot of every leaf, and build the 64-byte inner records:
    // record of internal node i = boxes and references of children i + 1 and
    // nodes[i].offset (BVH::RayCast's child1 / child2, bvh.h:139-151)
    std::vector<int32_t> inner_index(nodes.size(), -1);
    int32_t n_inner = 0;
    for (size_t i = 0; i < nodes.size(); ++i)
        if (nodes[i].count == 0) inner_index[i] = n_inner++;
        else {
            float4& last = tris[3 * size_t(nodes[i].offset + nodes[i].count - 1) + 2];
            last.w = i2f(f2i(last.w) | kSlotLast);
        }
    auto ref_of = [&](size_t c) { return nodes[c].count == 0 ? inner_index[c] : ~nodes[c].offset; };
    std::vector<float4> inner(size_t(4) * std::max(1, n_inner));
    for (size_t i = 0; i < nodes.size(); ++i) {
        if (nodes[i].count != 0) continue;
        const BbBvhNode& c1 = nodes[i + 1];
        const BbBvhNode& c2 = nodes[size_t(nodes[i].offset)];
        float4* r = &inner[size_t(4) * inner_index[i]];
        r[0] = make_float4(c1.bmin[0], c1.bmin[1], c1.bmin[2], c1.bmax[0]);
        r[1] = make_float4(c1.bmax[1], c1.bmax[2], c2.bmin[0], c2.bmin[1]);
        r[2] = make_float4(c2.bmin[2], c2.bmax[0], c2.bmax[1], c2.bmax[2]);
        r[3] = make_float4(i2f(ref_of(i + 1)), i2f(ref_of(size_t(nodes[i].offset))), i2f(int32_t(nodes[i].axis)), 0.0f);
    }
    // two-level records: one for the root and one for every inner node that
    // is an entry of another record (an inner grandchild), numbered in
    // depth-first order; quad_levels = the most records on a root-to-leaf walk
    std::vector<int32_t> quad_index(nodes.size(), -1);
    std::vector<int32_t> quad_order;
    int32_t quad_levels = 0;
    {
        struct Q
        {
            size_t node;
            int32_t level;
        };
        std::vector<Q> todo;
        if (nodes[0].count == 0) todo.push_back({ 0, 1 });
        while (!todo.empty()) {
            const Q q = todo.back();
            todo.pop_back();
            quad_index[q.node] = int32_t(quad_order.size());
            quad_order.push_back(int32_t(q.node));
            quad_levels = std::max(quad_levels, q.level);
            const size_t kids[2] = { q.node + 1, size_t(nodes[q.node].offset) };
            for (int c = 1; c >= 0; --c) { // child1's entries first in the numbering
                const size_t k = kids[c];
                if (nodes[k].count != 0) continue;
                const size_t gk[2] = { k + 1, size_t(nodes[k].offset) };
                for (int g = 1; g >= 0; --g)
                    if (nodes[gk[g]].count == 0) todo.push_back({ gk[g], q.level + 1 });
            }
        }
    }
    const int32_t n_quad = int32_t(quad_order.size());
    auto qref_of = [&](size_t c) { return nodes[c].count == 0 ? quad_index[c] : ~nodes[c].offset; };
    std::vector<float4> quads(size_t(8) * std::max(1, n_quad), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (int32_t qi = 0; qi < n_quad; ++qi) {
        const size_t nn = size_t(quad_order[qi]);
        const size_t kids[2] = { nn + 1, size_t(nodes[nn].offset) };
        size_t ent[4] = { 0, 0, 0, 0 };
        bool has[4] = { true, false, true, false };
        int32_t meta = int32_t(nodes[nn].axis);
        for (int c = 0; c < 2; ++c) {
            const size_t k = kids[c];
            if (nodes[k].count == 0) {
                ent[2 * c] = k + 1;
                ent[2 * c + 1] = size_t(nodes[k].offset);
                has[2 * c + 1] = true;
                meta |= int32_t(nodes[k].axis) << (2 + 2 * c);
                meta |= 1 << (6 + c);
            } else {
                ent[2 * c] = k;
            }
        }
        float box[4][6] = {};
        int32_t ref[4] = { 0, 0, 0, 0 };
        for (int e = 0; e < 4; ++e) {
            if (!has[e]) continue;
            const BbBvhNode& b = nodes[ent[e]];
            for (int a = 0; a < 3; ++a) {
                box[e][a] = b.bmin[a];
                box[e][3 + a] = b.bmax[a];
            }
            ref[e] = qref_of(ent[e]);
        }
        float4* r = &quads[size_t(8) * qi];
        for (int h = 0; h < 2; ++h) {
            const float* b0 = box[2 * h];
            const float* b1 = box[2 * h + 1];
            r[3 * h + 0] = make_float4(b0[0], b0[1], b0[2], b0[3]);
            r[3 * h + 1] = make_float4(b0[4], b0[5], b1[0], b1[1]);
            r[3 * h + 2] = make_float4(b1[2], b1[3], b1[4], b1[5]);
        }
        r[6] = make_float4(i2f(ref[0]), i2f(ref[1]), i2f(ref[2]), i2f(ref[3]));
        r[7] = make_float4(i2f(meta), 0.0f, 0.0f, 0.0f);
    }

    // every reference the kernels follow stays inside its array
    auto ref_ok = [&](int32_t ref) { return ref >= 0 ? ref < n_inner : (~ref >= 0 && ~ref < ns); };
    auto qref_ok = [&](int32_t ref) { return ref >= 0 ? ref < n_quad : (~ref >= 0 && ~ref < ns); };
    for (int32_t k = 0; k < n_quad; ++k) {
        const float4 rr = quads[size_t(8) * k + 6];
        const int32_t meta = f2i(quads[size_t(8) * k + 7].x);
        const int32_t rs[4] = { f2i(rr.x), f2i(rr.y), f2i(rr.z), f2i(rr.w) };
        const bool has1 = (meta >> 6) & 1, has3 = (meta >> 7) & 1;
        if (!qref_ok(rs[0]) || !qref_ok(rs[2]) || (has1 && !qref_ok(rs[1])) || (has3 && !qref_ok(rs[3])) ||
            (meta & 3) == 3 || ((meta >> 2) & 3) == 3 || ((meta >> 4) & 3) == 3) {
            err = "internal error: two-level BVH record out of range";
            return false;
        }
    }
    for (int32_t k = 0; k < n_inner; ++k) {
        const int32_t r1 = f2i(inner[size_t(4) * k + 3].x), r2 = f2i(inner[size_t(4) * k + 3].y);
        const int32_t ax = f2i(inner[size_t(4) * k + 3].z);
        if (!ref_ok(r1) || !ref_ok(r2) || ax < 0 || ax > 2) {
            err = "internal error: inner BVH record out of range";
            return false;
        }
    }
    if (!ref_ok(ref_of(0)) || !(f2i(tris[3 * size_t(ns - 1) + 2].w) & kSlotLast)) {
        err = "internal error: BVH root or last leaf slot malformed";
        return false;
    }

    std::vector<float4> normals(std::max(1, d.n_vertices)), tangents(std::max(1, d.n_vertices));
    for (int32_t i = 0; i < d.n_vertices; ++i) {
        normals[i] = make_float4(d.normals[3 * i], d.normals[3 * i + 1], d.normals[3 * i + 2], 0.0f);
        tangents[i] = make_float4(d.tangents[3 * i], d.tangents[3 * i + 1], d.tangents[3 * i + 2], 0.0f);
    }

    SceneView& V = s->view;
    hipGetDevice(&s->device);
    if (!upload(s, inner.data(), inner.size(), &V.inner, err)) return false;
    for (int k = 0; k < 3; ++k) {
        V.root_box[k] = nodes[0].bmin[k];
        V.root_box[3 + k] = nodes[0].bmax[k];
    }
    V.root_ref = ref_of(0);
    if (!upload(s, quads.data(), quads.size(), &V.quads, err)) return false;
    V.quad_root = qref_of(0);
    // each record pushes at most 3 entries
    V.quad_stack_depth = 3 * quad_levels + 1;
    if (!upload(s, tris.data(), tris.size(), &V.tris, err)) return false;
    if (!upload(s, slot_shade.data(), slot_shade.size(), &V.slot_shade, err)) return false;
    if (!upload(s, slot_light.data(), slot_light.size(), &V.slot_light, err)) return false;
    if (!upload(s, normals.data(), normals.size(), &V.normals, err)) return false;
    if (!upload(s, tangents.data(), tangents.size(), &V.tangents, err)) return false;
    std::vector<BbSphere> spheres(d.spheres, d.spheres + d.n_spheres);
    spheres.resize(std::max<size_t>(1, spheres.size()));
    if (!upload(s, spheres.data(), spheres.size(), &V.spheres, err)) return false;
    if (!upload(s, mats.data(), mats.size(), &V.materials, err)) return false;
    if (!upload(s, lights.data(), lights.size(), &V.lights, err)) return false;
    cdf.resize(std::max<size_t>(2, cdf.size()));
    if (!upload(s, cdf.data(), cdf.size(), &V.light_cdf, err)) return false;
    if (!upload(s, func.data(), func.size(), &V.light_func, err)) return false;
    if (!upload(s, alias_tab.data(), alias_tab.size(), &V.light_alias, err)) return false;
    inf.resize(std::max<size_t>(1, inf.size()));
    if (!upload(s, inf.data(), inf.size(), &V.inf_lights, err)) return false;
    const ReflectanceTables& tab = reflectance_tables();
    if (!upload(s, tab.ggx_E, size_t(kTableSize) * kTableSize, &V.ggx_E, err)) return false;
    if (!upload(s, tab.ggx_E_avg, size_t(kTableSize), &V.ggx_E_avg, err)) return false;
    texels.resize(std::max<size_t>(1, texels.size()));
    if (!upload(s, texs.data(), texs.size(), &V.textures, err)) return false;
    if (!upload(s, texels.data(), texels.size(), &V.texels, err)) return false;
    V.uvs = nullptr;
    if (d.uvs && d.n_vertices > 0) {
        std::vector<float2> uvs(d.n_vertices);
        for (int32_t i = 0; i < d.n_vertices; ++i) uvs[i] = make_float2(d.uvs[2 * i], d.uvs[2 * i + 1]);
        if (!upload(s, uvs.data(), uvs.size(), &V.uvs, err)) return false;
    }
    V.dielectric = nullptr;
    // compensated dielectrics need the tables; a textured roughness can make
    // any compensated dielectric rough at some hit
    if (std::any_of(mats.begin(), mats.end(), [](const GpuMaterial& g) {
            return g.kind == MK_DIELECTRIC_MS ||
                   (g.type == BB_MAT_DIELECTRIC && g.energy_compensation && (g.tex_u >= 0 || g.tex_v >= 0));
        })) {
        const DielectricTables& dt = dielectric_tables();
        static_assert(sizeof(DielectricTables) == sizeof(float) * kDielectricTableFloats, "packed tables");
        if (!upload(s, reinterpret_cast<const float*>(&dt), size_t(kDielectricTableFloats), &V.dielectric, err))
            return false;
    }
    V.has_alpha = 0;
    for (int32_t j = 0; j < ns; ++j)
        if (f2i(tris[3 * size_t(j) + 2].w) & kSlotAlpha) V.has_alpha = 1;
    V.kind_mask = 0;
    for (const GpuMaterial& g : mats) V.kind_mask |= 1 << g.kind;
    V.has_textures = std::any_of(mats.begin(), mats.end(),
                                 [](const GpuMaterial& g) { return g.textured || g.tex_normal >= 0; }) ? 1 : 0;
    V.n_lights = nl;
    V.n_inf = s->n_inf;
    V.func_integral = func_integral;
    V.stack_depth = s->max_depth + 2;
    V.has_null_material = has_null ? 1 : 0;
    return true;
}

void release_scene_workspace(BbScene* s);

void destroy_scene(BbScene* s)
{
    release_scene_workspace(s);
    for (void* p : s->allocations) hipFree(p);
    s->allocations.clear();
}

void build_alias_table(const float* w, int32_t n, float* prob, int32_t* alias)
{
    double sum = 0.0;
    for (int32_t i = 0; i < n; ++i) sum += std::max(0.0, double(w[i]));
    std::vector<double> q(n);
    for (int32_t i = 0; i < n; ++i) q[i] = sum > 0.0 ? std::max(0.0, double(w[i])) * n / sum : 1.0;
    std::vector<int32_t> small, large;
    for (int32_t i = 0; i < n; ++i) (q[i] < 1.0 ? small : large).push_back(i);
    while (!small.empty() && !large.empty()) {
        const int32_t l = small.back(), g = large.back();
        small.pop_back();
        prob[l] = float(q[l]);
        alias[l] = g;
        q[g] = (q[g] + q[l]) - 1.0;
        if (q[g] < 1.0) {
            large.pop_back();
            small.push_back(g);
        }
    }
    // what is left has probability 1 up to rounding
    for (int32_t i : large) {
        prob[i] = 1.0f;
        alias[i] = i;
    }
    for (int32_t i : small) {
        prob[i] = 1.0f;
        alias[i] = i;
    }
}

} // namespace bb

// =================================================================== C ABI

extern "C" {

int bbgpu_abi_version(void) { return BBGPU_ABI_VERSION; }

int bbgpu_alias_table(const float* weights, int32_t n, float* prob_out, int32_t* alias_out)
{
    if (n < 0 || (n > 0 && (!weights || !prob_out || !alias_out))) {
        bb::set_error("bbgpu_alias_table: invalid argument");
        return BBGPU_ERR_INVALID;
    }
    if (n > 0) bb::build_alias_table(weights, n, prob_out, alias_out);
    return BBGPU_OK;
}

int bbgpu_last_error(char* buf, size_t buf_size)
{
    if (!buf || buf_size == 0) return BBGPU_ERR_INVALID;
    snprintf(buf, buf_size, "%s", bb::g_last_error.c_str());
    return BBGPU_OK;
}

int bbgpu_device_count(int32_t* count)
{
    if (!count) return BBGPU_ERR_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return BBGPU_OK;
}

int bbgpu_build_bvh(const BbSceneDesc* desc, BbBvhNode* nodes_out, int32_t* ordered_out, int32_t* n_nodes_out)
{
    if (!desc || !nodes_out || !ordered_out || !n_nodes_out) {
        bb::set_error("bbgpu_build_bvh: null argument");
        return BBGPU_ERR_INVALID;
    }
    bb::HostBvh bvh;
    std::string err;
    if (!bb::build_bvh(*desc, bvh, err)) {
        bb::set_error("bbgpu_build_bvh: %s", err.c_str());
        return BBGPU_ERR_INVALID;
    }
    memcpy(nodes_out, bvh.nodes.data(), sizeof(BbBvhNode) * bvh.nodes.size());
    memcpy(ordered_out, bvh.ordered.data(), sizeof(int32_t) * bvh.ordered.size());
    *n_nodes_out = int32_t(bvh.nodes.size());
    return BBGPU_OK;
}

int bbgpu_reflectance_table(int32_t table, float* out, int32_t capacity, int32_t* length_out)
{
    if (!length_out) {
        bb::set_error("bbgpu_reflectance_table: null length_out");
        return BBGPU_ERR_INVALID;
    }
    const float* src = nullptr;
    int32_t len = 0;
    constexpr int32_t n = bb::kTableSize;
    switch (table) {
    case BB_TABLE_GGX_E:
        src = bb::reflectance_tables().ggx_E;
        len = n * n;
        break;
    case BB_TABLE_GGX_E_AVG:
        src = bb::reflectance_tables().ggx_E_avg;
        len = n;
        break;
    case BB_TABLE_DIELECTRIC_E:
        src = bb::dielectric_tables().E;
        len = n * n * n;
        break;
    case BB_TABLE_DIELECTRIC_E_INV:
        src = bb::dielectric_tables().E_inv;
        len = n * n * n;
        break;
    case BB_TABLE_DIELECTRIC_E_AVG:
        src = bb::dielectric_tables().E_avg;
        len = n * n;
        break;
    case BB_TABLE_DIELECTRIC_E_INV_AVG:
        src = bb::dielectric_tables().E_inv_avg;
        len = n * n;
        break;
    default:
        bb::set_error("bbgpu_reflectance_table: unknown table %d", int(table));
        return BBGPU_ERR_INVALID;
    }
    *length_out = len;
    if (out && capacity >= len) memcpy(out, src, sizeof(float) * size_t(len));
    return BBGPU_OK;
}

int bbgpu_scene_create(const BbSceneDesc* desc, BbScene** scene_out)
{
    if (!desc || !scene_out) {
        bb::set_error("bbgpu_scene_create: null argument");
        return BBGPU_ERR_INVALID;
    }
    *scene_out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        bb::set_error("bbgpu_scene_create: no HIP device");
        return BBGPU_ERR_DEVICE;
    }
    auto* s = new BbScene();
    std::string err;
    if (!bb::create_scene(*desc, s, err)) {
        bb::destroy_scene(s);
        delete s;
        bb::set_error("bbgpu_scene_create: %s", err.c_str());
        return err.find("not supported") != std::string::npos ? BBGPU_ERR_UNSUPPORTED : BBGPU_ERR_INVALID;
    }
    *scene_out = s;
    return BBGPU_OK;
}

int bbgpu_scene_destroy(BbScene* scene)
{
    if (!scene) return BBGPU_ERR_INVALID;
    bb::destroy_scene(scene);
    delete scene;
    return BBGPU_OK;
}

int bbgpu_scene_info(const BbScene* s, BbSceneInfo* info)
{
    if (!s || !info) {
        bb::set_error("bbgpu_scene_info: null argument");
        return BBGPU_ERR_INVALID;
    }
    info->n_nodes = s->n_nodes;
    info->n_primitives = s->n_prims;
    info->n_lights = s->n_lights;
    info->n_infinite_lights = s->n_inf;
    for (int k = 0; k < 3; ++k) {
        info->world_min[k] = s->world_min[k];
        info->world_max[k] = s->world_max[k];
    }
    info->build_seconds = s->build_seconds;
    info->device_bytes = s->device_bytes;
    return BBGPU_OK;
}

// Transform::LookAt (transform.h:112-122) with Quat(front, up) (quaternion.h:64-75)
// and Quat(const Mat3&) (quaternion.h:34-62)
int bbgpu_look_at(const float position[3], const float target[3], const float up[3], float q_out[4])
{
    using namespace bb;
    if (!position || !target || !up || !q_out) return BBGPU_ERR_INVALID;
    f3 w = mk3(target[0], target[1], target[2]) - mk3(position[0], position[1], position[2]);
    normalize_m(w);
    f3 ez = w;
    f3 ex = cross(mk3(up[0], up[1], up[2]), ez);
    normalize_m(ex);
    f3 ey = cross(ez, ex);
    float x, y, z, ww, t, sc;
    if (ez.z < 0.0f) {
        if (ex.x > ey.y) {
            t = 1.0f + ex.x - ey.y - ez.z;
            x = t, y = ex.y + ey.x, z = ez.x + ex.z, ww = ey.z - ez.y;
        } else {
            t = 1.0f - ex.x + ey.y - ez.z;
            x = ex.y + ey.x, y = t, z = ey.z + ez.y, ww = ez.x - ex.z;
        }
    } else {
        if (ex.x < -ey.y) {
            t = 1.0f - ex.x - ey.y + ez.z;
            x = ez.x + ex.z, y = ey.z + ez.y, z = t, ww = ex.y - ey.x;
        } else {
            t = 1.0f + ex.x + ey.y + ez.z;
            x = ey.z - ez.y, y = ez.x - ex.z, z = ex.y - ey.x, ww = t;
        }
    }
    sc = 0.5f / std::sqrt(t);
    q_out[0] = x * sc;
    q_out[1] = y * sc;
    q_out[2] = z * sc;
    q_out[3] = ww * sc;
    return BBGPU_OK;
}

} // extern "C"
