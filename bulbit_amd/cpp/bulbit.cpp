// bulbit.cpp — the C++ library API of bulbit/bulbit.h over libbbgpu.so.
//
// The scene objects are flattened into a BbSceneDesc in primitive order (the
// order of Scene::GetPrimitives(), which fixes hit identities and the BVH),
// with lights in the order handed to the integrator (which fixes the power
// sampler's CDF, src/light_sampler/power_light_sampler.cpp:8-23).  Errors of
// the C ABI stop the program with the library's message: the reference API
// has no error returns on these calls, and there is no CPU fallback.
#include "bulbit/bulbit.h"
#include "../../include/bbgpu.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <type_traits>

namespace bulbit
{

namespace
{

[[noreturn]] void fatal(const char* what, const std::string& msg)
{
    std::fprintf(stderr, "bulbit_amd: %s: %s\n", what, msg.c_str());
    std::fflush(stderr);
    std::abort();
}

void check(int status, const char* what)
{
    if (status == BBGPU_OK) return;
    char buf[1024];
    bbgpu_last_error(buf, sizeof(buf));
    fatal(what, buf);
}

template <typename T>
T constant_of(const Texture<T>* tex, T fallback, const char* what)
{
    if (!tex) return fallback;
    T v{};
    if (!tex->Constant(&v)) fatal(what, "only constant textures are supported on the GPU path");
    return v;
}

float component(Float v, int) { return v; }
float component(const Spectrum& v, int k) { return v[k]; }

void put3(float* dst, const Spectrum& s)
{
    dst[0] = s.r;
    dst[1] = s.g;
    dst[2] = s.b;
}

void put3(float* dst, const Vec3& v)
{
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

} // namespace

// ------------------------------------------------------------- flat scene

struct SceneExport
{
    std::vector<float> positions, normals, tangents, uvs;
    std::vector<int32_t> indices;
    std::vector<BbSphere> spheres;
    std::vector<BbPrimitive> prims;
    std::vector<BbMaterial> materials;
    std::vector<BbLight> lights;
    std::vector<BbTexture> textures;
    std::vector<std::unique_ptr<std::vector<float>>> texels; // stable storage of image texels
    std::map<const void*, int32_t> texture_slot;             // texture -> index + 1
    std::map<const Mesh*, int32_t> mesh_base;

    // a texture as a BbTexture; returns its slot (index + 1)
    template <typename T>
    int32_t add_texture(const Texture<T>* t)
    {
        auto it = texture_slot.find(t);
        if (it != texture_slot.end()) return it->second;
        constexpr int32_t channels = std::is_same<T, Float>::value ? 1 : 3;
        BbTexture b;
        std::memset(&b, 0, sizeof(b));
        b.channels = channels;
        T v{};
        if (t->Constant(&v)) {
            b.type = BB_TEX_CONSTANT;
            for (int k = 0; k < channels; ++k) b.value[k] = component(v, k);
        } else if (auto* im = dynamic_cast<const ImageTexture<T>*>(t)) {
            if (!im->image) fatal("ImageTexture", "empty image");
            b.type = BB_TEX_IMAGE;
            b.width = im->image.width;
            b.height = im->image.height;
            b.filter = im->texcoord_filter == TexCoordFilter::clamp ? BB_TEXCOORD_CLAMP : BB_TEXCOORD_REPEAT;
            auto store = std::make_unique<std::vector<float>>();
            store->reserve(im->image.data.size() * channels);
            for (const T& x : im->image.data)
                for (int k = 0; k < channels; ++k) store->push_back(component(x, k));
            b.texels = store->data();
            texels.push_back(std::move(store));
        } else if (auto* ch = dynamic_cast<const CheckerTexture<T>*>(t)) {
            b.type = BB_TEX_CHECKER;
            b.checker[0] = add_texture(ch->a);
            b.checker[1] = add_texture(ch->b);
            b.resolution[0] = ch->resolution.x;
            b.resolution[1] = ch->resolution.y;
        } else {
            fatal("scene export", "texture type not supported on the GPU path");
        }
        textures.push_back(b);
        const int32_t slot = int32_t(textures.size());
        texture_slot[t] = slot;
        return slot;
    }

    // a material parameter: its constant value (slot 0), or its texture's slot
    template <typename T>
    int32_t param(const Texture<T>* t, T fallback, T& value)
    {
        value = fallback;
        if (!t || t->Constant(&value)) return 0;
        return add_texture(t);
    }

    void param3(const SpectrumTexture* t, Spectrum fallback, float* dst, int32_t& slot)
    {
        Spectrum v;
        slot = param(t, fallback, v);
        put3(dst, v);
    }
    std::map<const Material*, int32_t> material_index;
    std::map<const Primitive*, int32_t> prim_index;

    int32_t add_material(const Material* m)
    {
        if (!m) return -1;
        auto it = material_index.find(m);
        if (it != material_index.end()) return it->second;
        BbMaterial b;
        std::memset(&b, 0, sizeof(b));
        if (auto* d = dynamic_cast<const DiffuseMaterial*>(m)) {
            b.type = BB_MAT_DIFFUSE;
            param3(d->reflectance, Spectrum(0), b.color, b.tex_color);
            // roughness > 0 selects the EON BxDF (per hit for a texture)
            b.tex_u_roughness = param(d->roughness, 0.0f, b.u_roughness);
            b.v_roughness = b.u_roughness;
            if (d->normal) b.tex_normal = add_texture(d->normal);
            if (d->alpha) b.tex_alpha = add_texture(d->alpha);
        } else if (auto* e = dynamic_cast<const DielectricMaterial*>(m)) {
            b.type = BB_MAT_DIELECTRIC;
            b.energy_compensation = e->energy_compensation ? 1 : 0;
            b.eta[0] = b.eta[1] = b.eta[2] = e->eta;
            if (!e->reflectance) fatal("DielectricMaterial", "reflectance texture is required");
            param3(e->reflectance, Spectrum(1), b.color, b.tex_color);
            b.tex_u_roughness = param(e->u_roughness, 0.0f, b.u_roughness);
            b.tex_v_roughness = param(e->v_roughness, 0.0f, b.v_roughness);
            if (e->normal) b.tex_normal = add_texture(e->normal);
        } else if (auto* c = dynamic_cast<const ConductorMaterial*>(m)) {
            b.energy_compensation = c->energy_compensation ? 1 : 0;
            if (c->eta) {
                b.type = BB_MAT_CONDUCTOR;
                put3(b.eta, constant_of<Spectrum>(c->eta, Spectrum(0), "ConductorMaterial eta"));
                put3(b.k, constant_of<Spectrum>(c->k, Spectrum(0), "ConductorMaterial k"));
            } else {
                b.type = BB_MAT_CONDUCTOR_R;
                put3(b.k, constant_of<Spectrum>(c->k, Spectrum(0), "ConductorMaterial R"));
            }
            param3(c->reflectance, Spectrum(1), b.color, b.tex_color);
            b.tex_u_roughness = param(c->u_roughness, 0.0f, b.u_roughness);
            b.tex_v_roughness = param(c->v_roughness, 0.0f, b.v_roughness);
            if (c->normal) b.tex_normal = add_texture(c->normal);
            if (c->alpha) b.tex_alpha = add_texture(c->alpha);
        } else if (auto* mi = dynamic_cast<const MirrorMaterial*>(m)) {
            b.type = BB_MAT_MIRROR;
            if (!mi->reflectance) fatal("MirrorMaterial", "reflectance texture is required");
            param3(mi->reflectance, Spectrum(1), b.color, b.tex_color);
            if (mi->normal) b.tex_normal = add_texture(mi->normal);
            if (mi->alpha) b.tex_alpha = add_texture(mi->alpha);
        } else if (auto* r = dynamic_cast<const MetallicRoughnessMaterial*>(m)) {
            b.type = BB_MAT_METALLIC_ROUGHNESS;
            param3(r->basecolor, Spectrum(0), b.color, b.tex_color);
            b.tex_metallic = param(r->metallic, 0.0f, b.metallic);
            b.tex_u_roughness = param(r->u_roughness, 0.0f, b.u_roughness);
            b.tex_v_roughness = param(r->v_roughness, 0.0f, b.v_roughness);
            if (r->normal) b.tex_normal = add_texture(r->normal);
            if (r->alpha) b.tex_alpha = add_texture(r->alpha);
        } else {
            fatal("scene export", "material type not supported on the GPU path");
        }
        const int32_t idx = int32_t(materials.size());
        materials.push_back(b);
        material_index[m] = idx;
        return idx;
    }

    int32_t add_mesh(const Mesh* mesh)
    {
        auto it = mesh_base.find(mesh);
        if (it != mesh_base.end()) return it->second;
        const int32_t base = int32_t(positions.size() / 3);
        const auto& P = mesh->Positions();
        const auto& N = mesh->Normals();
        const auto& T = mesh->Tangents();
        const auto& UV = mesh->TexCoords();
        for (size_t i = 0; i < P.size(); ++i) {
            positions.insert(positions.end(), { P[i].x, P[i].y, P[i].z });
            const Vec3 n = i < N.size() ? N[i] : Vec3(0, 0, 1);
            const Vec3 t = i < T.size() ? T[i] : Vec3(1, 0, 0);
            const Point2 uv = i < UV.size() ? UV[i] : Point2(0, 0);
            normals.insert(normals.end(), { n.x, n.y, n.z });
            tangents.insert(tangents.end(), { t.x, t.y, t.z });
            uvs.insert(uvs.end(), { uv.x, uv.y });
        }
        mesh_base[mesh] = base;
        return base;
    }

    void add_primitive(const Primitive* p)
    {
        const MediumInterface mi = p->GetMediumInterface();
        if (mi.inside || mi.outside) fatal("scene export", "participating media are not supported on the GPU path");
        BbPrimitive b{};
        b.material = add_material(p->GetMaterial());
        const Shape* s = p->GetShape();
        if (auto* t = dynamic_cast<const Triangle*>(s)) {
            const int32_t base = add_mesh(t->GetMesh());
            const auto& I = t->GetMesh()->Indices();
            const size_t f = size_t(t->GetIndex()) * 3;
            if (f + 2 >= I.size()) fatal("scene export", "triangle index out of range");
            b.shape_kind = BB_SHAPE_TRIANGLE;
            b.shape_index = int32_t(indices.size() / 3);
            indices.insert(indices.end(), { base + I[f], base + I[f + 1], base + I[f + 2] });
        } else if (auto* sp = dynamic_cast<const Sphere*>(s)) {
            BbSphere q{};
            put3(q.p, sp->transform.p);
            q.q[0] = sp->transform.q.x;
            q.q[1] = sp->transform.q.y;
            q.q[2] = sp->transform.q.z;
            q.q[3] = sp->transform.q.w;
            put3(q.s, sp->transform.s);
            q.radius = sp->radius;
            b.shape_kind = BB_SHAPE_SPHERE;
            b.shape_index = int32_t(spheres.size());
            spheres.push_back(q);
        } else {
            fatal("scene export", "shape type not supported on the GPU path");
        }
        prim_index[p] = int32_t(prims.size());
        prims.push_back(b);
    }

    void add_light(const Light* l)
    {
        BbLight b;
        std::memset(&b, 0, sizeof(b));
        if (auto* s = dynamic_cast<const SpotLight*>(l)) {
            if (s->medium) fatal("SpotLight", "participating media are not supported on the GPU path");
            b.type = BB_LIGHT_SPOT;
            put3(b.position, s->position);
            put3(b.direction, s->direction);
            put3(b.spectrum, s->intensity);
            b.angle_max = s->angle_max;
            b.angle_falloff_start = s->angle_falloff_start;
        } else if (auto* p = dynamic_cast<const PointLight*>(l)) {
            if (p->medium) fatal("PointLight", "participating media are not supported on the GPU path");
            b.type = BB_LIGHT_POINT;
            put3(b.position, p->position);
            put3(b.spectrum, p->intensity);
        } else if (auto* dl = dynamic_cast<const DirectionalLight*>(l)) {
            b.type = BB_LIGHT_DIRECTIONAL;
            put3(b.direction, dl->direction);
            put3(b.spectrum, dl->intensity);
        } else if (auto* u = dynamic_cast<const UniformInfiniteLight*>(l)) {
            b.type = BB_LIGHT_UNIFORM_INFINITE;
            put3(b.spectrum, u->l);
            b.scale = u->scale;
        } else if (auto* a = dynamic_cast<const DiffuseAreaLight*>(l)) {
            b.type = BB_LIGHT_DIFFUSE_AREA;
            auto it = prim_index.find(a->primitive);
            if (it == prim_index.end()) fatal("DiffuseAreaLight", "its primitive is not in the BVH");
            b.primitive = it->second;
            b.two_sided = a->two_sided ? 1 : 0;
            put3(b.spectrum, constant_of<Spectrum>(a->emission, Spectrum(0), "DiffuseAreaLight emission"));
        } else {
            fatal("scene export", "light type not supported on the GPU path");
        }
        lights.push_back(b);
    }

    BbSceneDesc desc() const
    {
        BbSceneDesc d;
        std::memset(&d, 0, sizeof(d));
        d.n_vertices = int32_t(positions.size() / 3);
        d.positions = positions.data();
        d.normals = normals.data();
        d.tangents = tangents.data();
        d.uvs = uvs.empty() ? nullptr : uvs.data();
        d.n_triangles = int32_t(indices.size() / 3);
        d.indices = indices.data();
        d.n_spheres = int32_t(spheres.size());
        d.spheres = spheres.data();
        d.n_primitives = int32_t(prims.size());
        d.primitives = prims.data();
        d.n_materials = int32_t(materials.size());
        d.materials = materials.data();
        d.n_lights = int32_t(lights.size());
        d.lights = lights.data();
        d.n_textures = int32_t(textures.size());
        d.textures = textures.data();
        return d;
    }
};

// ------------------------------------------------------------------ math

Transform Transform::LookAt(const Vec3& position, const Vec3& target, const Vec3& up)
{
    const float p[3] = { position.x, position.y, position.z };
    const float t[3] = { target.x, target.y, target.z };
    const float u[3] = { up.x, up.y, up.z };
    float q[4];
    check(bbgpu_look_at(p, t, u, q), "Transform::LookAt");
    return Transform(position, Quat(q[0], q[1], q[2], q[3]));
}

// ------------------------------------------------------------------ mesh

Mesh::Mesh(std::vector<Point3> _positions, std::vector<Vec3> _normals, std::vector<Vec3> _tangents,
           std::vector<Point2> _texCoords, std::vector<int32> _indices, const Mat4& transform)
    : positions{ std::move(_positions) }, normals{ std::move(_normals) }, tangents{ std::move(_tangents) },
      texCoords{ std::move(_texCoords) }, indices{ std::move(_indices) }
{
    if (normals.size() != positions.size() || tangents.size() != positions.size())
        fatal("Mesh", "normals and tangents must be given per vertex");
    for (size_t i = 0; i < positions.size(); ++i) {
        Vec4 vp = Mul(transform, Vec4(positions[i], 1));
        Vec4 vn = Mul(transform, Vec4(normals[i], 0));
        Vec4 vt = Mul(transform, Vec4(tangents[i], 0));
        vn.Normalize();
        vt.Normalize();
        positions[i].Set(vp.x, vp.y, vp.z);
        normals[i].Set(vn.x, vn.y, vn.z);
        tangents[i].Set(vt.x, vt.y, vt.z);
    }
    triangle_count = int32(indices.size() / 3);
}

Mesh::Mesh(const std::vector<MeshVertex>& vertices, std::vector<int32> _indices, const Mat4& transform)
    : indices{ std::move(_indices) }
{
    const size_t n = vertices.size();
    positions.resize(n);
    normals.resize(n);
    tangents.resize(n);
    texCoords.resize(n);
    for (size_t i = 0; i < n; ++i) {
        Vec4 vp = Mul(transform, Vec4(vertices[i].position, 1));
        Vec4 vn = Mul(transform, Vec4(vertices[i].normal, 0));
        Vec4 vt = Mul(transform, Vec4(vertices[i].tangent, 0));
        vn.Normalize();
        vt.Normalize();
        positions[i].Set(vp.x, vp.y, vp.z);
        normals[i].Set(vn.x, vn.y, vn.z);
        tangents[i].Set(vt.x, vt.y, vt.z);
        texCoords[i] = vertices[i].texCoord;
    }
    triangle_count = int32(indices.size() / 3);
}

// ------------------------------------------------------------------- BVH

BVH::BVH(const std::vector<Primitive*>& prims) : primitives{ prims }, desc{ new SceneExport() }
{
    if (primitives.empty()) fatal("BVH", "no primitives");
    for (const Primitive* p : primitives) desc->add_primitive(p);
    // the host build alone (no device needed): node count and world bounds
    const BbSceneDesc d = desc->desc();
    std::vector<BbBvhNode> nodes(size_t(2) * primitives.size());
    std::vector<int32_t> order(primitives.size());
    check(bbgpu_build_bvh(&d, nodes.data(), order.data(), &n_nodes), "BVH");
    bounds.min = Point3(nodes[0].bmin[0], nodes[0].bmin[1], nodes[0].bmin[2]);
    bounds.max = Point3(nodes[0].bmax[0], nodes[0].bmax[1], nodes[0].bmax[2]);
}

BVH::~BVH()
{
    if (device) bbgpu_scene_destroy(device);
}

AABB BVH::GetAABB() const { return bounds; }

BbScene* BVH::DeviceScene() const
{
    if (!device) {
        const BbSceneDesc d = desc->desc();
        check(bbgpu_scene_create(&d, &device), "BVH");
    }
    return device;
}

void BVH::IntersectBatch(const Ray* rays, const Float* t_min, const Float* t_max, int32 n, bool any_hit,
                         Intersection* out, bool* hit) const
{
    if (n <= 0) return;
    std::vector<BbRay> r(static_cast<size_t>(n));
    for (int32 i = 0; i < n; ++i) {
        put3(r[i].o, rays[i].o);
        put3(r[i].d, rays[i].d);
        r[i].t_min = t_min[i];
        r[i].t_max = t_max[i];
    }
    std::vector<BbHit> h(static_cast<size_t>(n));
    check(bbgpu_trace(DeviceScene(), r.data(), n, any_hit ? 1 : 0, h.data(), nullptr, nullptr), "BVH::Intersect");
    for (int32 i = 0; i < n; ++i) {
        hit[i] = h[i].prim >= 0;
        if (!out || !hit[i] || any_hit) continue;
        out[i].primitive = primitives[size_t(h[i].prim)];
        out[i].t = h[i].t;
        out[i].point = rays[i].At(h[i].t);
        out[i].uv = Point2(h[i].u, h[i].v);
    }
}

bool BVH::Intersect(Intersection* out_isect, const Ray& ray, Float t_min, Float t_max) const
{
    bool hit = false;
    IntersectBatch(&ray, &t_min, &t_max, 1, false, out_isect, &hit);
    return hit;
}

bool BVH::IntersectAny(const Ray& ray, Float t_min, Float t_max) const
{
    bool hit = false;
    IntersectBatch(&ray, &t_min, &t_max, 1, true, nullptr, &hit);
    return hit;
}

// ------------------------------------------------------------------ film

Film::Film(const Camera* camera) : camera{ camera }
{
    const Point2i r = camera->GetScreenResolution();
    const size_t n = size_t(std::max(0, r.x)) * size_t(std::max(0, r.y));
    sums.assign(3 * n, 0.0f);
    counts.assign(n, 0);
}

Image3 Film::GetRenderedImage() const
{
    const Point2i r = camera->GetScreenResolution();
    Image3 image(r.x, r.y);
    for (size_t i = 0; i < counts.size(); ++i) {
        // samples[i] / sample_counts[i] (film.cpp:80-100); no splats on this path
        const Float c = Float(counts[i]);
        image.data[i] = Spectrum(sums[3 * i] / c, sums[3 * i + 1] / c, sums[3 * i + 2] / c);
    }
    return image;
}

// ------------------------------------------------------------- rendering

Rendering::~Rendering()
{
    if (worker.joinable()) worker.join();
}

void Rendering::LogProgress() const
{
    const int64 d = samples_done.load(std::memory_order_relaxed);
    std::printf("\rRendering.. %.2f%%", samples_total > 0 ? 100.0 * double(d) / double(samples_total) : 100.0);
    std::fflush(stdout);
}

void Rendering::Wait() const
{
    if (worker.joinable()) worker.join();
}

void Rendering::WaitAndLogProgress() const
{
    while (!IsDone()) {
        LogProgress();
        std::this_thread::sleep_for(std::chrono::milliseconds(100));
    }
    LogProgress();
    std::printf("\nRendering done\n");
    Wait();
}

// ------------------------------------------------------------ integrator

PathIntegrator::PathIntegrator(const Intersectable* accel_, std::vector<Light*> lights_, const Sampler* sampler_,
                               int32 max_bounces_, int32 rr_min_bounces_, bool regularize_bsdf_)
    : accel{ dynamic_cast<const BVH*>(accel_) }, lights{ std::move(lights_) }, sampler{ sampler_ },
      max_bounces{ max_bounces_ }, rr_min_bounces{ rr_min_bounces_ }, regularize_bsdf{ regularize_bsdf_ }
{
    if (!accel) fatal("PathIntegrator", "the accelerator must be a bulbit::BVH");
    if (!dynamic_cast<const IndependentSampler*>(sampler)) fatal("PathIntegrator", "an IndependentSampler is required");
    if (regularize_bsdf) fatal("PathIntegrator", "regularize_bsdf is not supported on the GPU path");
    // the BVH's flat scene plus the lights, uploaded once
    SceneExport x;
    for (const Primitive* p : accel->GetPrimitives()) x.add_primitive(p);
    for (const Light* l : lights) x.add_light(l);
    const BbSceneDesc d = x.desc();
    check(bbgpu_scene_create(&d, &device), "PathIntegrator");
}

PathIntegrator::~PathIntegrator()
{
    if (device) bbgpu_scene_destroy(device);
}

Rendering* PathIntegrator::Render(Allocator& alloc, const Camera* camera)
{
    auto* pc = dynamic_cast<const PerspectiveCamera*>(camera);
    if (!pc) fatal("PathIntegrator::Render", "a PerspectiveCamera is required");
    const Filter* filter = pc->GetFilter();
    if (!filter) fatal("PathIntegrator::Render", "the camera has no filter");
    if (pc->GetMedium()) fatal("PathIntegrator::Render", "participating media are not supported on the GPU path");

    BbCameraDesc cam;
    std::memset(&cam, 0, sizeof(cam));
    put3(cam.position, pc->transform.p);
    cam.orientation[0] = pc->transform.q.x;
    cam.orientation[1] = pc->transform.q.y;
    cam.orientation[2] = pc->transform.q.z;
    cam.orientation[3] = pc->transform.q.w;
    cam.vfov_deg = pc->vfov;
    cam.aperture_radius = pc->aperture_radius;
    cam.focus_distance = pc->focus_distance;
    cam.width = pc->GetScreenResolution().x;
    cam.height = pc->GetScreenResolution().y;
    cam.filter_extent = filter->extent;
    if (dynamic_cast<const BoxFilter*>(filter)) {
        cam.filter_type = BB_FILTER_BOX;
    } else if (dynamic_cast<const TentFilter*>(filter)) {
        cam.filter_type = BB_FILTER_TENT;
    } else if (auto* gf = dynamic_cast<const GaussianFilter*>(filter)) {
        cam.filter_type = BB_FILTER_GAUSSIAN;
        cam.filter_sigma = gf->sigma;
    } else {
        fatal("PathIntegrator::Render", "filter type not supported on the GPU path");
    }

    const auto* is = static_cast<const IndependentSampler*>(sampler);
    BbRenderParams prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.seed = is->seed;
    prm.max_bounces = max_bounces;
    prm.rr_min_bounces = rr_min_bounces;
    const int32 spp = is->samples_per_pixel;

    Rendering* r = alloc.new_object<Rendering>(camera);
    r->samples_total = int64(spp) * cam.width * cam.height;
    BbScene* dev = device;
    // samples are accumulated in order, so rendering them in chunks (for
    // progress) gives the same film as one call
    r->worker = std::thread([r, dev, cam, prm, spp]() {
        const int32 chunk = std::max(1, (spp + 15) / 16);
        for (int32 s0 = 0; s0 < spp; s0 += chunk) {
            BbRenderParams p = prm;
            p.sample_begin = s0;
            p.spp = std::min(chunk, spp - s0);
            check(bbgpu_render(dev, &cam, &p, r->film.sums.data(), r->film.counts.data(), nullptr, nullptr, 0),
                  "PathIntegrator::Render");
            r->samples_done.store(int64(s0 + p.spp) * cam.width * cam.height, std::memory_order_relaxed);
        }
        r->done.store(true, std::memory_order_release);
    });
    return r;
}

} // namespace bulbit
