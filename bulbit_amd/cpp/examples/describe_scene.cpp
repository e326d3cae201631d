// Loads a Mitsuba 3 XML scene with LoadScene and prints what it built as
// JSON: the camera, film, filter, sampler and integrator settings, the
// materials with their textures, every primitive (sphere placement or
// triangle corners with normals, tangents and texture coordinates) and the
// lights.  Floats are printed with 9 significant digits, which round-trips
// float32.
//   describe_scene scene.xml
#include "bulbit/bulbit.h"

#include <cstdio>
#include <map>
#include <string>

using namespace bulbit;

namespace
{

std::string num(Float v)
{
    char b[32];
    std::snprintf(b, sizeof b, "%.9g", double(v));
    return b;
}
std::string vec(const Vec3& v) { return "[" + num(v.x) + "," + num(v.y) + "," + num(v.z) + "]"; }
std::string spec(const Spectrum& v) { return "[" + num(v.r) + "," + num(v.g) + "," + num(v.b) + "]"; }
std::string quat(const Quat& q) { return "[" + num(q.x) + "," + num(q.y) + "," + num(q.z) + "," + num(q.w) + "]"; }
std::string transform(const Transform& t)
{
    return "{\"p\":" + vec(t.p) + ",\"q\":" + quat(t.q) + ",\"s\":" + vec(t.s) + "}";
}

template <typename T>
std::string value(const T& v);
template <>
std::string value<Float>(const Float& v) { return num(v); }
template <>
std::string value<Spectrum>(const Spectrum& v) { return spec(v); }

template <typename T>
std::string texture(const Texture<T>* t)
{
    if (!t) return "null";
    if (auto* c = dynamic_cast<const ConstantTexture<T>*>(t)) return "{\"constant\":" + value<T>(c->value) + "}";
    if (auto* c = dynamic_cast<const CheckerTexture<T>*>(t))
        return "{\"checker\":[" + texture<T>(c->a) + "," + texture<T>(c->b) + "],\"resolution\":[" +
               num(c->resolution.x) + "," + num(c->resolution.y) + "]}";
    if (auto* c = dynamic_cast<const ImageTexture<T>*>(t)) {
        std::string s = "{\"image\":[" + std::to_string(c->image.width) + "," + std::to_string(c->image.height) +
                        "],\"texels\":[";
        for (int32 i = 0; i < c->image.width * c->image.height; ++i)
            s += (i ? "," : "") + value<T>(c->image[i]);
        return s + "]}";
    }
    return "\"unknown\"";
}

std::string material(const Material* m)
{
    if (auto* d = dynamic_cast<const DiffuseMaterial*>(m))
        return "{\"type\":\"diffuse\",\"reflectance\":" + texture(d->reflectance) + ",\"roughness\":" +
               texture(d->roughness) + ",\"normal\":" + texture(d->normal) + ",\"alpha\":" + texture(d->alpha) + "}";
    if (auto* d = dynamic_cast<const DielectricMaterial*>(m))
        return "{\"type\":\"dielectric\",\"eta\":" + num(d->eta) + ",\"u_roughness\":" + texture(d->u_roughness) +
               ",\"v_roughness\":" + texture(d->v_roughness) + ",\"reflectance\":" + texture(d->reflectance) +
               ",\"energy_compensation\":" + (d->energy_compensation ? "true" : "false") + ",\"normal\":" +
               texture(d->normal) + "}";
    if (auto* c = dynamic_cast<const ConductorMaterial*>(m))
        return "{\"type\":\"conductor\",\"eta\":" + texture(c->eta) + ",\"k\":" + texture(c->k) +
               ",\"u_roughness\":" + texture(c->u_roughness) + ",\"v_roughness\":" + texture(c->v_roughness) +
               ",\"reflectance\":" + texture(c->reflectance) + ",\"energy_compensation\":" +
               (c->energy_compensation ? "true" : "false") + ",\"normal\":" + texture(c->normal) +
               ",\"alpha\":" + texture(c->alpha) + "}";
    if (auto* r = dynamic_cast<const MirrorMaterial*>(m))
        return "{\"type\":\"mirror\",\"reflectance\":" + texture(r->reflectance) + ",\"normal\":" +
               texture(r->normal) + ",\"alpha\":" + texture(r->alpha) + "}";
    if (auto* r = dynamic_cast<const MetallicRoughnessMaterial*>(m))
        return "{\"type\":\"metallic_roughness\",\"basecolor\":" + texture(r->basecolor) + ",\"metallic\":" +
               texture(r->metallic) + ",\"u_roughness\":" + texture(r->u_roughness) + ",\"v_roughness\":" +
               texture(r->v_roughness) + ",\"normal\":" + texture(r->normal) + ",\"alpha\":" + texture(r->alpha) + "}";
    return "null";
}

const char* filter_name(FilterType t)
{
    return t == FilterType::box ? "box" : (t == FilterType::tent ? "tent" : "gaussian");
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: describe_scene scene.xml\n");
        return 2;
    }
    RendererInfo ri;
    if (!LoadScene(&ri, argv[1])) return 1;
    const CameraInfo& ci = ri.camera_info;
    const FilterInfo& fi = ci.film_info.filter_info;
    std::printf("{\"camera\":{\"transform\":%s,\"fov\":%s,\"aperture_radius\":%s,\"focus_distance\":%s,"
                "\"resolution\":[%d,%d],\"filename\":\"%s\",\"filter\":{\"type\":\"%s\",\"extent\":%s,\"stddev\":%s},"
                "\"spp\":%d},\n",
                transform(ci.transform).c_str(), num(ci.fov).c_str(), num(ci.aperture_radius).c_str(),
                num(ci.focus_distance).c_str(), ci.film_info.resolution.x, ci.film_info.resolution.y,
                ci.film_info.filename.c_str(), filter_name(fi.type), num(fi.extent).c_str(),
                num(fi.gaussian_stddev).c_str(), ci.sampler_info.spp);
    std::printf("\"integrator\":{\"max_bounces\":%d,\"rr_min_bounces\":%d},\n", ri.integrator_info.max_bounces,
                ri.integrator_info.rr_min_bounces);

    std::map<const Material*, int> mat_id;
    std::string mats;
    const auto& prims = ri.scene.GetPrimitives();
    for (const Primitive* p : prims) {
        const Material* m = p->GetMaterial();
        if (m && !mat_id.count(m)) {
            const int id = int(mat_id.size());
            mat_id[m] = id;
            mats += (id ? ",\n" : "") + material(m);
        }
    }
    std::printf("\"materials\":[%s],\n\"primitives\":[", mats.c_str());
    std::map<const Primitive*, int> prim_id;
    for (size_t i = 0; i < prims.size(); ++i) {
        const Primitive* p = prims[i];
        prim_id[p] = int(i);
        const int mid = p->GetMaterial() ? mat_id[p->GetMaterial()] : -1;
        std::string s;
        if (auto* sp = dynamic_cast<const Sphere*>(p->GetShape())) {
            s = "{\"sphere\":" + transform(sp->transform) + ",\"radius\":" + num(sp->radius);
        } else if (auto* tr = dynamic_cast<const Triangle*>(p->GetShape())) {
            const Mesh* m = tr->GetMesh();
            s = "{\"triangle\":[";
            for (int k = 0; k < 3; ++k) {
                const int32 v = m->Indices()[3 * tr->GetIndex() + k];
                s += std::string(k ? "," : "") + "{\"p\":" + vec(m->Positions()[v]) + ",\"n\":" +
                     (m->Normals().empty() ? "null" : vec(m->Normals()[v])) + ",\"t\":" +
                     (m->Tangents().empty() ? "null" : vec(m->Tangents()[v])) + ",\"uv\":" +
                     (m->TexCoords().empty()
                          ? "null"
                          : "[" + num(m->TexCoords()[v].x) + "," + num(m->TexCoords()[v].y) + "]") +
                     "}";
            }
            s += "]";
        } else {
            s = "{\"unknown\":true";
        }
        std::printf("%s%s,\"material\":%d}", i ? ",\n" : "\n", s.c_str(), mid);
    }
    std::printf("],\n\"lights\":[");
    const auto& lights = ri.scene.GetLights();
    for (size_t i = 0; i < lights.size(); ++i) {
        std::string s;
        if (auto* l = dynamic_cast<const SpotLight*>(lights[i]))
            s = "{\"spot\":{\"position\":" + vec(l->position) + ",\"direction\":" + vec(l->direction) +
                ",\"intensity\":" + spec(l->intensity) + ",\"angle_max\":" + num(l->angle_max) +
                ",\"angle_falloff_start\":" + num(l->angle_falloff_start) + "}}";
        else if (auto* l = dynamic_cast<const DirectionalLight*>(lights[i]))
            s = "{\"directional\":{\"direction\":" + vec(l->direction) + ",\"intensity\":" + spec(l->intensity) + "}}";
        else if (auto* l = dynamic_cast<const PointLight*>(lights[i]))
            s = "{\"point\":{\"position\":" + vec(l->position) + ",\"intensity\":" + spec(l->intensity) + "}}";
        else if (auto* l = dynamic_cast<const UniformInfiniteLight*>(lights[i]))
            s = "{\"infinite\":{\"l\":" + spec(l->l) + ",\"scale\":" + num(l->scale) + "}}";
        else if (auto* l = dynamic_cast<const DiffuseAreaLight*>(lights[i]))
            s = "{\"area\":{\"primitive\":" + std::to_string(prim_id[l->primitive]) + ",\"emission\":" +
                texture(l->emission) + ",\"two_sided\":" + (l->two_sided ? "true" : "false") + "}}";
        else
            s = "{\"unknown\":true}";
        std::printf("%s%s", i ? ",\n" : "\n", s.c_str());
    }
    std::printf("]}\n");
    return 0;
}
