// model_loader.cpp — LoadOBJ / LoadModel of Bulbit's CLI (cli/model_loader.cpp)
// for the library API: an OBJ + MTL reader and the reference's conversion of
// its shapes, faces and materials into meshes, MetallicRoughness materials
// and area lights.
//
// The reference parses with tinyobjloader (pinned at 7dbc543d, not vendored
// in the reference) with its default ObjReaderConfig.  What is restated here
// is the published OBJ/MTL format as that loader reads it:
//   v / vn / vt (extra components ignored), f with v, v/vt, v//vn, v/vt/vn
//   and negative (relative) indices, usemtl, mtllib, o / g (a new shape
//   once the current one has faces); newmtl, Kd, Ke, map_Kd, map_Ke,
//   bump / map_bump / map_Bump (options before the file name skipped).
// Polygons with more than three vertices are split into a fan from their
// first vertex, and numbers are parsed correctly rounded; tinyobjloader's
// own polygon triangulation and number parser are not restated.
#include "bulbit/bulbit.h"

#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace bulbit
{

namespace
{

// CoordinateSystem (include/bulbit/frame.h:105-112)
void coordinate_system(const Vec3& v1, Vec3* v2, Vec3* v3)
{
    const Float sign = std::copysign(Float(1), v1.z);
    const Float a = -1 / (sign + v1.z);
    const Float b = v1.x * v1.y * a;
    *v2 = Vec3(1 + sign * (v1.x * v1.x) * a, sign * b, -sign * v1.x);
    *v3 = Vec3(b, sign + (v1.y * v1.y) * a, -v1.y);
}

struct Index
{
    int v = -1, vt = -1, vn = -1;
};

struct ObjShape
{
    std::vector<int> face_vertices; // vertices per face (all 3: faces are split on read)
    std::vector<Index> indices;
    std::vector<int> material_ids;
};

struct ObjMaterial
{
    std::string name;
    Float kd[3] = { 0, 0, 0 }; // colours a material does not state are 0
    Float ke[3] = { 0, 0, 0 };
    std::string map_kd, map_ke, bump;
};

struct Obj
{
    std::vector<Float> vertices, normals, texcoords;
    std::vector<ObjShape> shapes;
    std::vector<ObjMaterial> materials;
};

// the file name after a map statement's options (-bm 1, -o u v w, ...)
std::string texture_name(std::istringstream& in)
{
    std::vector<std::string> t;
    std::string w;
    while (in >> w) t.push_back(w);
    return t.empty() ? std::string() : t.back();
}

void parse_mtl(const std::string& path, Obj& obj, std::map<std::string, int>& names)
{
    std::ifstream f(path);
    if (!f) {
        std::fprintf(stderr, "LoadOBJ: material library %s not found\n", path.c_str());
        return;
    }
    std::string line;
    ObjMaterial* m = nullptr;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string key;
        if (!(in >> key) || key[0] == '#') continue;
        if (key == "newmtl") {
            ObjMaterial nm;
            in >> nm.name;
            names[nm.name] = int(obj.materials.size());
            obj.materials.push_back(nm);
            m = &obj.materials.back();
        } else if (!m) {
            continue;
        } else if (key == "Kd") {
            in >> m->kd[0] >> m->kd[1] >> m->kd[2];
        } else if (key == "Ke") {
            in >> m->ke[0] >> m->ke[1] >> m->ke[2];
        } else if (key == "map_Kd") {
            m->map_kd = texture_name(in);
        } else if (key == "map_Ke") {
            m->map_ke = texture_name(in);
        } else if (key == "bump" || key == "map_bump" || key == "map_Bump") {
            m->bump = texture_name(in);
        }
    }
}

// OBJ index (1-based, negative: relative to the current count) -> 0-based
int fix_index(int i, size_t count) { return i > 0 ? i - 1 : (i < 0 ? int(count) + i : -1); }

bool parse_obj(const std::string& filename, const std::string& base, Obj& obj)
{
    std::ifstream f(filename);
    if (!f) return false;
    std::map<std::string, int> names;
    int material = -1;
    obj.shapes.emplace_back();
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string key;
        if (!(in >> key) || key[0] == '#') continue;
        if (key == "v") {
            Float x = 0, y = 0, z = 0;
            in >> x >> y >> z;
            obj.vertices.insert(obj.vertices.end(), { x, y, z });
        } else if (key == "vn") {
            Float x = 0, y = 0, z = 0;
            in >> x >> y >> z;
            obj.normals.insert(obj.normals.end(), { x, y, z });
        } else if (key == "vt") {
            Float u = 0, v = 0;
            in >> u >> v;
            obj.texcoords.insert(obj.texcoords.end(), { u, v });
        } else if (key == "f") {
            std::vector<Index> face;
            std::string tok;
            while (in >> tok) {
                Index ix;
                int a = 0, b = 0, c = 0;
                if (std::sscanf(tok.c_str(), "%d/%d/%d", &a, &b, &c) == 3) {
                    ix = { fix_index(a, obj.vertices.size() / 3), fix_index(b, obj.texcoords.size() / 2),
                           fix_index(c, obj.normals.size() / 3) };
                } else if (std::sscanf(tok.c_str(), "%d//%d", &a, &c) == 2) {
                    ix = { fix_index(a, obj.vertices.size() / 3), -1, fix_index(c, obj.normals.size() / 3) };
                } else if (std::sscanf(tok.c_str(), "%d/%d", &a, &b) == 2) {
                    ix = { fix_index(a, obj.vertices.size() / 3), fix_index(b, obj.texcoords.size() / 2), -1 };
                } else if (std::sscanf(tok.c_str(), "%d", &a) == 1) {
                    ix = { fix_index(a, obj.vertices.size() / 3), -1, -1 };
                } else {
                    return false;
                }
                if (ix.v < 0 || size_t(ix.v) >= obj.vertices.size() / 3) return false;
                face.push_back(ix);
            }
            ObjShape& s = obj.shapes.back();
            for (size_t k = 1; k + 1 < face.size(); ++k) {
                s.face_vertices.push_back(3);
                s.indices.insert(s.indices.end(), { face[0], face[k], face[k + 1] });
                s.material_ids.push_back(material);
            }
        } else if (key == "usemtl") {
            std::string name;
            in >> name;
            auto it = names.find(name);
            material = it == names.end() ? -1 : it->second;
        } else if (key == "mtllib") {
            std::string name;
            while (in >> name) parse_mtl(base + name, obj, names);
        } else if (key == "o" || key == "g") {
            if (!obj.shapes.back().face_vertices.empty()) obj.shapes.emplace_back();
        }
    }
    if (obj.shapes.back().face_vertices.empty() && obj.shapes.size() > 1) obj.shapes.pop_back();
    return true;
}

// HashMap<int, group> (include/bulbit/hash_map.h): Robin Hood hashing with
// linear probing, std::hash<int> (the identity), 32 buckets to start,
// doubled (reinserting in bucket order) once 70 % are occupied; iteration
// visits buckets in order.  LoadOBJ creates its meshes in this order.
template <typename V>
class RobinHoodMap
{
public:
    RobinHoodMap() { resize(32); }

    bool contains(int key) const
    {
        const size_t mask = keys.size() - 1;
        size_t index = size_t(key) & mask, dist = 0;
        for (;;) {
            if (dists[index] == -1 || dist > size_t(dists[index])) return false;
            if (keys[index] == key) return true;
            index = (index + 1) & mask;
            ++dist;
        }
    }

    void insert(int key, V value)
    {
        if (float(occupied) >= float(keys.size()) * 0.7f) rehash(keys.size() * 2);
        const size_t mask = keys.size() - 1;
        size_t index = size_t(key) & mask;
        int32 dist = 0;
        for (;;) {
            if (dists[index] == -1) {
                keys[index] = key;
                values[index] = std::move(value);
                dists[index] = dist;
                ++occupied;
                return;
            }
            if (keys[index] == key) {
                values[index] = std::move(value);
                return;
            }
            if (dist > dists[index]) {
                std::swap(key, keys[index]);
                std::swap(value, values[index]);
                std::swap(dist, dists[index]);
            }
            ++dist;
            index = (index + 1) & mask;
        }
    }

    V& at(int key)
    {
        const size_t mask = keys.size() - 1;
        size_t index = size_t(key) & mask;
        while (dists[index] == -1 || keys[index] != key) index = (index + 1) & mask;
        return values[index];
    }

    template <typename F>
    void for_each(F&& f)
    {
        for (size_t i = 0; i < keys.size(); ++i)
            if (dists[i] != -1) f(keys[i], values[i]);
    }

private:
    void resize(size_t n)
    {
        keys.assign(n, 0);
        values.assign(n, V{});
        dists.assign(n, -1);
        occupied = 0;
    }

    void rehash(size_t n)
    {
        std::vector<int> old_keys = std::move(keys);
        std::vector<V> old_values = std::move(values);
        std::vector<int32> old_dists = std::move(dists);
        resize(n);
        for (size_t i = 0; i < old_keys.size(); ++i)
            if (old_dists[i] != -1) insert(old_keys[i], std::move(old_values[i]));
    }

    std::vector<int> keys;
    std::vector<V> values;
    std::vector<int32> dists;
    size_t occupied = 0;
};

struct Group
{
    std::vector<Point3> positions;
    std::vector<Vec3> normals;
    std::vector<Vec3> tangents;
    std::vector<Point2> texcoords;
    std::vector<int32> indices;
};

template <typename T>
const ImageTexture<T>* image_texture(Scene& scene, Image<T> image)
{
    // the texture pool returns no texture for an image that failed to load
    if (!image) return nullptr;
    return scene.CreateTexture<ImageTexture, T>(std::move(image));
}

// CreateOBJMaterial (cli/model_loader.cpp:742-790)
std::pair<const Material*, const SpectrumTexture*> create_material(Scene& scene, const ObjMaterial& m,
                                                                   const std::string& root)
{
    const Spectrum basecolor_factor(m.kd[0], m.kd[1], m.kd[2]);
    const Spectrum emission_factor(m.ke[0], m.ke[1], m.ke[2]);
    const SpectrumTexture* basecolor = nullptr;
    const FloatTexture* alpha = nullptr;
    if (!m.map_kd.empty()) {
        basecolor = image_texture(scene, ReadImage3(root + m.map_kd));
        alpha = image_texture(scene, ReadImage1(root + m.map_kd, 3, true));
    } else {
        basecolor = scene.CreateTexture<ConstantTexture, Spectrum>(basecolor_factor);
    }
    const FloatTexture* metallic = scene.CreateTexture<ConstantTexture, Float>(Float(0));
    const FloatTexture* roughness = scene.CreateTexture<ConstantTexture, Float>(Float(1));
    const SpectrumTexture* normal = nullptr;
    if (!m.bump.empty()) normal = image_texture(scene, ReadImage3(root + m.bump, true));
    const SpectrumTexture* emission = nullptr;
    if (!m.map_ke.empty())
        emission = image_texture(scene, ReadImage3(root + m.map_ke));
    else
        emission = scene.CreateTexture<ConstantTexture, Spectrum>(emission_factor);
    const Material* mat =
        scene.CreateMaterial<MetallicRoughnessMaterial>(basecolor, metallic, roughness, roughness, normal, alpha);
    return { mat, emission };
}

// get_emission_texture (cli/primitive_builder.cpp:14-34)
const SpectrumTexture* emission_texture(Scene& scene, const AreaLightInfo& a)
{
    if (auto* f = std::get_if<Float>(&a.emission)) return scene.CreateTexture<ConstantTexture, Spectrum>(Spectrum(*f));
    if (auto* s = std::get_if<Spectrum>(&a.emission)) return scene.CreateTexture<ConstantTexture, Spectrum>(*s);
    return std::get<const SpectrumTexture*>(a.emission);
}

} // namespace

void LoadOBJ(Scene& scene, std::filesystem::path filename, const Transform& transform, const ModelLoaderOptions& options)
{
    if (options.use_fallback_material && options.fallback_material == nullptr) {
        std::fprintf(stderr, "Fallback material not provided!\n");
        return;
    }
    std::string base = filename.parent_path().string();
    if (!base.empty() && base.back() != '/') base += "/";
    Obj obj;
    if (!parse_obj(filename.string(), base, obj)) {
        std::fprintf(stderr, "LoadOBJ: failed to read %s\n", filename.string().c_str());
        return;
    }
    std::vector<std::pair<const Material*, const SpectrumTexture*>> materials;
    if (!options.use_fallback_material) {
        materials.resize(obj.materials.size());
        for (size_t i = 0; i < obj.materials.size(); ++i) materials[i] = create_material(scene, obj.materials[i], base);
    }
    const auto vtx = [&](int i) {
        return Vec3(obj.vertices[3 * size_t(i)], obj.vertices[3 * size_t(i) + 1], obj.vertices[3 * size_t(i) + 2]);
    };
    for (const ObjShape& shape : obj.shapes) {
        RobinHoodMap<Group> groups;
        std::vector<Vec3> smooth;
        if (options.gen_smooth_normal) {
            smooth.assign(obj.vertices.size() / 3, Vec3(0));
            for (size_t f = 0; f < shape.face_vertices.size(); ++f) {
                const Index& i0 = shape.indices[3 * f];
                const Index& i1 = shape.indices[3 * f + 1];
                const Index& i2 = shape.indices[3 * f + 2];
                const Vec3 v0 = vtx(i0.v), v1 = vtx(i1.v), v2 = vtx(i2.v);
                Vec3 face_normal = Cross(v1 - v0, v2 - v0);
                face_normal.Normalize();
                smooth[i0.v] = smooth[i0.v] + face_normal;
                smooth[i1.v] = smooth[i1.v] + face_normal;
                smooth[i2.v] = smooth[i2.v] + face_normal;
            }
            for (Vec3& n : smooth) n.Normalize();
        }
        for (size_t f = 0; f < shape.face_vertices.size(); ++f) {
            const int material_id = shape.material_ids[f];
            if (!groups.contains(material_id)) groups.insert(material_id, {});
            Group& g = groups.at(material_id);
            int32 face[3];
            for (int k = 0; k < 3; ++k) {
                const Index& ix = shape.indices[3 * f + k];
                const Vec3 position = vtx(ix.v);
                Vec3 normal(0);
                if (ix.vn >= 0)
                    normal = Vec3(obj.normals[3 * size_t(ix.vn)], obj.normals[3 * size_t(ix.vn) + 1],
                                  obj.normals[3 * size_t(ix.vn) + 2]);
                else if (options.gen_smooth_normal)
                    normal = smooth[ix.v];
                if (options.flip_normal) normal = -normal;
                Point2 uv(0, 0);
                if (ix.vt >= 0) {
                    uv = Point2(obj.texcoords[2 * size_t(ix.vt)], obj.texcoords[2 * size_t(ix.vt) + 1]);
                    if (options.flip_tex_coord) uv.y = 1 - uv.y;
                }
                Vec3 tangent(0);
                if (!(normal == Vec3(0))) {
                    Vec3 bitangent;
                    coordinate_system(normal, &tangent, &bitangent);
                }
                g.positions.push_back(position);
                g.normals.push_back(normal);
                g.tangents.push_back(tangent);
                g.texcoords.push_back(uv);
                face[k] = int32(g.positions.size() - 1);
            }
            g.indices.insert(g.indices.end(), { face[0], face[1], face[2] });
        }
        groups.for_each([&](int material_id, Group& g) {
            Mesh* mesh = scene.CreateMesh(std::move(g.positions), std::move(g.normals), std::move(g.tangents),
                                          std::move(g.texcoords), std::move(g.indices), Mat4(transform));
            const Material* material = nullptr;
            const SpectrumTexture* emission = nullptr;
            if (material_id < 0 || size_t(material_id) >= materials.size() || options.use_fallback_material) {
                if (options.fallback_material) {
                    material = options.fallback_material;
                } else {
                    std::fprintf(stderr, "No valid material found and no fallback specified\n");
                    material = scene.CreateMaterial<DiffuseMaterial>(
                        scene.CreateTexture<ConstantTexture, Spectrum>(Spectrum(1, 0, 1)));
                }
            } else if (options.fallback_material) {
                material = options.fallback_material;
            } else {
                material = materials[size_t(material_id)].first;
                emission = materials[size_t(material_id)].second;
            }
            std::optional<AreaLightInfo> ali = options.area_light;
            if (!ali && emission) {
                ali.emplace();
                ali->emission = emission;
            }
            // CreateTriangles (cli/primitive_builder.cpp:70-103)
            for (int32 i = 0; i < mesh->GetTriangleCount(); ++i) {
                Shape* tri = scene.CreateShape<Triangle>(mesh, i);
                Primitive* prim = scene.CreatePrimitive(tri, material, options.fallback_medium_interface);
                if (ali) scene.CreateLight<DiffuseAreaLight>(prim, emission_texture(scene, *ali), ali->two_sided);
            }
        });
    }
}

void LoadModel(Scene& scene, std::filesystem::path filename, const Transform& transform,
               const ModelLoaderOptions& options)
{
    const std::string ext = filename.extension().string();
    if (ext == ".obj") {
        LoadOBJ(scene, filename, transform, options);
    } else {
        std::fprintf(stderr, "Failed to load model: %s\nSupported extensions: .obj (.gltf and .glb are not built)\n",
                     filename.string().c_str());
    }
}

} // namespace bulbit
