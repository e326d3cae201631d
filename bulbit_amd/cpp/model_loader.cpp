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

#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
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
            const size_t nv = obj.vertices.size() / 3, nt = obj.texcoords.size() / 2, nn = obj.normals.size() / 3;
            // an index the corner names must exist: 0, or one beyond either end, fails the read
            const auto bad = [](int i, size_t n) { return i < 0 || size_t(i) >= n; };
            while (in >> tok) {
                Index ix;
                int a = 0, b = 0, c = 0;
                bool ok;
                if (std::sscanf(tok.c_str(), "%d/%d/%d", &a, &b, &c) == 3) {
                    ix = { fix_index(a, nv), fix_index(b, nt), fix_index(c, nn) };
                    ok = !bad(ix.vt, nt) && !bad(ix.vn, nn);
                } else if (std::sscanf(tok.c_str(), "%d//%d", &a, &c) == 2) {
                    ix = { fix_index(a, nv), -1, fix_index(c, nn) };
                    ok = !bad(ix.vn, nn);
                } else if (std::sscanf(tok.c_str(), "%d/%d", &a, &b) == 2) {
                    ix = { fix_index(a, nv), fix_index(b, nt), -1 };
                    ok = !bad(ix.vt, nt);
                } else if (std::sscanf(tok.c_str(), "%d", &a) == 1) {
                    ix = { fix_index(a, nv), -1, -1 };
                    ok = true;
                } else {
                    return false;
                }
                if (!ok || bad(ix.v, nv)) return false;
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

// ------------------------------------------------------------------ glTF
//
// LoadGLTF (cli/model_loader.cpp:405-740) for models drawn with a fallback
// material.  The reference parses with tinygltf (its image loader replaced
// by a no-op) and turns every glTF material into a PrincipledMaterial, a
// BSDF outside this path: a model that needs its own materials is reported
// and not loaded.  Restated here is the glTF 2.0 container as the reference
// reads it: buffers (a file next to the .gltf, a base64 data URI, or a
// .glb's binary chunk), buffer views and accessors read as one contiguous
// block from byteOffset (byteStride is not applied, as in the reference),
// meshes of triangle primitives with POSITION, NORMAL, TEXCOORD_0 (all
// required), TANGENT (optional: w = 1 keeps, else negates xyz; absent,
// CoordinateSystem of the normal) and 8/16/32-bit indices; nodes with a
// 16-value column-major matrix (translation/rotation/scale are not applied,
// as in the reference) composed as Mul(parent, node); the default scene's
// root nodes ("scene", or 0 when absent).

namespace
{

struct Json
{
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json* get(const std::string& k) const
    {
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    double number(const std::string& k, double dflt) const
    {
        const Json* v = get(k);
        return v && v->kind == Number ? v->num : dflt;
    }
    std::string string(const std::string& k) const
    {
        const Json* v = get(k);
        return v && v->kind == String ? v->str : std::string();
    }
};

struct JsonParser
{
    const char* p;
    const char* end;

    [[noreturn]] void fail(const char* what) { throw std::runtime_error(std::string("glTF JSON: ") + what); }
    void ws()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if (size_t(end - p) >= n && std::strncmp(p, w, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    static void utf8(std::string& out, unsigned c)
    {
        if (c < 0x80) {
            out += char(c);
        } else if (c < 0x800) {
            out += char(0xC0 | (c >> 6));
            out += char(0x80 | (c & 0x3F));
        } else if (c < 0x10000) {
            out += char(0xE0 | (c >> 12));
            out += char(0x80 | ((c >> 6) & 0x3F));
            out += char(0x80 | (c & 0x3F));
        } else {
            out += char(0xF0 | (c >> 18));
            out += char(0x80 | ((c >> 12) & 0x3F));
            out += char(0x80 | ((c >> 6) & 0x3F));
            out += char(0x80 | (c & 0x3F));
        }
    }
    unsigned hex4()
    {
        if (end - p < 4) fail("short \\u escape");
        unsigned c = 0;
        for (int i = 0; i < 4; ++i, ++p) {
            const char h = *p;
            c <<= 4;
            if (h >= '0' && h <= '9') c |= unsigned(h - '0');
            else if (h >= 'a' && h <= 'f') c |= unsigned(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') c |= unsigned(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return c;
    }
    std::string str()
    {
        if (p >= end || *p != '"') fail("expected a string");
        ++p;
        std::string out;
        while (p < end && *p != '"') {
            if (*p != '\\') {
                out += *p++;
                continue;
            }
            if (++p >= end) fail("unterminated escape");
            const char e = *p++;
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                unsigned c = hex4();
                if (c >= 0xD800 && c < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                    p += 2;
                    const unsigned lo = hex4();
                    c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
                }
                utf8(out, c);
                break;
            }
            default: fail("bad escape");
            }
        }
        if (p >= end) fail("unterminated string");
        ++p;
        return out;
    }
    Json value(int depth)
    {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (p >= end) fail("unexpected end");
        Json v;
        if (*p == '{') {
            ++p;
            v.kind = Json::Object;
            ws();
            if (p < end && *p == '}') {
                ++p;
                return v;
            }
            for (;;) {
                ws();
                std::string k = str();
                ws();
                if (p >= end || *p != ':') fail("expected ':'");
                ++p;
                v.obj.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == '}') {
                    ++p;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (*p == '[') {
            ++p;
            v.kind = Json::Array;
            ws();
            if (p < end && *p == ']') {
                ++p;
                return v;
            }
            for (;;) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == ']') {
                    ++p;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (*p == '"') {
            v.kind = Json::String;
            v.str = str();
            return v;
        }
        if (lit("true")) {
            v.kind = Json::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Json::Bool;
            return v;
        }
        if (lit("null")) return v;
        // a number: strtod over its characters (correctly rounded)
        const char* q = p;
        while (q < end &&
               (std::isdigit((unsigned char)*q) || *q == '-' || *q == '+' || *q == '.' || *q == 'e' || *q == 'E'))
            ++q;
        if (q == p) fail("unexpected character");
        const std::string t(p, q);
        char* stop = nullptr;
        v.kind = Json::Number;
        v.num = std::strtod(t.c_str(), &stop);
        if (stop != t.c_str() + t.size()) fail("bad number");
        p = q;
        return v;
    }
};

Json parse_json(const std::string& text)
{
    JsonParser ps{ text.data(), text.data() + text.size() };
    Json v = ps.value(0);
    ps.ws();
    if (ps.p != ps.end) ps.fail("trailing characters");
    return v;
}

bool read_file(const std::string& path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool base64(const std::string& in, size_t from, std::vector<unsigned char>& out)
{
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    unsigned acc = 0;
    int bits = 0;
    for (size_t i = from; i < in.size(); ++i) {
        if (in[i] == '=') break;
        const int v = val(in[i]);
        if (v < 0) return false;
        acc = ((acc << 6) | unsigned(v)) & 0xFFFFFFu;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((unsigned char)((acc >> bits) & 0xFF));
        }
    }
    return true;
}

std::string uri_decode(const std::string& u)
{
    std::string out;
    for (size_t i = 0; i < u.size(); ++i) {
        if (u[i] == '%' && i + 2 < u.size() && std::isxdigit((unsigned char)u[i + 1]) &&
            std::isxdigit((unsigned char)u[i + 2])) {
            out += char(std::stoi(u.substr(i + 1, 2), nullptr, 16));
            i += 2;
        } else {
            out += u[i];
        }
    }
    return out;
}

struct GltfModel
{
    Json doc;
    std::vector<std::vector<unsigned char>> buffers;
};

// the bytes of accessor `index`: count elements of type `type` (comps
// components), componentType want_ct (0: any integer or float type), read as
// one block from the view's and the accessor's byteOffset
bool accessor_bytes(const GltfModel& m, int index, const char* type, int comps, int want_ct, const unsigned char*& data,
                    size_t& count, int* ct_out = nullptr)
{
    const Json* accs = m.doc.get("accessors");
    if (!accs || index < 0 || size_t(index) >= accs->arr.size()) return false;
    const Json& a = accs->arr[size_t(index)];
    const int ct = int(a.number("componentType", -1));
    if (a.string("type") != type) return false;
    int csize = 0;
    switch (ct) {
    case 5120: case 5121: csize = 1; break;
    case 5122: case 5123: csize = 2; break;
    case 5125: case 5126: csize = 4; break;
    default: return false;
    }
    if (want_ct && ct != want_ct) return false;
    if (ct_out) *ct_out = ct;
    const Json* views = m.doc.get("bufferViews");
    const int vi = int(a.number("bufferView", -1));
    if (!views || vi < 0 || size_t(vi) >= views->arr.size()) return false;
    const Json& v = views->arr[size_t(vi)];
    const int bi = int(v.number("buffer", -1));
    if (bi < 0 || size_t(bi) >= m.buffers.size()) return false;
    const double voff = v.number("byteOffset", 0), aoff = a.number("byteOffset", 0), n = a.number("count", 0);
    if (voff < 0 || aoff < 0 || n < 0) return false;
    const size_t off = size_t(voff) + size_t(aoff);
    count = size_t(n);
    const std::vector<unsigned char>& buf = m.buffers[size_t(bi)];
    if (off > buf.size() || count > (buf.size() - off) / (size_t(comps) * size_t(csize))) return false;
    data = buf.data() + off;
    return true;
}

template <typename T>
T load_le(const unsigned char* p)
{
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

// LoadMesh (:405-629); false (the mesh's remaining primitives are left out,
// as in the reference) when a primitive lacks what it needs
bool load_gltf_mesh(Scene& scene, const GltfModel& m, const Json& mesh, const Mat4& transform,
                    const ModelLoaderOptions& options)
{
    const Json* prims = mesh.get("primitives");
    if (!prims) return true;
    for (const Json& prim : prims->arr) {
        if (int(prim.number("mode", 4)) != 4) continue; // TRIANGLES only
        const Json* attrs = prim.get("attributes");
        if (!attrs) return false;
        const int pos_i = int(attrs->number("POSITION", -1)), nrm_i = int(attrs->number("NORMAL", -1));
        const int tan_i = int(attrs->number("TANGENT", -1)), uv_i = int(attrs->number("TEXCOORD_0", -1));
        const int idx_i = int(prim.number("indices", -1));
        if (pos_i == -1 || nrm_i == -1 || uv_i == -1) return false;
        const unsigned char *pd = nullptr, *nd = nullptr, *ud = nullptr, *id = nullptr;
        size_t pn = 0, nn = 0, un = 0, in = 0;
        int ict = 0;
        if (!accessor_bytes(m, pos_i, "VEC3", 3, 5126, pd, pn) || !accessor_bytes(m, nrm_i, "VEC3", 3, 5126, nd, nn) ||
            !accessor_bytes(m, uv_i, "VEC2", 2, 5126, ud, un) ||
            !accessor_bytes(m, idx_i, "SCALAR", 1, 0, id, in, &ict))
            return false;
        if (ict != 5121 && ict != 5123 && ict != 5125) return false;
        std::vector<Point3> positions(pn);
        std::vector<Vec3> normals(nn);
        std::vector<Vec3> tangents(nn);
        std::vector<Point2> texcoords(un);
        std::vector<int32> indices(in);
        for (size_t i = 0; i < pn; ++i)
            positions[i] =
                Point3(load_le<float>(pd + 12 * i), load_le<float>(pd + 12 * i + 4), load_le<float>(pd + 12 * i + 8));
        for (size_t i = 0; i < nn; ++i)
            normals[i] =
                Vec3(load_le<float>(nd + 12 * i), load_le<float>(nd + 12 * i + 4), load_le<float>(nd + 12 * i + 8));
        for (size_t i = 0; i < un; ++i) texcoords[i] = Point2(load_le<float>(ud + 8 * i), load_le<float>(ud + 8 * i + 4));
        if (tan_i != -1) {
            const unsigned char* td = nullptr;
            size_t tn = 0;
            if (!accessor_bytes(m, tan_i, "VEC4", 4, 5126, td, tn)) return false;
            for (size_t i = 0; i < tn && i < tangents.size(); ++i) {
                const float x = load_le<float>(td + 16 * i), y = load_le<float>(td + 16 * i + 4),
                            z = load_le<float>(td + 16 * i + 8), w = load_le<float>(td + 16 * i + 12);
                tangents[i] = w == 1 ? Vec3(x, y, z) : Vec3(-x, -y, -z);
            }
        } else {
            for (size_t i = 0; i < normals.size(); ++i) {
                Vec3 bitangent;
                coordinate_system(normals[i], &tangents[i], &bitangent);
            }
        }
        for (size_t i = 0; i < in; ++i)
            indices[i] = ict == 5121 ? int32(id[i])
                                     : (ict == 5123 ? int32(load_le<uint16_t>(id + 2 * i))
                                                    : int32(load_le<uint32_t>(id + 4 * i)));
        // every vertex attribute per position, and every index inside them
        if (normals.size() != positions.size() || texcoords.size() != positions.size()) return false;
        for (int32 ix : indices)
            if (ix < 0 || size_t(ix) >= positions.size()) return false;
        // post-processes
        if (options.flip_normal)
            for (Vec3& n : normals) n = -n;
        if (!options.flip_tex_coord)
            for (Point2& t : texcoords) t.y = -t.y;
        Mesh* msh = scene.CreateMesh(std::move(positions), std::move(normals), std::move(tangents), std::move(texcoords),
                                     std::move(indices), transform);
        const Material* material = options.fallback_material;
        if (!material) { // a primitive without material and no fallback
            std::fprintf(stderr, "No valid material found and no fallback specified\n");
            material =
                scene.CreateMaterial<DiffuseMaterial>(scene.CreateTexture<ConstantTexture, Spectrum>(Spectrum(1, 0, 1)));
        }
        const std::optional<AreaLightInfo>& ali = options.area_light;
        // CreateTriangles (cli/primitive_builder.cpp:70-103)
        for (int32 i = 0; i < msh->GetTriangleCount(); ++i) {
            Shape* tri = scene.CreateShape<Triangle>(msh, i);
            Primitive* p = scene.CreatePrimitive(tri, material, options.fallback_medium_interface);
            if (ali) scene.CreateLight<DiffuseAreaLight>(p, emission_texture(scene, *ali), ali->two_sided);
        }
    }
    return true;
}

// ProcessNode (:631-668)
void process_gltf_node(Scene& scene, const GltfModel& m, int index, const Mat4& parent,
                       const ModelLoaderOptions& options, int depth)
{
    const Json* nodes = m.doc.get("nodes");
    if (!nodes || index < 0 || size_t(index) >= nodes->arr.size() || depth > 1024) return;
    const Json& node = nodes->arr[size_t(index)];
    Mat4 local = identity;
    const Json* mx = node.get("matrix");
    if (mx && mx->arr.size() == 16) {
        Float v[16];
        for (int i = 0; i < 16; ++i) v[i] = Float(mx->arr[size_t(i)].num);
        local = Mat4(Vec4(v[0], v[1], v[2], v[3]), Vec4(v[4], v[5], v[6], v[7]), Vec4(v[8], v[9], v[10], v[11]),
                     Vec4(v[12], v[13], v[14], v[15]));
    }
    const Mat4 transform = Mul(parent, local);
    const int mi = int(node.number("mesh", -1));
    const Json* meshes = m.doc.get("meshes");
    if (meshes && mi >= 0 && size_t(mi) < meshes->arr.size())
        load_gltf_mesh(scene, m, meshes->arr[size_t(mi)], transform, options);
    if (const Json* ch = node.get("children"))
        for (const Json& c : ch->arr) process_gltf_node(scene, m, int(c.num), transform, options, depth + 1);
}

bool read_gltf(const std::filesystem::path& filename, GltfModel& m)
{
    std::string data;
    if (!read_file(filename.string(), data)) return false;
    const unsigned char* u = reinterpret_cast<const unsigned char*>(data.data());
    std::vector<unsigned char> bin;
    bool has_bin = false;
    std::string json;
    if (filename.extension() == ".glb") {
        // header: magic "glTF", version, length; chunks: length, type, data
        if (data.size() < 12 || load_le<uint32_t>(u) != 0x46546C67u) return false;
        size_t off = 12;
        while (off + 8 <= data.size()) {
            const uint32_t len = load_le<uint32_t>(u + off);
            const uint32_t type = load_le<uint32_t>(u + off + 4);
            off += 8;
            if (len > data.size() - off) return false;
            if (type == 0x4E4F534Au && json.empty()) {
                json.assign(data, off, len);
            } else if (type == 0x004E4942u && !has_bin) {
                bin.assign(u + off, u + off + len);
                has_bin = true;
            }
            off += len;
        }
        if (json.empty()) return false;
    } else {
        json = data;
    }
    m.doc = parse_json(json);
    std::string base = filename.parent_path().string();
    if (!base.empty() && base.back() != '/') base += "/";
    if (const Json* bufs = m.doc.get("buffers")) {
        for (const Json& b : bufs->arr) {
            const std::string uri = b.string("uri");
            std::vector<unsigned char> bytes;
            if (uri.empty()) {
                if (!has_bin) return false;
                bytes = bin;
            } else if (uri.rfind("data:", 0) == 0) {
                const size_t at = uri.find(";base64,");
                if (at == std::string::npos || !base64(uri, at + 8, bytes)) return false;
            } else {
                std::string raw;
                if (!read_file(base + uri_decode(uri), raw)) return false;
                bytes.assign(raw.begin(), raw.end());
            }
            if (double(bytes.size()) < b.number("byteLength", 0)) return false;
            m.buffers.push_back(std::move(bytes));
        }
    }
    return true;
}

bool gltf_needs_model_materials(const GltfModel& m)
{
    if (const Json* meshes = m.doc.get("meshes"))
        for (const Json& mesh : meshes->arr)
            if (const Json* prims = mesh.get("primitives"))
                for (const Json& p : prims->arr)
                    if (p.number("material", -1) >= 0) return true;
    return false;
}

} // namespace

void LoadGLTF(Scene& scene, std::filesystem::path filename, const Transform& transform,
              const ModelLoaderOptions& options)
{
    if (options.use_fallback_material && options.fallback_material == nullptr) {
        std::fprintf(stderr, "Fallback material not provided!\n");
        return;
    }
    const std::string ext = filename.extension().string();
    if (ext != ".gltf" && ext != ".glb") {
        std::fprintf(stderr, "Failed to load model: %s\n", filename.string().c_str());
        return;
    }
    GltfModel m;
    try {
        if (!read_gltf(filename, m)) {
            std::fprintf(stderr, "gltf error: cannot read %s\n", filename.string().c_str());
            return;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "gltf error: %s (%s)\n", e.what(), filename.string().c_str());
        return;
    }
    // the model's own materials would be PrincipledMaterials
    if (!options.fallback_material && gltf_needs_model_materials(m)) {
        std::fprintf(stderr,
                     "LoadGLTF: %s: glTF materials (PrincipledMaterial) are not supported on the GPU path; "
                     "pass a fallback material\n",
                     filename.string().c_str());
        return;
    }
    const Json* scenes = m.doc.get("scenes");
    const int si = int(m.doc.number("scene", 0));
    if (!scenes || si < 0 || size_t(si) >= scenes->arr.size()) return;
    const Mat4 root(transform);
    if (const Json* roots = scenes->arr[size_t(si)].get("nodes"))
        for (const Json& r : roots->arr) process_gltf_node(scene, m, int(r.num), root, options, 0);
}

// LoadModel (:1045-1063): dispatch on the extension
void LoadModel(Scene& scene, std::filesystem::path filename, const Transform& transform,
               const ModelLoaderOptions& options)
{
    const std::string ext = filename.extension().string();
    if (ext == ".obj") {
        LoadOBJ(scene, filename, transform, options);
    } else if (ext == ".gltf" || ext == ".glb") {
        LoadGLTF(scene, filename, transform, options);
    } else {
        std::fprintf(stderr, "Failed to load model: %s\nSupported extensions: .obj, .gltf, .glb\n",
                     filename.string().c_str());
    }
}

} // namespace bulbit
