// bulbit/bulbit.h — the library API of Bulbit's README example
// (README.md:136-190), rendered on the GPU through the C ABI of libbbgpu.so
// (include/bbgpu.h).
//
// A program written against Bulbit's Scene / BVH / PathIntegrator /
// IndependentSampler / PerspectiveCamera / BoxFilter / Film / Rendering API
// compiles against this header unchanged and renders with the HIP wavefront
// path tracer.  The covered surface is exactly the path the GPU library
// implements: constant, image (bilinear) and checker textures in every
// material parameter the path reads, normal maps and alpha tests; Diffuse
// (Lambertian), Dielectric, Conductor and MetallicRoughness materials; spheres and triangle meshes; spot, uniform
// infinite and diffuse area lights; box, tent and Gaussian filters.  Anything outside it is
// reported by the C ABI as unsupported and stops the program with a message
// (there is no CPU fallback).
//
// Numeric conventions follow the reference so that scene data is bit-identical:
// float32 throughout, Dot = (x*x + y*y) + z*z, free Normalize multiplies by
// 1 / Length (include/bulbit/math.h:56-60), the member Normalize returns early
// below FLT_EPSILON (include/bulbit/vectors.h:121-133), and Mesh applies
// Mul(Mat4, Vec4) (include/bulbit/matrix.h:481-488) to its vertices.
#pragma once

#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <functional>
#include <optional>
#include <variant>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

struct BbScene;

namespace bulbit
{

using Float = float;
using int8 = int8_t;
using int32 = int32_t;
using int64 = int64_t;
using uint8 = uint8_t;
using uint32 = uint32_t;
using uint64 = uint64_t;

// ------------------------------------------------------------------ vectors

struct Point2
{
    Float x = 0, y = 0;
    constexpr Point2() = default;
    explicit constexpr Point2(Float v) : x{ v }, y{ v } {}
    constexpr Point2(Float x, Float y) : x{ x }, y{ y } {}
};

struct Point2i
{
    int32 x = 0, y = 0;
    constexpr Point2i() = default;
    explicit constexpr Point2i(int32 v) : x{ v }, y{ v } {}
    constexpr Point2i(int32 x, int32 y) : x{ x }, y{ y } {}
};

struct Vec3
{
    Float x = 0, y = 0, z = 0;
    constexpr Vec3() = default;
    explicit constexpr Vec3(Float v) : x{ v }, y{ v }, z{ v } {}
    constexpr Vec3(Float x, Float y, Float z) : x{ x }, y{ y }, z{ z } {}

    void Set(Float nx, Float ny, Float nz)
    {
        x = nx;
        y = ny;
        z = nz;
    }
    Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Float Normalize(); // member form: early out below FLT_EPSILON
};
using Point3 = Vec3;

constexpr Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
constexpr Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
constexpr Vec3 operator-(const Vec3& a) { return Vec3(-a.x, -a.y, -a.z); }
constexpr Vec3 operator*(const Vec3& a, Float s) { return Vec3(a.x * s, a.y * s, a.z * s); }
constexpr Vec3 operator*(Float s, const Vec3& a) { return Vec3(a.x * s, a.y * s, a.z * s); }
constexpr Vec3 operator/(const Vec3& a, Float s) { return Vec3(a.x / s, a.y / s, a.z / s); }
constexpr Vec3 operator*(const Vec3& a, const Vec3& b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
constexpr bool operator==(const Vec3& a, const Vec3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

inline Float Dot(const Vec3& a, const Vec3& b)
{
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
inline Vec3 Cross(const Vec3& a, const Vec3& b)
{
    return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline Float Length2(const Vec3& v) { return Dot(v, v); }
inline Float Length(const Vec3& v) { return std::sqrt(Length2(v)); }
inline Vec3 Normalize(const Vec3& v)
{
    const Float inv_length = Float(1) / Length(v);
    return v * inv_length;
}
inline Float Dist(const Point3& a, const Point3& b) { return Length(b - a); }

inline Float Vec3::Normalize()
{
    const Float length = Length(*this);
    if (length < std::numeric_limits<Float>::epsilon()) return Float(0);
    const Float inv_length = Float(1) / length;
    x *= inv_length;
    y *= inv_length;
    z *= inv_length;
    return length;
}

constexpr Vec3 x_axis(1, 0, 0);
constexpr Vec3 y_axis(0, 1, 0);
constexpr Vec3 z_axis(0, 0, 1);

struct Vec4
{
    Float x = 0, y = 0, z = 0, w = 0;
    constexpr Vec4() = default;
    constexpr Vec4(Float x, Float y, Float z, Float w) : x{ x }, y{ y }, z{ z }, w{ w } {}
    constexpr Vec4(const Vec3& v, Float w) : x{ v.x }, y{ v.y }, z{ v.z }, w{ w } {}
    Float Normalize(); // member form: early out below FLT_EPSILON
};

inline Float Vec4::Normalize()
{
    const Float length = std::sqrt(((x * x + y * y) + z * z) + w * w);
    if (length < std::numeric_limits<Float>::epsilon()) return Float(0);
    const Float inv_length = Float(1) / length;
    x *= inv_length;
    y *= inv_length;
    z *= inv_length;
    w *= inv_length;
    return length;
}

struct Identity
{
};
constexpr Identity identity{};

// column-major 4x4: ex, ey, ez, ew are the columns
struct Mat4
{
    Vec4 ex, ey, ez, ew;
    constexpr Mat4() = default;
    constexpr Mat4(Identity) : ex{ 1, 0, 0, 0 }, ey{ 0, 1, 0, 0 }, ez{ 0, 0, 1, 0 }, ew{ 0, 0, 0, 1 } {}
    explicit constexpr Mat4(Float d) : ex{ d, 0, 0, 0 }, ey{ 0, d, 0, 0 }, ez{ 0, 0, d, 0 }, ew{ 0, 0, 0, d } {}
    constexpr Mat4(const Vec4& x, const Vec4& y, const Vec4& z, const Vec4& w) : ex{ x }, ey{ y }, ez{ z }, ew{ w } {}
    Mat4(const struct Transform& t); // src/util/math.cpp:32-38
};

// M * v, summed left to right per row
inline Vec4 Mul(const Mat4& m, const Vec4& v)
{
    return Vec4(m.ex.x * v.x + m.ey.x * v.y + m.ez.x * v.z + m.ew.x * v.w,
                m.ex.y * v.x + m.ey.y * v.y + m.ez.y * v.z + m.ew.y * v.w,
                m.ex.z * v.x + m.ey.z * v.y + m.ez.z * v.z + m.ew.z * v.w,
                m.ex.w * v.x + m.ey.w * v.y + m.ez.w * v.z + m.ew.w * v.w);
}

// A * B (include/bulbit/matrix.h:498-501): B's columns through A
inline Mat4 Mul(const Mat4& a, const Mat4& b) { return Mat4(Mul(a, b.ex), Mul(a, b.ey), Mul(a, b.ez), Mul(a, b.ew)); }

// 3x3 matrix; ex, ey, ez are the columns
struct Mat3
{
    Vec3 ex, ey, ez;
    constexpr Mat3() = default;
    constexpr Mat3(const Vec3& x, const Vec3& y, const Vec3& z) : ex{ x }, ey{ y }, ez{ z } {}
};

struct Quat
{
    Float x = 0, y = 0, z = 0, w = 1;
    constexpr Quat() = default;
    constexpr Quat(Identity) {}
    constexpr Quat(Float x, Float y, Float z, Float w) : x{ x }, y{ y }, z{ z }, w{ w } {}
    explicit constexpr Quat(Float w) : x{ 0 }, y{ 0 }, z{ 0 }, w{ w } {}
    // the rotation of an orthonormal basis (include/bulbit/quaternion.h:34-62)
    Quat(const Mat3& m);
    // the basis whose z is `front`, x = normalized up x front (quaternion.h:64-72)
    Quat(const Vec3& front, const Vec3& up);
    // `angle` radians about `unit_axis` (quaternion.h:75-83)
    Quat(Float angle, const Vec3& unit_axis);

    Vec3 Rotate(const Vec3& v) const; // quaternion.h:142-155
    Vec3 GetBasisZ() const;           // quaternion.h:199-205
};

constexpr Quat operator*(const Quat& q, Float s) { return Quat(q.x * s, q.y * s, q.z * s, q.w * s); }

// Hamilton product, terms in the order of include/bulbit/quaternion.h:273-281
constexpr Quat operator*(const Quat& a, const Quat& b)
{
    return Quat(a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z, a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x,
                a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}

inline Quat::Quat(const Mat3& m)
{
    // the largest of 4w^2, 4x^2, 4y^2, 4z^2 picked by the signs of the
    // diagonal, then the other three components from the off-diagonal sums
    Float t;
    if (m.ez.z < 0) {
        if (m.ex.x > m.ey.y) {
            t = 1 + m.ex.x - m.ey.y - m.ez.z;
            *this = Quat(t, m.ex.y + m.ey.x, m.ez.x + m.ex.z, m.ey.z - m.ez.y);
        } else {
            t = 1 - m.ex.x + m.ey.y - m.ez.z;
            *this = Quat(m.ex.y + m.ey.x, t, m.ey.z + m.ez.y, m.ez.x - m.ex.z);
        }
    } else {
        if (m.ex.x < -m.ey.y) {
            t = 1 - m.ex.x - m.ey.y + m.ez.z;
            *this = Quat(m.ez.x + m.ex.z, m.ey.z + m.ez.y, t, m.ex.y - m.ey.x);
        } else {
            t = 1 + m.ex.x + m.ey.y + m.ez.z;
            *this = Quat(m.ey.z - m.ez.y, m.ez.x - m.ex.z, m.ex.y - m.ey.x, t);
        }
    }
    *this = *this * (0.5f / std::sqrt(t));
}

inline Quat::Quat(const Vec3& front, const Vec3& up)
{
    Mat3 basis;
    basis.ez = front;
    basis.ex = Cross(up, basis.ez);
    basis.ex.Normalize();
    basis.ey = Cross(basis.ez, basis.ex);
    *this = Quat(basis);
}

inline Quat::Quat(Float angle, const Vec3& unit_axis)
{
    const Float half = angle * 0.5f;
    const Float s = std::sin(half);
    x = unit_axis.x * s;
    y = unit_axis.y * s;
    z = unit_axis.z * s;
    w = std::cos(half);
}

// q v q*, in the arrangement of include/bulbit/quaternion.h:142-155
inline Vec3 Quat::Rotate(const Vec3& v) const
{
    const Float vx = 2 * v.x, vy = 2 * v.y, vz = 2 * v.z;
    const Float w2 = w * w - 0.5f;
    const Float dot2 = x * vx + y * vy + z * vz;
    return Vec3(vx * w2 + (y * vz - z * vy) * w + x * dot2, vy * w2 + (z * vx - x * vz) * w + y * dot2,
                vz * w2 + (x * vy - y * vx) * w + z * dot2);
}

inline Vec3 Quat::GetBasisZ() const
{
    const Float z2 = z * 2, w2 = w * 2;
    return Vec3((y * w2) + x * z2, (-x * w2) + y * z2, (w * w2) - 1 + z * z2);
}

// rigid transform with scale: position, orientation, scale (include/bulbit/transform.h)
struct Transform
{
    Vec3 p;
    Quat q;
    Vec3 s{ 1, 1, 1 };


    constexpr Transform() = default;
    constexpr Transform(Identity) {}
    constexpr Transform(const Vec3& position) : p{ position } {}
    constexpr Transform(const Quat& orientation) : q{ orientation } {}
    constexpr Transform(const Vec3& position, const Quat& orientation) : p{ position }, q{ orientation } {}
    constexpr Transform(const Vec3& position, const Quat& orientation, const Vec3& scale)
        : p{ position }, q{ orientation }, s{ scale }
    {
    }
    constexpr Transform(Float x, Float y, Float z, const Quat& orientation = Quat(1), const Vec3& scale = Vec3(1))
        : p{ x, y, z }, q{ orientation }, s{ scale }
    {
    }
    // scale = column lengths, rotation from the normalized columns,
    // position = last column (include/bulbit/transform.h:59-69)
    Transform(const Mat4& m);

    // camera transform at `position` looking at `target` (transform.h:117-122)
    static Transform LookAt(const Vec3& position, const Vec3& target, const Vec3& up);
};

// A * B: B's position moved by A, rotations and scales multiplied
// (include/bulbit/transform.h:147-157)
inline Transform Mul(const Transform& a, const Transform& b)
{
    return Transform(a.q.Rotate(a.s * b.p) + a.p, a.q * b.q, a.s * b.s);
}
inline Transform operator*(const Transform& a, const Transform& b) { return Mul(a, b); }
inline Vec3 Mul(const Transform& t, const Vec3& v) { return t.q.Rotate(t.s * v) + t.p; }

// Mat4(Mat3(q), p) with the columns scaled by s (src/util/math.cpp:7-38)
inline Mat4::Mat4(const Transform& t)
{
    const Quat& q = t.q;
    const Float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const Float xz = q.x * q.z, xy = q.x * q.y, yz = q.y * q.z;
    const Float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    ex = Vec4(1 - 2 * (yy + zz), 2 * (xy + wz), 2 * (xz - wy), 0);
    ey = Vec4(2 * (xy - wz), 1 - 2 * (xx + zz), 2 * (yz + wx), 0);
    ez = Vec4(2 * (xz + wy), 2 * (yz - wx), 1 - 2 * (xx + yy), 0);
    ew = Vec4(t.p, 1);
    ex = Vec4(ex.x * t.s.x, ex.y * t.s.x, ex.z * t.s.x, ex.w * t.s.x);
    ey = Vec4(ey.x * t.s.y, ey.y * t.s.y, ey.z * t.s.y, ey.w * t.s.y);
    ez = Vec4(ez.x * t.s.z, ez.y * t.s.z, ez.z * t.s.z, ez.w * t.s.z);
}

inline Transform::Transform(const Mat4& m)
{
    s.x = std::sqrt(m.ex.x * m.ex.x + m.ex.y * m.ex.y + m.ex.z * m.ex.z);
    s.y = std::sqrt(m.ey.x * m.ey.x + m.ey.y * m.ey.y + m.ey.z * m.ey.z);
    s.z = std::sqrt(m.ez.x * m.ez.x + m.ez.y * m.ez.y + m.ez.z * m.ez.z);
    q = Quat(Mat3(Vec3(m.ex.x, m.ex.y, m.ex.z) / s.x, Vec3(m.ey.x, m.ey.y, m.ey.z) / s.y,
                  Vec3(m.ez.x, m.ez.y, m.ez.z) / s.z));
    p = Vec3(m.ew.x, m.ew.y, m.ew.z);
}

// ----------------------------------------------------------------- spectrum

struct Spectrum
{
    Float r = 0, g = 0, b = 0;
    constexpr Spectrum() = default;
    explicit constexpr Spectrum(Float v) : r{ v }, g{ v }, b{ v } {}
    constexpr Spectrum(Float r, Float g, Float b) : r{ r }, g{ g }, b{ b } {}
    Float operator[](int i) const { return i == 0 ? r : (i == 1 ? g : b); }
    Float& operator[](int i) { return i == 0 ? r : (i == 1 ? g : b); }
};

// -------------------------------------------------------------------- image

template <typename T>
class Image
{
public:
    using Type = T;
    Image() = default;
    Image(int32 width, int32 height) : width{ width }, height{ height }, data(size_t(width) * size_t(height)) {}

    T& operator[](int32 i) { return data[size_t(i)]; }
    const T& operator[](int32 i) const { return data[size_t(i)]; }
    T& operator()(int32 x, int32 y) { return data[size_t(y) * size_t(width) + size_t(x)]; }
    const T& operator()(int32 x, int32 y) const { return data[size_t(y) * size_t(width) + size_t(x)]; }
    explicit operator bool() const { return !data.empty(); }

    int32 width = 0, height = 0;
    std::vector<T> data;
};

// ----------------------------------------------------------------- textures

// TexCoordFilter (include/bulbit/texture.h:20-24)
enum TexCoordFilter
{
    repeat,
    clamp,
};

template <typename T>
class Texture
{
public:
    virtual ~Texture() = default;
    // the value of a constant texture; false for any other kind
    virtual bool Constant(T* value) const = 0;
};

// ImageTexture (include/bulbit/textures.h:35-107): bilinear lookups of the
// image, texel (x, y) at x + y * width; the scene's texture pool creates it
// with TexCoordFilter::repeat (textures.h:238-251)
template <typename T>
class ImageTexture : public Texture<T>
{
public:
    ImageTexture(Image<T> image, TexCoordFilter texcoord_filter = TexCoordFilter::repeat)
        : image{ std::move(image) }, texcoord_filter{ texcoord_filter }
    {
    }
    int32 GetWidth() const { return image.width; }
    int32 GetHeight() const { return image.height; }
    bool Constant(T*) const override { return false; }

    Image<T> image;
    TexCoordFilter texcoord_filter;
};

// CheckerTexture (include/bulbit/textures.h:109-142): a where the integer
// parts of uv * resolution sum to an odd number, else b
template <typename T>
class CheckerTexture : public Texture<T>
{
public:
    CheckerTexture(const Texture<T>* a, const Texture<T>* b, const Point2& resolution)
        : a{ a }, b{ b }, resolution{ resolution }
    {
    }
    bool Constant(T*) const override { return false; }

    const Texture<T>* a;
    const Texture<T>* b;
    Point2 resolution;
};

template <typename T>
class ConstantTexture : public Texture<T>
{
public:
    ConstantTexture(T value) : value{ std::move(value) } {}
    T Evaluate(const Point2&) const { return value; }
    bool Constant(T* out) const override
    {
        *out = value;
        return true;
    }

    T value;
};

using FloatTexture = Texture<Float>;
using SpectrumTexture = Texture<Spectrum>;

// ---------------------------------------------------------------- materials

class Material
{
public:
    virtual ~Material() = default;
};

// DiffuseMaterial (include/bulbit/materials.h:8-30): Lambertian when roughness is absent or 0
class DiffuseMaterial : public Material
{
public:
    DiffuseMaterial(const SpectrumTexture* reflectance, const FloatTexture* roughness = nullptr,
                    const SpectrumTexture* normal = nullptr, const FloatTexture* alpha = nullptr)
        : reflectance{ reflectance }, roughness{ roughness }, normal{ normal }, alpha{ alpha }
    {
    }

    const SpectrumTexture* reflectance;
    const FloatTexture* roughness;
    const SpectrumTexture* normal;
    const FloatTexture* alpha;
};

// MirrorMaterial (include/bulbit/materials.h:33-51): specular reflection scaled by reflectance
class MirrorMaterial : public Material
{
public:
    MirrorMaterial(const SpectrumTexture* reflectance, const SpectrumTexture* normal = nullptr,
                   const FloatTexture* alpha = nullptr)
        : reflectance{ reflectance }, normal{ normal }, alpha{ alpha }
    {
    }

    const SpectrumTexture* reflectance;
    const SpectrumTexture* normal;
    const FloatTexture* alpha;
};

// DielectricMaterial (include/bulbit/materials.h:52-84)
class DielectricMaterial : public Material
{
public:
    DielectricMaterial(Float eta, const FloatTexture* u_roughness, const FloatTexture* v_roughness,
                       const SpectrumTexture* reflectance, bool energy_compensation = true,
                       const SpectrumTexture* normal = nullptr)
        : eta{ eta }, u_roughness{ u_roughness }, v_roughness{ v_roughness }, reflectance{ reflectance },
          energy_compensation{ energy_compensation }, normal{ normal }
    {
    }

    Float eta;
    const FloatTexture* u_roughness;
    const FloatTexture* v_roughness;
    const SpectrumTexture* reflectance;
    bool energy_compensation;
    const SpectrumTexture* normal;
};

// ConductorMaterial (include/bulbit/materials.h:98-142): (eta, k) or physical reflectance R
class ConductorMaterial : public Material
{
public:
    ConductorMaterial(const SpectrumTexture* eta, const SpectrumTexture* k, const FloatTexture* u_roughness,
                      const FloatTexture* v_roughness, const SpectrumTexture* reflectance = nullptr,
                      bool energy_compensation = true, const SpectrumTexture* normal = nullptr,
                      const FloatTexture* alpha = nullptr)
        : eta{ eta }, k{ k }, u_roughness{ u_roughness }, v_roughness{ v_roughness }, reflectance{ reflectance },
          energy_compensation{ energy_compensation }, normal{ normal }, alpha{ alpha }
    {
    }
    ConductorMaterial(const SpectrumTexture* R, const FloatTexture* u_roughness, const FloatTexture* v_roughness,
                      const SpectrumTexture* reflectance = nullptr, bool energy_compensation = true,
                      const SpectrumTexture* normal = nullptr, const FloatTexture* alpha = nullptr)
        : eta{ nullptr }, k{ R }, u_roughness{ u_roughness }, v_roughness{ v_roughness }, reflectance{ reflectance },
          energy_compensation{ energy_compensation }, normal{ normal }, alpha{ alpha }
    {
    }

    const SpectrumTexture* eta; // null: k holds the physical reflectance R
    const SpectrumTexture* k;
    const FloatTexture* u_roughness;
    const FloatTexture* v_roughness;
    const SpectrumTexture* reflectance;
    bool energy_compensation;
    const SpectrumTexture* normal;
    const FloatTexture* alpha;
};

// MetallicRoughnessMaterial (include/bulbit/materials.h:142-170)
class MetallicRoughnessMaterial : public Material
{
public:
    MetallicRoughnessMaterial(const SpectrumTexture* basecolor, const FloatTexture* metallic,
                              const FloatTexture* u_roughness, const FloatTexture* v_roughness,
                              const SpectrumTexture* normal = nullptr, const FloatTexture* alpha = nullptr)
        : basecolor{ basecolor }, metallic{ metallic }, u_roughness{ u_roughness }, v_roughness{ v_roughness },
          normal{ normal }, alpha{ alpha }
    {
    }

    const SpectrumTexture* basecolor;
    const FloatTexture* metallic;
    const FloatTexture* u_roughness;
    const FloatTexture* v_roughness;
    const SpectrumTexture* normal;
    const FloatTexture* alpha;
};

// -------------------------------------------------------------------- media

class Medium
{
public:
    virtual ~Medium() = default;
};

struct MediumInterface
{
    const Medium* inside = nullptr;
    const Medium* outside = nullptr;
};

// ------------------------------------------------------------------- shapes

struct MeshVertex
{
    Point3 position;
    Vec3 normal;
    Vec3 tangent;
    Point2 texCoord;
};

// Mesh (include/bulbit/mesh.h:19-44, src/geometry/mesh.cpp:8-68): vertices
// are moved to world space by `transform` at construction.
class Mesh
{
public:
    Mesh(std::vector<Point3> positions, std::vector<Vec3> normals, std::vector<Vec3> tangents,
         std::vector<Point2> texCoords, std::vector<int32> indices, const Mat4& transform);
    Mesh(const std::vector<MeshVertex>& vertices, std::vector<int32> indices, const Mat4& transform);

    int32 GetTriangleCount() const { return triangle_count; }

    const std::vector<Point3>& Positions() const { return positions; }
    const std::vector<Vec3>& Normals() const { return normals; }
    const std::vector<Vec3>& Tangents() const { return tangents; }
    const std::vector<Point2>& TexCoords() const { return texCoords; }
    const std::vector<int32>& Indices() const { return indices; }

private:
    int32 triangle_count = 0;
    std::vector<Point3> positions;
    std::vector<Vec3> normals;
    std::vector<Vec3> tangents;
    std::vector<Point2> texCoords;
    std::vector<int32> indices;
};

class Shape
{
public:
    virtual ~Shape() = default;
};

class Sphere : public Shape
{
public:
    Sphere(const Point3& center, Float radius) : transform{ center, identity }, radius{ radius } {}
    Sphere(const Transform& transform, Float radius) : transform{ transform }, radius{ radius } {}

    Transform transform;
    Float radius;
};

class Triangle : public Shape
{
public:
    Triangle(const Mesh* mesh, size_t tri_index) : mesh{ mesh }, index{ int32(tri_index) } {}

    const Mesh* GetMesh() const { return mesh; }
    int32 GetIndex() const { return index; }

private:
    const Mesh* mesh;
    int32 index;
};

class Primitive
{
public:
    Primitive(const Shape* shape, const Material* material, const MediumInterface& medium_interface)
        : shape{ shape }, material{ material }, medium_interface{ medium_interface }
    {
    }

    const Shape* GetShape() const { return shape; }
    const Material* GetMaterial() const { return material; }
    MediumInterface GetMediumInterface() const { return medium_interface; }

private:
    const Shape* shape;
    const Material* material;
    MediumInterface medium_interface;
};

// ------------------------------------------------------------------- lights

class Light
{
public:
    virtual ~Light() = default;
};

// SpotLight (include/bulbit/lights.h:41-63); angles in degrees
class SpotLight : public Light
{
public:
    SpotLight(const Point3& position, const Vec3& direction, const Spectrum& intensity, Float angle_max,
              Float angle_falloff_start, const Medium* medium)
        : position{ position }, direction{ direction }, intensity{ intensity }, angle_max{ angle_max },
          angle_falloff_start{ angle_falloff_start }, medium{ medium }
    {
    }

    Point3 position;
    Vec3 direction;
    Spectrum intensity;
    Float angle_max, angle_falloff_start;
    const Medium* medium;
};

// PointLight (include/bulbit/lights.h:15-40)
class PointLight : public Light
{
public:
    PointLight(const Point3& position, const Spectrum& intensity, const Medium* medium)
        : position{ position }, intensity{ intensity }, medium{ medium }
    {
    }

    Point3 position;
    Spectrum intensity;
    const Medium* medium;
};

// DirectionalLight (include/bulbit/lights.h:65-95): light arriving from `direction`
class DirectionalLight : public Light
{
public:
    DirectionalLight(const Vec3& direction, const Spectrum& intensity) : direction{ direction }, intensity{ intensity } {}

    Vec3 direction;
    Spectrum intensity;
};

// UniformInfiniteLight (include/bulbit/lights.h:236-260)
class UniformInfiniteLight : public Light
{
public:
    UniformInfiniteLight(const Spectrum& l, Float scale = 1) : l{ l }, scale{ scale } {}

    Spectrum l;
    Float scale;
};

// DiffuseAreaLight (include/bulbit/lights.h:108-134) over a triangle primitive
class DiffuseAreaLight : public Light
{
public:
    DiffuseAreaLight(const Primitive* primitive, const SpectrumTexture* emission, bool two_sided)
        : primitive{ primitive }, emission{ emission }, two_sided{ two_sided }
    {
    }

    const Primitive* primitive;
    const SpectrumTexture* emission;
    bool two_sided;
};

// ---------------------------------------------------------------- allocator

// Allocator::new_object / delete_object (include/bulbit/allocator.h)
class Allocator
{
public:
    template <typename T, typename... Args>
    T* new_object(Args&&... args)
    {
        return new T(std::forward<Args>(args)...);
    }
    template <typename T>
    void delete_object(T* p)
    {
        delete p;
    }
};

// --------------------------------------------------------------------- scene

// Scene (include/bulbit/scene.h:14-163): owns every object it creates; the
// pointers stay valid for its lifetime.
class Scene
{
public:
    Scene() = default;
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    template <typename... Args>
    Mesh* CreateMesh(Args&&... args)
    {
        meshes.emplace_back(new Mesh(std::forward<Args>(args)...));
        return meshes.back().get();
    }

    template <typename ShapeType, typename... Args>
    ShapeType* CreateShape(Args&&... args)
    {
        ShapeType* s = new ShapeType(std::forward<Args>(args)...);
        shapes.emplace_back(s);
        return s;
    }

    template <typename... Args>
    Primitive* CreatePrimitive(Args&&... args)
    {
        owned_primitives.emplace_back(new Primitive(std::forward<Args>(args)...));
        primitives.push_back(owned_primitives.back().get());
        return primitives.back();
    }

    template <typename LightType, typename... Args>
    LightType* CreateLight(Args&&... args)
    {
        LightType* l = new LightType(std::forward<Args>(args)...);
        owned_lights.emplace_back(l);
        lights.push_back(l);
        return l;
    }

    template <typename MediumType, typename... Args>
    MediumType* CreateMedium(Args&&... args)
    {
        MediumType* m = new MediumType(std::forward<Args>(args)...);
        media.emplace_back(m);
        return m;
    }

    template <template <typename> class TextureType, typename T, typename... Args>
    TextureType<T>* CreateTexture(Args&&... args)
    {
        TextureType<T>* t = new TextureType<T>(std::forward<Args>(args)...);
        if constexpr (std::is_same_v<T, Float>)
            float_textures.emplace_back(t);
        else
            spectrum_textures.emplace_back(t);
        return t;
    }

    template <typename MaterialType, typename... Args>
    MaterialType* CreateMaterial(Args&&... args)
    {
        MaterialType* m = new MaterialType(std::forward<Args>(args)...);
        materials.emplace_back(m);
        return m;
    }

    const std::vector<Primitive*>& GetPrimitives() const { return primitives; }
    const std::vector<Light*>& GetLights() const { return lights; }

private:
    std::vector<std::unique_ptr<Mesh>> meshes;
    std::vector<std::unique_ptr<Shape>> shapes;
    std::vector<std::unique_ptr<Primitive>> owned_primitives;
    std::vector<Primitive*> primitives;
    std::vector<std::unique_ptr<Light>> owned_lights;
    std::vector<Light*> lights;
    std::vector<std::unique_ptr<Medium>> media;
    std::vector<std::unique_ptr<FloatTexture>> float_textures;
    std::vector<std::unique_ptr<SpectrumTexture>> spectrum_textures;
    std::vector<std::unique_ptr<Material>> materials;
};

// ---------------------------------------------------------- model loading

// AreaLightInfo (cli/primitive_builder.h:23-34), diffuse area lights only
struct AreaLightInfo
{
    bool two_sided = false;
    std::variant<Float, Spectrum, const SpectrumTexture*> emission = 0.0f;
};

// ModelLoaderOptions (cli/model_loader.h:8-19)
struct ModelLoaderOptions
{
    bool flip_normal = false;
    bool gen_smooth_normal = false;
    bool flip_tex_coord = false;
    bool use_fallback_material = false;

    const Material* fallback_material = nullptr;
    MediumInterface fallback_medium_interface = {};

    std::optional<AreaLightInfo> area_light = {};
};

// LoadOBJ (cli/model_loader.cpp:803-1043): meshes per shape and material,
// MetallicRoughness materials from the MTL file (CreateOBJMaterial,
// :742-790), every triangle of a material group an area light with the
// material's emission.  LoadGLTF (:405-740): the meshes of a .gltf or .glb
// model under its node matrices, drawn with the fallback material (the
// model's own materials would be PrincipledMaterials, which the GPU path
// does not have: such a model is reported and not loaded).  LoadModel
// dispatches on the extension.
void LoadOBJ(Scene& scene, std::filesystem::path filename, const Transform& transform,
             const ModelLoaderOptions& options = {});
void LoadGLTF(Scene& scene, std::filesystem::path filename, const Transform& transform,
              const ModelLoaderOptions& options = {});
void LoadModel(Scene& scene, std::filesystem::path filename, const Transform& transform,
               const ModelLoaderOptions& options = {});

// --------------------------------------------------------------- threading

// The GPU render does not use host worker threads; the pool is kept so that
// programs written for the reference compile and run unchanged.
class ThreadPool
{
public:
    explicit ThreadPool(uint32 num_threads) : num_threads{ num_threads } {}
    uint32 NumThreads() const { return num_threads; }

    static inline std::unique_ptr<ThreadPool> global_thread_pool;

private:
    uint32 num_threads;
};

// ---------------------------------------------------------- acceleration

struct Ray
{
    Point3 o;
    Vec3 d;
    Ray() = default;
    Ray(const Point3& o, const Vec3& d) : o{ o }, d{ d } {}
    Point3 At(Float t) const { return o + d * t; }
};

struct AABB
{
    Point3 min, max;
};

// Closest-hit record: the hit primitive, its ray parameter, the point o + t d
// and the shape's (u, v) (triangle barycentrics; zero for spheres).
struct Intersection
{
    const Primitive* primitive = nullptr;
    Float t = 0;
    Point3 point;
    Point2 uv;
};

class Intersectable
{
public:
    virtual ~Intersectable() = default;
    virtual AABB GetAABB() const = 0;
    virtual bool Intersect(Intersection* out_isect, const Ray& ray, Float t_min, Float t_max) const = 0;
    virtual bool IntersectAny(const Ray& ray, Float t_min, Float t_max) const = 0;
};

struct SceneExport; // flat scene description handed to the C ABI

// BVH (include/bulbit/bvh.h:11-20): the SAH BVH over the primitives, built on
// the host exactly as the reference builds it and kept in device memory.
// Intersect / IntersectAny trace on the GPU (one ray per call; batches use
// IntersectBatch).  The BVH does not own the primitives.
class BVH : public Intersectable
{
public:
    explicit BVH(const std::vector<Primitive*>& primitives);
    ~BVH() override;
    BVH(const BVH&) = delete;
    BVH& operator=(const BVH&) = delete;

    AABB GetAABB() const override;
    bool Intersect(Intersection* out_isect, const Ray& ray, Float t_min, Float t_max) const override;
    bool IntersectAny(const Ray& ray, Float t_min, Float t_max) const override;

    // n rays in one GPU batch; hit[i] = false on a miss
    void IntersectBatch(const Ray* rays, const Float* t_min, const Float* t_max, int32 n, bool any_hit,
                        Intersection* out_isects, bool* hit) const;

    const std::vector<Primitive*>& GetPrimitives() const { return primitives; }
    int32 NodeCount() const { return n_nodes; }

private:
    BbScene* DeviceScene() const;

    std::vector<Primitive*> primitives;
    std::unique_ptr<SceneExport> desc;
    mutable BbScene* device = nullptr;
    int32 n_nodes = 0;
    AABB bounds;
};

// ------------------------------------------------------------ renderer info

// Render settings read from a scene file (include/bulbit/renderer_info.h:8-120)
enum class FilterType
{
    box,
    tent,
    gaussian,
};

struct FilterInfo
{
    FilterType type = FilterType::gaussian;
    Float extent = 3;
    Float gaussian_stddev = 0.5f;
};

struct FilmInfo
{
    std::string filename = "";
    Point2i resolution = { 1280, 720 };
    FilterInfo filter_info;
};

enum class SamplerType
{
    independent,
    stratified,
};

struct SamplerInfo
{
    SamplerType type = SamplerType::independent;
    int32 spp = 64;
};

enum class CameraType
{
    perspective,
    orthographic,
    spherical,
};

struct CameraInfo
{
    CameraType type = CameraType::perspective;
    Transform transform = identity;
    Float fov = 35.0f;
    Float aperture_radius = 0.0f;
    Float focus_distance = 1.0f;
    Point2 viewport_size = { 1, 1 };
    FilmInfo film_info;
    SamplerInfo sampler_info;
    const Medium* medium = nullptr;
};

// the reference's integrators; only `path` runs here
enum class IntegratorType
{
    path,
    vol_path,
    light_path,
    light_vol_path,
    bdpt,
    vol_bdpt,
    pm,
    vol_pm,
    sppm,
    vol_sppm,
    vcm,
    restir_di,
    restir_pt,
    naive_path,
    naive_vol_path,
    random_walk,
    ao,
    albedo,
    debug,
    count,
};

struct IntegratorInfo
{
    IntegratorType type = IntegratorType::path;
    int32 max_bounces = 16;
    int32 rr_min_bounces = 1;
    bool regularize_bsdf = false;
};

// ----------------------------------------------------------------- sampling

class Sampler
{
public:
    explicit Sampler(int32 samples_per_pixel) : samples_per_pixel{ samples_per_pixel } {}
    virtual ~Sampler() = default;
    // IndependentSampler for SamplerType::independent, else null (the
    // stratified sampler is not on the GPU path) (src/sampler/sampler.cpp:7-23)
    static Sampler* Create(Allocator& alloc, const SamplerInfo& sampler_info);
    int32 samples_per_pixel;
};

// IndependentSampler (include/bulbit/samplers.h:10-25)
class IndependentSampler : public Sampler
{
public:
    explicit IndependentSampler(int32 samples_per_pixel, int32 seed = 0) : Sampler(samples_per_pixel), seed{ seed } {}
    int32 seed;
};

// ------------------------------------------------------------------ filters

class Filter
{
public:
    explicit Filter(Float extent) : extent{ extent } {}
    virtual ~Filter() = default;
    // src/reconstruction/filter.cpp:7-21
    static Filter* Create(Allocator& alloc, const FilterInfo& filter_info);
    Float extent;
};

class BoxFilter : public Filter
{
public:
    explicit BoxFilter(Float extent = 1) : Filter(extent) {}
};

class TentFilter : public Filter
{
public:
    explicit TentFilter(Float extent = 2) : Filter(extent) {}
};

// sampled from a 32 x 32 table of the Gaussian (gaussian_filter.cpp:8-43),
// tabulated by the GPU library exactly as the reference does
class GaussianFilter : public Filter
{
public:
    explicit GaussianFilter(Float sigma = 0.5f, Float extent = 3) : Filter(extent), sigma{ sigma } {}
    const Float sigma;
};

// ------------------------------------------------------------------ cameras

class Camera
{
public:
    Camera(const Point2i& resolution, const Medium* medium, const Filter* filter)
        : resolution{ resolution }, medium{ medium }, filter{ filter }
    {
    }
    virtual ~Camera() = default;
    // a PerspectiveCamera for CameraType::perspective, else null (src/camera/camera.cpp:7-23)
    static Camera* Create(Allocator& alloc, const CameraInfo& camera_info, const Filter* filter);

    // the reference's default camera filter (include/bulbit/camera.h:48)
    static inline std::unique_ptr<Filter> default_filter = std::make_unique<GaussianFilter>();

    const Point2i& GetScreenResolution() const { return resolution; }
    const Filter* GetFilter() const { return filter; }
    const Medium* GetMedium() const { return medium; }

protected:
    Point2i resolution;
    const Medium* medium;
    const Filter* filter;
};

// PerspectiveCamera (include/bulbit/cameras.h:36-48)
class PerspectiveCamera : public Camera
{
public:
    PerspectiveCamera(const Transform& tf, Float vfov, Float aperture_radius, Float focus_distance,
                      const Point2i& resolution, const Medium* medium = nullptr,
                      const Filter* filter = Camera::default_filter.get())
        : Camera(resolution, medium, filter), transform{ tf }, vfov{ vfov }, aperture_radius{ aperture_radius },
          focus_distance{ focus_distance }
    {
    }

    Transform transform;
    Float vfov, aperture_radius, focus_distance;
};

using Image3 = Image<Spectrum>;
using Image3f = Image3;

// sRGB_from_RGB (include/bulbit/color.h:31-50): the sRGB transfer curve
Vec3 sRGB_from_RGB(const Vec3& rgb);

// Tone mapping applied when writing an 8-bit image (include/bulbit/image.h:89-99)
using ToneMappingCallback = Vec3(const Vec3&);
inline ToneMappingCallback* default_tonemapping_callback = [](const Vec3& rgb) -> Vec3 { return sRGB_from_RGB(rgb); };

// WriteImage (src/util/image.cpp:219-266): .png and .jpg tone-mapped to 8
// bits (callback, then clamp to [0, 1] and scale by 256), linear .hdr
// (Radiance RGBE) and .pfm; false for any other extension or an I/O error.
// The film's row 0 is the bottom of the picture.
bool WriteImage(const Image3& image, const std::string& filename,
                ToneMappingCallback* callback = default_tonemapping_callback);

using Image1 = Image<Float>;

// ReadImage1 / ReadImage3 (include/bulbit/image.h:76-87): channel `channel`,
// or RGB, of an image file as linear floats, bottom row first; 8-bit colour
// data decoded with gamma 2.2 unless non_color.  Reads binary PNM (.ppm,
// .pgm, .pnm), PNG, JPEG (.jpg, .jpeg; baseline and progressive), Radiance
// .hdr and .pfm; an empty image on failure.
Image1 ReadImage1(const std::string& filename, int32 channel, bool non_color = false,
                  std::function<Float(Float)> transform = {});
Image3 ReadImage3(const std::string& filename, bool non_color = false,
                  std::function<Spectrum(Spectrum)> transform = {});

// --------------------------------------------------------- film, rendering

// Film (include/bulbit/film.h:10-33): per-pixel sum of samples and sample
// count; GetRenderedImage divides (src/reconstruction/film.cpp:80-100).
class Film
{
public:
    explicit Film(const Camera* camera);

    Image3 GetRenderedImage() const;

    const Camera* camera;
    std::vector<Float> sums;    // 3 per pixel, row-major
    std::vector<int32> counts;  // 1 per pixel
};

class Rendering
{
public:
    explicit Rendering(const Camera* camera) : camera{ camera }, film(camera) {}
    virtual ~Rendering();

    bool IsDone() const { return done.load(std::memory_order_acquire); }
    void LogProgress() const;
    void Wait() const;
    void WaitAndLogProgress() const;
    const Film& GetFilm() const { return film; }

    const Camera* camera;

private:
    friend class PathIntegrator;

    Film film;
    std::atomic<bool> done{ false };
    std::atomic<int64> samples_done{ 0 };
    int64 samples_total = 0;
    mutable std::thread worker;
};

// --------------------------------------------------------------- integrators

class Integrator
{
public:
    virtual ~Integrator() = default;
    virtual Rendering* Render(Allocator& alloc, const Camera* camera) = 0;
    // a PathIntegrator for IntegratorType::path, else null (src/integrator/integrator.cpp:15-30)
    static Integrator* Create(Allocator& alloc, const IntegratorInfo& integrator_info, const Intersectable* accel,
                              const std::vector<Light*>& lights, const Sampler* sampler);
};

// PathIntegrator (include/bulbit/integrators.h:92-114): unidirectional path
// tracing with NEE and continuation-ray MIS on the GPU.  `accel` must be a
// BVH over the scene's primitives.  Render returns at once; the returned
// Rendering completes asynchronously.
class PathIntegrator : public Integrator
{
public:
    PathIntegrator(const Intersectable* accel, std::vector<Light*> lights, const Sampler* sampler, int32 max_bounces,
                   int32 rr_min_bounces = 1, bool regularize_bsdf = false);
    ~PathIntegrator() override;
    PathIntegrator(const PathIntegrator&) = delete;
    PathIntegrator& operator=(const PathIntegrator&) = delete;

    Rendering* Render(Allocator& alloc, const Camera* camera) override;

private:
    const BVH* accel;
    std::vector<Light*> lights;
    const Sampler* sampler;
    int32 max_bounces, rr_min_bounces;
    bool regularize_bsdf;
    BbScene* device = nullptr;
};

// ------------------------------------------------------------ scene files

// RendererInfo (include/bulbit/renderer_info.h:114-120)
struct RendererInfo
{
    Scene scene;
    CameraInfo camera_info;
    IntegratorInfo integrator_info;
};

// LoadScene (cli/scene_loader.cpp:2147-2164): a Mitsuba 3 XML scene into
// `render_info`, with the reference loader's element rules.  Relative file
// names resolve against the scene file's folder.  Elements the reference
// reads but the GPU path does not render (media, envmap emitters, principled, plastic, layered, blend, thin
// dielectric and subsurface BSDFs, other integrators) make it print
// `bulbit_amd: LoadScene: ...` and return false; elements the reference
// itself skips with a message are skipped with a message.
bool LoadScene(RendererInfo* render_info, std::filesystem::path filename);

} // namespace bulbit
