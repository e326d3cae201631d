// bb_math.h — float32 vector, spectrum, frame and sampling arithmetic shared by
// the host set-up code and the gfx950 kernels.
//
// Every expression keeps Bulbit's evaluation order so that results match the
// reference bit for bit where the reference is exact (traversal, hit ids) and
// to a few ulps elsewhere.  The whole library is compiled with
// -ffp-contract=off: no multiply-add is fused unless written as fmaf().
//
// Reference definitions restated here (Bulbit source tree):
//   include/bulbit/floats.h:16-24          constants
//   include/bulbit/vectors.h:121-133       member Normalize (epsilon early-out)
//   include/bulbit/vectors.h:363-373       Dot, Cross
//   include/bulbit/math.h:47-83,109-113    Clamp, Normalize, SmoothStep, Reflect
//   include/bulbit/frame.h:9-203           shading-frame helpers, Frame
//   include/bulbit/spectrum.h:49-83        RGBSpectrum helpers
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstring>
#include <cmath>
#include <cstdint>

#define BB_HD __host__ __device__ __forceinline__
#define BB_D __device__ __forceinline__

namespace bb
{

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kFourPi = 4.0f * kPi;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kInvTwoPi = 1.0f / (2.0f * kPi);
constexpr float kInvFourPi = 1.0f / (4.0f * kPi);
constexpr float kEps = FLT_EPSILON;
constexpr float kRayEps = 1e-4f; // Ray::epsilon, include/bulbit/ray.h:19

struct f2
{
    float x, y;
};
struct f3
{
    float x, y, z;
};

BB_HD f2 mk2(float x, float y) { return f2{ x, y }; }
BB_HD f3 mk3(float x, float y, float z) { return f3{ x, y, z }; }
BB_HD f3 mk3(float s) { return f3{ s, s, s }; }

// std::min / std::max semantics
BB_HD float smin(float a, float b) { return (b < a) ? b : a; }
BB_HD float smax(float a, float b) { return (a < b) ? b : a; }
BB_HD float clampf_(float v, float l, float r) { return v < l ? l : (v > r ? r : v); }
BB_HD float sqr(float v) { return v * v; }
BB_HD float safe_sqrt(float x) { return sqrtf(smax(0.0f, x)); }

BB_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
BB_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
BB_HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
BB_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
BB_HD f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
BB_HD f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
BB_HD f3 operator/(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
BB_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BB_HD float absdot(f3 a, f3 b) { return fabsf(dot(a, b)); }
BB_HD f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
BB_HD float length2(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
BB_HD float length(f3 a) { return sqrtf(length2(a)); }
BB_HD f3 vmin(f3 a, f3 b) { return mk3(smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)); }
BB_HD f3 vmax(f3 a, f3 b) { return mk3(smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)); }
BB_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Normalize(v) = v * (1 / Length(v))                     math.h:56-60
BB_HD f3 normalize(f3 v)
{
    float inv_length = 1.0f / length(v);
    return v * inv_length;
}

// Vector3::Normalize: returns the length, v unchanged below epsilon   vectors.h:121-133
BB_HD float normalize_m(f3& v)
{
    float l = length(v);
    if (l < kEps) return 0.0f;
    float inv = 1.0f / l;
    v.x *= inv;
    v.y *= inv;
    v.z *= inv;
    return l;
}

// Dot(v, z_axis) evaluated as written
BB_HD float dot_z(f3 v) { return v.x * 0.0f + v.y * 0.0f + v.z * 1.0f; }

// Reflect(v, n) = -v + 2 * Dot(v, n) * n                 math.h:109-113
BB_HD f3 reflect(f3 v, f3 n) { return (-v) + n * (2.0f * dot(v, n)); }

// SmoothStep                                              math.h:73-83
BB_HD float smooth_step(float a, float b, float x)
{
    if (a == b) return (x < a) ? 0.0f : 1.0f;
    float t = clampf_((x - a) / (b - a), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// ---------------------------------------------------------------- spectrum
struct spec
{
    float r, g, b;
};
BB_HD spec mks(float r, float g, float b) { return spec{ r, g, b }; }
BB_HD spec mks(float s) { return spec{ s, s, s }; }
BB_HD spec operator+(spec a, spec b) { return mks(a.r + b.r, a.g + b.g, a.b + b.b); }
BB_HD spec operator-(spec a, spec b) { return mks(a.r - b.r, a.g - b.g, a.b - b.b); }
BB_HD spec operator*(spec a, spec b) { return mks(a.r * b.r, a.g * b.g, a.b * b.b); }
BB_HD spec operator*(spec a, float s) { return mks(a.r * s, a.g * s, a.b * s); }
BB_HD spec operator/(spec a, float s) { return mks(a.r / s, a.g / s, a.b / s); }
BB_HD spec operator/(spec a, spec b) { return mks(a.r / b.r, a.g / b.g, a.b / b.b); }
BB_HD bool is_black(spec a) { return a.r == 0.0f && a.g == 0.0f && a.b == 0.0f; }
BB_HD float luminance(spec a) { return a.r * 0.2126f + a.g * 0.7152f + a.b * 0.0722f; }
BB_HD float max_component(spec a)
{
    float m = a.r;
    if (m < a.g) m = a.g;
    if (m < a.b) m = a.b;
    return m;
}
BB_HD bool is_nullish(spec a)
{
    return isnan(a.r) || isinf(a.r) || isnan(a.g) || isinf(a.g) || isnan(a.b) || isinf(a.b);
}
BB_HD spec sqrt_s(spec a) { return mks(sqrtf(a.r), sqrtf(a.g), sqrtf(a.b)); }

// ------------------------------------------------------- shading coordinates
BB_HD float cos2_theta(f3 w) { return sqr(w.z); }
BB_HD float abs_cos_theta(f3 w) { return fabsf(w.z); }
BB_HD float sin2_theta(f3 w) { return smax(0.0f, 1.0f - cos2_theta(w)); }
BB_HD float sin_theta(f3 w) { return sqrtf(sin2_theta(w)); }
BB_HD float tan2_theta(f3 w) { return sin2_theta(w) / cos2_theta(w); }
BB_HD float cos_phi(f3 w)
{
    float st = sin_theta(w);
    return (st == 0.0f) ? 1.0f : clampf_(w.x / st, -1.0f, 1.0f);
}
BB_HD float sin_phi(f3 w)
{
    float st = sin_theta(w);
    return (st == 0.0f) ? 0.0f : clampf_(w.y / st, -1.0f, 1.0f);
}
BB_HD bool same_hemisphere(f3 a, f3 b) { return a.z * b.z > 0.0f; }

// CoordinateSystem                                        frame.h:107-114
BB_HD void coordinate_system(f3 v1, f3& v2, f3& v3)
{
    float sign = copysignf(1.0f, v1.z);
    float a = -1.0f / (sign + v1.z);
    float b = v1.x * v1.y * a;
    v2 = mk3(1.0f + sign * (v1.x * v1.x) * a, sign * b, -sign * v1.x);
    v3 = mk3(b, sign + (v1.y * v1.y) * a, -v1.y);
}

// Frame::FromXZ(x, z) = (x, Cross(z, x), z)               frame.h:139-142
struct frame3
{
    f3 x, y, z;
};
BB_HD frame3 frame_from_xz(f3 x, f3 z) { return frame3{ x, cross(z, x), z }; }
BB_HD f3 to_local(const frame3& f, f3 v) { return mk3(dot(v, f.x), dot(v, f.y), dot(v, f.z)); }
BB_HD f3 from_local(const frame3& f, f3 v) { return f.x * v.x + f.y * v.y + f.z * v.z; }

// --------------------------------------------------------------- quaternions
struct quat
{
    float x, y, z, w;
};

// Quat::Rotate / RotateInv                                quaternion.h:142-170
BB_HD f3 quat_rotate(quat q, f3 v)
{
    float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
    float w2 = q.w * q.w - 0.5f;
    float dot2 = (q.x * vx + q.y * vy + q.z * vz);
    return mk3((vx * w2 + (q.y * vz - q.z * vy) * q.w + q.x * dot2), (vy * w2 + (q.z * vx - q.x * vz) * q.w + q.y * dot2),
               (vz * w2 + (q.x * vy - q.y * vx) * q.w + q.z * dot2));
}
BB_HD f3 quat_rotate_inv(quat q, f3 v)
{
    float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
    float w2 = q.w * q.w - 0.5f;
    float dot2 = (q.x * vx + q.y * vy + q.z * vz);
    return mk3((vx * w2 - (q.y * vz - q.z * vy) * q.w + q.x * dot2), (vy * w2 - (q.z * vx - q.x * vz) * q.w + q.y * dot2),
               (vz * w2 - (q.x * vy - q.y * vx) * q.w + q.z * dot2));
}

// ------------------------------------------------------------ RNG / hashing

// MurmurHash64A over the 16-byte key (px, py, seed, sample)   hash.h:17-70,111-120
BB_HD uint64_t pixel_sample_hash(int32_t px, int32_t py, int32_t seed, int32_t sample)
{
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    uint64_t h = 0ull ^ (16ull * m);
    uint64_t k0 = (uint64_t)(uint32_t)px | ((uint64_t)(uint32_t)py << 32);
    uint64_t k1 = (uint64_t)(uint32_t)seed | ((uint64_t)(uint32_t)sample << 32);
    k0 *= m;
    k0 ^= k0 >> r;
    k0 *= m;
    h ^= k0;
    h *= m;
    k1 *= m;
    k1 ^= k1 >> r;
    k1 *= m;
    h ^= k1;
    h *= m;
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

// HashFloat(ray, point) (hash.h:111-126): MurmurHash64A, seed 0, over the
// 36 bytes of Ray{o, d} followed by the point, low 32 bits times 2^-32
BB_HD float hash_float_ray_point(f3 o, f3 d, f3 p)
{
    const uint64_t m = 0xc6a4a7935bd1e995ull;
    const int r = 47;
    const float w[9] = { o.x, o.y, o.z, d.x, d.y, d.z, p.x, p.y, p.z };
    uint32_t b[9];
    for (int i = 0; i < 9; ++i) memcpy(&b[i], &w[i], 4);
    uint64_t h = 0ull ^ (36ull * m);
    for (int i = 0; i < 4; ++i) {
        uint64_t k = uint64_t(b[2 * i]) | (uint64_t(b[2 * i + 1]) << 32);
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    h ^= uint64_t(b[8]); // the 4-byte tail, little-endian
    h *= m;
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return float(uint32_t(h)) * 0x1p-32f;
}

// PCG32 (random.h:27-46, 87-90); inc = (1 << 1) | 1 = 3 for initseq 1
constexpr uint64_t kPcgMult = 0x5851f42d4c957f2dULL;
constexpr uint64_t kPcgInc = 3ull;

BB_HD uint32_t pcg_next(uint64_t& state)
{
    uint64_t old = state;
    state = old * kPcgMult + kPcgInc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
}

BB_HD uint64_t pcg_seed(uint64_t initstate)
{
    uint64_t state = 0u;
    pcg_next(state);
    state += initstate;
    pcg_next(state);
    return state;
}

BB_HD float pcg_float(uint64_t& state) { return fminf(1.0f - FLT_EPSILON, (float)pcg_next(state) * 0x1p-32f); }

} // namespace bb
