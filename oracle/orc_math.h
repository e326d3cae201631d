/*
 * orc_math.h — float32 vector, spectrum and sampling arithmetic of the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of Bulbit's path
 * tracer used as the parity checker for the GPU library; nothing shipped links
 * it.  Every operation keeps the reference's evaluation order, because the GPU
 * results are compared with it bit for bit (hit ids) or to 1e-3 (pixels).
 * Build with -ffp-contract=off and without -ffast-math.
 *
 * Cited reference files (paths relative to the Bulbit source tree):
 *   include/bulbit/floats.h:16-24      constants
 *   include/bulbit/vectors.h:121-133   member Normalize (epsilon early-out)
 *   include/bulbit/vectors.h:363-373   Dot, Cross
 *   include/bulbit/math.h:56-60        free Normalize = v * (1 / Length(v))
 *   include/bulbit/math.h:47-51,73-83  Clamp, SmoothStep
 *   include/bulbit/spectrum.h:49-83    RGBSpectrum IsBlack, Luminance, MaxComponent, IsNullish
 *   include/bulbit/frame.h:9-116       CosTheta ... SameHemisphere, CoordinateSystem
 */
#ifndef ORC_MATH_H
#define ORC_MATH_H

#include <float.h>
#include <math.h>
#include <stdint.h>

/* floats.h:16-22: pi is a float constant; derived constants are float ops */
#define ORC_PI ((float)3.14159265358979323846)
static const float orc_pi = ORC_PI;
static const float orc_two_pi = 2.0f * ORC_PI;
static const float orc_four_pi = 4.0f * ORC_PI;
static const float orc_inv_pi = 1.0f / ORC_PI;
static const float orc_inv_two_pi = 1.0f / (2.0f * ORC_PI);
static const float orc_inv_four_pi = 1.0f / (4.0f * ORC_PI);
#define ORC_EPS FLT_EPSILON

typedef struct { float x, y; } v2;
typedef struct { float x, y, z; } v3;
typedef struct { float r, g, b; } spec;

static inline v2 V2(float x, float y) { v2 v = { x, y }; return v; }
static inline v3 V3(float x, float y, float z) { v3 v = { x, y, z }; return v; }
static inline spec S3(float r, float g, float b) { spec s = { r, g, b }; return s; }
static inline spec S1(float s) { return S3(s, s, s); }

/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a, max(a,b) = (a < b) ? b : a */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }
/* math.h:47-51 */
static inline float clampf_(float v, float l, float r) { return v < l ? l : (v > r ? r : v); }

static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vneg(v3 a) { return V3(-a.x, -a.y, -a.z); }
static inline v3 vmul(v3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vmulv(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vdiv(v3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline v3 vdivv(v3 a, v3 b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float vlen2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline float vlen(v3 a) { return sqrtf(vlen2(a)); }
static inline float vabsdot(v3 a, v3 b) { return fabsf(vdot(a, b)); }
static inline v3 vmin(v3 a, v3 b) { return V3(smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return V3(smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)); }
static inline float vcomp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* math.h:56-60 */
static inline v3 vnormalize(v3 v)
{
    float inv_length = 1.0f / vlen(v);
    return vmul(v, inv_length);
}

/* vectors.h:121-133: returns the length, leaves v unchanged below epsilon */
static inline float vnormalize_m(v3* v)
{
    float length = vlen(*v);
    if (length < ORC_EPS) return 0.0f;
    float inv_length = 1.0f / length;
    v->x *= inv_length;
    v->y *= inv_length;
    v->z *= inv_length;
    return length;
}

static inline spec sadd(spec a, spec b) { return S3(a.r + b.r, a.g + b.g, a.b + b.b); }
static inline spec ssub(spec a, spec b) { return S3(a.r - b.r, a.g - b.g, a.b - b.b); }
static inline spec smul(spec a, spec b) { return S3(a.r * b.r, a.g * b.g, a.b * b.b); }
static inline spec smulf(spec a, float s) { return S3(a.r * s, a.g * s, a.b * s); }
static inline spec sdivf(spec a, float s) { return S3(a.r / s, a.g / s, a.b / s); }
static inline spec sdiv(spec a, spec b) { return S3(a.r / b.r, a.g / b.g, a.b / b.b); }
static inline int sblack(spec a) { return a.r == 0.0f && a.g == 0.0f && a.b == 0.0f; }
static inline float slum(spec a) { return a.r * 0.2126f + a.g * 0.7152f + a.b * 0.0722f; }
static inline float smaxc(spec a)
{
    float m = a.r;
    if (m < a.g) m = a.g;
    if (m < a.b) m = a.b;
    return m;
}
static inline int snullish(spec a)
{
    return isnan(a.r) || isinf(a.r) || isnan(a.g) || isinf(a.g) || isnan(a.b) || isinf(a.b);
}
static inline spec ssqrt(spec a) { return S3(sqrtf(a.r), sqrtf(a.g), sqrtf(a.b)); }

static inline float sqrf(float v) { return v * v; }
/* math.h:24-27 */
static inline float safe_sqrt(float x) { return sqrtf(smax(0.0f, x)); }

/* math.h:73-83 */
static inline float smooth_step(float a, float b, float x)
{
    if (a == b) return (x < a) ? 0.0f : 1.0f;
    float t = clampf_((x - a) / (b - a), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

/* frame.h:9-116, shading coordinate helpers */
static inline float cos2_theta(v3 w) { return sqrf(w.z); }
static inline float abs_cos_theta(v3 w) { return fabsf(w.z); }
static inline float sin2_theta(v3 w) { return smax(0.0f, 1.0f - cos2_theta(w)); }
static inline float sin_theta(v3 w) { return sqrtf(sin2_theta(w)); }
static inline float tan2_theta(v3 w) { return sin2_theta(w) / cos2_theta(w); }
static inline float cos_phi(v3 w)
{
    float st = sin_theta(w);
    return (st == 0.0f) ? 1.0f : clampf_(w.x / st, -1.0f, 1.0f);
}
static inline float sin_phi(v3 w)
{
    float st = sin_theta(w);
    return (st == 0.0f) ? 0.0f : clampf_(w.y / st, -1.0f, 1.0f);
}
static inline int same_hemisphere(v3 a, v3 b) { return a.z * b.z > 0.0f; }

/* frame.h:107-114 */
static inline void coordinate_system(v3 v1, v3* v2o, v3* v3o)
{
    float sign = copysignf(1.0f, v1.z);
    float a = -1.0f / (sign + v1.z);
    float b = v1.x * v1.y * a;
    *v2o = V3(1.0f + sign * (v1.x * v1.x) * a, sign * b, -sign * v1.x);
    *v3o = V3(b, sign + (v1.y * v1.y) * a, -v1.y);
}

/* Frame (frame.h:117-203): FromXZ(x, z) = (x, Cross(z, x), z) */
typedef struct { v3 x, y, z; } frame3;
static inline frame3 frame_from_xz(v3 x, v3 z)
{
    frame3 f = { x, vcross(z, x), z };
    return f;
}
static inline frame3 frame_from_z(v3 z)
{
    frame3 f;
    f.z = z;
    coordinate_system(z, &f.x, &f.y);
    return f;
}
static inline v3 frame_to_local(const frame3* f, v3 v) { return V3(vdot(v, f->x), vdot(v, f->y), vdot(v, f->z)); }
/* v.x * x + v.y * y + v.z * z, each term Vector * scalar */
static inline v3 frame_from_local(const frame3* f, v3 v)
{
    return vadd(vadd(vmul(f->x, v.x), vmul(f->y, v.y)), vmul(f->z, v.z));
}

/* Dot(v, z_axis), evaluated as written */
static inline float dot_z(v3 v) { return v.x * 0.0f + v.y * 0.0f + v.z * 1.0f; }

#endif /* ORC_MATH_H */
