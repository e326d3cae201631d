// bb_shade.h — device-side BSDFs, microfacet sampling, Fresnel terms, light
// sampling and MIS for the wavefront path tracer (gfx950).
//
// Restates, in the reference's evaluation order:
//   include/bulbit/sampling.h:42-48,64-74,114-169,203-232   sampling, MIS
//   include/bulbit/scattering.h:9-132                        Refract, Fresnel
//   include/bulbit/microfacet.h:13-93                        Trowbridge-Reitz (GGX)
//   src/bxdf/{lambertian,dielectric,conductor,metallic_roughness}_bxdf.cpp
//   include/bulbit/bsdf.h:75-125                             BSDF wrapper
//   src/light/{spot_light,uniform_infinite_light,diffuse_area_light}.cpp
//
// Transcendentals: sin, cos and pow are evaluated in double precision and
// rounded once to float, which reproduces a correctly rounded float result;
// the reference's glibc sinf/cosf agree with it except in ~1% of arguments
// (by one ulp).  Division and sqrt are IEEE (correctly rounded) on gfx950
// with the library's compile flags.
#pragma once

#include "bb_scene.h"
#include "bb_texcoord.h"

namespace bb
{

BB_D float dsin(float x) { return (float)::sin((double)x); }
BB_D float dcos(float x) { return (float)::cos((double)x); }
BB_D float dpow(float x, float y) { return (float)::pow((double)x, (double)y); }
BB_D float dacos(float x) { return (float)::acos((double)x); }

// ------------------------------------------------------------------ sampling

// SampleUniformUnitDiskConcentric, sampling.h:114-134
BB_D f2 sample_concentric(f2 u)
{
    f2 uo = mk2(u.x * 2.0f - 1.0f, u.y * 2.0f - 1.0f);
    if (uo.x == 0.0f && uo.y == 0.0f) return mk2(0.0f, 0.0f);
    float theta, r;
    if (fabsf(uo.x) > fabsf(uo.y)) {
        r = uo.x;
        theta = (kPi / 4.0f) * (uo.y / uo.x);
    } else {
        r = uo.y;
        theta = (kPi / 2.0f) - (kPi / 4.0f) * (uo.x / uo.y);
    }
    return mk2(dcos(theta) * r, dsin(theta) * r);
}

// SampleUniformSphere, sampling.h:64-74
BB_D f3 sample_uniform_sphere(f2 u)
{
    float z = 1.0f - 2.0f * u.x;
    float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float phi = kTwoPi * u.y;
    return mk3(r * dcos(phi), r * dsin(phi), z);
}

// SampleCosineHemisphere (Malley), sampling.h:151-169
BB_D f3 sample_cosine_hemisphere(f2 u)
{
    f2 d = sample_concentric(u);
    float z = safe_sqrt(1.0f - sqr(d.x) - sqr(d.y));
    return mk3(d.x, d.y, z);
}

// PowerHeuristic(1, f, 1, g), sampling.h:42-48
BB_D float power_heuristic(float pdf_f, float pdf_g)
{
    float f = 1.0f * pdf_f;
    float g = 1.0f * pdf_g;
    return (f * f) / (f * f + g * g);
}

// ------------------------------------------------------------------ Fresnel

// FresnelDielectric, scattering.h:51-77
BB_D float fresnel_dielectric(float cos_theta_i, float eta)
{
    cos_theta_i = clampf_(cos_theta_i, -1.0f, 1.0f);
    if (cos_theta_i < 0.0f) {
        eta = 1.0f / eta;
        cos_theta_i = -cos_theta_i;
    }
    float sin2_theta_i = 1.0f - sqr(cos_theta_i);
    float sin2_theta_t = sin2_theta_i / sqr(eta);
    if (sin2_theta_t >= 1.0f) return 1.0f;
    float cos_theta_t = sqrtf(smax(0.0f, 1.0f - sin2_theta_t));
    float r_parl = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    float r_perp = (cos_theta_i - eta * cos_theta_t) / (cos_theta_i + eta * cos_theta_t);
    return (sqr(r_parl) + sqr(r_perp)) / 2.0f;
}

// std::complex<float> as the reference's compiler evaluates it: float
// multiplication, division by the naive formula in double, |z| = hypot.
struct cpx
{
    float re, im;
};
BB_D cpx cmul(cpx a, cpx b) { return cpx{ a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re }; }
BB_D cpx cdiv(cpx a, cpx b)
{
    double ar = a.re, ai = a.im, br = b.re, bi = b.im;
    double den = br * br + bi * bi;
    return cpx{ (float)((ar * br + ai * bi) / den), (float)((ai * br - ar * bi) / den) };
}
BB_D float cabs_(cpx z)
{
    double r = z.re, i = z.im;
    return (float)::sqrt(r * r + i * i);
}
BB_D float cnorm(cpx z) { return z.re * z.re + z.im * z.im; }

// SqrtComplex, scattering.h:81-100
BB_D cpx csqrt_(cpx z)
{
    float n = cabs_(z);
    float t1 = sqrtf(0.5f * (n + fabsf(z.re)));
    float t2 = 0.5f * z.im / t1;
    if (n == 0.0f) return cpx{ 0.0f, 0.0f };
    if (z.re >= 0.0f) return cpx{ t1, t2 };
    return cpx{ fabsf(t2), copysignf(t1, z.im) };
}

// FresnelComplex, scattering.h:102-121
BB_D float fresnel_complex(float cos_theta_i, cpx eta)
{
    cos_theta_i = clampf_(cos_theta_i, 0.0f, 1.0f);
    float sin2_theta_i = 1.0f - sqr(cos_theta_i);
    cpx sin2_theta_t = cdiv(cpx{ sin2_theta_i, 0.0f }, cmul(eta, eta));
    cpx cos_theta_t = csqrt_(cpx{ 1.0f - sin2_theta_t.re, 0.0f - sin2_theta_t.im });
    cpx ec = cpx{ eta.re * cos_theta_i, eta.im * cos_theta_i };
    cpx r_parl = cdiv(cpx{ ec.re - cos_theta_t.re, ec.im - cos_theta_t.im }, cpx{ ec.re + cos_theta_t.re, ec.im + cos_theta_t.im });
    cpx et = cmul(eta, cos_theta_t);
    cpx r_perp = cdiv(cpx{ cos_theta_i - et.re, -et.im }, cpx{ cos_theta_i + et.re, et.im });
    return (cnorm(r_parl) + cnorm(r_perp)) / 2.0f;
}

BB_D spec fresnel_complex_s(float c, spec eta, spec k)
{
    return mks(fresnel_complex(c, cpx{ eta.r, k.r }), fresnel_complex(c, cpx{ eta.g, k.g }),
               fresnel_complex(c, cpx{ eta.b, k.b }));
}

// Refract, scattering.h:9-39
BB_D bool refract(f3& wt, f3 wi, f3 n, float eta, float& eta_p)
{
    float cos_theta_i = dot(n, wi);
    if (cos_theta_i < 0.0f) {
        eta = 1.0f / eta;
        cos_theta_i = -cos_theta_i;
        n = -n;
    }
    float sin2_theta_i = smax(0.0f, 1.0f - sqr(cos_theta_i));
    float sin2_theta_t = sin2_theta_i / sqr(eta);
    if (sin2_theta_t >= 1.0f) return false;
    float cos_theta_t = sqrtf(1.0f - sin2_theta_t);
    float inv_eta = 1.0f / eta;
    wt = (-wi) * inv_eta + n * (cos_theta_i * inv_eta - cos_theta_t);
    eta_p = eta;
    return true;
}

// ---------------------------------------------------- Trowbridge-Reitz (GGX)
struct ggx
{
    float ax, ay;
};
BB_D bool ggx_smooth(ggx m) { return smax(m.ax, m.ay) < 1e-3f; }

BB_D float ggx_D(ggx m, f3 wm)
{
    float tan2 = tan2_theta(wm);
    if (isinf(tan2)) return 0.0f;
    float cos4 = sqr(cos2_theta(wm));
    if (cos4 < 1e-16f) return 0.0f;
    float e = tan2 * (sqr(cos_phi(wm) / m.ax) + sqr(sin_phi(wm) / m.ay));
    return 1.0f / (kPi * m.ax * m.ay * cos4 * sqr(1.0f + e));
}
BB_D float ggx_lambda(ggx m, f3 w)
{
    float tan2 = tan2_theta(w);
    if (isinf(tan2)) return 0.0f;
    float alpha2 = sqr(cos_phi(w) * m.ax) + sqr(sin_phi(w) * m.ay);
    return (sqrtf(1.0f + alpha2 * tan2) - 1.0f) / 2.0f;
}
BB_D float ggx_G1(ggx m, f3 w) { return 1.0f / (1.0f + ggx_lambda(m, w)); }
BB_D float ggx_G(ggx m, f3 wo, f3 wi) { return 1.0f / (1.0f + ggx_lambda(m, wo) + ggx_lambda(m, wi)); }
BB_D float ggx_pdf(ggx m, f3 w, f3 wm) { return ggx_G1(m, w) / abs_cos_theta(w) * ggx_D(m, wm) * absdot(w, wm); }

// Sample_GGX_VNDF_Dupuy_Benyoub, sampling.h:203-232 (explicit fma at :210)
BB_D f3 ggx_sample_wm(ggx m, f3 wo, f2 u)
{
    f3 wo_std = normalize(mk3(wo.x * m.ax, wo.y * m.ay, wo.z));
    float phi = kTwoPi * u.x;
    float z = fmaf((1.0f - u.y), (1.0f + wo_std.z), -wo_std.z);
    float sin_t = sqrtf(clampf_(1.0f - z * z, 0.0f, 1.0f));
    float x = sin_t * dcos(phi);
    float y = sin_t * dsin(phi);
    f3 h = mk3(x, y, z) + wo_std;
    return normalize(mk3(h.x * m.ax, h.y * m.ay, h.z));
}

// -------------------------------------------------------------------- BxDFs

struct bxdf
{
    int32_t kind;
    int32_t flags;
    float eta;
    ggx mf;
    spec r, ceta, ck, color;
    float metallic;
    const float* E;     // GGX albedo table (compensated conductor)
    const float* E_avg; // its average over directions
    const float* dt;    // DielectricTables (compensated dielectric)
};

struct bsample
{
    spec f;
    f3 wi;
    float pdf;
    int32_t flags;
    float eta;
};

BB_D bxdf make_bxdf(const GpuMaterial& m, bool front_face, const SceneView& S)
{
    bxdf b;
    b.E = S.ggx_E;
    b.E_avg = S.ggx_E_avg;
    b.dt = S.dielectric;
    b.kind = m.kind;
    b.flags = m.flags_front;
    const bool dielectric = m.kind == MK_DIELECTRIC || m.kind == MK_DIELECTRIC_MS;
    b.eta = dielectric ? (front_face ? m.eta : 1.0f / m.eta) : 1.0f;
    b.mf = ggx{ m.ax, m.ay };
    b.r = m.r;
    b.ceta = m.ceta;
    b.ck = m.ck;
    b.color = m.color;
    b.metallic = m.metallic;
    return b;
}

// ------------------------------------------------------------- textures

// ImageTexture::Evaluate, bilinear (include/bulbit/textures.h:60-88): channel c of the image;
// filter_coord and tex_floor_index (bb_texcoord.h) keep every tap inside the image
BB_D void image_eval(const GpuTexture& t, const float* texels, f2 uv, float* out)
{
    const float w = uv.x * float(t.width) - 0.5f;
    const float h = uv.y * float(t.height) - 0.5f;
    int32_t i0 = tex_floor_index(w, t.width, t.filter), i1 = i0 + 1;
    int32_t j0 = tex_floor_index(h, t.height, t.filter), j1 = j0 + 1;
    i0 = filter_coord(i0, t.width, t.filter);
    j0 = filter_coord(j0, t.height, t.filter);
    i1 = filter_coord(i1, t.width, t.filter);
    j1 = filter_coord(j1, t.height, t.filter);
    const float fu = w - floorf(w);
    const float fv = h - floorf(h);
    const float* img = texels + t.texel_offset;
    const int c = t.channels;
    const float* v00 = img + size_t(i0 + j0 * t.width) * c;
    const float* v10 = img + size_t(i1 + j0 * t.width) * c;
    const float* v01 = img + size_t(i0 + j1 * t.width) * c;
    const float* v11 = img + size_t(i1 + j1 * t.width) * c;
    const float w00 = (1.0f - fu) * (1.0f - fv), w01 = (1.0f - fu) * fv, w10 = fu * (1.0f - fv), w11 = fu * fv;
    for (int k = 0; k < c; ++k) out[k] = w00 * v00[k] + w01 * v01[k] + w10 * v10[k] + w11 * v11[k];
}

// Texture<T>::Evaluate of a constant or image texture
BB_D void leaf_texture_eval(const GpuTexture& t, const float* texels, f2 uv, float* out)
{
    if (t.type == BB_TEX_IMAGE) {
        image_eval(t, texels, uv, out);
    } else {
        out[0] = t.value[0];
        out[1] = t.value[1];
        out[2] = t.value[2];
    }
}

// Texture<T>::Evaluate; CheckerTexture (textures.h:109-130): Point2i of
// uv * resolution (truncation), a when the coordinate sum is odd, else b
BB_D void texture_eval(const SceneView& S, int32_t idx, f2 uv, float* out)
{
    const GpuTexture& t = S.textures[idx];
    if (t.type != BB_TEX_CHECKER) {
        leaf_texture_eval(t, S.texels, uv, out);
        return;
    }
    const int32_t sx = int32_t(uv.x * t.res_x), sy = int32_t(uv.y * t.res_y);
    const int32_t c = sx + sy;
    leaf_texture_eval(S.textures[(c % 2) ? t.a : t.b], S.texels, uv, out);
}

BB_D float texture1(const SceneView& S, int32_t idx, f2 uv)
{
    float v[3];
    texture_eval(S, idx, uv, v);
    return v[0];
}

BB_D spec texture3(const SceneView& S, int32_t idx, f2 uv)
{
    float v[3];
    texture_eval(S, idx, uv, v);
    return mks(v[0], v[1], v[2]);
}

// The material's constants with its textured parameters evaluated at uv:
// the per-hit part of the *Material::GetBSDF functions
// (diffuse_material.cpp:21-36, dielectric_material.cpp:28-52,
// conductor_material.cpp:52-99, metallic_roughness_material.cpp:28-39),
// with the same derived values as the constant set-up (kind, alphas, flags)
BB_D GpuMaterial material_at(const GpuMaterial& m0, const SceneView& S, f2 uv)
{
    GpuMaterial m = m0;
    switch (m.type) {
    case BB_MAT_DIFFUSE:
        if (m.tex_color >= 0) m.r = texture3(S, m.tex_color, uv);
        if (m.tex_u >= 0) {
            // DiffuseMaterial::GetBSDF: EON when the roughness is > 0, else Lambertian
            const float r = texture1(S, m.tex_u, uv);
            m.kind = r > 0.0f ? MK_EON : MK_LAMBERT;
            m.ax = m.ay = r;
        }
        m.flags_front = !is_black(m.r) ? (BX_DIFFUSE | BX_REFLECTION) : 0;
        break;
    case BB_MAT_DIELECTRIC: {
        if (m.tex_color >= 0) m.r = sqrt_s(texture3(S, m.tex_color, uv));
        if (m.tex_u >= 0) m.ax = sqr(texture1(S, m.tex_u, uv));
        if (m.tex_v >= 0) m.ay = sqr(texture1(S, m.tex_v, uv));
        const bool smooth = smax(m.ax, m.ay) < 1e-3f;
        m.kind = (m.energy_compensation && !smooth && m.eta != 1.0f) ? MK_DIELECTRIC_MS : MK_DIELECTRIC;
        const int32_t f = (m.eta == 1.0f) ? BX_TRANSMISSION : (BX_REFLECTION | BX_TRANSMISSION);
        m.flags_front = f | (smooth ? BX_SPECULAR : BX_GLOSSY);
        break;
    }
    case BB_MAT_CONDUCTOR:
    case BB_MAT_CONDUCTOR_R: {
        if (m.tex_color >= 0) m.r = texture3(S, m.tex_color, uv);
        if (m.tex_u >= 0) m.ax = sqr(texture1(S, m.tex_u, uv));
        if (m.tex_v >= 0) m.ay = sqr(texture1(S, m.tex_v, uv));
        const bool smooth = smax(m.ax, m.ay) < 1e-3f;
        m.flags_front = smooth ? (BX_SPECULAR | BX_REFLECTION) : (BX_GLOSSY | BX_REFLECTION);
        break;
    }
    case BB_MAT_MIRROR:
        if (m.tex_color >= 0) m.r = texture3(S, m.tex_color, uv);
        m.flags_front = !is_black(m.r) ? (BX_SPECULAR | BX_REFLECTION) : 0;
        break;
    case BB_MAT_METALLIC_ROUGHNESS:
        if (m.tex_color >= 0) m.color = texture3(S, m.tex_color, uv);
        if (m.tex_metallic >= 0) m.metallic = texture1(S, m.tex_metallic, uv);
        if (m.tex_u >= 0) m.ax = fmaxf(sqr(texture1(S, m.tex_u, uv)), 1e-3f);
        if (m.tex_v >= 0) m.ay = fmaxf(sqr(texture1(S, m.tex_v, uv)), 1e-3f);
        break;
    }
    return m;
}

// NormalMapping (src/geometry/intersection.cpp:9-30): the shading normal
// from the map in the (tangent, normal) frame, reflected about the
// geometric normal when it faces away from wo; the tangent projected
// (not renormalized)
BB_D void normal_map(const SceneView& S, int32_t idx, f2 uv, f3 wo, f3 normal, f3& ns, f3& ts)
{
    const spec sp = texture3(S, idx, uv);
    f3 sn = mk3(sp.r * 2.0f - 1.0f, sp.g * 2.0f - 1.0f, sp.b * 2.0f - 1.0f);
    normalize_m(sn);
    const frame3 tbn = frame_from_xz(ts, ns);
    f3 n = from_local(tbn, sn);
    if (dot(n, wo) < 0.0f) n = -n + normal * (2.0f * dot(n, normal));
    const f3 t = ts - n * dot(ts, n);
    ns = n;
    ts = t;
}

// --------------------------------------------------------------- EON
// EONBxDF (src/bxdf/eon_bxdf.cpp), the energy-preserving Oren-Nayar BRDF of
// Portsmouth et al., "EON: A Practical Energy-Preserving Rough Diffuse BRDF"
// (JCGT 14(1), 2025), exact variant: a single-scattering FON lobe plus a
// multiple-scattering lobe that restores the energy FON loses, sampled by
// mixing a clipped linearly transformed cosine (CLTC) with the uniform
// hemisphere.  The fitted constants are the paper's.  b.r = rho, b.mf.ax = r.

constexpr float kFON1 = 0.5f - 2.0f / (3.0f * kPi);           // constant1_FON
constexpr float kFON2 = 2.0f / 3.0f - 28.0f / (15.0f * kPi);  // constant2_FON

// E_FON_exact: the FON lobe's directional albedo
BB_D float eon_E(float mu, float r)
{
    const float AF = 1.0f / (1.0f + kFON1 * r);
    const float BF = r * AF;
    const float Si = sqrtf(1.0f - (mu * mu));
    const float G = Si * (dacos(mu) - Si * mu) + (2.0f / 3.0f) * ((Si / mu) * (1.0f - (Si * Si * Si)) - Si);
    return AF + (BF * kInvPi) * G;
}

// ltc_coeffs (the paper's Listing 2)
BB_D void eon_ltc(float mu, float r, float& a, float& b, float& c, float& d)
{
    a = 1.0f + r * (0.303392f + (-0.518982f + 0.111709f * mu) * mu + (-0.276266f + 0.335918f * mu) * r);
    b = r * (-1.16407f + 1.15859f * mu + (0.150815f - 0.150105f * mu) * r) / (mu * mu * mu - 1.43545f);
    c = 1.0f + (0.20013f + (-0.506373f + 0.261777f * mu) * mu) * r;
    d = ((0.540852f + (-1.01625f + 0.475392f * mu) * mu) * r) / (-1.0743f + mu * (0.0725628f + mu));
}

// orthonormal_basis_ltc: X = wo's direction in the tangent plane, Y = Z x X, Z = z
BB_D void eon_basis(f3 w, f3& X, f3& Y)
{
    const float len2 = w.x * w.x + w.y * w.y;
    X = len2 > 0.0f ? mk3(w.x, w.y, 0.0f) / sqrtf(len2) : mk3(1.0f, 0.0f, 0.0f);
    Y = mk3(-X.y, X.x, 0.0f);
}

BB_D float eon_cltc_pdf(f3 wo, f3 wi_local, float r)
{
    f3 X, Y;
    eon_basis(wo, X, Y);
    // Mul(transpose(X, Y, Z), wi)
    const f3 wi = mk3(X.x * wi_local.x + X.y * wi_local.y + X.z * wi_local.z,
                      Y.x * wi_local.x + Y.y * wi_local.y + Y.z * wi_local.z,
                      0.0f * wi_local.x + 0.0f * wi_local.y + 1.0f * wi_local.z);
    float a, b, c, d;
    eon_ltc(wo.z, r, a, b, c, d);
    const float detM = c * (a - b * d);
    const f3 wh = mk3(c * (wi.x - b * wi.z), (a - b * d) * wi.y, -c * (d * wi.x - a * wi.z));
    const float len2 = dot(wh, wh);
    const float vz = 1.0f / sqrtf(d * d + 1.0f);
    const float s = 0.5f * (1.0f + vz);
    return detM * detM / (len2 * len2) * fmaxf(wh.z, 0.0f) / (kPi * s);
}

BB_D f3 eon_cltc_sample(f3 wo, float r, float u1, float u2, float& pdf)
{
    float a, b, c, d;
    eon_ltc(wo.z, r, a, b, c, d);
    const float R = sqrtf(u1);
    const float phi = kTwoPi * u2;
    float x = R * dcos(phi);
    const float y = R * dsin(phi);
    const float vz = 1.0f / sqrtf(d * d + 1.0f);
    const float s = 0.5f * (1.0f + vz);
    x = -(sqrtf(1.0f - y * y) * (1.0f - s) + x * s); // -Lerp(sqrt(1 - y^2), x, s)
    const f3 wh = mk3(x, y, sqrtf(fmaxf(1.0f - (x * x + y * y), 0.0f)));
    const float pdf_wh = wh.z / (kPi * s);
    const f3 wi = mk3(a * wh.x + b * wh.z, c * wh.y, d * wh.x + wh.z);
    const float len = length(wi);
    const float detM = c * (a - b * d);
    pdf = pdf_wh * len * len * len / detM;
    f3 X, Y;
    eon_basis(wo, X, Y);
    // Normalize(Mul((X, Y, Z), wi))
    return normalize(mk3(X.x * wi.x + Y.x * wi.y + 0.0f * wi.z, X.y * wi.x + Y.y * wi.y + 0.0f * wi.z,
                         X.z * wi.x + Y.z * wi.y + 1.0f * wi.z));
}

BB_D spec eon_f(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi)) return mks(0.0f);
    const float r = b.mf.ax;
    const float mu_i = wi.z, mu_o = wo.z;
    const float s = dot(wi, wo) - mu_i * mu_o;
    const float sovertF = s > 0.0f ? s / smax(mu_i, mu_o) : s;
    const float AF = 1.0f / (1.0f + kFON1 * r);
    const spec f_ss = ((b.r * kInvPi) * AF) * (1.0f + r * sovertF);
    const float EFo = eon_E(mu_o, r), EFi = eon_E(mu_i, r);
    const float avgEF = AF * (1.0f + kFON2 * r);
    const spec rho_ms = ((b.r * b.r) * avgEF) / (mks(1.0f) - b.r * (1.0f - avgEF));
    const float eps = 1.0e-7f;
    const spec f_ms = (((rho_ms * kInvPi) * smax(eps, 1.0f - EFo)) * smax(eps, 1.0f - EFi)) / smax(eps, 1.0f - avgEF);
    return f_ss + f_ms;
}

// probability of the uniform lobe
BB_D float eon_P_u(float mu, float r)
{
    return dpow(r, 0.1f) * (0.162925f + mu * (-0.372058f + (0.538233f - 0.290822f * mu) * mu));
}

BB_D float eon_pdf(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi)) return 0.0f;
    const float P_u = eon_P_u(wo.z, b.mf.ax);
    const float P_c = 1.0f - P_u;
    return P_u * kInvTwoPi + P_c * eon_cltc_pdf(wo, wi, b.mf.ax);
}

BB_D bool eon_sample(const bxdf& b, f3 wo, float u0, f2 u12, bsample& out)
{
    const float r = b.mf.ax;
    const float P_u = eon_P_u(wo.z, r);
    const float P_c = 1.0f - P_u;
    f3 wi;
    float pdf_c;
    if (u0 <= P_u) {
        // SampleUniformHemisphere (include/bulbit/sampling.h:50-57)
        const float z = u12.x;
        const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        const float phi = kTwoPi * u12.y;
        wi = mk3(rr * dcos(phi), rr * dsin(phi), z);
        pdf_c = eon_cltc_pdf(wo, wi, r);
    } else {
        wi = eon_cltc_sample(wo, r, u12.x, u12.y, pdf_c);
    }
    out.pdf = P_u * kInvTwoPi + P_c * pdf_c;
    out.f = eon_f(b, wo, wi);
    out.wi = wi;
    out.flags = BX_DIFFUSE | BX_REFLECTION;
    return true;
}

// Lambertian, lambertian_bxdf.cpp:8-53
BB_D spec lambert_f(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi)) return mks(0.0f);
    return b.r * kInvPi;
}
BB_D float lambert_pdf(f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi)) return 0.0f;
    return abs_cos_theta(wi) * kInvPi;
}

// Conductor, conductor_bxdf.cpp:7-112
BB_D spec conductor_f(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b.mf)) return mks(0.0f);
    float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
    if (ci == 0.0f || co == 0.0f) return mks(0.0f);
    f3 wm = wo + wi;
    if (length2(wm) == 0.0f) return mks(0.0f);
    normalize_m(wm);
    spec F = fresnel_complex_s(absdot(wo, wm), b.ceta, b.ck);
    return (((b.r * F) * ggx_D(b.mf, wm)) * ggx_G(b.mf, wo, wi)) / (4.0f * ci * co);
}
BB_D float conductor_pdf(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b.mf)) return 0.0f;
    f3 wm = wo + wi;
    if (length2(wm) == 0.0f) return 0.0f;
    normalize_m(wm);
    if (dot_z(wm) < 0.0f) wm = -wm;
    return ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
}

// ImageTexture<Float>::Evaluate, bilinear with TexCoordFilter::clamp
// (textures.h:60-88, texture.h:26-46)
BB_D float tex_bilinear(const float* img, int32_t w, int32_t h, float u, float v)
{
    const float fw = u * float(w) - 0.5f;
    const float fh = v * float(h) - 0.5f;
    int32_t i0 = int32_t(floorf(fw)), i1 = i0 + 1;
    int32_t j0 = int32_t(floorf(fh)), j1 = j0 + 1;
    i0 = min(max(i0, 0), w - 1);
    j0 = min(max(j0, 0), h - 1);
    i1 = min(max(i1, 0), w - 1);
    j1 = min(max(j1, 0), h - 1);
    const float fu = fw - floorf(fw);
    const float fv = fh - floorf(fh);
    const float v00 = img[i0 + j0 * w], v10 = img[i1 + j0 * w];
    const float v01 = img[i0 + j1 * w], v11 = img[i1 + j1 * w];
    return (1.0f - fu) * (1.0f - fv) * v00 + (1.0f - fu) * fv * v01 + fu * (1.0f - fv) * v10 + fu * fv * v11;
}

// TrowbridgeReitzDistribution::E / E_avg (microfacet.h:133-145) with
// GetMeanAlpha = sqrt(alpha_x alpha_y) (microfacet.h:157-160)
BB_D float ggx_E(const bxdf& b, f3 wo)
{
    return tex_bilinear(b.E, kTableSize, kTableSize, wo.z, sqrtf(b.mf.ax * b.mf.ay));
}
BB_D float ggx_E_avg(const bxdf& b) { return tex_bilinear(b.E_avg, kTableSize, 1, sqrtf(b.mf.ax * b.mf.ay), 0.0f); }

// Conductor with multiple-scattering compensation,
// conductor_multi_scattering_bxdf.cpp:7-148; FresnelConductorAverage
// (scattering.h:174-185)
BB_D spec conductor_ms_lobe(const bxdf& b, f3 wo, f3 wi)
{
    const float E_avg = ggx_E_avg(b);
    const spec F0 = fresnel_complex_s(1.0f, b.ceta, b.ck);
    const spec F90 = fresnel_complex_s(0.0f, b.ceta, b.ck);
    const spec fa = (F0 * 20.0f + F90) / 21.0f;
    const spec fresnel_ms = ((fa * fa) * E_avg) / (mks(1.0f) - fa * (1.0f - E_avg));
    return (fresnel_ms * ((1.0f - ggx_E(b, wo)) * (1.0f - ggx_E(b, wi)))) / (kPi * (1.0f - E_avg));
}
BB_D spec conductor_ms_f(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b.mf)) return mks(0.0f);
    float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
    if (ci == 0.0f || co == 0.0f) return mks(0.0f);
    f3 wm = wo + wi;
    if (length2(wm) == 0.0f) return mks(0.0f);
    normalize_m(wm);
    spec F = fresnel_complex_s(absdot(wo, wm), b.ceta, b.ck);
    spec f_ss = ((F * ggx_D(b.mf, wm)) * ggx_G(b.mf, wo, wi)) / (4.0f * ci * co);
    return b.r * (f_ss + conductor_ms_lobe(b, wo, wi));
}
BB_D float conductor_ms_pdf(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b.mf)) return 0.0f;
    f3 wm = wo + wi;
    if (length2(wm) == 0.0f) return 0.0f;
    normalize_m(wm);
    if (dot_z(wm) < 0.0f) wm = -wm;
    const float pdf_ss = ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
    const float pdf_ms = wi.z * kInvPi; // CosineHemispherePDF(wi.z)
    const float e = ggx_E(b, wo);
    return pdf_ms * (1.0f - e) + pdf_ss * e; // Lerp(pdf_ms, pdf_ss, E(wo)), math.h:87-90
}

// Dielectric, dielectric_bxdf.cpp:9-129
BB_D spec dielectric_f(const bxdf& b, f3 wo, f3 wi)
{
    if (b.eta == 1.0f || ggx_smooth(b.mf)) return mks(0.0f);
    float co = wo.z, ci = wi.z;
    float eta_p = 1.0f;
    bool refl = ci * co > 0.0f;
    if (!refl) eta_p = co > 0.0f ? b.eta : (1.0f / b.eta);
    f3 wm = wi * eta_p + wo;
    if (ci == 0.0f || co == 0.0f || length2(wm) == 0.0f) return mks(0.0f);
    normalize_m(wm);
    if (dot_z(wm) < 0.0f) wm = -wm;
    if (dot(wm, wi) * ci < 0.0f || dot(wm, wo) * co < 0.0f) return mks(0.0f);
    float F = fresnel_dielectric(dot(wo, wm), b.eta);
    if (refl) return mks(F * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) / fabsf(4.0f * ci * co));
    float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta_p) * ci * co;
    float ft = (1.0f - F) * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / denom);
    ft /= sqr(eta_p);
    return b.r * ft;
}
BB_D float dielectric_pdf(const bxdf& b, f3 wo, f3 wi)
{
    if (b.eta == 1.0f || ggx_smooth(b.mf)) return 0.0f;
    float co = wo.z, ci = wi.z;
    float eta_p = 1.0f;
    bool refl = ci * co > 0.0f;
    if (!refl) eta_p = co > 0.0f ? b.eta : (1.0f / b.eta);
    f3 wm = wi * eta_p + wo;
    if (ci == 0.0f || co == 0.0f || length2(wm) == 0.0f) return 0.0f;
    normalize_m(wm);
    if (dot_z(wm) < 0.0f) wm = -wm;
    if (dot(wm, wi) * ci < 0.0f || dot(wm, wo) * co < 0.0f) return 0.0f;
    float R = fresnel_dielectric(dot(wo, wm), b.eta);
    float T = 1.0f - R;
    float pr = R, pt = T;
    if (pr == 0.0f && pt == 0.0f) return 0.0f;
    if (refl) return ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm)) * pr / (pr + pt);
    float dwm_dwi = absdot(wi, wm) / sqr(dot(wi, wm) + dot(wo, wm) / eta_p);
    return ggx_pdf(b.mf, wo, wm) * dwm_dwi * pt / (pr + pt);
}

// Metallic-roughness, metallic_roughness_bxdf.cpp:8-80, bxdfs.h:466-489
BB_D spec mr_F0(const bxdf& b) { return mks(0.04f) * (1.0f - b.metallic) + b.color * b.metallic; }
BB_D spec mr_schlick(spec f0, float c) { return f0 + (mks(1.0f) - f0) * dpow(1.0f - c, 5.0f); }
BB_D spec mr_f(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi)) return mks(0.0f);
    float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
    if (ci == 0.0f || co == 0.0f) return mks(0.0f);
    f3 wm = wo + wi;
    if (length2(wm) == 0.0f) return mks(0.0f);
    normalize_m(wm);
    spec F = mr_schlick(mr_F0(b), dot(wi, wm));
    spec f_s = ((F * ggx_D(b.mf, wm)) * ggx_G(b.mf, wo, wi)) / (4.0f * ci * co);
    spec f_d = ((mks(1.0f) - F) * (1.0f - b.metallic)) * (b.color * kInvPi);
    return f_d + f_s;
}
BB_D float mr_pdf(const bxdf& b, f3 wo, f3 wi)
{
    if (!same_hemisphere(wo, wi) || ggx_smooth(b.mf)) return 0.0f;
    f3 wm = wo + wi;
    if (length2(wm) == 0.0f) return 0.0f;
    normalize_m(wm);
    if (dot_z(wm) < 0.0f) wm = -wm;
    float R = fresnel_dielectric(dot(wo, wm), 1.5f);
    float T = 1.0f - R;
    float pr = R, pt = T * (1.0f - b.metallic);
    if (pr == 0.0f && pt == 0.0f) return 0.0f;
    float p_sum = pr + pt;
    pr /= p_sum;
    pt /= p_sum;
    float pdf_r = ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
    float pdf_d = abs_cos_theta(wi) * kInvPi;
    return pr * pdf_r + pt * pdf_d;
}

// Dielectric with multiple-scattering compensation,
// dielectric_multi_scattering_bxdf.cpp:9-433, bxdfs.h:145-247.  The host
// assigns this kind only to a rough interface with eta != 1, so the specular
// branches (which equal the plain dielectric's) are not needed here.

// ImageTexture3D<Float>::Evaluate, trilinear with TexCoordFilter::clamp
// (textures3d.h:36-88, texture3d.h:16-43), image(x, y, z) at
// [(z * n + y) * n + x] (image3d.h:54-62)
BB_D float tex_trilinear(const float* img, int32_t n, float u, float v, float w)
{
    const float fw = u * float(n) - 0.5f;
    const float fh = v * float(n) - 0.5f;
    const float fd = w * float(n) - 0.5f;
    int32_t i0 = int32_t(floorf(fw)), i1 = i0 + 1;
    int32_t j0 = int32_t(floorf(fh)), j1 = j0 + 1;
    int32_t k0 = int32_t(floorf(fd)), k1 = k0 + 1;
    i0 = min(max(i0, 0), n - 1);
    j0 = min(max(j0, 0), n - 1);
    k0 = min(max(k0, 0), n - 1);
    i1 = min(max(i1, 0), n - 1);
    j1 = min(max(j1, 0), n - 1);
    k1 = min(max(k1, 0), n - 1);
    const float fu = fw - floorf(fw);
    const float fv = fh - floorf(fh);
    const float fz = fd - floorf(fd);
    auto at = [&](int32_t i, int32_t j, int32_t k) { return img[(k * n + j) * n + i]; };
    const float v00 = (1.0f - fu) * at(i0, j0, k0) + fu * at(i1, j0, k0);
    const float v01 = (1.0f - fu) * at(i0, j0, k1) + fu * at(i1, j0, k1);
    const float v10 = (1.0f - fu) * at(i0, j1, k0) + fu * at(i1, j1, k0);
    const float v11 = (1.0f - fu) * at(i0, j1, k1) + fu * at(i1, j1, k1);
    const float v0 = (1.0f - fv) * v00 + fv * v10;
    const float v1 = (1.0f - fv) * v01 + fv * v11;
    return (1.0f - fz) * v0 + fz * v1;
}

// MapIORtoF0 (scattering.h:205-208)
BB_D float map_ior_to_f0(float ior) { return sqrtf((ior - 1.0f) / (ior + 1.0f)); }

// E(wo, eta) / E_avg(eta) (bxdfs.h:180-212): tables of the interface seen
// from outside (eta >= 1) or from inside, at GetMeanAlpha
BB_D float dms_E(const bxdf& b, f3 w, float eta)
{
    constexpr int32_t n = kTableSize;
    const float a = sqrtf(b.mf.ax * b.mf.ay);
    if (eta >= 1.0f) return tex_trilinear(b.dt, n, map_ior_to_f0(eta), fabsf(w.z), a);
    return tex_trilinear(b.dt + n * n * n, n, map_ior_to_f0(1.0f / eta), fabsf(w.z), a);
}
BB_D float dms_E_avg(const bxdf& b, float eta)
{
    constexpr int32_t n = kTableSize;
    const float a = sqrtf(b.mf.ax * b.mf.ay);
    const float* avg = b.dt + 2 * n * n * n;
    if (eta >= 1.0f) return tex_bilinear(avg, n, n, map_ior_to_f0(eta), a);
    return tex_bilinear(avg + n * n, n, n, map_ior_to_f0(1.0f / eta), a);
}

// FresnelDielectricAverage (scattering.h:154-172), eta_p = eta_i / eta_t
BB_D float fresnel_dielectric_average(float eta_p)
{
    if (eta_p < 1.0f) {
        const float eta2 = eta_p * eta_p;
        const float eta3 = eta_p * eta2;
        return 0.997118f + 0.1014f * eta_p - 0.965241f * eta2 - 0.130607f * eta3;
    }
    return (eta_p - 1.0f) / (4.08567f + 1.00071f * eta_p);
}

// ComputeScatteringRatio (bxdfs.h:222-242): the share of the
// multiple-scattering energy that leaves on wo's side
BB_D float dms_ratio(const bxdf& b, float eta_o)
{
    const float eta_i = b.eta >= 1.0f ? b.eta : 1.0f / b.eta;
    const float eta_t = 1.0f / eta_i;
    const float a = (1.0f - fresnel_dielectric_average(eta_i)) / (1.0f - dms_E_avg(b, eta_t));
    const float c = sqr(eta_i) * (1.0f - fresnel_dielectric_average(eta_t)) / (1.0f - dms_E_avg(b, eta_i));
    const float x = 1.0f / (a / c + 1.0f);
    if (eta_o >= 1.0f) return 1.0f - x * (1.0f - fresnel_dielectric_average(eta_o));
    return 1.0f - (1.0f - x) * (1.0f - fresnel_dielectric_average(eta_o));
}

// the multiple-scattering lobes of reflection and transmission
BB_D float dms_fr_ms(const bxdf& b, f3 wo, f3 wi, float ratio, float eta_o)
{
    return ratio * (1.0f - dms_E(b, wi, eta_o)) * (1.0f - dms_E(b, wo, eta_o)) /
           fmaxf(1e-4f, kPi * (1.0f - dms_E_avg(b, eta_o)));
}
BB_D float dms_ft_ms(const bxdf& b, f3 wo, f3 wi, float ratio, float eta_o, float eta_i)
{
    return (1.0f - ratio) * (1.0f - dms_E(b, wi, eta_i)) * (1.0f - dms_E(b, wo, eta_o)) /
           fmaxf(1e-4f, kPi * (1.0f - dms_E_avg(b, eta_i)));
}

// f, TransportDirection::ToLight (dielectric_multi_scattering_bxdf.cpp:9-89)
BB_D spec dielectric_ms_f(const bxdf& b, f3 wo, f3 wi)
{
    const float co = wo.z, ci = wi.z;
    float eta_p = 1.0f;
    const bool refl = ci * co > 0.0f;
    if (!refl) eta_p = co > 0.0f ? b.eta : (1.0f / b.eta);
    const float eta_o = co > 0.0f ? b.eta : 1.0f / b.eta;
    const float eta_i = co > 0.0f ? 1.0f / b.eta : b.eta;
    f3 wm = wi * eta_p + wo;
    if (ci == 0.0f || co == 0.0f || length2(wm) == 0.0f) return mks(0.0f);
    normalize_m(wm);
    if (dot_z(wm) < 0.0f) wm = -wm;
    if (dot(wm, wi) * ci < 0.0f || dot(wm, wo) * co < 0.0f) return mks(0.0f);
    const float ratio = dms_ratio(b, eta_o);
    const float F = fresnel_dielectric(dot(wo, wm), b.eta);
    if (refl) {
        const float fr_ss = F * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) / (4.0f * ci * co);
        return mks(fr_ss + dms_fr_ms(b, wo, wi, ratio, eta_o));
    }
    const float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta_p);
    float ft_ss = (1.0f - F) * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / (ci * co * denom));
    const float ft_ms = dms_ft_ms(b, wo, wi, ratio, eta_o, eta_i);
    ft_ss /= sqr(eta_p);
    return b.r * (ft_ss + ft_ms);
}

// PDF, all sampling flags (dielectric_multi_scattering_bxdf.cpp:91-187)
BB_D float dielectric_ms_pdf(const bxdf& b, f3 wo, f3 wi)
{
    const float co = wo.z, ci = wi.z;
    float eta_p = 1.0f;
    const bool refl = ci * co > 0.0f;
    if (!refl) eta_p = co > 0.0f ? b.eta : (1.0f / b.eta);
    const float eta_o = co > 0.0f ? b.eta : 1.0f / b.eta;
    f3 wm = wi * eta_p + wo;
    if (ci == 0.0f || co == 0.0f || length2(wm) == 0.0f) return 0.0f;
    normalize_m(wm);
    if (dot_z(wm) < 0.0f) wm = -wm;
    if (dot(wm, wi) * ci < 0.0f || dot(wm, wo) * co < 0.0f) return 0.0f;
    const float ratio = dms_ratio(b, eta_o);
    const float E_o = clampf_(dms_E(b, wo, eta_o), 1e-4f, 1.0f - 1e-4f);
    const float R = fresnel_dielectric(dot(wo, wm), b.eta);
    const float T = 1.0f - R;
    float pr_ss = R, pt_ss = T, pr_ms = ratio, pt_ms = 1.0f - ratio;
    if (pr_ss == 0.0f && pt_ss == 0.0f && pr_ms == 0.0f && pt_ms == 0.0f) return 0.0f;
    const float p_ss_sum = pr_ss + pt_ss, p_ms_sum = pr_ms + pt_ms;
    pr_ss /= p_ss_sum;
    pt_ss /= p_ss_sum;
    pr_ms /= p_ms_sum;
    pt_ms /= p_ms_sum;
    float pdf_ss, pdf_ms;
    if (refl) {
        pdf_ss = pr_ss * ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
        pdf_ms = pr_ms * (abs_cos_theta(wi) * kInvPi);
    } else {
        const float dwm_dwi = absdot(wi, wm) / sqr(dot(wi, wm) + dot(wo, wm) / eta_p);
        pdf_ss = pt_ss * ggx_pdf(b.mf, wo, wm) * dwm_dwi;
        pdf_ms = pt_ms * (abs_cos_theta(wi) * kInvPi);
    }
    return pdf_ms * (1.0f - E_o) + pdf_ss * E_o; // Lerp(pdf_ms, pdf_ss, E_o), math.h:87-90
}

// Sample_f, rough branch, ToLight, all sampling flags
// (dielectric_multi_scattering_bxdf.cpp:246-433)
BB_D bool dielectric_ms_sample(const bxdf& b, f3 wo, float u0, f2 u12, bsample& s)
{
    const float eta_o = wo.z > 0.0f ? b.eta : 1.0f / b.eta;
    const float eta_i = wo.z > 0.0f ? 1.0f / b.eta : b.eta;
    const float ratio = dms_ratio(b, eta_o);
    const float E_o = clampf_(dms_E(b, wo, eta_o), 1e-4f, 1.0f - 1e-4f);
    float pr_ms = ratio, pt_ms = 1.0f - ratio;
    if (u0 < E_o) {
        // single scattering at the rough interface
        f3 wm = ggx_sample_wm(b.mf, wo, u12);
        normalize_m(wm);
        if (dot_z(wm) < 0.0f) wm = -wm;
        const float R = fresnel_dielectric(dot(wo, wm), b.eta);
        const float T = 1.0f - R;
        float pr_ss = R, pt_ss = T;
        if (pr_ss == 0.0f && pt_ss == 0.0f) return false;
        const float p_ss_sum = pr_ss + pt_ss, p_ms_sum = pr_ms + pt_ms;
        pr_ss /= p_ss_sum;
        pt_ss /= p_ss_sum;
        pr_ms /= p_ms_sum;
        pt_ms /= p_ms_sum;
        u0 /= E_o;
        if (u0 < pr_ss) {
            const f3 wi = reflect(wo, wm);
            if (!same_hemisphere(wo, wi)) return false;
            const float pdf_ss = pr_ss * ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
            const float pdf_ms = pr_ms * (abs_cos_theta(wi) * kInvPi);
            const float fr_ss = R * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) / (4.0f * wi.z * wo.z);
            s.f = mks(fr_ss + dms_fr_ms(b, wo, wi, ratio, eta_o));
            s.wi = wi;
            s.pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
            s.flags = BX_GLOSSY | BX_REFLECTION;
            return true;
        }
        f3 wi = mk3(0.0f);
        float eta_unused = 0.0f;
        const bool tir = !refract(wi, wo, wm, b.eta, eta_unused);
        if (same_hemisphere(wo, wi) || wi.z == 0.0f || tir) return false;
        const float eta_p = wo.z > 0.0f ? b.eta : (1.0f / b.eta);
        const float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta_p);
        const float dwm_dwi = absdot(wi, wm) / denom;
        const float pdf_ss = pt_ss * ggx_pdf(b.mf, wo, wm) * dwm_dwi;
        const float pdf_ms = pt_ms * (abs_cos_theta(wi) * kInvPi);
        float ft_ss = T * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom));
        const float ft_ms = dms_ft_ms(b, wo, wi, ratio, eta_o, eta_i);
        ft_ss /= sqr(eta_p);
        s.f = b.r * (ft_ss + ft_ms);
        s.wi = wi;
        s.pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
        s.flags = BX_GLOSSY | BX_TRANSMISSION;
        s.eta = eta_p;
        return true;
    }
    // the multiple-scattering lobe: cosine-distributed reflection or transmission
    if (pr_ms == 0.0f && pt_ms == 0.0f) return false;
    const float p_ms_sum = pr_ms + pt_ms;
    pr_ms /= p_ms_sum;
    pt_ms /= p_ms_sum;
    u0 = (u0 - E_o) / (1.0f - E_o);
    f3 wi = sample_cosine_hemisphere(u12);
    if (u0 < pr_ms) {
        if (wo.z < 0.0f) wi = -wi;
        const f3 wm = normalize(wi + wo);
        const float R = fresnel_dielectric(dot(wo, wm), b.eta);
        float pr_ss = R, pt_ss = 1.0f - R;
        const float p_ss_sum = pr_ss + pt_ss;
        pr_ss /= p_ss_sum;
        pt_ss /= p_ss_sum;
        const float pdf_ss = pr_ss * ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
        const float pdf_ms = pr_ms * (abs_cos_theta(wi) * kInvPi);
        const float fr_ss = R * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) / (4.0f * wi.z * wo.z);
        s.f = mks(fr_ss + dms_fr_ms(b, wo, wi, ratio, eta_o));
        s.wi = wi;
        s.pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
        s.flags = BX_DIFFUSE | BX_REFLECTION;
        return true;
    }
    if (wo.z > 0.0f) wi = -wi;
    f3 wm = normalize(wi * eta_o + wo);
    if (dot_z(wm) < 0.0f) wm = -wm;
    const float R = fresnel_dielectric(dot(wo, wm), b.eta);
    const float T = 1.0f - R;
    float pr_ss = R, pt_ss = 1.0f - R;
    const float p_ss_sum = pr_ss + pt_ss;
    pr_ss /= p_ss_sum;
    pt_ss /= p_ss_sum;
    const float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta_o);
    const float dwm_dwi = absdot(wi, wm) / denom;
    const float pdf_ss = pt_ss * ggx_pdf(b.mf, wo, wm) * dwm_dwi;
    const float pdf_ms = pt_ms * (abs_cos_theta(wi) * kInvPi);
    float ft_ss = T * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom));
    const float ft_ms = dms_ft_ms(b, wo, wi, ratio, eta_o, eta_i);
    ft_ss /= sqr(eta_o);
    s.f = b.r * (ft_ss + ft_ms);
    s.wi = wi;
    s.pdf = pdf_ms * (1.0f - E_o) + pdf_ss * E_o;
    s.flags = BX_DIFFUSE | BX_TRANSMISSION;
    s.eta = eta_o;
    return true;
}

// BSDF::f / BSDF::PDF on local directions (bsdf.h:75-99)
// Kind masks: a shading kernel built for the BxDF kinds of bit mask KM only
// (the kinds a scene's materials use) leaves the other BxDFs' code out and
// needs fewer registers.  kAllKinds keeps every BxDF.
constexpr int kAllKinds = -1;
template <int KM, int K>
constexpr bool kind_on()
{
    return KM == kAllKinds || ((KM >> K) & 1) != 0;
}
// a kind outside the mask never reaches here (the host picks the kernel from
// the scene's materials)
#define BB_KIND_CASE(K, ...)                     \
    case K:                                      \
        if constexpr (kind_on<KM, K>()) {        \
            __VA_ARGS__;                         \
        } else {                                 \
            __builtin_unreachable();             \
        }

template <int KM = kAllKinds>
BB_D spec bxdf_f(const bxdf& b, f3 wo, f3 wi)
{
    switch (b.kind) {
        BB_KIND_CASE(MK_LAMBERT, return lambert_f(b, wo, wi))
        BB_KIND_CASE(MK_EON, return eon_f(b, wo, wi))
        BB_KIND_CASE(MK_DIELECTRIC, return dielectric_f(b, wo, wi))
        BB_KIND_CASE(MK_DIELECTRIC_MS, return dielectric_ms_f(b, wo, wi))
        BB_KIND_CASE(MK_CONDUCTOR, return conductor_f(b, wo, wi))
        BB_KIND_CASE(MK_CONDUCTOR_MS, return conductor_ms_f(b, wo, wi))
        BB_KIND_CASE(MK_MIRROR, return mks(0.0f)) // specular_reflection_bxdf.cpp:7-13
    default:
        if constexpr (kind_on<KM, MK_METALLIC_ROUGHNESS>())
            return mr_f(b, wo, wi);
        else
            __builtin_unreachable();
    }
}
template <int KM = kAllKinds>
BB_D float bxdf_pdf(const bxdf& b, f3 wo, f3 wi)
{
    switch (b.kind) {
        BB_KIND_CASE(MK_LAMBERT, return lambert_pdf(wo, wi))
        BB_KIND_CASE(MK_EON, return eon_pdf(b, wo, wi))
        BB_KIND_CASE(MK_DIELECTRIC, return dielectric_pdf(b, wo, wi))
        BB_KIND_CASE(MK_DIELECTRIC_MS, return dielectric_ms_pdf(b, wo, wi))
        BB_KIND_CASE(MK_CONDUCTOR, return conductor_pdf(b, wo, wi))
        BB_KIND_CASE(MK_CONDUCTOR_MS, return conductor_ms_pdf(b, wo, wi))
        BB_KIND_CASE(MK_MIRROR, return 0.0f)
    default:
        if constexpr (kind_on<KM, MK_METALLIC_ROUGHNESS>())
            return mr_pdf(b, wo, wi);
        else
            __builtin_unreachable();
    }
}

// BxDF::Sample_f on local wo (returns false for a failed sample)
template <int KM = kAllKinds>
BB_D bool bxdf_sample(const bxdf& b, f3 wo, float u0, f2 u12, bsample& s)
{
    s.eta = 1.0f;
    switch (b.kind) {
    case MK_LAMBERT: {
        if constexpr (!kind_on<KM, MK_LAMBERT>()) __builtin_unreachable();
        f3 wi = sample_cosine_hemisphere(u12);
        float pdf = wi.z * kInvPi;
        if (wo.z < 0.0f) wi.z = -wi.z;
        s.f = b.r * kInvPi;
        s.wi = wi;
        s.pdf = pdf;
        s.flags = BX_DIFFUSE | BX_REFLECTION;
        return true;
    }
    case MK_EON:
        if constexpr (!kind_on<KM, MK_EON>()) __builtin_unreachable();
        return eon_sample(b, wo, u0, u12, s);
    case MK_MIRROR: {
        if constexpr (!kind_on<KM, MK_MIRROR>()) __builtin_unreachable();
        // specular_reflection_bxdf.cpp:24-38: wo mirrored about the normal, weight r / |cos|
        const f3 wi = mk3(-wo.x, -wo.y, wo.z);
        s.f = b.r / fabsf(wi.z);
        s.wi = wi;
        s.pdf = 1.0f;
        s.flags = BX_SPECULAR | BX_REFLECTION;
        return true;
    }
    case MK_DIELECTRIC_MS:
        if constexpr (!kind_on<KM, MK_DIELECTRIC_MS>()) __builtin_unreachable();
        return dielectric_ms_sample(b, wo, u0, u12, s);
    case MK_CONDUCTOR_MS:
        if constexpr (!kind_on<KM, MK_CONDUCTOR_MS>()) __builtin_unreachable();
        if (!ggx_smooth(b.mf)) {
            if (wo.z == 0.0f) return false;
            f3 wm, wi;
            int32_t flags;
            const float e = ggx_E(b, wo);
            if (u0 < e) {
                wm = ggx_sample_wm(b.mf, wo, u12);
                wi = reflect(wo, wm);
                flags = BX_GLOSSY | BX_REFLECTION;
            } else {
                wi = sample_cosine_hemisphere(u12);
                wm = normalize(wo + wi);
                flags = BX_DIFFUSE | BX_REFLECTION;
            }
            if (!same_hemisphere(wo, wi)) return false;
            const float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
            if (ci == 0.0f || co == 0.0f) return false;
            const spec F = fresnel_complex_s(absdot(wo, wm), b.ceta, b.ck);
            const spec f_ss = ((F * ggx_D(b.mf, wm)) * ggx_G(b.mf, wo, wi)) / (4.0f * ci * co);
            const spec f_ms = conductor_ms_lobe(b, wo, wi);
            const float pdf_ss = ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
            const float pdf_ms = wi.z * kInvPi;
            s.f = b.r * (f_ss + f_ms);
            s.wi = wi;
            s.pdf = pdf_ms * (1.0f - e) + pdf_ss * e;
            s.flags = flags;
            return true;
        }
        [[fallthrough]]; // effectively smooth: the perfect specular conductor
    case MK_CONDUCTOR: {
        if constexpr (!kind_on<KM, MK_CONDUCTOR>() && !kind_on<KM, MK_CONDUCTOR_MS>()) __builtin_unreachable(); // also the smooth compensated conductor
        if (ggx_smooth(b.mf)) {
            f3 wi = mk3(-wo.x, -wo.y, wo.z);
            spec f = fresnel_complex_s(abs_cos_theta(wi), b.ceta, b.ck) / abs_cos_theta(wi);
            s.f = b.r * f;
            s.wi = wi;
            s.pdf = 1.0f;
            s.flags = BX_SPECULAR | BX_REFLECTION;
            return true;
        }
        if (wo.z == 0.0f) return false;
        f3 wm = ggx_sample_wm(b.mf, wo, u12);
        f3 wi = reflect(wo, wm);
        if (!same_hemisphere(wo, wi)) return false;
        float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
        if (ci == 0.0f || co == 0.0f) return false;
        float pdf = ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
        spec F = fresnel_complex_s(absdot(wo, wm), b.ceta, b.ck);
        spec f = ((F * ggx_D(b.mf, wm)) * ggx_G(b.mf, wo, wi)) / (4.0f * ci * co);
        s.f = b.r * f;
        s.wi = wi;
        s.pdf = pdf;
        s.flags = BX_GLOSSY | BX_REFLECTION;
        return true;
    }
    case MK_DIELECTRIC: {
        if constexpr (!kind_on<KM, MK_DIELECTRIC>()) __builtin_unreachable();
        if (b.eta == 1.0f || ggx_smooth(b.mf)) {
            float R = fresnel_dielectric(wo.z, b.eta);
            float T = 1.0f - R;
            float pr = R, pt = T;
            if (pr == 0.0f && pt == 0.0f) return false;
            if (u0 < pr / (pr + pt)) {
                f3 wi = mk3(-wo.x, -wo.y, wo.z);
                s.f = mks(R / abs_cos_theta(wi));
                s.wi = wi;
                s.pdf = pr / (pr + pt);
                s.flags = BX_SPECULAR | BX_REFLECTION;
            } else {
                f3 wi;
                float eta_p;
                if (!refract(wi, wo, mk3(0.0f, 0.0f, 1.0f), b.eta, eta_p)) return false;
                float ft = T / abs_cos_theta(wi);
                ft /= sqr(eta_p);
                s.f = b.r * ft;
                s.wi = wi;
                s.pdf = pt / (pr + pt);
                s.flags = BX_SPECULAR | BX_TRANSMISSION;
                s.eta = eta_p;
            }
            return true;
        }
        f3 wm = ggx_sample_wm(b.mf, wo, u12);
        float R = fresnel_dielectric(dot(wo, wm), b.eta);
        float T = 1.0f - R;
        float pr = R, pt = T;
        if (pr == 0.0f && pt == 0.0f) return false;
        if (u0 < pr / (pr + pt)) {
            f3 wi = reflect(wo, wm);
            if (!same_hemisphere(wo, wi)) return false;
            float pdf = ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm)) * pr / (pr + pt);
            s.f = mks(R * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) / (4.0f * wi.z * wo.z));
            s.wi = wi;
            s.pdf = pdf;
            s.flags = BX_GLOSSY | BX_REFLECTION;
        } else {
            float eta_p = 0.0f;
            f3 wi = mk3(0.0f);
            bool tir = !refract(wi, wo, wm, b.eta, eta_p);
            if (same_hemisphere(wo, wi) || wi.z == 0.0f || tir) return false;
            float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta_p);
            float dwm_dwi = absdot(wi, wm) / denom;
            float pdf = ggx_pdf(b.mf, wo, wm) * dwm_dwi * pt / (pr + pt);
            float ft = T * ggx_D(b.mf, wm) * ggx_G(b.mf, wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom));
            ft /= sqr(eta_p);
            s.f = b.r * ft;
            s.wi = wi;
            s.pdf = pdf;
            s.flags = BX_GLOSSY | BX_TRANSMISSION;
            s.eta = eta_p;
        }
        return true;
    }
    default: {
        if constexpr (!kind_on<KM, MK_METALLIC_ROUGHNESS>()) __builtin_unreachable();
        if (wo.z == 0.0f) return false;
        f3 wm = ggx_sample_wm(b.mf, wo, u12);
        f3 wi = reflect(wo, wm);
        float R = fresnel_dielectric(dot(wo, wm), 1.5f);
        float T = 1.0f - R;
        float pr = R, pt = T * (1.0f - b.metallic);
        if (pr == 0.0f && pt == 0.0f) return false;
        float p_sum = pr + pt;
        pr /= p_sum;
        pt /= p_sum;
        int32_t flag;
        if (u0 < pr) {
            if (!same_hemisphere(wo, wi)) return false;
            flag = BX_GLOSSY | BX_REFLECTION;
        } else {
            wi = sample_cosine_hemisphere(u12);
            wm = normalize(wi + wo);
            flag = BX_DIFFUSE | BX_REFLECTION;
        }
        float co = abs_cos_theta(wo), ci = abs_cos_theta(wi);
        if (ci == 0.0f || co == 0.0f) return false;
        spec F = mr_schlick(mr_F0(b), dot(wi, wm));
        spec f_r = ((F * ggx_D(b.mf, wm)) * ggx_G(b.mf, wo, wi)) / (4.0f * ci * co);
        spec f_d = ((mks(1.0f) - F) * (1.0f - b.metallic)) * (b.color * kInvPi);
        float pdf_r = ggx_pdf(b.mf, wo, wm) / (4.0f * absdot(wo, wm));
        float pdf_d = ci * kInvPi;
        s.f = f_r + f_d;
        s.wi = wi;
        s.pdf = pr * pdf_r + pt * pdf_d;
        s.flags = flag;
        return true;
    }
    }
}

// ------------------------------------------------------------------- lights

// Distribution1D::SampleDiscrete + FindInterval (sampling.h:383-397, math.h:117-135)
// cdf: the light CDF (S.light_cdf, or a copy of it in LDS)
BB_D int32_t light_select(const SceneView& S, float u, float& pmf, const float* cdf)
{
    int32_t size = S.n_lights + 1;
    int32_t first = 0, len = size;
    while (len > 0) {
        int32_t half = len >> 1, middle = first + half;
        if (cdf[middle] <= u) {
            first = middle + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    int32_t offset = first - 1;
    if (offset < 0) offset = 0;
    if (offset > size - 2) offset = size - 2;
    pmf = S.light_func[offset] / (S.func_integral * float(S.n_lights));
    return offset;
}

// Opt-in alias-table selection over the same light powers (no counterpart
// in the reference, whose PowerLightSampler searches the CDF): i from the
// integer part of u n, kept when the fractional part is below its keep
// probability, else its alias.  The returned pmf is the power sampler's, so
// the estimator is the same and only the draw-to-light map differs.
BB_D int32_t light_select_alias(const SceneView& S, float u, float& pmf)
{
    const int32_t n = S.n_lights;
    const float x = u * float(n);
    int32_t i = int32_t(x);
    if (i > n - 1) i = n - 1;
    if (i < 0) i = 0;
    const float2 e = S.light_alias[i];
    if (!(x - float(i) < e.x)) i = __float_as_int(e.y);
    pmf = S.light_func[i] / (S.func_integral * float(n));
    return i;
}

struct light_sample
{
    f3 wi;
    float visibility;
    spec Li;
    float pdf;
};

// Light::Sample_Li for spot, point, uniform infinite and diffuse area (triangle) lights
BB_D bool light_sample_li(const GpuLight& L, f3 ref, f2 u, light_sample& ls)
{
    if (L.type == BB_LIGHT_DIRECTIONAL) {
        // directional_light.cpp:29-38
        ls.wi = L.z;
        ls.visibility = 2.0f * L.world_radius;
        ls.pdf = 1.0f;
        ls.Li = L.L;
        return true;
    }
    if (L.type == BB_LIGHT_POINT) {
        // point_light.cpp:30-42
        f3 wi = L.p - ref;
        const float distance = normalize_m(wi);
        ls.wi = wi;
        ls.visibility = distance;
        ls.pdf = 1.0f;
        ls.Li = L.L / (distance * distance);
        return true;
    }
    if (L.type == BB_LIGHT_SPOT) {
        // spot_light.cpp:46-68
        f3 wi = L.p - ref;
        float distance = normalize_m(wi);
        float cos_t = -dot(wi, L.z);
        spec l = L.L * smooth_step(L.cos_max, L.cos_min, cos_t);
        if (is_black(l)) return false;
        ls.wi = wi;
        ls.visibility = distance;
        ls.pdf = 1.0f;
        ls.Li = l / (distance * distance);
        return true;
    }
    if (L.type == BB_LIGHT_UNIFORM_INFINITE) {
        // uniform_infinite_light.cpp:30-43
        ls.wi = sample_uniform_sphere(u);
        ls.pdf = kInvFourPi;
        ls.visibility = 2.0f * L.world_radius;
        ls.Li = L.L;
        return true;
    }
    // diffuse_area_light.cpp:42-72 with Triangle::Sample(ref, u), triangle.cpp:159-229
    float uu = u.x, vv = u.y;
    if (uu + vv > 1.0f) {
        uu = 1.0f - uu;
        vv = 1.0f - vv;
    }
    f3 point = (L.p + L.e1 * uu) + L.e2 * vv;
    float pdf = 1.0f / L.area;
    f3 d = point - ref;
    float dist2 = dot(d, d);
    float cosine = dot(d, L.n) / sqrtf(dist2);
    if (cosine < 0.0f) cosine = -cosine;
    pdf *= dist2 / cosine;
    f3 wi = point - ref;
    ls.visibility = normalize_m(wi) - kRayEps;
    ls.wi = wi;
    ls.pdf = pdf;
    bool front = dot(L.n, wi) < 0.0f;
    ls.Li = (front || L.two_sided) ? L.L : mks(0.0f);
    return true;
}

} // namespace bb
