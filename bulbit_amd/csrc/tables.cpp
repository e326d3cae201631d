// tables.cpp — host computation of the reflectance tables that Bulbit's
// Integrator prepares before rendering (PrepareReflectanceTextures,
// src/bxdf/microfacet.cpp:239-270, called with texture_size 32 and 1024
// samples by src/integrator/integrator.cpp:149-151).
//
// The tables are computed once per process on the host, with the float
// arithmetic and the C library's sinf/cosf that the reference uses.  Two sets
// are needed on this path: the Trowbridge-Reitz albedo tables
// (ConductorMultiScatteringBxDF), uploaded with every scene, and the rough
// dielectric's albedo tables (DielectricMultiScatteringBxDF), computed on
// first use and uploaded with scenes that have a compensated rough dielectric.
#include "bb_scene.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <thread>

namespace bb
{

namespace
{

// PermutationElement (hash.h:128-159): Kensler's permutation of [0, l)
int32_t permutation_element(uint32_t i, uint32_t l, uint32_t p)
{
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return int32_t((i + p) % l);
}

// Trowbridge-Reitz pieces on the host (include/bulbit/microfacet.h:37-86,
// sampling.h:203-232), isotropic alpha
float tr_D(float a, f3 wm)
{
    const float tan2 = tan2_theta(wm);
    if (std::isinf(tan2)) return 0.0f;
    const float cos4 = sqr(cos2_theta(wm));
    if (cos4 < 1e-16f) return 0.0f;
    const float e = tan2 * (sqr(cos_phi(wm) / a) + sqr(sin_phi(wm) / a));
    return 1.0f / (kPi * a * a * cos4 * sqr(1.0f + e));
}

float tr_lambda(float a, f3 w)
{
    const float tan2 = tan2_theta(w);
    if (std::isinf(tan2)) return 0.0f;
    const float alpha2 = sqr(cos_phi(w) * a) + sqr(sin_phi(w) * a);
    return (std::sqrt(1.0f + alpha2 * tan2) - 1.0f) / 2.0f;
}

float tr_G1(float a, f3 w) { return 1.0f / (1.0f + tr_lambda(a, w)); }
float tr_G(float a, f3 wo, f3 wi) { return 1.0f / (1.0f + tr_lambda(a, wo) + tr_lambda(a, wi)); }
float tr_pdf(float a, f3 w, f3 wm) { return tr_G1(a, w) / abs_cos_theta(w) * tr_D(a, wm) * absdot(w, wm); }

f3 tr_sample_wm(float a, f3 wo, f2 u)
{
    const f3 wo_std = normalize(mk3(wo.x * a, wo.y * a, wo.z));
    const float phi = kTwoPi * u.x;
    const float z = std::fma((1.0f - u.y), (1.0f + wo_std.z), -wo_std.z);
    const float sin_t = std::sqrt(clampf_(1.0f - z * z, 0.0f, 1.0f));
    const float x = sin_t * std::cos(phi);
    const float y = sin_t * std::sin(phi);
    const f3 h = mk3(x, y, z) + wo_std;
    return normalize(mk3(h.x * a, h.y * a, h.z));
}

// TrowbridgeReitzDistribution::rho (microfacet.h:109-131)
float tr_rho(float a, f3 wo, const f2* u, int32_t n)
{
    float r = 0.0f;
    for (int32_t i = 0; i < n; ++i) {
        const f3 wm = tr_sample_wm(a, wo, u[i]);
        const f3 wi = reflect(wo, wm);
        if (!same_hemisphere(wo, wi)) continue;
        const float pdf = tr_pdf(a, wo, wm) / (4.0f * absdot(wo, wm));
        const float f = tr_D(a, wm) * tr_G(a, wo, wi) / (4.0f * abs_cos_theta(wo) * abs_cos_theta(wi));
        r += f * abs_cos_theta(wi) / pdf;
    }
    return r / float(n);
}

// FresnelDielectric (scattering.h:51-77)
float fr_dielectric(float cos_theta_i, float eta)
{
    cos_theta_i = clampf_(cos_theta_i, -1.0f, 1.0f);
    if (cos_theta_i < 0.0f) {
        eta = 1.0f / eta;
        cos_theta_i = -cos_theta_i;
    }
    const float sin2_theta_i = 1.0f - sqr(cos_theta_i);
    const float sin2_theta_t = sin2_theta_i / sqr(eta);
    if (sin2_theta_t >= 1.0f) return 1.0f;
    const float cos_theta_t = std::sqrt(smax(0.0f, 1.0f - sin2_theta_t));
    const float r_parl = (eta * cos_theta_i - cos_theta_t) / (eta * cos_theta_i + cos_theta_t);
    const float r_perp = (cos_theta_i - eta * cos_theta_t) / (cos_theta_i + eta * cos_theta_t);
    return (sqr(r_parl) + sqr(r_perp)) / 2.0f;
}

// Refract (scattering.h:9-39)
bool refract_dir(f3& wt, f3 wi, f3 n, float eta, float& eta_p)
{
    float cos_theta_i = dot(n, wi);
    if (cos_theta_i < 0.0f) {
        eta = 1.0f / eta;
        cos_theta_i = -cos_theta_i;
        n = -n;
    }
    const float sin2_theta_i = smax(0.0f, 1.0f - sqr(cos_theta_i));
    const float sin2_theta_t = sin2_theta_i / sqr(eta);
    if (sin2_theta_t >= 1.0f) return false;
    const float cos_theta_t = std::sqrt(1.0f - sin2_theta_t);
    const float inv_eta = 1.0f / eta;
    wt = (-wi) * inv_eta + n * (cos_theta_i * inv_eta - cos_theta_t);
    eta_p = eta;
    return true;
}

// DielectricBxDF::Sample_f (dielectric_bxdf.cpp:140-254), rough branch, with
// TransportDirection::ToCamera (no 1 / eta'^2 on transmission) and reflectance
// 1; returns f |cos theta_i| / pdf, or 0 for a failed sample (BxDF::rho adds
// nothing then)
float dielectric_weight(float eta, float a, f3 wo, float u0, f2 u12)
{
    const f3 wm = tr_sample_wm(a, wo, u12);
    const float R = fr_dielectric(dot(wo, wm), eta);
    const float T = 1.0f - R;
    const float pr = R, pt = T;
    if (pr == 0.0f && pt == 0.0f) return 0.0f;
    if (u0 < pr / (pr + pt)) {
        const f3 wi = reflect(wo, wm);
        if (!same_hemisphere(wo, wi)) return 0.0f;
        const float pdf = tr_pdf(a, wo, wm) / (4.0f * absdot(wo, wm)) * pr / (pr + pt);
        const float f = R * tr_D(a, wm) * tr_G(a, wo, wi) / (4.0f * wi.z * wo.z);
        return f * abs_cos_theta(wi) / pdf;
    }
    float eta_p = 0.0f;
    f3 wi = mk3(0.0f);
    const bool tir = !refract_dir(wi, wo, wm, eta, eta_p);
    if (same_hemisphere(wo, wi) || wi.z == 0.0f || tir) return 0.0f;
    const float denom = sqr(dot(wi, wm) + dot(wo, wm) / eta_p);
    const float dwm_dwi = absdot(wi, wm) / denom;
    const float pdf = tr_pdf(a, wo, wm) * dwm_dwi * pt / (pr + pt);
    const float ft = T * tr_D(a, wm) * tr_G(a, wo, wi) * std::fabs(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom));
    return ft * abs_cos_theta(wi) / pdf;
}

// BxDF::rho (bxdf.cpp:8-21) and Spectrum::Average (spectrum.h:59-71) of its
// three equal channels
float dielectric_rho(float eta, float a, f3 wo, const float* u0, const f2* u12, int32_t n)
{
    float r = 0.0f;
    for (int32_t i = 0; i < n; ++i) r += dielectric_weight(eta, a, wo, u0[i], u12[i]);
    r = r / float(n);
    return (r + r + r) / 3.0f;
}

// MapF0toIOR (scattering.h:188-203)
float map_f0_to_ior(float f)
{
    const float sqrt_f0 = std::sqrt(sqr(sqr(f)));
    return (1.0f + sqrt_f0) / (1.0f - sqrt_f0);
}

// the stratified samples of PrepareReflectanceTextures (microfacet.cpp:239-270)
struct TableSamples
{
    std::vector<float> u0;
    std::vector<f2> u12;
};

TableSamples table_samples()
{
    const int32_t samples = kTableSamples, xs = 32, ys = 32;
    TableSamples t;
    t.u0.resize(samples);
    t.u12.resize(samples);
    uint64_t rng = pcg_seed(3); // RNG rng(3): initstate 3, initseq 1
    for (int32_t i = 0; i < samples; ++i) {
        int32_t stratum = permutation_element(uint32_t(i), uint32_t(samples), 1u);
        t.u0[i] = (float(stratum) + pcg_float(rng)) / float(samples);
        stratum = permutation_element(uint32_t(i), uint32_t(samples), 2u);
        const int32_t x = stratum % xs, y = stratum / xs;
        const float ux = (float(x) + pcg_float(rng)) / float(xs);
        const float uy = (float(y) + pcg_float(rng)) / float(ys);
        t.u12[i] = mk2(ux, uy);
    }
    return t;
}

ReflectanceTables g_tables;
std::once_flag g_tables_once;
DielectricTables g_dielectric;
std::once_flag g_dielectric_once;

void build_tables()
{
    ReflectanceTables& t = g_tables;
    const int32_t samples = kTableSamples;
    const std::vector<f2> u12 = table_samples().u12;
    // TrowbridgeReitzDistribution::PrepareReflectanceTexture (microfacet.cpp:31-75)
    const int32_t n = kTableSize;
    const float d = 1.0f / float(n);
    for (int32_t j = 0; j < n; ++j) {
        const float a = d / 2.0f + d * float(j);
        float r_sum = 0.0f;
        for (int32_t i = 0; i < n; ++i) {
            const float cos_theta = d / 2.0f + d * float(i);
            const float sin_theta = std::sqrt(1.0f - sqr(cos_theta));
            const f3 wo = mk3(sin_theta, 0.0f, cos_theta);
            const float r = tr_rho(a, wo, u12.data(), samples);
            r_sum += r * cos_theta * d;
            t.ggx_E[j * n + i] = r;
        }
        t.ggx_E_avg[j] = 2.0f * r_sum;
    }
}

// DielectricMultiScatteringBxDF::PrepareReflectanceTexture
// (microfacet.cpp:173-236); the 32 f0 slices are independent (the
// reference's ParallelFor over them) and run on host threads
void build_dielectric_tables()
{
    DielectricTables& t = g_dielectric;
    const TableSamples smp = table_samples();
    const int32_t n = kTableSize;
    const float d = 1.0f / float(n);
    auto slice = [&](int32_t k) {
        const float f = d / 2.0f + d * float(k);
        const float ior = map_f0_to_ior(f);
        for (int32_t j = 0; j < n; ++j) {
            const float a = d / 2.0f + d * float(j);
            float r_sum = 0.0f, r_inv_sum = 0.0f;
            for (int32_t i = 0; i < n; ++i) {
                const float cos_theta = d / 2.0f + d * float(i);
                const float sin_theta = std::sqrt(1.0f - sqr(cos_theta));
                const f3 wo = mk3(sin_theta, 0.0f, cos_theta);
                const float r = dielectric_rho(ior, a, wo, smp.u0.data(), smp.u12.data(), kTableSamples);
                const float r_inv = dielectric_rho(1.0f / ior, a, wo, smp.u0.data(), smp.u12.data(), kTableSamples);
                r_sum += r * cos_theta * d;
                r_inv_sum += r_inv * cos_theta * d;
                t.E[(j * n + i) * n + k] = r;
                t.E_inv[(j * n + i) * n + k] = r_inv;
            }
            t.E_avg[j * n + k] = 2.0f * r_sum;
            t.E_inv_avg[j * n + k] = 2.0f * r_inv_sum;
        }
    };
    const int32_t nt = int32_t(std::clamp(std::thread::hardware_concurrency(), 1u, 16u));
    std::vector<std::thread> pool;
    for (int32_t w = 0; w < nt; ++w)
        pool.emplace_back([&, w] {
            for (int32_t k = w; k < n; k += nt) slice(k);
        });
    for (std::thread& th : pool) th.join();
}

} // namespace

const ReflectanceTables& reflectance_tables()
{
    std::call_once(g_tables_once, build_tables);
    return g_tables;
}

const DielectricTables& dielectric_tables()
{
    std::call_once(g_dielectric_once, build_dielectric_tables);
    return g_dielectric;
}

} // namespace bb
