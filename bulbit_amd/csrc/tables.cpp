// tables.cpp — host computation of the reflectance tables that Bulbit's
// Integrator prepares before rendering (PrepareReflectanceTextures,
// src/bxdf/microfacet.cpp:239-270, called with texture_size 32 and 1024
// samples by src/integrator/integrator.cpp:149-151).
//
// The tables are computed once per process on the host, with the float
// arithmetic and the C library's sinf/cosf that the reference uses, and are
// uploaded with every scene.  Only the Trowbridge-Reitz albedo tables are
// needed on this path (ConductorMultiScatteringBxDF).
#include "bb_scene.h"

#include <cmath>
#include <mutex>

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

ReflectanceTables g_tables;
std::once_flag g_tables_once;

void build_tables()
{
    ReflectanceTables& t = g_tables;
    const int32_t samples = kTableSamples, xs = 32, ys = 32;
    std::vector<f2> u12(samples);
    uint64_t rng = pcg_seed(3); // RNG rng(3): initstate 3, initseq 1
    for (int32_t i = 0; i < samples; ++i) {
        int32_t stratum = permutation_element(uint32_t(i), uint32_t(samples), 1u);
        const float u0 = (float(stratum) + pcg_float(rng)) / float(samples); // drawn, unused by these tables
        (void)u0;
        stratum = permutation_element(uint32_t(i), uint32_t(samples), 2u);
        const int32_t x = stratum % xs, y = stratum / xs;
        const float ux = (float(x) + pcg_float(rng)) / float(xs);
        const float uy = (float(y) + pcg_float(rng)) / float(ys);
        u12[i] = mk2(ux, uy);
    }
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

} // namespace

const ReflectanceTables& reflectance_tables()
{
    std::call_once(g_tables_once, build_tables);
    return g_tables;
}

} // namespace bb
