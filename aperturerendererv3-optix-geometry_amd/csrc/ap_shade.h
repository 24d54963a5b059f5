// Shading math of the integrator (device and host inline), single-wavelength
// spectral form, with the reference's operand order:
//   FresnelSchlick, GGX_D, GGX_G_Correlated, F0_UEStyle,
//   GGXLambert_BRDF_Spectrum             core/render/brdf.cuh:30-129
//   SampleGGXBoundedVNDF, GGXBoundedVNDF_PDF core/render/sample.cuh:179-248
//   SampleBSDF / EvaluateBSDFWithCos / BSDF_PDF (diffuse, light, GGX)
//                                         core/render/material.cuh:48-387
//   PowerHeuristic                        core/render/render.cuh:185-196
//   SkyLighting                           core/render/render.cuh:58-73
//   AreaToSolidAnglePDF, box-facet and sphere-cap light sampling and their
//   paired solid-angle pdfs               core/nee/nee.cuh:24-327
//   Glass: IORFromAbbeNM, SigmaAFromTiltNM  core/spectrum/spectrum.cuh:7-57
//          BeerLambert_Spectrum, FresnelDielectricExact  core/render/brdf.cuh:20-32
//          refract                         core/vector.cuh:267-275
//          SampleBSDF / EvaluateBSDFWithCos (MAT_GLASS)  core/render/material.cuh:122-169, 221-272
#pragma once

#include "ap_sample.h"
#include "ap_spectrum.h"
#include "ap_fmath.h"
#include "ap_vec.h"

namespace ap {

enum : int { SM_DIFFUSE = 0, SM_LIGHT = 1, SM_GLASS = 2, SM_GGX = 3 };

// Per-object material as the shading stage sees it. `albedo_basis` and
// `emissive_basis` are QueryRGBBasis of the colour (divided by its maximum
// when that exceeds 1, the scale kept in *_scale), so Interact reduces to one
// clamped dot product with the wavelength basis.
struct alignas(16) MatRec {
    f4 albedo_basis;
    f4 emissive_basis;
    float albedo_scale, emissive_scale;
    float specular, metallic;
    float roughness;
    int32_t shading_model;
    int32_t light_index;
    int32_t albedo_tex;      // texture cache index of the albedo texture, or -1
    float emissive_rgb[3];   // the GBuffer emissive (guiding endpoints query it with a MIS weight)
    int32_t pad2;
    float albedo_rgb[3];     // the GBuffer albedo (radiance-cache query records)
    int32_t pad3;
    float ior, sigma_a;      // glass: IOR at the d line, absorption coefficient
    float abbe, tilt;        // glass: Abbe number, absorption tilt over wavelength
};

AP_HD float interact_basis(f4 rgb_basis, float scale, f4 lambda_basis) {
    return scale > 1.0f ? spectral_interact(rgb_basis, lambda_basis) * scale : spectral_interact(rgb_basis, lambda_basis);
}

AP_HD float fresnel_schlick(float F0, float cosx) {
    const float c = clampf(1.0f - cosx, 0.0f, 1.0f);
    const float c5 = c * c * c * c * c;
    return F0 + (1.0f - F0) * c5;
}

AP_HD float ggx_d(float alpha, float NoH) {
    const float a2 = alpha * alpha;
    const float d = NoH * NoH * (a2 - 1.0f) + 1.0f;
    return a2 / fmaxf(kPiF * d * d, 1e-12f);
}

AP_HD float ggx_g_correlated(float alpha, float NoL, float NoV) {
    const float a2 = alpha * alpha;
    const float gv = NoL * sqrtf(a2 + (1.0f - a2) * NoV * NoV);
    const float gl = NoV * sqrtf(a2 + (1.0f - a2) * NoL * NoL);
    const float denom = gv + gl;
    return (denom > 0.0f) ? (2.0f * NoL * NoV) / denom : 0.0f;
}

AP_HD float f0_ue(float specular) { return 0.08f * clampf(specular, 0.0f, 1.0f); }

// GGXLambert_BRDF_Spectrum: returns SpecCos + DiffCos.
AP_HD float ggx_lambert(f3 V, f3 L, f3 N, float F0, float NoV, float kD, float alpha, float albedo_l) {
    const float NoL = fmaxf(0.0f, dot(N, L));
    const f3 H = normalize(V + L);
    const float NoH = fmaxf(0.0f, dot(N, H));
    const float LoH = fmaxf(0.0f, dot(L, H));
    const float F = fresnel_schlick(F0, LoH);
    const float D = ggx_d(alpha, NoH);
    const float G = ggx_g_correlated(alpha, NoL, NoV);
    const float spec = (NoL > 0.0f) ? (F * D * G) / fmaxf(4.0f * NoV, 1e-7f) : 0.0f;
    const float diff = kD * (1.0f - F) * albedo_l * (NoL / kPiF);
    return spec + diff;
}

AP_HD DirSample sample_ggx_vndf(f3 V, f3 N, float alpha, f3 rnd) {
    f3 X, Y;
    tangent_basis_random(N, X, Y, rnd.z);
    const float XoV = dot(X, V), YoV = dot(Y, V), NoV = dot(N, V);
    const f3 lv{XoV, YoV, NoV};
    const float NoV2 = fminf(1.0f, NoV * NoV);
    const float a2 = alpha * alpha;
    const f3 vs = f3{XoV * alpha, YoV * alpha, NoV} / sqrtf(a2 + (1.0f - a2) * NoV2);
    const float phi = 2.0f * kPiF * rnd.x;
    const float s = 1.0f + sqrtf(1.0f - NoV2);
    const float s2 = s * s;
    const float k = (1.0f - a2) * s2 / (s2 + a2 * NoV2);
    const float b = NoV > 0 ? k * vs.z : vs.z;
    const float z = fmaf(1.0f - rnd.y, 1.0f + b, -b);
    const float sinT = sqrtf(saturatef(1.0f - z * z));
    const f3 os{sinT * apfm::cos(phi), sinT * apfm::sin(phi), z};
    const f3 ms = vs + os;
    const f3 H = normalize(f3{ms.x * alpha, ms.y * alpha, ms.z});
    const f3 lL = reflect(-lv, H);
    const f3 dir = lL.x * X + lL.y * Y + lL.z * N;
    const float ndf = ggx_d(alpha, H.z);
    const float len2 = (1.0f - NoV2) * a2;
    const float t = sqrtf(len2 + NoV2);
    return {dir, ndf / (2.0f * (k * NoV + t))};
}

AP_HD float ggx_vndf_pdf(f3 V, f3 N, f3 L, float alpha) {
    const float NoV = dot(N, V), NoL = dot(N, L);
    if (NoV <= 0.0f || NoL <= 0.0f) return 0.0f;
    const f3 H = normalize(V + L);
    const float NoV2 = fminf(1.0f, NoV * NoV);
    const float a2 = alpha * alpha;
    const float s = 1.0f + sqrtf(1.0f - NoV2);
    const float s2 = s * s;
    const float k = (1.0f - a2) * s2 / (s2 + a2 * NoV2);
    const float NoH = fmaxf(0.0f, dot(N, H));
    const float ndf = ggx_d(alpha, NoH);
    const float len2 = (1.0f - NoV2) * a2;
    const float t = sqrtf(len2 + NoV2);
    return fmaxf(ndf / (2.0f * (k * NoV + t)), 1e-12f);
}

// Wavelength response of the albedo (Interact(lambdaBasis, albedo)).
AP_HD float albedo_response(const MatRec& m, f4 lb) { return interact_basis(m.albedo_basis, m.albedo_scale, lb); }

// EvaluateBSDFWithCos (material.cuh:48-175) for DIFFUSE / LIGHT / GGX.
AP_HD float eval_bsdf_cos(const MatRec& m, f4 lb, f3 ray_dir, f3 L, f3 N) {
    const float NoL = fmaxf(0.0f, dot(N, L));
    if (m.shading_model == SM_DIFFUSE) return albedo_response(m, lb) * (NoL / kPiF);
    if (m.shading_model == SM_GGX) {
        const f3 V = normalize(-ray_dir);
        const float NoV = fmaxf(0.0f, dot(N, V));
        if (NoV <= 0.0f || NoL <= 0.0f) return 0.0f;
        const float alpha = m.roughness * m.roughness;
        const float F0d = f0_ue(m.specular);
        const float al = albedo_response(m, lb);
        const float F0 = lerpf(F0d, al, m.metallic);
        const float kD = 1.0f - m.metallic;
        return ggx_lambert(V, L, N, F0, NoV, kD, alpha, al);
    }
    return 0.0f;
}

// BSDF_PDF (material.cuh:332-387).
AP_HD float bsdf_pdf(const MatRec& m, f4 lb, f3 N, f3 ray_dir, f3 L) {
    const f3 V = -ray_dir;
    const float NoL = fmaxf(0.0f, dot(N, L));
    if (m.shading_model == SM_DIFFUSE) return NoL > 0.0f ? (NoL * (1.0f / kPiF)) : 0.0f;
    if (m.shading_model == SM_GGX) {
        const float NoV = fmaxf(0.0f, dot(N, V));
        if (NoV <= 0.0f || NoL <= 0.0f) return 0.0f;
        const float alpha = m.roughness * m.roughness;
        const float F0 = lerpf(f0_ue(m.specular), albedo_response(m, lb), m.metallic);
        const float pSpec = clampf(F0, 0.1f, 0.9f);
        const float pdfSpec = ggx_vndf_pdf(V, N, L, alpha);
        const float pdfDiff = NoL * (1.0f / kPiF);
        return fmaxf(0.0f, pSpec * pdfSpec + (1.0f - pSpec) * pdfDiff);
    }
    return 0.0f;
}

// ---- glass ----
// Fraunhofer F, C and d lines (nm) as reciprocals.
constexpr float kInvF = 1.0f / 486.1327f;
constexpr float kInvC = 1.0f / 656.2725f;
constexpr float kInvD = 1.0f / 587.5618f;
constexpr float kPerfectDelta = 0.99f;       // PERFECT_DELTA_SURFACE
constexpr int kMaxDeltaDepth = 1024;         // MAX_DELTA_SURFACE_DEPTH

// IORLinearKFromAbbe: the slope of n over 1/lambda from the Abbe number.
AP_HD float abbe_slope(float nd, float vd) {
    if (!(vd > 0.0f)) return 0.0f;
    return (nd - 1.0f) / (vd * (kInvF - kInvC));
}

// IORFromAbbeNM: n(lambda) = nd + K (1/lambda - 1/lambda_d).
AP_HD float ior_at(float lambda, float nd, float vd) {
    if (!(lambda > 0.0f)) return nd;
    const float K = abbe_slope(nd, vd);
    const float invL = 1.0f / lambda;
    return nd + K * (invL - kInvD);
}

// SigmaAFromTiltNM with a tilt gain of 1.
AP_HD float sigma_a_at(float sigma_a, float lambda, float tilt) {
    const float u = 1.0f / fmaxf(lambda, 1.0e-6f);
    const float du = (u - kInvD) / (kInvF - kInvC);
    const float sA = sigma_a * (1.0f + 1.0f * tilt * du);
    return fmaxf(0.0f, sA);
}

// BeerLambert_Spectrum: albedo_l^(distance sigma_a).
AP_HD float beer_lambert(float distance, float albedo_l, float sigma_a) {
    return apfm::exp(distance * sigma_a * apfm::log(fminf(fmaxf(albedo_l, 1.0e-12f), 1.0f)));
}

// FresnelDielectricExact, eta = n_in / n_out.
AP_HD float fresnel_dielectric(float cosI, float cosT, float eta) {
    const float invE = 1.0f / eta;
    const float rPar = (invE * cosI - cosT) / (invE * cosI + cosT);
    const float rPerp = (cosI - invE * cosT) / (cosI + invE * cosT);
    return 0.5f * (rPar * rPar + rPerp * rPerp);
}

AP_HD f3 refract_dir(f3 in, f3 normal, float eta) {
    const f3 n = (dot(normal, in) > 0.0f) ? -normal : normal;
    const float cosI = -dot(n, in);
    const float k = 1.0f - (eta * eta) * fmaxf(0.0f, 1.0f - cosI * cosI);
    const float cosT = sqrtf(fmaxf(0.0f, k));
    return eta * in + (eta * cosI - cosT) * n;
}

// The interface a ray meets: the normal on its side, n_in / n_out, the
// incident cosine and 1 - sin^2(theta_t) (<= 0: total internal reflection).
struct Interface {
    f3 n_in;
    float eta, cosI, k;
};

AP_HD Interface glass_interface(float ior, bool inside, f3 ray_dir, f3 N) {
    Interface f;
    f.n_in = (dot(N, ray_dir) > 0.0f) ? -N : N;
    f.eta = inside ? ior : 1.0f / ior;
    f.cosI = -dot(f.n_in, ray_dir);
    const float sin2I = fmaxf(0.0f, 1.0f - f.cosI * f.cosI);
    f.k = 1.0f - sin2I * (f.eta * f.eta);
    return f;
}

// EvaluateBSDFWithCos for MAT_GLASS at one wavelength: the absorption over
// `distance` when the ray ran inside the medium, times the Fresnel weight
// of the branch (reflection or refraction) that L follows.
AP_HD float eval_glass(const MatRec& m, float lambda, f4 lb, bool inside, float distance, f3 ray_dir, f3 L, f3 N) {
    const float ior = ior_at(lambda, m.ior, m.abbe);
    float transmittance = 1.0f;
    if (inside) transmittance = beer_lambert(distance, albedo_response(m, lb), sigma_a_at(m.sigma_a, lambda, m.tilt));
    const Interface f = glass_interface(ior, inside, ray_dir, N);
    if (f.k <= 0.0f) return transmittance * kPerfectDelta;
    const float cosT = sqrtf(f.k);
    const float Fr = fresnel_dielectric(f.cosI, cosT, f.eta);
    const f3 R = reflect(ray_dir, f.n_in);
    const f3 T = refract_dir(ray_dir, f.n_in, f.eta);
    const float dR = dot(L, R);
    const float dT = dot(L, T);
    if (dR >= dT) return transmittance * Fr;
    const float weight = (f.eta * f.eta) * (fabsf(cosT) / fmaxf(1e-6f, f.cosI));
    return transmittance * (1.0f - Fr) * weight;
}

AP_HD float power_heuristic(float pa, float pb) {
    const float a = pa * pa, b = pb * pb;
    return a / fmaxf(1e-24f, a + b);
}

AP_HD f3 sky_lighting(f3 d) {
    const f3 white{1.0f, 1.0f, 1.0f};
    const f3 blue{0.05f, 0.45f, 1.0f};
    const f3 down = f3{0.2f, 0.6f, 1.0f} * 0.125f;
    return 0.5f * lerp3(lerp3(blue, white, apfm::pow075(fmaxf(d.y, 0.0f))), down,
                        fminf(fmaxf(-d.y * 0.5f + 0.5f, 0.0f), 1.0f));
}

AP_HD float area_to_solid_angle_pdf(f3 surf, f3 lp, f3 ln, float pa_total) {
    const f3 sl = lp - surf;
    const float d2 = fmaxf(dot(sl, sl), 1e-20f);
    const f3 dir = sl * (1.0f / sqrtf(d2));
    const float cosY = fmaxf(0.0f, dot(ln, -dir));
    if (cosY <= 0.0f) return 0.0f;
    return pa_total * (d2 / cosY);
}

// Light geometry for NEE (a cube or a sphere object); `mat` indexes the
// light object's MatRec for its emission.
struct alignas(16) LightRec {
    float center[3]; int32_t type;      // AP_OBJ_CUBE, AP_OBJ_SPHERE, or -1 (cannot be sampled)
    float extent[3]; int32_t object;
    float ux[3], uy[3], uz[3];          // normalised OBB axes (GetOBBAxes, nee.cuh:24-31)
    float select_prob;
    int32_t mat;
    int32_t pad[3];
};

struct LightSample {
    f3 point, normal;
    float p_facet, p_area;
    int32_t object, mat;
    bool valid;
};

// One facet of a box light: outward normal n, centre c, in-plane axes t and
// b with half extents s1 and s2. Facets 0..5 are +x, -x, +y, -y, +z, -z of
// the OBB frame. Built by value from a facet index so that a loop over the
// six facets, or a facet picked at run time, needs no indexed local arrays.
struct BoxFacet {
    f3 n, c, t, b;
    float s1, s2;
};

AP_HD BoxFacet box_facet(const LightRec& L, int f) {
    const f3 ux{L.ux[0], L.ux[1], L.ux[2]}, uy{L.uy[0], L.uy[1], L.uy[2]}, uz{L.uz[0], L.uz[1], L.uz[2]};
    const f3 bc{L.center[0], L.center[1], L.center[2]};
    const f3 e{L.extent[0], L.extent[1], L.extent[2]};
    BoxFacet F;
    f3 axis;
    float ea;
    const int a = f >> 1;
    if (a == 0) {
        axis = ux; ea = e.x; F.t = uy; F.b = uz; F.s1 = e.y; F.s2 = e.z;
    } else if (a == 1) {
        axis = uy; ea = e.y; F.t = ux; F.b = uz; F.s1 = e.x; F.s2 = e.z;
    } else {
        axis = uz; ea = e.z; F.t = ux; F.b = uy; F.s1 = e.x; F.s2 = e.y;
    }
    if (f & 1) {
        F.n = -axis;
        F.c = bc - axis * ea;
    } else {
        F.n = axis;
        F.c = bc + axis * ea;
    }
    return F;
}

AP_HD float facing_weight(f3 fn, f3 fc, f3 p, float area) {
    const f3 l2s = normalize(p - fc);
    return area * fmaxf(0.0f, dot(fn, l2s));
}

// NEEPDFOmegaFromHit (nee.cuh:64-151, 254-327) for a light hit by a BSDF ray.
AP_HD float nee_pdf_from_hit(const LightRec& L, f3 surf, f3 lp, f3 ln) {
    if (L.type == 0) {  // cube
        int hit = -1;
#pragma unroll
        for (int f = 0; f < 6; ++f)
            if (hit < 0 && dot(box_facet(L, f).n, ln) > 0.999f) hit = f;
        if (hit < 0) {
            float best = 1e30f;
#pragma unroll
            for (int f = 0; f < 6; ++f) {
                const BoxFacet F = box_facet(L, f);
                const float d = fabsf(dot(F.n, lp - F.c));
                if (d < best) { best = d; hit = f; }
            }
        }
        float sum = 0.0f, whit = 0.0f, ahit = 0.0f;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const BoxFacet F = box_facet(L, f);
            const float area = 4.0f * F.s1 * F.s2;
            const float w = facing_weight(F.n, F.c, surf, area);
            sum += w;
            if (f == hit) { whit = w; ahit = area; }
        }
        if (sum <= 0.0f || whit <= 0.0f) return 0.0f;
        const float pa = L.select_prob * (1.0f / ahit) * (whit / sum);
        return area_to_solid_angle_pdf(surf, lp, ln, pa);
    }
    if (L.type == 1) {  // sphere
        const f3 sc{L.center[0], L.center[1], L.center[2]};
        const float r = fmaxf(L.extent[0], fmaxf(L.extent[1], L.extent[2]));
        const float dist = length(surf - sc);
        if (dist <= r) return 0.0f;
        const float cosC = clampf(r / dist, 0.0f, 1.0f);
        const float cap = 2.0f * kPiF * r * r * (1.0f - cosC);
        if (cap <= 1e-20f) return 0.0f;
        return area_to_solid_angle_pdf(surf, lp, ln, L.select_prob / cap);
    }
    return 0.0f;
}

}  // namespace ap
