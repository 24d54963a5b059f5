// ORACLE — test infrastructure only. Spectral tables and functions of the
// CPU restatement (see or_spectrum.cpp for the reference sources).
#pragma once

#include <vector>

#include "or_math.h"

namespace orc {

struct SpectrumTables {
    std::vector<float> storage;
    const float* rgb_in_w = nullptr;
    const float* rgb_in_b = nullptr;
    const float* rgb_blk_w[7] = {};
    const float* rgb_blk_b[7] = {};
    const float* rgb_out_w = nullptr;
    const float* lam_blk_w[7] = {};
    const float* lam_blk_b[7] = {};
    const float* lam_out_w = nullptr;
    float d65[421] = {};
    float cmf[421 * 3] = {};
    bool loaded = false;
};
extern SpectrumTables g_spec;

void RGBEncode(const float in3[3], float outK[4]);
void LambdaEncode(float lambda, float outK[4]);
float SpectralInteract(const float* rgbK, const float* lamK);
void QueryRGBBasis(float3 rgb, float out[4]);
void QueryLambda(float lambda, float out[4]);
float Interact(const float lamBasis[4], float3 rgb);
float ndtrif(float u);
void SampleCIE1931(float r0, float r1, float r2, float* lambda, float* pdf);
float PdfCIE1931Mix(float lambda);
float D65Norm(float lambda);
float3 LambdaToXYZ(float lambda);
float3 XYZ2SRGB(float3 xyz);

}  // namespace orc
