// Texture cache (include/aperture/ap_mesh.h: ap_texture_cache_*): RGBA8
// images on the device and a fixed table of their descriptors that scenes
// point their shading at (the images behind LoadCUDATexture's texture
// objects, core/helper/texture_loader.cuh:10-55).
#include <cstring>
#include <vector>

#include "aperture/ap_mesh.h"
#include "ap_internal.h"
#include "ap_texture.h"

namespace ap {
namespace {

constexpr uint32_t kTexCapacity = AP_TEXTURE_CACHE_CAPACITY;
constexpr int32_t kTexMaxSide = 32768;

__global__ void sample_kernel(const TexRec* __restrict__ table, uint32_t index, const float* __restrict__ uv,
                              uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float c[4];
    tex_sample(table[index], uv[2 * i], uv[2 * i + 1], c);
    for (int k = 0; k < 4; ++k) out[4 * i + k] = c[k];
}

}  // namespace
}  // namespace ap

using namespace ap;

struct ap_texture_cache {
    int device = 0;
    TexRec* table = nullptr;            // kTexCapacity device descriptors, zeroed past count
    std::vector<TexRec> host_table;     // same descriptors (device texel pointers)
};

namespace ap {
const TexRec* texture_cache_table(const ap_texture_cache* c) { return c ? c->table : nullptr; }
uint32_t texture_cache_count(const ap_texture_cache* c) {
    return c ? static_cast<uint32_t>(c->host_table.size()) : 0u;
}
}  // namespace ap

extern "C" {

int ap_texture_cache_create(ap_texture_cache** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_texture_cache_create: NULL out");
    *out = nullptr;
    auto* c = new ap_texture_cache();
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->table), kTexCapacity * sizeof(TexRec));
    if (e == hipSuccess) e = hipMemset(c->table, 0, kTexCapacity * sizeof(TexRec));
    if (e != hipSuccess) {
        if (c->table) (void)hipFree(c->table);
        delete c;
        return hip_fail(e, "ap_texture_cache_create");
    }
    *out = c;
    return AP_OK;
}

int ap_texture_cache_destroy(ap_texture_cache* c) {
    if (!c) return AP_OK;
    for (const TexRec& t : c->host_table)
        if (t.texels) (void)hipFree(const_cast<uint32_t*>(t.texels));
    if (c->table) (void)hipFree(c->table);
    delete c;
    return AP_OK;
}

int ap_texture_cache_add_rgba8(ap_texture_cache* c, const uint8_t* rgba, int32_t width, int32_t height,
                               uint64_t* handle) {
    if (!c || !rgba || !handle) return fail(AP_ERR_INVALID_ARG, "ap_texture_cache_add_rgba8: NULL argument");
    if (width <= 0 || height <= 0 || width > kTexMaxSide || height > kTexMaxSide)
        return fail(AP_ERR_INVALID_ARG, "ap_texture_cache_add_rgba8: size out of range");
    if (c->host_table.size() >= kTexCapacity) return fail(AP_ERR_INVALID_ARG, "ap_texture_cache_add_rgba8: cache full");
    const size_t bytes = static_cast<size_t>(width) * static_cast<size_t>(height) * 4u;
    uint32_t* texels = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&texels), bytes);
    if (e == hipSuccess) e = hipMemcpy(texels, rgba, bytes, hipMemcpyHostToDevice);
    TexRec t{};
    t.texels = texels;
    t.width = width;
    t.height = height;
    const size_t k = c->host_table.size();
    if (e == hipSuccess) e = hipMemcpy(c->table + k, &t, sizeof(TexRec), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (texels) (void)hipFree(texels);
        return hip_fail(e, "ap_texture_cache_add_rgba8");
    }
    c->host_table.push_back(t);
    *handle = static_cast<uint64_t>(k) + 1u;
    return AP_OK;
}

int ap_texture_cache_count(const ap_texture_cache* c, uint32_t* count) {
    if (!c || !count) return fail(AP_ERR_INVALID_ARG, "ap_texture_cache_count: NULL argument");
    *count = static_cast<uint32_t>(c->host_table.size());
    return AP_OK;
}

int ap_texture_sample(const ap_texture_cache* c, uint64_t handle, const float* uv, size_t n, float* out) {
    if (!c || (n && (!uv || !out))) return fail(AP_ERR_INVALID_ARG, "ap_texture_sample: NULL argument");
    if (handle == 0 || handle > c->host_table.size())
        return fail(AP_ERR_INVALID_ARG, "ap_texture_sample: unknown texture handle");
    if (n == 0) return AP_OK;
    if (n > (1u << 30)) return fail(AP_ERR_INVALID_ARG, "ap_texture_sample: too many coordinates");
    float *d_uv = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_uv), n * 2 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_out), n * 4 * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d_uv, uv, n * 2 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sample_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, nullptr,
                           c->table, static_cast<uint32_t>(handle - 1), d_uv, static_cast<uint32_t>(n), d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_out, n * 4 * sizeof(float), hipMemcpyDeviceToHost);
    if (d_uv) (void)hipFree(d_uv);
    if (d_out) (void)hipFree(d_out);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_texture_sample");
}

}  // extern "C"
