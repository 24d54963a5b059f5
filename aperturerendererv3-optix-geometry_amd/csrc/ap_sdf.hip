// SDF volume cache (include/aperture/ap_core.h: ap_sdf_cache_*): volumes of
// signed distances on grids, filled on the device from the built-in
// functors or uploaded, and a fixed device table of their descriptors that
// scenes point their traversal at (SDFCacheManager,
// core/object/sdf/sdf_cache_manager.cuh:36-114).
#include <cstring>
#include <vector>

#include "aperture/ap_core.h"
#include "ap_internal.h"
#include "ap_sdf.h"

namespace ap {
namespace {

constexpr uint32_t kVolumeCapacity = 128;

// FillSDFVolume (sdf_cache_manager.cuh:15-34): one thread per sample.
__global__ void fill_volume_kernel(float* __restrict__ out, int32_t nx, int32_t ny, int32_t nz, int32_t functor,
                                   float i0, float i1, float i2, float i3) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    if (x >= nx || y >= ny || z >= nz) return;
    const float fx = nx > 1 ? static_cast<float>(x) / static_cast<float>(nx - 1) : 0.0f;
    const float fy = ny > 1 ? static_cast<float>(y) / static_cast<float>(ny - 1) : 0.0f;
    const float fz = nz > 1 ? static_cast<float>(z) / static_cast<float>(nz - 1) : 0.0f;
    const float info[4] = {i0, i1, i2, i3};
    out[static_cast<size_t>(z) * nx * ny + static_cast<size_t>(y) * nx + x] =
        sdf_functor(functor, f3{fx - 0.5f, fy - 0.5f, fz - 0.5f}, info);
}

}  // namespace
}  // namespace ap

using namespace ap;

struct ap_sdf_cache {
    int device = 0;
    SdfVolume* table = nullptr;              // kVolumeCapacity device descriptors, zeroed past count
    std::vector<SdfVolume> host_table;       // same descriptors (device value pointers)
    std::vector<std::vector<float>> values;  // host copies of the samples
};

namespace ap {

const SdfVolume* sdf_cache_table(const ap_sdf_cache* c) { return c ? c->table : nullptr; }
int32_t sdf_cache_capacity(const ap_sdf_cache* c) { return c ? static_cast<int32_t>(kVolumeCapacity) : 0; }

// Host view of volume k for host-side distance queries (camera inside an object).
bool sdf_cache_host_volume(const ap_sdf_cache* c, int32_t k, SdfVolume& v) {
    if (!c || k < 0 || k >= static_cast<int32_t>(c->values.size())) return false;
    v = c->host_table[k];
    v.values = c->values[k].data();
    return true;
}

}  // namespace ap

namespace {

int add_device_volume(ap_sdf_cache* c, float* dev_values, const int32_t dims[3], uint32_t* index) {
    SdfVolume v{};
    v.values = dev_values;
    for (int a = 0; a < 3; ++a) v.dim[a] = dims[a];
    v.voxel_size_max = sdf_voxel_size_max(v.dim);
    const size_t k = c->host_table.size();
    std::vector<float> host(static_cast<size_t>(dims[0]) * dims[1] * dims[2]);
    hipError_t e = hipMemcpy(host.data(), dev_values, host.size() * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(c->table + k, &v, sizeof(SdfVolume), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dev_values);
        return hip_fail(e, "ap_sdf_cache: publish volume");
    }
    c->host_table.push_back(v);
    c->values.push_back(std::move(host));
    if (index) *index = static_cast<uint32_t>(k);
    return AP_OK;
}

int check_dims(const ap_sdf_cache* c, const int32_t dims[3], const char* who) {
    if (!c || !dims) return fail(AP_ERR_INVALID_ARG, std::string(who) + ": NULL argument");
    if (c->host_table.size() >= kVolumeCapacity)
        return fail(AP_ERR_INVALID_ARG, std::string(who) + ": capacity of 128 volumes exceeded");
    for (int a = 0; a < 3; ++a)
        if (dims[a] <= 0 || dims[a] > 4096) return fail(AP_ERR_INVALID_ARG, std::string(who) + ": dims must be 1..4096");
    if (static_cast<uint64_t>(dims[0]) * dims[1] * dims[2] > (uint64_t{1} << 31))
        return fail(AP_ERR_INVALID_ARG, std::string(who) + ": at most 2^31 samples");
    return AP_OK;
}

}  // namespace

extern "C" {

int ap_sdf_cache_create(ap_sdf_cache** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_sdf_cache_create: NULL output");
    *out = nullptr;
    ap_sdf_cache* c = new ap_sdf_cache();
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->table), kVolumeCapacity * sizeof(SdfVolume));
    if (e == hipSuccess) e = hipMemset(c->table, 0, kVolumeCapacity * sizeof(SdfVolume));
    if (e != hipSuccess) {
        if (c->table) (void)hipFree(c->table);
        delete c;
        return hip_fail(e, "ap_sdf_cache_create");
    }
    *out = c;
    return AP_OK;
}

int ap_sdf_cache_destroy(ap_sdf_cache* c) {
    if (!c) return AP_OK;
    for (const SdfVolume& v : c->host_table) (void)hipFree(const_cast<float*>(v.values));
    (void)hipFree(c->table);
    delete c;
    return AP_OK;
}

int ap_sdf_cache_add_functor(ap_sdf_cache* c, int32_t functor, const float info[4], const int32_t dims[3],
                             uint32_t* index) {
    int rc = check_dims(c, dims, "ap_sdf_cache_add_functor");
    if (rc != AP_OK) return rc;
    if (!info) return fail(AP_ERR_INVALID_ARG, "ap_sdf_cache_add_functor: NULL info");
    if (functor != AP_SDF_FUNCTOR_TORUS && functor != AP_SDF_FUNCTOR_DIAMOND)
        return fail(AP_ERR_INVALID_ARG, "ap_sdf_cache_add_functor: functor must be AP_SDF_FUNCTOR_TORUS or _DIAMOND");
    const size_t n = static_cast<size_t>(dims[0]) * dims[1] * dims[2];
    float* d = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), n * sizeof(float));
    if (e != hipSuccess) return hip_fail(e, "ap_sdf_cache_add_functor: hipMalloc");
    const dim3 block(8, 8, 8);
    const dim3 grid((dims[0] + 7) / 8, (dims[1] + 7) / 8, (dims[2] + 7) / 8);
    hipLaunchKernelGGL(fill_volume_kernel, grid, block, 0, nullptr, d, dims[0], dims[1], dims[2], functor, info[0],
                       info[1], info[2], info[3]);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "ap_sdf_cache_add_functor: fill");
    }
    return add_device_volume(c, d, dims, index);
}

int ap_sdf_cache_add_volume(ap_sdf_cache* c, const float* host_values, const int32_t dims[3], uint32_t* index) {
    int rc = check_dims(c, dims, "ap_sdf_cache_add_volume");
    if (rc != AP_OK) return rc;
    if (!host_values) return fail(AP_ERR_INVALID_ARG, "ap_sdf_cache_add_volume: NULL values");
    const size_t n = static_cast<size_t>(dims[0]) * dims[1] * dims[2];
    float* d = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&d), n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d, host_values, n * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (d) (void)hipFree(d);
        return hip_fail(e, "ap_sdf_cache_add_volume");
    }
    return add_device_volume(c, d, dims, index);
}

int ap_sdf_cache_get_volume(const ap_sdf_cache* c, uint32_t index, int32_t dims[3], float* host_values) {
    if (!c || !dims) return fail(AP_ERR_INVALID_ARG, "ap_sdf_cache_get_volume: NULL argument");
    if (index >= c->host_table.size()) return fail(AP_ERR_INVALID_ARG, "ap_sdf_cache_get_volume: no such volume");
    for (int a = 0; a < 3; ++a) dims[a] = c->host_table[index].dim[a];
    if (host_values) std::memcpy(host_values, c->values[index].data(), c->values[index].size() * sizeof(float));
    return AP_OK;
}

}  // extern "C"
