// Per-frame build of the guiding probes (SceneRenderer::UpdatePathGuidingSample,
// core/scene.cu:583-744, over core/guiding/guiding.cuh:70-517).
//
// Input: the frame's guiding samples, PG_COLLECT slots per path in pixel
// order (written by the accumulate kernel, each path's valid samples packed
// to the front of its slots) and each path's number of valid samples. Steps,
// all on the render stream:
//   0. an exclusive scan of the per-path counts gives the valid slots in
//      slot order, and their voxel bounds;
//   1. key every valid slot by its voxel (x, y, z, face) relative to those
//      bounds, so that the unsigned order is the reference's Int4Less order
//      and the key fits 32 bits (absolute 62-bit keys when it does not);
//      radix-sort (key, slot) pairs; mark run heads and scan them: the unique
//      voxels in sorted order (thrust sort + unique in the reference);
//   2. grow the index volume to the union of its bounds and the voxels'
//      (CopyOldIntoNew), then number the voxels without a probe in sorted
//      order and append zeroed probes for them (FilterUnallocatedGrids +
//      ScatterAssignIndices);
//   3. key every valid slot by (probe, leaf texel), stable radix sort, and
//      let the head of each run sum radiance * 4pi over the run in slot
//      order (the reference adds with float atomics in arbitrary order, so
//      its sums are not reproducible; this order is, and the oracle uses it);
//   4. one workgroup per probe with samples merges the sums into the
//      running mean in FP64 and rebuilds the mips (MergeProbeKernel).
// The scans and the stable radix sorts are the project's own wave64
// primitives (ap_sort.hip).
// The host waits once per frame, for the unique voxels' count and bounds
// (the index volume is reallocated on the host when they grow it); the
// valid-slot count stays on the device, and the new-probe count is read
// back without a wait and taken in at the next use of the probes.
#include <cstring>

#include "aperture/ap_guiding.h"
#include "ap_guide.h"
#include "ap_internal.h"
#include "ap_sort.h"

namespace ap {

// Voxel coordinates must lie in [-2^18, 2^18): 19 bits each, face in 5 bits.
constexpr int kCoordBits = 19;
constexpr int32_t kCoordBias = 1 << (kCoordBits - 1);
constexpr uint64_t kNoKey = ~0ull;
// Build counters on the device (GuideDevice::info): [0] unique voxels,
// [1..3] / [4..6] the frame's voxel bounds, [7] new probes, [8] error, [9]
// valid slots, [10] a voxel range too wide for 32-bit keys, and:
constexpr int kInfoCount = 11;  // probes on the device (carried across frames)
constexpr int kInfoHost = 12;   // 1: the host finishes this build (the speculative kernels stand down)
constexpr int kInfoBase = 13;   // probe count this frame's new probes are numbered from

struct GuideDevice {
    ap_octahedron* probes = nullptr;
    ap_octahedron* temp = nullptr;
    int32_t count = 0, cap = 0;
    int32_t* volume = nullptr;
    int32_t start[3] = {0, 0, 0}, end[3] = {0, 0, 0}, size[3] = {0, 0, 0};
    GuideRec* recs = nullptr;
    size_t rec_cap = 0;
    GuideRec* recs2 = nullptr;    // the second frame set's path records (pipelined frames)
    size_t rec2_cap = 0;
    ap_path_guiding_sample* slots = nullptr;
    size_t slot_cap = 0;
    uint32_t* nvalid = nullptr;   // per path: its valid samples (packed to the front of its slots)
    size_t nvalid_cap = 0;
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* flags = nullptr;
    uint32_t* ranks = nullptr;
    uint32_t* list = nullptr;     // valid slots of the frame, in slot order
    uint64_t* uniq = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    size_t sort_cap = 0;
    int32_t* bounds_part = nullptr;  // per compaction block: its voxel bounds (8 ints)
    int32_t* info = nullptr;      // device: [0] unique count, [1..3] min xyz, [4..6] max xyz, [7] new probes, [8] error,
                                  // [9] valid slots, [10] voxel range too wide for 32-bit keys
    int32_t* info_host = nullptr; // pinned
    // Launch size of the voxel sort: a bound on the frame's valid slots from
    // the frames before it; a frame with more sorts again at its real count.
    uint32_t sort_hint = 0;
    // The new-probe count of the last update is read back without a wait;
    // `count` takes it in at the next use (settle).
    bool pending = false;
    hipEvent_t ev_count = nullptr;
    // The host's read-back of an update's counters and check (info[12]),
    // recorded right after the check so the host waits only that long.
    hipEvent_t ev_flag = nullptr;
    // An update between guide_update_begin and guide_update_end: its slot
    // count and voxel-sort launch size.
    bool open = false;
    uint32_t upd_n = 0, upd_cap = 0;
    bool wide = false;            // the update's voxel keys are absolute 62-bit keys
};

int settle(GuideDevice* g);

namespace {

constexpr int kBlock = 256;
inline dim3 grid_for(size_t n) { return dim3(static_cast<unsigned>((n + kBlock - 1) / kBlock)); }

__device__ __forceinline__ bool in_range(int v) { return v >= -kCoordBias && v < kCoordBias; }

__device__ __forceinline__ uint64_t pack_voxel(const ap_int4& v) {
    const uint64_t x = static_cast<uint64_t>(v.x + kCoordBias), y = static_cast<uint64_t>(v.y + kCoordBias),
                   z = static_cast<uint64_t>(v.z + kCoordBias);
    return (x << (2 * kCoordBits + 5)) | (y << (kCoordBits + 5)) | (z << 5) | static_cast<uint64_t>(v.w);
}

__device__ __forceinline__ ap_int4 unpack_voxel(uint64_t k) {
    const int32_t mask = (1 << kCoordBits) - 1;
    ap_int4 v;
    v.x = static_cast<int32_t>((k >> (2 * kCoordBits + 5)) & mask) - kCoordBias;
    v.y = static_cast<int32_t>((k >> (kCoordBits + 5)) & mask) - kCoordBias;
    v.z = static_cast<int32_t>((k >> 5) & mask) - kCoordBias;
    v.w = static_cast<int32_t>(k & 31u);
    return v;
}

struct VolumeView {
    int32_t* volume;
    int32_t start[3], size[3];
    __device__ int64_t slot(const ap_int4& v) const {
        const int lx = v.x - start[0], ly = v.y - start[1], lz = v.z - start[2];
        if (v.w < 0 || v.w >= PG_FACES) return -1;
        if (lx < 0 || ly < 0 || lz < 0 || lx >= size[0] || ly >= size[1] || lz >= size[2]) return -1;
        const size_t lin = static_cast<size_t>(lx) +
                           static_cast<size_t>(size[0]) * (static_cast<size_t>(ly) + static_cast<size_t>(size[1]) * lz);
        return static_cast<int64_t>(lin * PG_FACES + static_cast<size_t>(v.w));
    }
};

// The voxel bounds of an index volume (inclusive).
struct VolumeBox {
    int32_t start[3], end[3];
};

// Wave-wide minimum / maximum (every lane of the wave takes part).
__device__ __forceinline__ int32_t wave_min(int32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int32_t wave_max(int32_t v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// The valid slots in slot order (FilterValidSamples, guiding.cuh:70-99):
// path p's nvalid[p] samples sit at the front of its PG_COLLECT slots and
// start at offs[p] in the list. Also their count (info[9]) and each block's
// voxel bounds (part[block][0..5], reduced by bounds_reduce_kernel); a valid
// slot whose voxel is outside the packable range is reported in info[8].
// Grid-stride over the paths.
constexpr unsigned kCompactBlocks = 4096;
__global__ __launch_bounds__(kBlock) void compact_valid_kernel(const ap_path_guiding_sample* __restrict__ slots,
                                                               const uint32_t* __restrict__ nvalid,
                                                               const uint32_t* __restrict__ offs, uint32_t n_paths,
                                                               uint32_t* __restrict__ list,
                                                               int32_t* __restrict__ info,
                                                               int32_t* __restrict__ part) {
    __shared__ int32_t red[6][kBlock / 64];
    int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    bool bad = false;
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < n_paths; p += gridDim.x * kBlock) {
        const uint32_t c = min(nvalid[p], static_cast<uint32_t>(PG_COLLECT));
        const uint32_t o = offs[p];
        if (p == n_paths - 1) info[9] = static_cast<int32_t>(o + c);
        for (uint32_t k = 0; k < c; ++k) {
            const uint32_t slot = p * PG_COLLECT + k;
            const ap_int4 v = slots[slot].voxel_grid;
            list[o + k] = slot;
            if (in_range(v.x) && in_range(v.y) && in_range(v.z) && v.w >= 0 && v.w < PG_FACES) {
                lo[0] = min(lo[0], v.x); lo[1] = min(lo[1], v.y); lo[2] = min(lo[2], v.z);
                hi[0] = max(hi[0], v.x); hi[1] = max(hi[1], v.y); hi[2] = max(hi[2], v.z);
            } else {
                bad = true;
            }
        }
    }
    if (bad) atomicOr(&info[8], 1);
    const int w = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
        if ((threadIdx.x & 63) == 0) {
            red[a][w] = lo[a];
            red[3 + a][w] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k)
            for (int a = 0; a < 3; ++a) {
                red[a][0] = min(red[a][0], red[a][k]);
                red[3 + a][0] = max(red[3 + a][0], red[3 + a][k]);
            }
        for (int a = 0; a < 6; ++a) part[blockIdx.x * 8 + a] = red[a][0];
    }
}

// info[1..6]: the voxel bounds over the `blocks` partial bounds (one block).
__global__ __launch_bounds__(kBlock) void bounds_reduce_kernel(const int32_t* __restrict__ part, uint32_t blocks,
                                                               int32_t* __restrict__ info) {
    __shared__ int32_t red[6][kBlock / 64];
    int32_t v[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint32_t b = threadIdx.x; b < blocks; b += kBlock)
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], part[b * 8 + a]);
            v[3 + a] = max(v[3 + a], part[b * 8 + 3 + a]);
        }
    const int w = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        v[a] = wave_min(v[a]);
        v[3 + a] = wave_max(v[3 + a]);
        if ((threadIdx.x & 63) == 0) {
            red[a][w] = v[a];
            red[3 + a][w] = v[3 + a];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / 64; ++k)
            for (int a = 0; a < 3; ++a) {
                red[a][0] = min(red[a][0], red[a][k]);
                red[3 + a][0] = max(red[3 + a][0], red[3 + a][k]);
            }
        for (int a = 0; a < 6; ++a) info[1 + a] = red[a][0];
    }
}
// Item count of a launch sized for m: min(*dyn, m) (the valid slots).
__device__ __forceinline__ uint32_t live_count(uint32_t m, const int32_t* dyn) {
    return min(static_cast<uint32_t>(*dyn), m);
}

// (voxel, face) key of each valid slot, the slot index as value. Narrow
// keys are relative to the frame's voxel bounds (info[1..6]):
// ((x - x0) * ny + (y - y0)) * nz + (z - z0)) * 32 + face, the Int4Less order
// in 32 bits; a range too wide for that sets info[10] (the build then sorts
// the absolute keys of pack_voxel). Wide keys are pack_voxel's.
template <bool WIDE>
__global__ void voxel_keys_kernel(const ap_path_guiding_sample* __restrict__ slots,
                                  const uint32_t* __restrict__ list, uint32_t m, int32_t* __restrict__ info,
                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= live_count(m, &info[9])) return;
    const uint32_t j = list[i];
    const ap_int4 v = slots[j].voxel_grid;
    uint64_t k;
    if constexpr (WIDE) {
        k = pack_voxel(v);
    } else {
        const uint64_t ny = static_cast<uint64_t>(static_cast<int64_t>(info[5]) - info[2] + 1);
        const uint64_t nz = static_cast<uint64_t>(static_cast<int64_t>(info[6]) - info[3] + 1);
        const uint64_t nx = static_cast<uint64_t>(static_cast<int64_t>(info[4]) - info[1] + 1);
        if (nx * ny * nz * 32ull > 0xffffffffull) {
            if (i == 0) info[10] = 1;
            k = 0;
        } else {
            k = ((static_cast<uint64_t>(v.x - info[1]) * ny + static_cast<uint64_t>(v.y - info[2])) * nz +
                 static_cast<uint64_t>(v.z - info[3])) * 32ull + static_cast<uint64_t>(v.w);
        }
    }
    keys[i] = k;
    vals[i] = j;
}

__global__ void run_heads_kernel(const uint64_t* __restrict__ keys, uint32_t n, const int32_t* __restrict__ dyn,
                                 uint32_t* __restrict__ flags) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= live_count(n, dyn)) return;
    const uint64_t k = keys[i];
    flags[i] = (k != kNoKey && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
}

// The unique voxel keys in sorted order and their count (info[0]).
__global__ void unique_voxels_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
                                     const uint32_t* __restrict__ ranks, uint32_t n,
                                     const int32_t* __restrict__ dyn, uint64_t* __restrict__ uniq,
                                     int32_t* __restrict__ info) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    n = live_count(n, dyn);
    if (i >= n) return;
    if (i == n - 1) info[0] = static_cast<int32_t>(ranks[i] + flags[i]);
    if (flags[i]) uniq[ranks[i]] = keys[i];
}

// Starts a build's counters: unique count 0, empty bounds, no new probes,
// no error, no valid slots; the device's probe count (info[11]) carries over.
__global__ void init_info_kernel(int32_t* __restrict__ info) {
    const int k = threadIdx.x;
    if (k < 16 && k != kInfoCount) info[k] = (k >= 1 && k <= 3) ? INT32_MAX : (k >= 4 && k <= 6) ? INT32_MIN : 0;
}

// Sets the device's probe count (info[11]).
__global__ void set_count_kernel(int32_t* __restrict__ info, int32_t count) {
    if (threadIdx.x == 0) info[kInfoCount] = count;
}

// Whether the rest of the build can run without the host (info[12] = 0):
// no error, 32-bit keys, every valid slot sorted (m <= cap), and, with
// unique voxels, an index volume that already covers their bounds (so it is
// not reallocated) and room for a probe per unique voxel. info[13] takes the
// probe count the new probes are numbered from.
__global__ void check_kernel(int32_t* __restrict__ info, uint32_t cap, int32_t has_volume, VolumeBox box,
                             int32_t probe_cap) {
    if (threadIdx.x != 0) return;
    const int32_t n_uniq = info[0];
    bool host = info[8] != 0 || info[10] != 0 || static_cast<uint32_t>(info[9]) > cap;
    if (!host && n_uniq > 0) {
        host = !has_volume || static_cast<int64_t>(info[kInfoCount]) + n_uniq > probe_cap;
        for (int a = 0; a < 3 && !host; ++a)
            host = info[1 + a] < box.start[a] || info[4 + a] > box.end[a];
    }
    info[kInfoHost] = host ? 1 : 0;
    info[kInfoBase] = info[kInfoCount];
}
// A speculative kernel's work: none once the host finishes the build
// (info[12]) or without unique voxels.
__device__ __forceinline__ bool spec_off(const int32_t* __restrict__ info) {
    return info[kInfoHost] != 0 || info[0] <= 0;
}
// The narrow-key decoder of this frame's voxel bounds (info[1..6]).
__device__ __forceinline__ void narrow_decode(const int32_t* __restrict__ info, uint64_t k, ap_int4& v) {
    const uint32_t ny = static_cast<uint32_t>(info[5] - info[2] + 1), nz = static_cast<uint32_t>(info[6] - info[3] + 1);
    v.w = static_cast<int32_t>(k & 31u);
    k >>= 5;
    v.z = info[3] + static_cast<int32_t>(k % nz);
    k /= nz;
    v.y = info[2] + static_cast<int32_t>(k % ny);
    v.x = info[1] + static_cast<int32_t>(k / ny);
}
// The device's probe count after this frame's new probes.
__global__ void commit_kernel(int32_t* __restrict__ info, int32_t host_base) {
    if (threadIdx.x != 0) return;
    if (host_base >= 0) info[kInfoCount] = host_base + info[7];
    else if (!spec_off(info)) info[kInfoCount] = info[kInfoBase] + info[7];
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, size_t n, int32_t v) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// CopyOldIntoNew (guiding.cuh:171-244), one thread per old voxel.
__global__ void copy_volume_kernel(const int32_t* __restrict__ old_vol, VolumeView oldv, VolumeView newv) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    const size_t nvox = static_cast<size_t>(oldv.size[0]) * oldv.size[1] * oldv.size[2];
    if (i >= nvox) return;
    const int x = static_cast<int>(i % oldv.size[0]);
    const int y = static_cast<int>((i / oldv.size[0]) % oldv.size[1]);
    const int z = static_cast<int>(i / (static_cast<size_t>(oldv.size[0]) * oldv.size[1]));
    const int tx = x + oldv.start[0] - newv.start[0], ty = y + oldv.start[1] - newv.start[1],
              tz = z + oldv.start[2] - newv.start[2];
    const size_t dst = static_cast<size_t>(tx) +
                       static_cast<size_t>(newv.size[0]) * (static_cast<size_t>(ty) + static_cast<size_t>(newv.size[1]) * tz);
    for (int f = 0; f < PG_FACES; ++f) newv.volume[dst * PG_FACES + f] = old_vol[i * PG_FACES + f];
}

// Voxel of a unique key: narrow keys relative to start[] over ny x nz
// voxels, or pack_voxel's absolute keys.
struct KeyDecode {
    int32_t wide;
    int32_t start[3];
    uint32_t ny, nz;
    __device__ ap_int4 operator()(uint64_t k) const {
        if (wide) return unpack_voxel(k);
        ap_int4 v;
        v.w = static_cast<int32_t>(k & 31u);
        k >>= 5;
        v.z = start[2] + static_cast<int32_t>(k % nz);
        k /= nz;
        v.y = start[1] + static_cast<int32_t>(k % ny);
        v.x = start[0] + static_cast<int32_t>(k / ny);
        return v;
    }
};

// Flags the unique voxels without a probe. With spec (the speculative
// build), the count and the key decoder come from the device's counters.
__global__ void unindexed_kernel(const uint64_t* __restrict__ uniq, uint32_t n, KeyDecode dec, VolumeView vv,
                                 uint32_t* __restrict__ flags, const int32_t* __restrict__ spec) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    ap_int4 v;
    if (spec) {
        if (spec_off(spec) || i >= static_cast<uint32_t>(spec[0])) return;
        narrow_decode(spec, uniq[i], v);
    } else {
        if (i >= n) return;
        v = dec(uniq[i]);
    }
    const int64_t s = vv.slot(v);
    flags[i] = (s >= 0 && vv.volume[s] < 0) ? 1u : 0u;
}

// Numbers the unindexed voxels from base in sorted order (info[7] takes
// their count). With spec, n and base come from the device's counters.
__global__ void assign_kernel(const uint64_t* __restrict__ uniq, const uint32_t* __restrict__ flags,
                              const uint32_t* __restrict__ ranks, uint32_t n, KeyDecode dec, VolumeView vv,
                              int32_t base, int32_t* __restrict__ info, bool spec) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (spec) {
        if (spec_off(info)) return;
        n = static_cast<uint32_t>(info[0]);
        base = info[kInfoBase];
    }
    if (i >= n) return;
    if (i == n - 1) info[7] = static_cast<int32_t>(ranks[i] + flags[i]);
    if (!flags[i]) return;
    ap_int4 v;
    if (spec) narrow_decode(info, uniq[i], v);
    else v = dec(uniq[i]);
    const int64_t s = vv.slot(v);
    vv.volume[s] = base + static_cast<int32_t>(ranks[i]);
}

// Zeroes the new probes [info[13], info[13] + info[7]) of a speculative
// build (grid-stride).
__global__ void zero_new_spec_kernel(ap_octahedron* __restrict__ probes, const int32_t* __restrict__ info) {
    if (spec_off(info)) return;
    const size_t words = static_cast<size_t>(info[7]) * (sizeof(ap_octahedron) / 16);
    uint4* p = reinterpret_cast<uint4*>(probes + info[kInfoBase]);
    for (size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; i < words;
         i += static_cast<size_t>(gridDim.x) * kBlock)
        p[i] = uint4{0u, 0u, 0u, 0u};
}

// Zeroes probes [base, base + info[7]) (the probes assign_kernel created).
__global__ void zero_new_probes_kernel(ap_octahedron* __restrict__ p, size_t n_max, const int32_t* __restrict__ info) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    const size_t n = min(n_max, static_cast<size_t>(info[7]));
    if (i >= n * (sizeof(ap_octahedron) / 16)) return;
    reinterpret_cast<uint4*>(p)[i] = uint4{0u, 0u, 0u, 0u};
}

__global__ void zero_probes_kernel(ap_octahedron* __restrict__ p, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    const size_t words = n * (sizeof(ap_octahedron) / 16);
    if (i >= words) return;
    reinterpret_cast<uint4*>(p)[i] = uint4{0u, 0u, 0u, 0u};
}

// (probe, leaf texel) key of each valid slot; with spec, over the first
// min(n, m) slots and only while the build is the device's.
__global__ void texel_keys_kernel(const ap_path_guiding_sample* __restrict__ slots, const uint32_t* __restrict__ list,
                                  uint32_t n, VolumeView vv, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                  const int32_t* __restrict__ spec) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (spec) {
        if (spec_off(spec)) return;
        n = live_count(n, &spec[9]);
    }
    if (i >= n) return;
    const uint32_t slot = list[i];
    const ap_path_guiding_sample s = slots[slot];
    uint64_t k = kNoKey;
    if (s.valid) {
        const ap_int4 v = s.voxel_grid;
        if (in_range(v.x) && in_range(v.y) && in_range(v.z)) {
            const int64_t sl = vv.slot(v);
            const int32_t probe = sl >= 0 ? vv.volume[sl] : -1;
            if (probe >= 0) {
                const f3 d = normalize(f3{s.direction.x, s.direction.y, s.direction.z});
                k = (static_cast<uint64_t>(probe) << 6) | static_cast<uint64_t>(pg_leaf_texel(normalize(d)));
            }
        }
    }
    keys[i] = k;
    vals[i] = slot;
}

// Head of each (probe, texel) run sums its samples in slot order.
__global__ void texel_sums_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t n,
                                  const ap_path_guiding_sample* __restrict__ slots, ap_octahedron* __restrict__ temp,
                                  const int32_t* __restrict__ spec) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (spec) {
        if (spec_off(spec)) return;
        n = live_count(n, &spec[9]);
    }
    if (i >= n) return;
    const uint64_t k = keys[i];
    if (k == kNoKey || (i > 0 && keys[i - 1] == k)) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    unsigned long long cnt = 0;
    for (uint32_t j = i; j < n && keys[j] == k; ++j) {
        const ap_float3 r = slots[vals[j]].radiance;
        sx += r.x * PG_JACOBIAN;
        sy += r.y * PG_JACOBIAN;
        sz += r.z * PG_JACOBIAN;
        ++cnt;
    }
    ap_octahedron& t = temp[k >> 6];
    ap_float4& leaf = t.data[k & 63u];
    leaf.x = sx;
    leaf.y = sy;
    leaf.z = sz;
    atomicAdd(&t.count, cnt);
}

// Lane src's texel of a 64-lane probe block.
__device__ __forceinline__ ap_float4 texel_from(const ap_float4& t, int src) {
    return ap_float4{__shfl(t.x, src, 64), __shfl(t.y, src, 64), __shfl(t.z, src, 64), __shfl(t.w, src, 64)};
}
// (((a + b) + c) + d) * 0.25, as GenerateMipmaps averages four texels.
__device__ __forceinline__ ap_float4 quad_mean(const ap_float4& a, const ap_float4& b, const ap_float4& c,
                                               const ap_float4& d) {
    return ap_float4{(((a.x + b.x) + c.x) + d.x) * 0.25f, (((a.y + b.y) + c.y) + d.y) * 0.25f,
                     (((a.z + b.z) + c.z) + d.z) * 0.25f, (((a.w + b.w) + c.w) + d.w) * 0.25f};
}

// One probe's FP64 count-weighted merge of this frame's texel sums, then its
// mips (pg_generate_mipmaps: leaves clamped to PG_PROBE_MIN, each mip texel
// the mean of its four children), one leaf texel per lane and the mips
// reduced across lanes with the same operations in the same order.
__device__ void merge_probe(ap_octahedron& od, ap_octahedron& nw) {
    const unsigned long long cnew = nw.count;
    if (cnew == 0ull) return;
    const unsigned long long cold = od.count;
    const double inv_den = 1.0 / static_cast<double>(cold + cnew);
    const double alpha = static_cast<double>(cold) * inv_den;
    const int k = threadIdx.x;
    ap_float4 o = od.data[k];
    const ap_float4 n = nw.data[k];
    o.x = static_cast<float>(static_cast<double>(o.x) * alpha + static_cast<double>(n.x) * inv_den);
    o.y = static_cast<float>(static_cast<double>(o.y) * alpha + static_cast<double>(n.y) * inv_den);
    o.z = static_cast<float>(static_cast<double>(o.z) * alpha + static_cast<double>(n.z) * inv_den);
    nw.data[k] = ap_float4{0.0f, 0.0f, 0.0f, 0.0f};
    o.x = fmaxf(o.x, PG_PROBE_MIN);
    o.y = fmaxf(o.y, PG_PROBE_MIN);
    o.z = fmaxf(o.z, PG_PROBE_MIN);
    o.w = fmaxf(o.w, PG_PROBE_MIN);
    od.data[k] = o;
    // 8x8 -> 4x4 (lanes 0..15), 4x4 -> 2x2 (lanes 0..3), 2x2 -> 1x1 (lane 0).
    const int x1 = k & 3, y1 = (k >> 2) & 3;
    const int a1 = (2 * y1) * PG_RES + 2 * x1;
    const ap_float4 m1 = quad_mean(texel_from(o, a1), texel_from(o, a1 + 1), texel_from(o, a1 + PG_RES),
                                   texel_from(o, a1 + PG_RES + 1));
    if (k < 16) od.data[PG_RES2 + k] = m1;
    const int x2 = k & 1, y2 = (k >> 1) & 1;
    const int a2 = (2 * y2) * 4 + 2 * x2;
    const ap_float4 m2 = quad_mean(texel_from(m1, a2), texel_from(m1, a2 + 1), texel_from(m1, a2 + 4),
                                   texel_from(m1, a2 + 5));
    if (k < 4) od.data[PG_RES2 + 16 + k] = m2;
    const ap_float4 m3 = quad_mean(texel_from(m2, 0), texel_from(m2, 1), texel_from(m2, 2), texel_from(m2, 3));
    if (k == 0) {
        od.data[PG_RES2 + 20] = m3;
        od.data[PG_TOTAL] = ap_float4{0.0f, 0.0f, 0.0f, 0.0f};
        const unsigned long long sum = cold + cnew;
        od.count = sum < PG_MAX_COUNT ? sum : PG_MAX_COUNT;
        nw.count = 0ull;
    }
}

// MergeProbeKernel (guiding.cuh:436-506): one 64-lane workgroup per probe
// of [0, n), grid-stride; with spec, n is the device's probe count and the
// merge runs only while the build is the device's.
__global__ __launch_bounds__(64) void merge_kernel(ap_octahedron* __restrict__ probes, ap_octahedron* __restrict__ temp,
                                                   uint32_t n, const int32_t* __restrict__ spec) {
    if (spec) {
        if (spec_off(spec)) return;
        n = static_cast<uint32_t>(spec[kInfoCount]);
    }
    for (uint32_t p = blockIdx.x; p < n; p += gridDim.x) merge_probe(probes[p], temp[p]);
}


template <class T>
int grow(T** p, size_t* cap, size_t need, const char* what) {
    if (*cap >= need) return AP_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), need * sizeof(T));
    if (e != hipSuccess) return hip_fail(e, what);
    *cap = need;
    return AP_OK;
}

int ensure_sort(GuideDevice* g, size_t n) {
    if (g->sort_cap >= n) return AP_OK;
    for (int b = 0; b < 2; ++b) {
        if (g->keys[b]) (void)hipFree(g->keys[b]);
        if (g->vals[b]) (void)hipFree(g->vals[b]);
        g->keys[b] = nullptr;
        g->vals[b] = nullptr;
    }
    for (void* p : {static_cast<void*>(g->flags), static_cast<void*>(g->ranks), static_cast<void*>(g->list),
                    static_cast<void*>(g->uniq), g->tmp})
        if (p) (void)hipFree(p);
    g->flags = g->ranks = g->list = nullptr;
    g->uniq = nullptr;
    g->tmp = nullptr;
    g->sort_cap = 0;
    hipError_t e = hipSuccess;
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        e = hipMalloc(reinterpret_cast<void**>(&g->keys[b]), n * sizeof(uint64_t));
        if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&g->vals[b]), n * sizeof(uint32_t));
    }
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&g->flags), n * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&g->ranks), n * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&g->list), n * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&g->uniq), n * sizeof(uint64_t));
    if (e == hipSuccess) {
        g->tmp_bytes = std::max(sort_workspace_bytes(n), scan_workspace_bytes(n));
        e = hipMalloc(&g->tmp, g->tmp_bytes);
    }
    if (e != hipSuccess) return hip_fail(e, "guiding: sort buffers");
    g->sort_cap = n;
    return AP_OK;
}

int grow_probes(GuideDevice* g, int32_t need, hipStream_t st) {
    if (need <= g->cap) return AP_OK;
    // Room for four times the need at the first growth, then doubling: a
    // frame asks for room for all its unique voxels (count + n_uniq), which
    // at 1080p is about twice the probe count, and a reallocation costs a
    // wait and about a millisecond.
    const int64_t grown = std::max<int64_t>(g->cap > 0 ? 2ll * g->cap : 4ll * need, 1024);
    const int32_t cap = static_cast<int32_t>(std::min<int64_t>(std::max<int64_t>(need, grown), INT32_MAX));
    ap_octahedron *np = nullptr, *nt = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&np), static_cast<size_t>(cap) * sizeof(ap_octahedron));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&nt), static_cast<size_t>(cap) * sizeof(ap_octahedron));
    if (e != hipSuccess) {
        if (np) (void)hipFree(np);
        return hip_fail(e, "guiding: probes");
    }
    hipLaunchKernelGGL(zero_probes_kernel, grid_for(static_cast<size_t>(cap) * (sizeof(ap_octahedron) / 16)),
                       dim3(kBlock), 0, st, nt, static_cast<size_t>(cap));
    if (g->count > 0) {
        e = hipMemcpyAsync(np, g->probes, static_cast<size_t>(g->count) * sizeof(ap_octahedron),
                           hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "guiding: probe copy");
    }
    if (g->probes) (void)hipFree(g->probes);
    if (g->temp) (void)hipFree(g->temp);
    g->probes = np;
    g->temp = nt;
    g->cap = cap;
    return AP_OK;
}

VolumeView view_of(const GuideDevice* g) {
    VolumeView v;
    v.volume = g->volume;
    for (int a = 0; a < 3; ++a) {
        v.start[a] = g->start[a];
        v.size[a] = g->size[a];
    }
    return v;
}

}  // namespace

void guide_free(GuideDevice* g) {
    if (!g) return;
    void* ptrs[] = {g->probes, g->temp, g->volume, g->recs, g->recs2, g->slots, g->nvalid, g->bounds_part, g->keys[0], g->keys[1], g->vals[0],
                    g->vals[1], g->flags, g->ranks, g->list, g->uniq, g->tmp, g->info};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (g->info_host) (void)hipHostFree(g->info_host);
    if (g->ev_count) (void)hipEventDestroy(g->ev_count);
    if (g->ev_flag) (void)hipEventDestroy(g->ev_flag);
    delete g;
}

int guide_prepare(GuideDevice** pg, size_t n_paths, int sets) {
    if (!*pg) *pg = new GuideDevice();
    GuideDevice* g = *pg;
    int rc;
    if ((rc = grow(&g->recs, &g->rec_cap, n_paths, "guiding: path records")) != AP_OK) return rc;
    if (sets > 1 && (rc = grow(&g->recs2, &g->rec2_cap, n_paths, "guiding: path records")) != AP_OK) return rc;
    if ((rc = grow(&g->slots, &g->slot_cap, n_paths * PG_COLLECT, "guiding: sample slots")) != AP_OK) return rc;
    if ((rc = grow(&g->nvalid, &g->nvalid_cap, n_paths, "guiding: sample counts")) != AP_OK) return rc;
    if (!g->info) {
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&g->info), 16 * sizeof(int32_t));
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&g->info_host), 16 * sizeof(int32_t));
        if (e == hipSuccess) e = hipMemset(g->info, 0, 16 * sizeof(int32_t));   // no probes on the device yet
        if (e == hipSuccess)
            e = hipMalloc(reinterpret_cast<void**>(&g->bounds_part), kCompactBlocks * 8 * sizeof(int32_t));
        if (e != hipSuccess) return hip_fail(e, "guiding: info");
    }
    return AP_OK;
}

// Probes, index volume and frame set 0's records. Needs no wait: the
// pointers and bounds change only on the host (a volume or probe growth),
// and an index volume exists once a probe does, so the probe count the
// kernels test for zero is max(count, 1) with a volume (the host's count
// may lag the device's by the last update).
GuideView guide_view(GuideDevice* g) {
    GuideView v{};
    if (!g) return v;
    v.probes = g->probes;
    v.probe_count = g->volume ? std::max(g->count, 1) : 0;
    v.volume = g->volume;
    for (int a = 0; a < 3; ++a) {
        v.start[a] = g->start[a];
        v.size[a] = g->size[a];
    }
    v.recs = g->recs;
    return v;
}

GuideRec* guide_recs(GuideDevice* g, int set) { return g ? (set == 1 ? g->recs2 : g->recs) : nullptr; }

ap_path_guiding_sample* guide_slots(GuideDevice* g) { return g ? g->slots : nullptr; }
uint32_t* guide_nvalid(GuideDevice* g) { return g ? g->nvalid : nullptr; }

int guide_reset(GuideDevice* g) {
    if (!g) return AP_OK;
    const int rc = settle(g);
    if (rc != AP_OK) return rc;
    if (g->volume) (void)hipFree(g->volume);
    g->volume = nullptr;
    for (int a = 0; a < 3; ++a) g->start[a] = g->end[a] = g->size[a] = 0;
    g->count = 0;
    if (g->info) {
        const int32_t zero = 0;
        const hipError_t e = hipMemcpy(g->info + kInfoCount, &zero, sizeof(zero), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "guiding: reset");
    }
    return AP_OK;
}

// Takes in the device's probe count after the last update (a wait on its
// read-back, which the stream has normally long passed).
int settle(GuideDevice* g) {
    if (!g || !g->pending) return AP_OK;
    const hipError_t e = hipEventSynchronize(g->ev_count);
    if (e != hipSuccess) return hip_fail(e, "guiding: read new probes");
    g->count = g->info_host[kInfoCount];
    g->pending = false;
    return AP_OK;
}

// Step 1 of a build over the first min(m, cap) valid slots (m on the
// device), then the read-back of the counts the host needs (unique voxels
// and bounds, the valid-slot count and the range flags), all enqueued
// without a wait. Narrow keys take four 8-bit radix passes, wide ones eight.
hipError_t unique_voxels(GuideDevice* g, uint32_t cap, bool wide, hipStream_t st) {
    const uint32_t* m_dev = reinterpret_cast<const uint32_t*>(g->info + 9);
    g->wide = wide;
    if (wide)
        hipLaunchKernelGGL(voxel_keys_kernel<true>, grid_for(cap), dim3(kBlock), 0, st, g->slots, g->list, cap,
                           g->info, g->keys[0], g->vals[0]);
    else
        hipLaunchKernelGGL(voxel_keys_kernel<false>, grid_for(cap), dim3(kBlock), 0, st, g->slots, g->list, cap,
                           g->info, g->keys[0], g->vals[0]);
    hipError_t err = radix_sort_pairs_u64(g->keys[0], g->keys[1], g->vals[0], g->vals[1], cap, 0, wide ? 64 : 32,
                                          g->tmp, st, m_dev);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(run_heads_kernel, grid_for(cap), dim3(kBlock), 0, st, g->keys[1], cap, g->info + 9, g->flags);
    err = scan_exclusive_u32(g->flags, g->ranks, cap, nullptr, g->tmp, st, m_dev);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(unique_voxels_kernel, grid_for(cap), dim3(kBlock), 0, st, g->keys[1], g->flags, g->ranks, cap,
                       g->info + 9, g->uniq, g->info);
    return hipGetLastError();
}

// Probes the texel sort must number: up to `bound`.
int texel_end_bit(int64_t bound) {
    int probe_bits = 1;
    while ((1ll << probe_bits) <= bound) ++probe_bits;
    return std::min(64, 6 + probe_bits + 1);
}

int guide_update_begin(GuideDevice* g, size_t n_paths, hipStream_t st, bool capture) {
    g->open = false;
    const size_t n_slots = n_paths * PG_COLLECT;
    if (n_slots == 0) return AP_OK;
    if (n_slots > 0xffffffffull) return fail(AP_ERR_UNSUPPORTED, "guiding: too many sample slots in one frame");
    const uint32_t n = static_cast<uint32_t>(n_slots);
    int rc;
    if ((rc = ensure_sort(g, n_slots)) != AP_OK) return rc;
    if (!g->ev_count) {
        hipError_t e = hipEventCreateWithFlags(&g->ev_count, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_flag, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "guiding: event");
    }
    hipLaunchKernelGGL(init_info_kernel, dim3(1), dim3(64), 0, st, g->info);
    // 0. the valid slots in slot order from the per-path counts (most of a
    // frame's slots are empty); their count m stays on the device (info[9]).
    const uint32_t np = static_cast<uint32_t>(n_paths);
    hipError_t e = scan_exclusive_u32(g->nvalid, g->ranks, np, nullptr, g->tmp, st);
    if (e != hipSuccess) return hip_fail(e, "guiding: slot scan");
    const unsigned cblocks = std::min(kCompactBlocks, grid_for(np).x);
    hipLaunchKernelGGL(compact_valid_kernel, dim3(cblocks), dim3(kBlock), 0, st, g->slots, g->nvalid, g->ranks, np,
                       g->list, g->info, g->bounds_part);
    hipLaunchKernelGGL(bounds_reduce_kernel, dim3(1), dim3(kBlock), 0, st, g->bounds_part, cblocks, g->info);
    // 1. unique voxels in sorted order, launched for `cap` slots and run
    // over the first min(m, cap).
    const uint32_t cap = std::min(n, std::max(g->sort_hint, 1u << 16));
    if ((e = unique_voxels(g, cap, false, st)) != hipSuccess) return hip_fail(e, "guiding: voxel sort");
    // 2-4 on the device alone when the index volume needs no growth and the
    // probes have room (check_kernel), with every count read on the device;
    // otherwise these kernels stand down and guide_update_end finishes the
    // build on the host's counts. The host reads the check's result back
    // right after it, while the device goes on.
    VolumeBox box;
    for (int a = 0; a < 3; ++a) {
        box.start[a] = g->start[a];
        box.end[a] = g->end[a];
    }
    hipLaunchKernelGGL(check_kernel, dim3(1), dim3(64), 0, st, g->info, cap, g->volume ? 1 : 0, box, g->cap);
    e = hipMemcpyAsync(g->info_host, g->info, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && !capture) e = hipEventRecord(g->ev_flag, st);
    if (e != hipSuccess) return hip_fail(e, "guiding: read counts");
    int32_t* info = g->info;
    const uint32_t* n_uniq_dev = reinterpret_cast<const uint32_t*>(info);
    const VolumeView vv = view_of(g);
    hipLaunchKernelGGL(unindexed_kernel, grid_for(cap), dim3(kBlock), 0, st, g->uniq, cap, KeyDecode{}, vv, g->flags,
                       info);
    if ((e = scan_exclusive_u32(g->flags, g->ranks, cap, nullptr, g->tmp, st, n_uniq_dev)) != hipSuccess)
        return hip_fail(e, "guiding: probe scan");
    hipLaunchKernelGGL(assign_kernel, grid_for(cap), dim3(kBlock), 0, st, g->uniq, g->flags, g->ranks, cap,
                       KeyDecode{}, vv, 0, info, true);
    hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(64), 0, st, info, -1);
    hipLaunchKernelGGL(zero_new_spec_kernel, dim3(1024), dim3(kBlock), 0, st, g->probes, info);
    hipLaunchKernelGGL(texel_keys_kernel, grid_for(cap), dim3(kBlock), 0, st, g->slots, g->list, cap, vv, g->keys[0],
                       g->vals[0], info);
    e = radix_sort_pairs_u64(g->keys[0], g->keys[1], g->vals[0], g->vals[1], cap, 0, texel_end_bit(g->cap), g->tmp,
                             st, reinterpret_cast<const uint32_t*>(info + 9));
    if (e != hipSuccess) return hip_fail(e, "guiding: texel sort");
    hipLaunchKernelGGL(texel_sums_kernel, grid_for(cap), dim3(kBlock), 0, st, g->keys[1], g->vals[1], cap, g->slots,
                       g->temp, info);
    const uint32_t merge_grid = static_cast<uint32_t>(std::max<int64_t>(
        1, std::min<int64_t>(g->cap, std::max<int64_t>(4096, 2ll * g->count + 1024))));
    hipLaunchKernelGGL(merge_kernel, dim3(merge_grid), dim3(64), 0, st, g->probes, g->temp, 0u, info);
    e = hipMemcpyAsync(g->info_host + kInfoCount, info + kInfoCount, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && !capture) e = hipEventRecord(g->ev_count, st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "guiding: update");
    // Captured (a graph replays these launches): no events; the host reads
    // the counters after it has waited for a replay (guide_update_end with
    // synced).
    g->pending = !capture;
    g->open = true;
    g->upd_n = n;
    g->upd_cap = cap;
    return AP_OK;
}

int guide_update_end(GuideDevice* g, hipStream_t st, bool synced) {
    if (!g->open) return AP_OK;
    g->open = false;
    int rc;
    hipError_t e = synced ? hipSuccess : hipEventSynchronize(g->ev_flag);
    if (e != hipSuccess) return hip_fail(e, "guiding: read counts");
    if (g->info_host[8]) return fail(AP_ERR_UNSUPPORTED, "guiding: a voxel coordinate is outside [-2^18, 2^18)");
    const uint32_t m = static_cast<uint32_t>(g->info_host[9]);
    g->sort_hint = static_cast<uint32_t>(std::min<uint64_t>(g->upd_n, 2ull * m));
    if (g->info_host[kInfoHost] == 0) {  // the device finished the build (or there was nothing to build)
        if (synced) {
            g->count = g->info_host[kInfoCount];
            g->pending = false;
        }
        return AP_OK;
    }
    // The host finishes it: wait for the stood-down kernels, take the
    // device's probe count (unchanged), then grow what must grow.
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "guiding: read counts");
    g->count = g->info_host[kInfoCount];
    g->pending = false;
    if (m == 0) return AP_OK;
    if (m > g->upd_cap || g->info_host[10]) {
        // More valid slots than the launch covered, or a voxel range too
        // wide for 32-bit keys: sort them all again (with absolute keys).
        if ((e = unique_voxels(g, m, g->info_host[10] != 0, st)) == hipSuccess)
            e = hipMemcpyAsync(g->info_host, g->info, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "guiding: voxel sort");
    }
    KeyDecode dec{};
    dec.wide = g->wide ? 1 : 0;
    for (int a = 0; a < 3; ++a) dec.start[a] = g->info_host[1 + a];
    dec.ny = static_cast<uint32_t>(g->info_host[5] - g->info_host[2] + 1);
    dec.nz = static_cast<uint32_t>(g->info_host[6] - g->info_host[3] + 1);
    const uint32_t n_uniq = static_cast<uint32_t>(g->info_host[0]);
    if (n_uniq == 0) return AP_OK;
    // 2. grow the index volume to the union of the bounds.
    int32_t ds[3], de[3];
    for (int a = 0; a < 3; ++a) {
        ds[a] = std::min(g->start[a], g->info_host[1 + a]);
        de[a] = std::max(g->end[a], g->info_host[4 + a]);
    }
    const bool changed = ds[0] != g->start[0] || ds[1] != g->start[1] || ds[2] != g->start[2] ||
                         de[0] != g->end[0] || de[1] != g->end[1] || de[2] != g->end[2];
    if (changed || !g->volume) {
        const int32_t nd[3] = {de[0] - ds[0] + 1, de[1] - ds[1] + 1, de[2] - ds[2] + 1};
        const size_t nvox = static_cast<size_t>(nd[0]) * nd[1] * nd[2];
        const size_t elems = nvox * PG_FACES;
        if (elems > (size_t{1} << 33)) return fail(AP_ERR_OUT_OF_MEMORY, "guiding: index volume too large");
        int32_t* nv = nullptr;
        e = hipMalloc(reinterpret_cast<void**>(&nv), elems * sizeof(int32_t));
        if (e != hipSuccess) return hip_fail(e, "guiding: index volume");
        hipLaunchKernelGGL(fill_i32_kernel, grid_for(elems), dim3(kBlock), 0, st, nv, elems, -1);
        if (g->volume) {
            VolumeView oldv = view_of(g), newv;
            newv.volume = nv;
            for (int a = 0; a < 3; ++a) {
                newv.start[a] = ds[a];
                newv.size[a] = nd[a];
            }
            const size_t old_vox = static_cast<size_t>(g->size[0]) * g->size[1] * g->size[2];
            hipLaunchKernelGGL(copy_volume_kernel, grid_for(old_vox), dim3(kBlock), 0, st, g->volume, oldv, newv);
            e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                (void)hipFree(nv);
                return hip_fail(e, "guiding: volume copy");
            }
            (void)hipFree(g->volume);
        }
        g->volume = nv;
        for (int a = 0; a < 3; ++a) {
            g->start[a] = ds[a];
            g->end[a] = de[a];
            g->size[a] = nd[a];
        }
    }
    // Probes for the voxels without one, numbered in sorted order: at most
    // n_uniq new ones, so room for them is made before their number is known.
    if ((rc = grow_probes(g, g->count + static_cast<int32_t>(n_uniq), st)) != AP_OK) return rc;
    hipLaunchKernelGGL(unindexed_kernel, grid_for(n_uniq), dim3(kBlock), 0, st, g->uniq, n_uniq, dec, view_of(g),
                       g->flags, nullptr);
    e = scan_exclusive_u32(g->flags, g->ranks, n_uniq, nullptr, g->tmp, st);
    if (e != hipSuccess) return hip_fail(e, "guiding: probe scan");
    hipLaunchKernelGGL(assign_kernel, grid_for(n_uniq), dim3(kBlock), 0, st, g->uniq, g->flags, g->ranks, n_uniq,
                       dec, view_of(g), g->count, g->info, false);
    hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(64), 0, st, g->info, g->count);
    hipLaunchKernelGGL(zero_new_probes_kernel,
                       grid_for(static_cast<size_t>(n_uniq) * (sizeof(ap_octahedron) / 16)), dim3(kBlock), 0, st,
                       g->probes + g->count, static_cast<size_t>(n_uniq), g->info);
    e = hipMemcpyAsync(g->info_host + kInfoCount, g->info + kInfoCount, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(g->ev_count, st);
    if (e != hipSuccess) return hip_fail(e, "guiding: new probes");
    // Probes [count, count + n_uniq) cover the new ones; those past them have
    // no samples and no texel keys.
    const int32_t bound = g->count + static_cast<int32_t>(n_uniq);
    // 3. per (probe, texel) sums in slot order.
    hipLaunchKernelGGL(texel_keys_kernel, grid_for(m), dim3(kBlock), 0, st, g->slots, g->list, m, view_of(g),
                       g->keys[0], g->vals[0], nullptr);
    e = radix_sort_pairs_u64(g->keys[0], g->keys[1], g->vals[0], g->vals[1], m, 0, texel_end_bit(bound), g->tmp, st);
    if (e != hipSuccess) return hip_fail(e, "guiding: texel sort");
    hipLaunchKernelGGL(texel_sums_kernel, grid_for(m), dim3(kBlock), 0, st, g->keys[1], g->vals[1], m, g->slots,
                       g->temp, nullptr);
    // 4. FP64 merge and mips (a probe without samples this frame is left as it is).
    hipLaunchKernelGGL(merge_kernel, dim3(static_cast<unsigned>(bound)), dim3(64), 0, st, g->probes, g->temp,
                       static_cast<uint32_t>(bound), nullptr);
    g->pending = true;
    e = hipGetLastError();
    return e == hipSuccess ? AP_OK : hip_fail(e, "guiding: update");
}

int guide_update(GuideDevice* g, size_t n_paths, hipStream_t st) {
    const int rc = guide_update_begin(g, n_paths, st, false);
    if (rc != AP_OK || !g->open) return rc;
    return guide_update_end(g, st, false);
}

// A captured update was replayed and waited for: its counters are on the host.
void guide_replayed(GuideDevice* g) {
    if (g) g->open = true;
}
bool guide_needs_host(const GuideDevice* g) { return g && g->info_host && g->info_host[kInfoHost] != 0; }

int guide_export(GuideDevice* g, ap_octahedron* probes, int32_t* volume, int32_t* bounds, int32_t* count) {
    const int rc = settle(g);
    if (rc != AP_OK) return rc;
    if (count) *count = g ? g->count : 0;
    if (bounds) {
        for (int a = 0; a < 3; ++a) {
            bounds[a] = g ? g->start[a] : 0;
            bounds[3 + a] = g ? g->end[a] : 0;
            bounds[6 + a] = g ? g->size[a] : 0;
        }
    }
    if (!g) return AP_OK;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && probes && g->count > 0)
        e = hipMemcpy(probes, g->probes, static_cast<size_t>(g->count) * sizeof(ap_octahedron), hipMemcpyDeviceToHost);
    if (e == hipSuccess && volume && g->volume)
        e = hipMemcpy(volume, g->volume,
                      static_cast<size_t>(g->size[0]) * g->size[1] * g->size[2] * PG_FACES * sizeof(int32_t),
                      hipMemcpyDeviceToHost);
    return e == hipSuccess ? AP_OK : hip_fail(e, "guiding: export");
}

}  // namespace ap

using namespace ap;

extern "C" {

int ap_guiding_reset(ap_scene* scene) {
    if (!scene) return fail(AP_ERR_INVALID_ARG, "ap_guiding_reset: NULL scene");
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(e, "ap_guiding_reset");
    return guide_reset(scene->guide);
}

int ap_guiding_export(const ap_scene* scene, ap_octahedron* probes, size_t probe_capacity, int32_t* volume,
                      size_t volume_capacity, int32_t* bounds, int32_t* probe_count) {
    if (!scene) return fail(AP_ERR_INVALID_ARG, "ap_guiding_export: NULL scene");
    int32_t b[9], count = 0;
    int rc = guide_export(scene->guide, nullptr, nullptr, b, &count);
    if (rc != AP_OK) return rc;
    const size_t vol = static_cast<size_t>(b[6]) * b[7] * b[8] * PG_FACES;
    if (probes && probe_capacity < static_cast<size_t>(count))
        return fail(AP_ERR_INVALID_ARG, "ap_guiding_export: probe array too small");
    if (volume && volume_capacity < vol) return fail(AP_ERR_INVALID_ARG, "ap_guiding_export: volume array too small");
    rc = guide_export(scene->guide, probes, volume, nullptr, nullptr);
    if (rc != AP_OK) return rc;
    if (bounds) std::memcpy(bounds, b, sizeof(b));
    if (probe_count) *probe_count = count;
    return AP_OK;
}

}  // extern "C"
