// Device-wide exclusive scan and stable LSD radix sort (ap_sort.h).
//
// Scan: three launches. Blocks of 256 lanes own 2048 consecutive items (8
// per lane, contiguous): (1) each block's total; (2) one 1024-lane block scans
// the totals in chunks, carrying the running sum; (3) each block scans its
// items (lane totals by a wave64 shuffle scan, the four waves' totals through
// LDS) and adds its offset. Up to 65,536 items (the radix sort's digit
// counts, for instance) one 1024-lane block scans them all in one launch.
//
// Radix sort, per 8-bit digit: (1) each block of 256 lanes histograms a tile
// of 4096 keys in LDS and writes its counts digit-major (count[d * blocks +
// b]); (2) the counts are scanned, giving each (digit, block) its first output
// position; (3) each wave of a block owns 1024 consecutive keys of the tile
// and walks them in 16 steps of 64: the lanes of a step holding the same digit
// find each other with 8 ballots (one per digit bit), the lowest of them adds
// their number to the wave's LDS counter of that digit and hands the old value
// to the others, so every key learns its position among the wave's keys of its
// digit in input order. The four waves' counts are then prefix-summed per
// digit and the keys, still in registers, are written out. Keys of one digit
// leave a block in input order and blocks follow each other, so every pass is
// stable.
#include "ap_sort.h"

#include "aperture/ap_guiding.h"
#include "ap_internal.h"

namespace ap {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 2048
constexpr int kSortThreads = 256;
constexpr int kSortSteps = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortSteps;  // 4096
constexpr int kWaves = kSortThreads / 64;

inline uint32_t div_up(uint64_t a, uint64_t b) { return static_cast<uint32_t>((a + b - 1) / b); }
inline size_t align256(size_t b) { return (b + 255) & ~size_t{255}; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per lane over the block; the block total in *sum.
template <int THREADS>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sum) {
    __shared__ uint32_t wtot[THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < THREADS / 64; ++k) {
        if (k < w) before += wtot[k];
        total += wtot[k];
    }
    __syncthreads();
    *sum = total;
    return before + inc - v;
}

__device__ __forceinline__ uint32_t count_of(uint32_t n, const uint32_t* dyn_n) {
    return dyn_n ? min(*dyn_n, n) : n;
}

__global__ __launch_bounds__(kScanThreads) void scan_totals_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                                                   const uint32_t* __restrict__ dyn_n,
                                                                   uint32_t* __restrict__ totals) {
    n = count_of(n, dyn_n);
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    (void)block_excl_scan<kScanThreads>(s, &total);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// Exclusive scan of the block totals in place; the grand total to *total.
__global__ __launch_bounds__(1024) void scan_offsets_kernel(uint32_t* __restrict__ totals, uint32_t nb,
                                                            uint32_t* __restrict__ total) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < nb; c += 1024) {
        const uint32_t i = c + threadIdx.x;
        const uint32_t v = i < nb ? totals[i] : 0u;
        uint32_t chunk;
        const uint32_t ex = block_excl_scan<1024>(v, &chunk);
        if (i < nb) totals[i] = carry + ex;
        carry += chunk;
    }
    if (total && threadIdx.x == 0) *total = carry;
}

// A scan of at most kSmallScan items in one launch: one 1024-lane block
// walks the items in chunks of 1024 x kSmallItems (kSmallItems consecutive
// items per lane), carrying the running sum; in place is allowed (a lane
// reads its items before it writes them).
constexpr int kSmallItems = 8;
constexpr uint32_t kSmallScan = 65536;
__global__ __launch_bounds__(1024) void scan_small_kernel(const uint32_t* in, uint32_t* out, uint32_t n,
                                                          const uint32_t* __restrict__ dyn_n,
                                                          uint32_t* __restrict__ total) {
    n = count_of(n, dyn_n);
    uint32_t carry = 0;
    for (uint32_t c = 0; c < n; c += 1024u * kSmallItems) {
        const uint32_t base = c + threadIdx.x * kSmallItems;
        uint32_t v[kSmallItems];
        uint32_t s = 0;
        for (int k = 0; k < kSmallItems; ++k) {
            v[k] = base + k < n ? in[base + k] : 0u;
            s += v[k];
        }
        uint32_t chunk;
        uint32_t run = carry + block_excl_scan<1024>(s, &chunk);
        for (int k = 0; k < kSmallItems; ++k) {
            if (base + k < n) out[base + k] = run;
            run += v[k];
        }
        carry += chunk;
    }
    if (total && threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                                                  const uint32_t* __restrict__ dyn_n,
                                                                  const uint32_t* __restrict__ offsets,
                                                                  uint32_t* __restrict__ out) {
    n = count_of(n, dyn_n);
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan<kScanThreads>(s, &total) + offsets[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint64_t* __restrict__ keys, uint32_t n,
                                                                  const uint32_t* __restrict__ dyn_n, int shift,
                                                                  uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[256];
    n = count_of(n, dyn_n);
    hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
    for (int k = 0; k < kSortSteps; ++k) {
        const uint64_t e = base + static_cast<uint64_t>(k) * kSortThreads + threadIdx.x;
        if (e < n) atomicAdd(&hist[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[static_cast<size_t>(threadIdx.x) * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint64_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint64_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, uint32_t n, const uint32_t* __restrict__ dyn_n, int shift,
    const uint32_t* __restrict__ offsets) {
    __shared__ uint32_t cnt[kWaves][256];
    n = count_of(n, dyn_n);
    __shared__ uint32_t base_of[kWaves][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 256; i += kSortThreads) (&cnt[0][0])[i] = 0u;
    __syncthreads();
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t sub = static_cast<uint64_t>(blockIdx.x) * kSortTile + static_cast<uint64_t>(w) * (kSortTile / kWaves);
    uint64_t key[kSortSteps];
    uint32_t val[kSortSteps], pos[kSortSteps];
    for (int k = 0; k < kSortSteps; ++k) {
        const uint64_t e = sub + static_cast<uint64_t>(k) * 64 + lane;
        const bool valid = e < n;
        key[k] = valid ? keys_in[e] : 0ull;
        val[k] = valid ? vals_in[e] : 0u;
        const uint32_t d = static_cast<uint32_t>(key[k] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const int leader = __ffsll(static_cast<unsigned long long>(peers)) - 1;
        uint32_t old = 0;
        if (valid && lane == leader) old = atomicAdd(&cnt[w][d], static_cast<uint32_t>(__popcll(peers)));
        old = static_cast<uint32_t>(__shfl(static_cast<int>(old), valid ? leader : lane, 64));
        pos[k] = old + static_cast<uint32_t>(__popcll(peers & below));
    }
    __syncthreads();
    {
        const int d = threadIdx.x;  // 256 lanes, one digit each
        uint32_t run = offsets[static_cast<size_t>(d) * gridDim.x + blockIdx.x];
        for (int v = 0; v < kWaves; ++v) {
            base_of[v][d] = run;
            run += cnt[v][d];
        }
    }
    __syncthreads();
    for (int k = 0; k < kSortSteps; ++k) {
        const uint64_t e = sub + static_cast<uint64_t>(k) * 64 + lane;
        if (e >= n) continue;
        const uint32_t d = static_cast<uint32_t>(key[k] >> shift) & 255u;
        const uint32_t o = base_of[w][d] + pos[k];
        keys_out[o] = key[k];
        vals_out[o] = val[k];
    }
}

__global__ void set_u32_kernel(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0) *p = v;
}

}  // namespace

size_t scan_workspace_bytes(size_t n) { return align256((div_up(n, kScanTile) + 1) * sizeof(uint32_t)); }

size_t sort_workspace_bytes(size_t n) {
    const size_t counts = static_cast<size_t>(256) * div_up(n, kSortTile) + 1;
    return align256(counts * sizeof(uint32_t)) + scan_workspace_bytes(counts) + align256(n * sizeof(uint64_t)) +
           align256(n * sizeof(uint32_t));
}

hipError_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* work,
                              hipStream_t stream, const uint32_t* dyn_n) {
    if (n == 0) {
        if (total) hipLaunchKernelGGL(set_u32_kernel, dim3(1), dim3(64), 0, stream, total, 0u);
        return hipGetLastError();
    }
    if (n <= kSmallScan) {
        hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(1024), 0, stream, in, out, n, dyn_n, total);
        return hipGetLastError();
    }
    const uint32_t nb = div_up(n, kScanTile);
    auto* totals = static_cast<uint32_t*>(work);
    hipLaunchKernelGGL(scan_totals_kernel, dim3(nb), dim3(kScanThreads), 0, stream, in, n, dyn_n, totals);
    hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, stream, totals, nb, total);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(kScanThreads), 0, stream, in, n, dyn_n, totals, out);
    return hipGetLastError();
}

hipError_t radix_sort_pairs_u64(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                                uint32_t* vals_out, uint32_t n, int begin_bit, int end_bit, void* work,
                                hipStream_t stream, const uint32_t* dyn_n) {
    if (n == 0) return hipSuccess;
    if (begin_bit < 0 || end_bit > 64 || end_bit <= begin_bit) return hipErrorInvalidValue;
    const uint32_t nb = div_up(n, kSortTile);
    const size_t n_counts = static_cast<size_t>(256) * nb;
    auto* p = static_cast<char*>(work);
    auto* counts = reinterpret_cast<uint32_t*>(p);
    p += align256((n_counts + 1) * sizeof(uint32_t));
    void* scan_work = p;
    p += scan_workspace_bytes(n_counts + 1);
    auto* tmp_keys = reinterpret_cast<uint64_t*>(p);
    p += align256(n * sizeof(uint64_t));
    auto* tmp_vals = reinterpret_cast<uint32_t*>(p);
    const int passes = (end_bit - begin_bit + 7) / 8;
    const uint64_t* src_k = keys_in;
    const uint32_t* src_v = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        // The last pass lands in the outputs; the ones before alternate so that it does.
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        uint64_t* dst_k = to_out ? keys_out : tmp_keys;
        uint32_t* dst_v = to_out ? vals_out : tmp_vals;
        const int shift = begin_bit + 8 * pass;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(kSortThreads), 0, stream, src_k, n, dyn_n, shift, counts);
        hipError_t e = scan_exclusive_u32(counts, counts, static_cast<uint32_t>(n_counts), nullptr, scan_work, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(kSortThreads), 0, stream, src_k, src_v, dst_k, dst_v,
                           n, dyn_n, shift, counts);
        src_k = dst_k;
        src_v = dst_v;
    }
    return hipGetLastError();
}

}  // namespace ap

extern "C" {

int ap_guiding_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total) {
    using namespace ap;
    if ((n > 0 && (!in || !out)) || n > 0xffffffffull) return fail(AP_ERR_INVALID_ARG, "ap_guiding_scan_u32: bad arguments");
    void* work = nullptr;
    hipError_t e = hipMalloc(&work, scan_workspace_bytes(n) + 256);
    if (e == hipSuccess) e = scan_exclusive_u32(in, out, static_cast<uint32_t>(n), total, work, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (work) (void)hipFree(work);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_guiding_scan_u32");
}

int ap_guiding_sort_pairs_u64(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in, uint32_t* vals_out,
                              size_t n, int32_t begin_bit, int32_t end_bit) {
    using namespace ap;
    if ((n > 0 && (!keys_in || !keys_out || !vals_in || !vals_out)) || n > 0xffffffffull || begin_bit < 0 ||
        end_bit > 64 || end_bit <= begin_bit)
        return fail(AP_ERR_INVALID_ARG, "ap_guiding_sort_pairs_u64: bad arguments");
    void* work = nullptr;
    hipError_t e = hipMalloc(&work, sort_workspace_bytes(n) + 256);
    if (e == hipSuccess)
        e = radix_sort_pairs_u64(keys_in, keys_out, vals_in, vals_out, static_cast<uint32_t>(n), begin_bit, end_bit,
                                 work, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (work) (void)hipFree(work);
    return e == hipSuccess ? AP_OK : hip_fail(e, "ap_guiding_sort_pairs_u64");
}

}  // extern "C"
