// Device-wide scan and stable radix sort for wave64 (gfx950), written for the
// guiding build: its slot compaction, voxel sort and unique, and the
// (probe, texel) sort whose runs are summed in slot order. All calls are
// asynchronous on `stream` and use only the caller's workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ap {

// Workspace bytes for scan_exclusive_u32 / radix_sort_pairs_u64 over n items.
size_t scan_workspace_bytes(size_t n);
size_t sort_workspace_bytes(size_t n);

// out[i] = in[0] + ... + in[i-1] for i < n (out may equal in); the total of
// all n inputs goes to *total when total is not null (device memory).
// With dyn_n (device memory), the item count is min(*dyn_n, n), read by the
// kernels: n sizes the launch and the workspace.
hipError_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total, void* work,
                              hipStream_t stream, const uint32_t* dyn_n = nullptr);

// Stable sort of n (key, value) pairs by key bits [begin_bit, end_bit),
// least significant digit first, 8 bits per pass: equal keys keep their
// input order. The inputs are left unchanged; the sorted pairs are written
// to keys_out / vals_out. dyn_n as for the scan: the first min(*dyn_n, n)
// pairs are sorted.
hipError_t radix_sort_pairs_u64(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                                uint32_t* vals_out, uint32_t n, int begin_bit, int end_bit, void* work,
                                hipStream_t stream, const uint32_t* dyn_n = nullptr);

}  // namespace ap
