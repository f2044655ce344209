// kp_compact.h -- device sorting helpers shared inside the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace kp {

// Stable LSD radix sort (8-bit digits) of n non-negative keys below 2^bits
// (bits in [1, 30]), giving the permutation: order[p] = index of the p-th key
// in ascending order.  Kernels only, on `stream` (capturable into a graph);
// ws needs kp_compact_workspace_bytes(n, 1) bytes; n < 2^31.  With n_dev,
// only the first min(n, *n_dev) keys are sorted (device-side count; order
// past them is left unwritten).
void sort_order(const int32_t *keys, uint64_t n, int bits, int32_t *order, void *ws,
                size_t ws_bytes, hipStream_t stream,
                const unsigned long long *n_dev = nullptr);

// sort_order without the copy of the permutation: *sorted_idx and
// *sorted_keys point into ws, at the permutation and at the keys in sorted
// order, valid until ws is used again.
void sort_in_place(const int32_t *keys, uint64_t n, int bits, void *ws, size_t ws_bytes,
                   hipStream_t stream, const unsigned long long *n_dev,
                   const int32_t **sorted_idx, const uint32_t **sorted_keys);

} // namespace kp
