// kp_compact.hip -- device-side packet-pool compaction and per-destination
// binning (wave ballot / prefix primitives, deterministic and stable).
//
// kp_pool_keep replaces pool_t::keep_pool (types/particle_pool.h:191-248):
// survivors end up in [0, n_keep), tail survivors fill the head's holes.
// kp_bin_by_key replaces comm_t::reg + cpy_d (boundary/comm.h:118-127,
// 226-239): slot claiming per destination, then the gather of the records.
// The reference draws positions with atomics (completion order); here they
// are ranks in index order, so every result is reproducible bit for bit.
//
// Tiles of kTile = 256 threads x 16 items.  Counts and offsets are 32-bit
// (pools below 2^31 records).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "kp_compact.h"
#include "kp_error.h"

#define KP_HIPC(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw kp::error(KP_ERR_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kScanChunk = 4096; // elements per block of the hierarchical scan

__device__ __forceinline__ uint64_t lanemask_lt()
{
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// exclusive scan of one uint32 per thread over the 256-thread block;
// returns the block total in *total
__device__ __forceinline__ uint32_t block_exscan(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t wsum[kThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        wsum[w] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) {
        if (k < w)
            base += wsum[k];
        tot += wsum[k];
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// ------------------------------------------------------------ scan (u32)
// hierarchical exclusive scan: chunk sums, scan of the sums, chunk scans
__global__ void __launch_bounds__(kThreads) k_chunk_sum(const uint32_t *in, uint64_t n,
                                                      uint32_t *sums)
{
    const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk;
    uint32_t s = 0;
    for (int k = threadIdx.x; k < kScanChunk; k += kThreads)
        if (c0 + k < n)
            s += in[c0 + k];
    uint32_t tot;
    (void)block_exscan(s, &tot);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = tot;
}

// single block: exclusive scan of m chunk sums in place, total to *total
__global__ void __launch_bounds__(kThreads) k_scan_sums(uint32_t *sums, uint64_t m,
                                                      uint32_t *total)
{
    uint32_t carry = 0;
    for (uint64_t c0 = 0; c0 < m; c0 += kThreads) {
        const uint64_t i = c0 + threadIdx.x;
        const uint32_t v = i < m ? sums[i] : 0;
        uint32_t tot;
        const uint32_t ex = block_exscan(v, &tot);
        if (i < m)
            sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total)
        *total = carry;
}

// out[i] = exclusive prefix of in[] (in and out may alias)
__global__ void __launch_bounds__(kThreads) k_chunk_scan(const uint32_t *in, uint32_t *out,
                                                       uint64_t n, const uint32_t *sums)
{
    const uint64_t c0 = (uint64_t)blockIdx.x * kScanChunk;
    constexpr int per = kScanChunk / kThreads; // consecutive items per thread
    uint32_t v[per];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const uint64_t i = c0 + (uint64_t)threadIdx.x * per + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + block_exscan(s, &tot);
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const uint64_t i = c0 + (uint64_t)threadIdx.x * per + k;
        if (i < n)
            out[i] = run;
        run += v[k];
    }
}

struct scan_ws {
    uint32_t *sums;
    uint64_t n_chunks;
};

constexpr uint64_t scan_chunks(uint64_t n) { return (n + kScanChunk - 1) / kScanChunk; }

void exscan(uint32_t *data, uint64_t n, const scan_ws &ws, uint32_t *total, hipStream_t s)
{
    const uint64_t m = scan_chunks(n);
    if (m == 0) {
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, s, ws.sums, (uint64_t)0,
                           total);
        return;
    }
    hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)m), dim3(kThreads), 0, s, data, n, ws.sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, s, ws.sums, m, total);
    hipLaunchKernelGGL(k_chunk_scan, dim3((unsigned)m), dim3(kThreads), 0, s, data, data, n,
                       ws.sums);
}

// ------------------------------------------------------------ pool keep
// survivors per tile
__global__ void __launch_bounds__(kThreads) k_keep_count(const uint8_t *keep, uint64_t n,
                                                       uint32_t *tile_cnt)
{
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = t0 + (uint64_t)k * kThreads + threadIdx.x;
        c += (i < n && keep[i]) ? 1u : 0u;
    }
    uint32_t tot;
    (void)block_exscan(c, &tot);
    if (threadIdx.x == 0)
        tile_cnt[blockIdx.x] = tot;
}

// holes of the head [0, n_keep) ascending -> hole_at[j];
// survivors of the tail descending -> mover_at[j]
__global__ void __launch_bounds__(kThreads) k_keep_index(const uint8_t *keep, uint64_t n,
                                                       const uint32_t *tile_off,
                                                       const uint32_t *n_keep_p,
                                                       uint32_t *hole_at, uint32_t *mover_at,
                                                       uint32_t *n_holes_p)
{
    const uint32_t n_keep = *n_keep_p;
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
    // items in index order: thread-major inside the tile
    bool f[kItems];
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = t0 + (uint64_t)threadIdx.x * kItems + k;
        f[k] = i < n && keep[i];
        c += f[k] ? 1u : 0u;
    }
    uint32_t tot;
    uint32_t before = tile_off[blockIdx.x] + block_exscan(c, &tot);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = t0 + (uint64_t)threadIdx.x * kItems + k;
        if (i < n) {
            if (i < n_keep && !f[k])
                hole_at[i - before] = (uint32_t)i;
            else if (i >= n_keep && f[k])
                mover_at[n_keep - before - 1] = (uint32_t)i;
            if (i + 1 == n_keep)
                *n_holes_p = n_keep - (before + (f[k] ? 1u : 0u));
        }
        before += f[k] ? 1u : 0u;
    }
}

template <class W>
__global__ void __launch_bounds__(kThreads) k_keep_move(W *pool, uint32_t words,
                                                      const uint32_t *hole_at,
                                                      const uint32_t *mover_at,
                                                      const uint32_t *n_holes_p)
{
    const uint64_t total = (uint64_t)(*n_holes_p) * words;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint64_t j = e / words, w = e - j * words;
        pool[(uint64_t)hole_at[j] * words + w] = pool[(uint64_t)mover_at[j] * words + w];
    }
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

__global__ void __launch_bounds__(kThreads) k_fill_i64(int64_t *p, uint64_t n, int64_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
        p[e] = v;
}

__global__ void k_u32_to_i64(const uint32_t *in, int64_t *out) { *out = (int64_t)*in; }

// ------------------------------------------------------------ binning
// LSD radix passes of 8 bits over (key, index) pairs; key n_bins marks
// "stays" and sorts last.
__device__ __forceinline__ uint32_t bin_key(const int32_t *keys, uint64_t i, int32_t n_bins)
{
    const int32_t k = keys[i];
    return (k >= 0 && k < n_bins) ? (uint32_t)k : (uint32_t)n_bins;
}

// n_dev (optional): a device-side count; only the first min(n, *n_dev) keys
// are sorted (the order past them is not written).
__device__ __forceinline__ uint64_t n_eff(uint64_t n, const unsigned long long *n_dev)
{
    if (!n_dev)
        return n;
    const uint64_t d = *n_dev;
    return d < n ? d : n;
}

template <bool FIRST>
__global__ void __launch_bounds__(kThreads) k_radix_hist(const int32_t *keys0,
                                                       const uint32_t *keys, uint64_t n,
                                                       int32_t n_bins, int shift,
                                                       uint32_t *hist, uint64_t n_tiles,
                                                       const unsigned long long *n_dev)
{
    n = n_eff(n, n_dev);
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = t0 + (uint64_t)k * kThreads + threadIdx.x;
        if (i < n) {
            const uint32_t key = FIRST ? bin_key(keys0, i, n_bins) : keys[i];
            atomicAdd(&h[(key >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * n_tiles + blockIdx.x] = h[threadIdx.x];
}

// stable scatter of one 8-bit digit; items in rounds of 256 in index order,
// ranks inside a wave from eight digit-bit ballots
template <bool FIRST>
__global__ void __launch_bounds__(kThreads) k_radix_scatter(
    const int32_t *keys0, const uint32_t *keys_in, const uint32_t *idx_in, uint32_t *keys_out,
    uint32_t *idx_out, uint64_t n, int32_t n_bins, int shift, const uint32_t *hist_off,
    uint64_t n_tiles, const unsigned long long *n_dev)
{
    n = n_eff(n, n_dev);
    if ((uint64_t)blockIdx.x * kTile >= n)
        return;
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kThreads / 64][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    run[threadIdx.x] = hist_off[(uint64_t)threadIdx.x * n_tiles + blockIdx.x];
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
    const uint64_t lt = lanemask_lt();
    for (int r = 0; r < kItems; ++r) {
#pragma unroll
        for (int q = 0; q < kThreads / 64; ++q)
            wcnt[q][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t i = t0 + (uint64_t)r * kThreads + threadIdx.x;
        const bool valid = i < n;
        uint32_t key = 0, idx = 0;
        if (valid) {
            key = FIRST ? bin_key(keys0, i, n_bins) : keys_in[i];
            idx = FIRST ? (uint32_t)i : idx_in[i];
        }
        const uint32_t d = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0)
            wcnt[w][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t base = run[d];
            for (int q = 0; q < w; ++q)
                base += wcnt[q][d];
            const uint32_t dst = base + rank;
            keys_out[dst] = key;
            idx_out[dst] = idx;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int q = 0; q < kThreads / 64; ++q)
            add += wcnt[q][threadIdx.x];
        run[threadIdx.x] += add;
        // the next round clears wcnt only after every thread read it
        __syncthreads();
    }
    (void)lane;
}

// offsets[b] = first sorted position with key >= b, for b in [0, n_bins]
__global__ void __launch_bounds__(kThreads) k_bin_offsets(const uint32_t *skeys, uint64_t n,
                                                        int32_t n_bins, int64_t *offsets)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p <= n; p += stride) {
        const int64_t kp = p > 0 ? (int64_t)skeys[p - 1] : -1;
        const int64_t kc = p < n ? (int64_t)skeys[p] : (int64_t)n_bins;
        for (int64_t b = kp + 1; b <= kc && b <= n_bins; ++b)
            offsets[b] = (int64_t)p;
    }
}

__global__ void __launch_bounds__(kThreads) k_bin_slots(const uint32_t *sidx, uint64_t n,
                                                      const int64_t *offsets, int32_t n_bins,
                                                      int32_t *order, int32_t *slot)
{
    const uint64_t n_sent = (uint64_t)offsets[n_bins];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const uint32_t i = sidx[p];
        if (p < n_sent) {
            if (order)
                order[p] = (int32_t)i;
            if (slot)
                slot[i] = (int32_t)p;
        } else if (slot) {
            slot[i] = -1;
        }
    }
}

template <class W>
__global__ void __launch_bounds__(kThreads) k_gather(const W *src, uint32_t words,
                                                   const int32_t *order, const int64_t *n_p,
                                                   uint64_t n_max, W *dst)
{
    const uint64_t n = (uint64_t)(*n_p) < n_max ? (uint64_t)(*n_p) : n_max;
    const uint64_t total = n * words;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint64_t p = e / words, w = e - p * words;
        dst[e] = src[(uint64_t)order[p] * words + w];
    }
}

// ------------------------------------------------------------ workspace
constexpr size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

struct keep_layout {
    size_t tile_cnt, sums, scalars, hole_at, mover_at, total;
};
keep_layout keep_ws(uint64_t n)
{
    const uint64_t nt = (n + kTile - 1) / kTile;
    keep_layout L{};
    size_t o = 0;
    L.tile_cnt = o;
    o += align_up(sizeof(uint32_t) * (nt + 1));
    L.sums = o;
    o += align_up(sizeof(uint32_t) * (scan_chunks(nt) + 1));
    L.scalars = o;
    o += align_up(sizeof(uint32_t) * 4);
    L.hole_at = o;
    o += align_up(sizeof(uint32_t) * (n / 2 + 1));
    L.mover_at = o;
    o += align_up(sizeof(uint32_t) * (n / 2 + 1));
    L.total = o;
    return L;
}

struct bin_layout {
    size_t ka, kb, ia, ib, hist, sums, total;
};
constexpr bin_layout bin_ws(uint64_t n)
{
    const uint64_t nt = (n + kTile - 1) / kTile;
    bin_layout L{};
    size_t o = 0;
    L.ka = o;
    o += align_up(sizeof(uint32_t) * (n + 1));
    L.kb = o;
    o += align_up(sizeof(uint32_t) * (n + 1));
    L.ia = o;
    o += align_up(sizeof(uint32_t) * (n + 1));
    L.ib = o;
    o += align_up(sizeof(uint32_t) * (n + 1));
    L.hist = o;
    o += align_up(sizeof(uint32_t) * (256 * nt + 1));
    L.sums = o;
    o += align_up(sizeof(uint32_t) * (scan_chunks(256 * nt) + 1));
    L.total = o;
    return L;
}

unsigned grid_for(uint64_t work)
{
    uint64_t b = (work + kThreads - 1) / kThreads;
    if (b > 8192)
        b = 8192;
    if (b < 1)
        b = 1;
    return (unsigned)b;
}

// widest word that divides the record size and both base addresses
template <class F> void by_word(int32_t unit, const void *a, const void *b, F &&f)
{
    const uintptr_t al = (uintptr_t)unit | (uintptr_t)a | (uintptr_t)b;
    if (al % 16 == 0)
        f((uint4 *)nullptr, unit / 16);
    else if (al % 8 == 0)
        f((uint2 *)nullptr, unit / 8);
    else
        f((uint32_t *)nullptr, unit / 4);
}

__global__ void __launch_bounds__(kThreads) k_copy_order(const uint32_t *sidx, uint64_t n,
                                                       int32_t *order,
                                                       const unsigned long long *n_dev)
{
    n = n_eff(n, n_dev);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride)
        order[p] = (int32_t)sidx[p];
}

// Body of kp_bin_by_key on a workspace of ws_bytes at ws (device memory).
void bin_by_key_run(const char *who, const int32_t *keys, int64_t n, int32_t n_bins,
                    int64_t *offsets, int32_t *order, int32_t *slot, void *ws, size_t ws_bytes,
                    hipStream_t s)
{
    const std::string w(who);
    if (n < 0 || n >= (1ll << 31))
        throw kp::error(KP_ERR_ARG, w + ": n must be in [0, 2^31)");
    if (n_bins < 1 || n_bins >= (1 << 30))
        throw kp::error(KP_ERR_ARG, w + ": n_bins must be in [1, 2^30)");
    if (!offsets || (n > 0 && (!keys || !ws)))
        throw kp::error(KP_ERR_ARG, w + ": null argument");
    const bin_layout L = bin_ws((uint64_t)n);
    if (n > 0 && ws_bytes < L.total)
        throw kp::error(KP_ERR_ARG, w + ": workspace too small (see "
                                        "kp_compact_workspace_bytes)");
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_i64, dim3(grid_for((uint64_t)n_bins + 1)), dim3(kThreads),
                           0, s, offsets, (uint64_t)n_bins + 1, (int64_t)0);
        KP_HIPC(hipGetLastError());
        return;
    }
    char *w8 = (char *)ws;
    uint32_t *ka = (uint32_t *)(w8 + L.ka), *kb = (uint32_t *)(w8 + L.kb);
    uint32_t *ia = (uint32_t *)(w8 + L.ia), *ib = (uint32_t *)(w8 + L.ib);
    uint32_t *hist = (uint32_t *)(w8 + L.hist);
    const uint64_t nt = ((uint64_t)n + kTile - 1) / kTile;
    scan_ws sw{(uint32_t *)(w8 + L.sums), scan_chunks(256 * nt)};
    // key n_bins ("stays") needs the bits of n_bins itself
    int bits = 1;
    while (bits < 31 && ((uint64_t)n_bins >> bits) != 0)
        ++bits;
    const int passes = (bits + 7) / 8;
    const uint32_t *kin = nullptr, *iin = nullptr;
    uint32_t *kout = ka, *iout = ia;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        if (p == 0)
            hipLaunchKernelGGL(k_radix_hist<true>, dim3((unsigned)nt), dim3(kThreads), 0, s,
                               keys, (const uint32_t *)nullptr, (uint64_t)n, n_bins, shift,
                               hist, nt, (const unsigned long long *)nullptr);
        else
            hipLaunchKernelGGL(k_radix_hist<false>, dim3((unsigned)nt), dim3(kThreads), 0, s,
                               keys, kin, (uint64_t)n, n_bins, shift, hist, nt,
                               (const unsigned long long *)nullptr);
        exscan(hist, 256 * nt, sw, nullptr, s);
        if (p == 0)
            hipLaunchKernelGGL(k_radix_scatter<true>, dim3((unsigned)nt), dim3(kThreads), 0, s,
                               keys, (const uint32_t *)nullptr, (const uint32_t *)nullptr,
                               kout, iout, (uint64_t)n, n_bins, shift, hist, nt,
                               (const unsigned long long *)nullptr);
        else
            hipLaunchKernelGGL(k_radix_scatter<false>, dim3((unsigned)nt), dim3(kThreads), 0, s,
                               keys, kin, iin, kout, iout, (uint64_t)n, n_bins, shift, hist,
                               nt, (const unsigned long long *)nullptr);
        kin = kout;
        iin = iout;
        kout = (kout == ka) ? kb : ka;
        iout = (iout == ia) ? ib : ia;
    }
    hipLaunchKernelGGL(k_bin_offsets, dim3(grid_for((uint64_t)n + 1)), dim3(kThreads), 0, s,
                       kin, (uint64_t)n, n_bins, offsets);
    if (order || slot)
        hipLaunchKernelGGL(k_bin_slots, dim3(grid_for((uint64_t)n)), dim3(kThreads), 0, s, iin,
                           (uint64_t)n, offsets, n_bins, order, slot);
    KP_HIPC(hipGetLastError());
}

} // namespace

// Workspace of kp_bin_by_key_static: a fixed device array, one copy per GPU,
// sized for KP_BIN_STATIC_MAX records (every part of it is written before it
// is read, so nothing needs clearing between calls).  Host code never uses
// its name as a device address: static_scratch() gives this device's copy,
// looked up once per device with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[bin_ws(KP_BIN_STATIC_MAX).total];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

void *static_scratch()
{
    int dev = 0;
    KP_HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw kp::error(KP_ERR_UNSUPPORTED, "kp_bin_by_key_static: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        KP_HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            throw kp::error(KP_ERR_HIP, "kp_bin_by_key_static: no device address for the "
                                        "static workspace");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

} // namespace

namespace {

// The digit passes of sort_order; returns the buffers holding the sorted
// keys and the permutation (inside ws).
void sort_passes(const int32_t *keys, uint64_t n, int bits, void *ws, size_t ws_bytes,
                 hipStream_t s, const unsigned long long *n_dev, const uint32_t **kin_out,
                 const uint32_t **iin_out)
{
    if (n >= (1ull << 31) || bits < 1 || bits > 30)
        throw kp::error(KP_ERR_ARG, "sort_order: n must be below 2^31 and bits in [1,30]");
    const bin_layout L = bin_ws(n);
    if (!keys || !ws || ws_bytes < L.total)
        throw kp::error(KP_ERR_ARG, "sort_order: bad buffers");
    char *w = (char *)ws;
    uint32_t *ka = (uint32_t *)(w + L.ka), *kb = (uint32_t *)(w + L.kb);
    uint32_t *ia = (uint32_t *)(w + L.ia), *ib = (uint32_t *)(w + L.ib);
    uint32_t *hist = (uint32_t *)(w + L.hist);
    const uint64_t nt = (n + kTile - 1) / kTile;
    scan_ws sw{(uint32_t *)(w + L.sums), scan_chunks(256 * nt)};
    const int32_t lim = (int32_t)(1u << 30); // every key passes bin_key unchanged
    const uint32_t *kin = nullptr, *iin = nullptr;
    uint32_t *kout = ka, *iout = ia;
    for (int p = 0; p < (bits + 7) / 8; ++p) {
        const int shift = 8 * p;
        if (p == 0)
            hipLaunchKernelGGL(k_radix_hist<true>, dim3((unsigned)nt), dim3(kThreads), 0, s, keys,
                               (const uint32_t *)nullptr, n, lim, shift, hist, nt, n_dev);
        else
            hipLaunchKernelGGL(k_radix_hist<false>, dim3((unsigned)nt), dim3(kThreads), 0, s,
                               keys, kin, n, lim, shift, hist, nt, n_dev);
        exscan(hist, 256 * nt, sw, nullptr, s);
        if (p == 0)
            hipLaunchKernelGGL(k_radix_scatter<true>, dim3((unsigned)nt), dim3(kThreads), 0, s,
                               keys, (const uint32_t *)nullptr, (const uint32_t *)nullptr, kout,
                               iout, n, lim, shift, hist, nt, n_dev);
        else
            hipLaunchKernelGGL(k_radix_scatter<false>, dim3((unsigned)nt), dim3(kThreads), 0, s,
                               keys, kin, iin, kout, iout, n, lim, shift, hist, nt, n_dev);
        kin = kout;
        iin = iout;
        kout = (kout == ka) ? kb : ka;
        iout = (iout == ia) ? ib : ia;
    }
    KP_HIPC(hipGetLastError());
    *kin_out = kin;
    *iin_out = iin;
}

} // namespace

void kp::sort_order(const int32_t *keys, uint64_t n, int bits, int32_t *order, void *ws,
                    size_t ws_bytes, hipStream_t s, const unsigned long long *n_dev)
{
    if (n == 0)
        return;
    if (!order)
        throw kp::error(KP_ERR_ARG, "sort_order: bad buffers");
    const uint32_t *kin = nullptr, *iin = nullptr;
    sort_passes(keys, n, bits, ws, ws_bytes, s, n_dev, &kin, &iin);
    hipLaunchKernelGGL(k_copy_order, dim3(grid_for(n)), dim3(kThreads), 0, s, iin, n, order,
                       n_dev);
    KP_HIPC(hipGetLastError());
}

void kp::sort_in_place(const int32_t *keys, uint64_t n, int bits, void *ws, size_t ws_bytes,
                       hipStream_t s, const unsigned long long *n_dev,
                       const int32_t **sorted_idx, const uint32_t **sorted_keys)
{
    if (n == 0 || !sorted_idx || !sorted_keys)
        throw kp::error(KP_ERR_ARG, "sort_in_place: empty sort or null output");
    const uint32_t *kin = nullptr, *iin = nullptr;
    sort_passes(keys, n, bits, ws, ws_bytes, s, n_dev, &kin, &iin);
    *sorted_idx = reinterpret_cast<const int32_t *>(iin);
    *sorted_keys = kin;
}

extern "C" {

size_t kp_compact_workspace_bytes(int64_t n, int32_t n_bins)
{
    (void)n_bins;
    if (n < 0)
        return 0;
    const keep_layout a = keep_ws((uint64_t)n);
    const bin_layout b = bin_ws((uint64_t)n);
    return a.total > b.total ? a.total : b.total;
}

kp_status kp_pool_keep(void *pool, int64_t n, int32_t unit_bytes, const uint8_t *keep,
                       int64_t *n_keep_dev, void *ws, size_t ws_bytes, void *stream_)
{
    return kp::guard([&] {
        if (n < 0 || n >= (1ll << 31))
            throw kp::error(KP_ERR_ARG, "kp_pool_keep: n must be in [0, 2^31)");
        if (unit_bytes <= 0 || unit_bytes % 4)
            throw kp::error(KP_ERR_ARG, "kp_pool_keep: unit_bytes must be a positive multiple of 4");
        if (!n_keep_dev || (n > 0 && (!pool || !keep || !ws)))
            throw kp::error(KP_ERR_ARG, "kp_pool_keep: null argument");
        const keep_layout L = keep_ws((uint64_t)n);
        if (n > 0 && ws_bytes < L.total)
            throw kp::error(KP_ERR_ARG, "kp_pool_keep: workspace too small (see "
                                        "kp_compact_workspace_bytes)");
        hipStream_t s = (hipStream_t)stream_;
        if (n == 0) {
            hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(kThreads), 0, s, n_keep_dev,
                               (uint64_t)1, (int64_t)0);
            KP_HIPC(hipGetLastError());
            return;
        }
        char *w = (char *)ws;
        uint32_t *tile_cnt = (uint32_t *)(w + L.tile_cnt);
        uint32_t *scal = (uint32_t *)(w + L.scalars); // [0] n_keep, [1] n_holes
        uint32_t *hole_at = (uint32_t *)(w + L.hole_at);
        uint32_t *mover_at = (uint32_t *)(w + L.mover_at);
        const uint64_t nt = ((uint64_t)n + kTile - 1) / kTile;
        scan_ws sw{(uint32_t *)(w + L.sums), scan_chunks(nt)};
        hipLaunchKernelGGL(k_keep_count, dim3((unsigned)nt), dim3(kThreads), 0, s, keep,
                           (uint64_t)n, tile_cnt);
        exscan(tile_cnt, nt, sw, scal + 0, s);
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, scal + 1, 0u);
        hipLaunchKernelGGL(k_keep_index, dim3((unsigned)nt), dim3(kThreads), 0, s, keep,
                           (uint64_t)n, tile_cnt, scal + 0, hole_at, mover_at, scal + 1);
        if ((uintptr_t)pool % 4)
            throw kp::error(KP_ERR_ARG, "kp_pool_keep: pool must be 4-byte aligned");
        by_word(unit_bytes, pool, pool, [&](auto *tag, int32_t words) {
            using W = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_keep_move<W>, dim3(grid_for((uint64_t)n * words / 2)),
                               dim3(kThreads), 0, s, (W *)pool, (uint32_t)words, hole_at,
                               mover_at, scal + 1);
        });
        hipLaunchKernelGGL(k_u32_to_i64, dim3(1), dim3(1), 0, s, scal + 0, n_keep_dev);
        KP_HIPC(hipGetLastError());
    });
}

kp_status kp_bin_by_key(const int32_t *keys, int64_t n, int32_t n_bins, int64_t *offsets,
                        int32_t *order, int32_t *slot, void *ws, size_t ws_bytes,
                        void *stream_)
{
    return kp::guard([&] {
        bin_by_key_run("kp_bin_by_key", keys, n, n_bins, offsets, order, slot, ws, ws_bytes,
                       (hipStream_t)stream_);
    });
}

kp_status kp_bin_by_key_static(const int32_t *keys, int64_t n, int32_t n_bins,
                               int64_t *offsets, int32_t *order, int32_t *slot, void *stream_)
{
    return kp::guard([&] {
        if (n < 0 || n > (int64_t)KP_BIN_STATIC_MAX)
            throw kp::error(KP_ERR_ARG, "kp_bin_by_key_static: n must be in [0, " +
                                            std::to_string(KP_BIN_STATIC_MAX) +
                                            "] (the static workspace's size); larger runs "
                                            "need kp_bin_by_key with a workspace");
        void *ws = static_scratch();
        bin_by_key_run("kp_bin_by_key_static", keys, n, n_bins, offsets, order, slot, ws,
                       sizeof(scratch), (hipStream_t)stream_);
    });
}

kp_status kp_gather(const void *src, int32_t unit_bytes, const int32_t *order,
                    const int64_t *n_dev, int64_t n_max, void *dst, void *stream_)
{
    return kp::guard([&] {
        if (unit_bytes <= 0 || unit_bytes % 4)
            throw kp::error(KP_ERR_ARG, "kp_gather: unit_bytes must be a positive multiple of 4");
        if (n_max < 0)
            throw kp::error(KP_ERR_ARG, "kp_gather: n_max must be >= 0");
        if (n_max == 0)
            return;
        if (!src || !order || !n_dev || !dst)
            throw kp::error(KP_ERR_ARG, "kp_gather: null argument");
        hipStream_t s = (hipStream_t)stream_;
        if ((uintptr_t)src % 4 || (uintptr_t)dst % 4)
            throw kp::error(KP_ERR_ARG, "kp_gather: src and dst must be 4-byte aligned");
        by_word(unit_bytes, src, dst, [&](auto *tag, int32_t words) {
            using W = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL(k_gather<W>, dim3(grid_for((uint64_t)n_max * words)),
                               dim3(kThreads), 0, s, (const W *)src, (uint32_t)words, order,
                               n_dev, (uint64_t)n_max, (W *)dst);
        });
        KP_HIPC(hipGetLastError());
    });
}

} // extern "C"
