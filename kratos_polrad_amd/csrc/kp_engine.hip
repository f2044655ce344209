// kp_engine.hip -- transport engine: device copies of the block table and
// medium, and the forward launch.
//
// One launch replaces the reference's per-cycle orchestration for photons
// (driver.h:59-79): the generator (generate_par.h:51-62) and the integrator
// (integrator.h:47-106) are fused -- each thread draws its packet's initial
// state from Philox keyed by the packet id instead of reading an emitted
// pool, then runs the whole history (photon.h:104-125 runs the whole march
// inside one ker_itg thread too).  Bands are the n_hv photon-energy bins of
// rad_t::flx (radiation.h:24-35).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "kp_compact.h"
#include "kp_device.h"
#include "kp_error.h"
#include "kp_adjoint.h"
#include "kp_persist.h"
#include "kp_blockq.h"
#include "kp_sph.h"

#define KP_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw kp::error(KP_ERR_HIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    } while (0)

// Device scratch of forward launches: the persistent slot counter and the
// tally / image replicas (zero between launches).  The engine owns one set;
// each captured graph owns its own, so replays never share scratch with
// eager launches.
struct fwd_bufs {
    unsigned long long *work = nullptr;
    double *j_rep = nullptr, *img_rep = nullptr;
    size_t j_cap = 0, img_cap = 0; // elements
    // emission-direction order: keys, permutation, sort workspace
    int32_t *keys = nullptr, *order = nullptr;
    void *ws = nullptr;
    size_t ord_cap = 0, ws_cap = 0; // elements / bytes
    // block queue: two record sets and their keys, the sort permutation,
    // sorted keys, per-key segment starts, sort workspace, record counters,
    // emission peel-off crossings per band
    void *bq_rec[2] = {nullptr, nullptr};
    uint32_t *bq_key[2] = {nullptr, nullptr};
    uint32_t *bq_seg = nullptr, *bq_np = nullptr;
    uint32_t *bq_toff = nullptr; // [n_keys + 1] first task of each key
    uint2 *bq_task = nullptr;    // tasks of a pass: (first sorted position, key)
    void *bq_ws = nullptr;
    // [0], [1] records, [2] queued peel-off rays, [3] tasks of the pass,
    // [4] next task to claim
    unsigned long long *bq_cnt = nullptr;
    size_t bq_cap = 0, bq_rec_bytes = 0, bq_seg_cap = 0, bq_np_cap = 0, bq_ws_cap = 0;
    size_t bq_task_cap = 0;
    void *bq_rays = nullptr;
    size_t bq_rays_bytes = 0;
};

struct kp_engine {
    int device = 0;
    int n_blocks = 0, n_cell = 0, n3 = 0;
    int periodic[3] = {0, 0, 0};
    double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0};
    std::vector<double> xf0_h, dx0_h;
    // device block table
    double *xf0 = nullptr, *dx0 = nullptr, *vol0 = nullptr;
    int32_t *nb_id = nullptr;
    uint8_t *nb_shift = nullptr, *phybnd = nullptr;
    int8_t *d_lvl = nullptr;
    float *h32 = nullptr, *org32 = nullptr, *vol32 = nullptr;
    // spherical mesh: per-block face/volume table (device + host copy)
    int coord = KP_COORD_CARTESIAN;
    double *sph = nullptr;
    std::vector<double> sph_h;
    // medium
    float *kappa = nullptr, *albedo = nullptr, *g = nullptr;
    bool media_set = false;
    // band table (device + pinned staging + host shadow)
    kp::dband *bands_d = nullptr;
    uint64_t *band_start_d = nullptr;
    void *stage = nullptr;
    int bands_cap = 0;
    std::vector<unsigned char> shadow;
    // launch shape
    kp_schedule sched{KP_SCHED_PERSISTENT, 256, 4, 64, 0, 0, 8, 0, 0, 1, 0, 0};
    int n_cu = 0;
    fwd_bufs fb;                        // forward launch scratch
    // adjoint: per-thread rows of peel-off contributions, gradient replicas
    double *adj_scratch = nullptr;
    size_t adj_cap = 0;
    double *g_rep = nullptr;
    size_t g_rep_cap = 0;
    // one-pass adjoint: queue of peel-off rays (records of the run's
    // precision), its sort keys, order and workspace; queue of main-path
    // vertices; per-slot A_k sums; packets left to the two-pass kernel.
    // Queue capacities in bytes (keys and order in records).
    void *pq = nullptr;
    unsigned long long *pq_count = nullptr;
    int32_t *pq_keys = nullptr, *pq_order = nullptr;
    void *pq_ws = nullptr;
    size_t pq_bytes = 0, pq_kcap = 0, pq_ws_cap = 0;
    void *vq = nullptr;
    unsigned long long *vq_count = nullptr;
    size_t vq_bytes = 0;
    double *pk_atot = nullptr;
    int32_t *fb_list = nullptr;
    unsigned int *fb_count = nullptr;
    size_t fb_cap = 0;
    std::map<std::string, kp_schedule> tuned; // autotuner cache
};

// A captured forward run.  Holds its own copy of the band table and its own
// launch scratch; reads the engine's block table and medium (kp_set_media
// rewrites those in place, so a replay sees the current medium).
struct kp_graph {
    kp_engine *e = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    kp::dband *bands = nullptr;
    uint64_t *band_start = nullptr;
    fwd_bufs fb;
};

namespace {

template <class T> T *dmalloc(size_t n)
{
    void *p = nullptr;
    if (n == 0)
        n = 1;
    KP_HIP(hipMalloc(&p, n * sizeof(T)));
    return (T *)p;
}

template <class T> void dfree(T *&p)
{
    if (p)
        (void)hipFree((void *)p);
    p = nullptr;
}

template <class T> void upload(T *dst, const T *src, size_t n)
{
    KP_HIP(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
}

struct device_guard {
    int prev = 0;
    explicit device_guard(int dev)
    {
        KP_HIP(hipGetDevice(&prev));
        if (prev != dev)
            KP_HIP(hipSetDevice(dev));
    }
    ~device_guard() { (void)hipSetDevice(prev); }
};

// wave-level sum of a 64-bit counter, then one atomic per wave
__device__ __forceinline__ void wave_add_u64(unsigned long long *dst, uint64_t v)
{
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    uint64_t sum = 0;
    // reduce in 64-bit using two 32-bit shuffles per step
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t olo = __shfl_xor(lo, off, 64);
        const uint32_t ohi = __shfl_xor(hi, off, 64);
        const uint64_t s = ((uint64_t)hi << 32 | lo) + ((uint64_t)ohi << 32 | olo);
        lo = (uint32_t)s;
        hi = (uint32_t)(s >> 32);
    }
    sum = ((uint64_t)hi << 32) | lo;
    if ((threadIdx.x & 63) == 0 && sum)
        atomicAdd(dst, (unsigned long long)sum);
}

template <class C> __device__ __forceinline__ void flush_counters(const kp::kparams &P, const C &c)
{
    if (!P.counters)
        return;
    wave_add_u64(P.counters + KP_C_CROSS_MAIN, c.cross);
    wave_add_u64(P.counters + KP_C_CROSS_PEEL, c.peel);
    wave_add_u64(P.counters + KP_C_EVENTS, c.events);
    wave_add_u64(P.counters + KP_C_ESCAPED, c.escaped);
    wave_add_u64(P.counters + KP_C_KILLED_RR, c.rr);
    wave_add_u64(P.counters + KP_C_KILLED_STEPS, c.steps);
    wave_add_u64(P.counters + KP_C_LOST, c.lost);
    wave_add_u64(P.counters + KP_C_PACKETS, c.packets);
}

__global__ void __launch_bounds__(256) k_forward_f64(kp::kparams P)
{
    kp::f64::counters_t C;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < P.total;
         k += stride) {
        const uint64_t slot = kp::slot_at(P, k);
        const int band = kp::band_of(P, slot);
        const uint64_t pid = P.packet_offset + (slot - P.band_start[band]);
        kp::f64::packet(P, band, pid, slot, C);
    }
    flush_counters(P, C);
}

__global__ void __launch_bounds__(256) k_forward_f32(kp::kparams P)
{
    kp::f32::counters_t C;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < P.total;
         k += stride) {
        const uint64_t slot = kp::slot_at(P, k);
        const int band = kp::band_of(P, slot);
        const uint64_t pid = P.packet_offset + (slot - P.band_start[band]);
        kp::f32::packet(P, band, pid, slot, C);
    }
    flush_counters(P, C);
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }
__global__ void k_zero_u64x2(unsigned long long *p, unsigned long long *q)
{
    *p = 0ull;
    *q = 0ull;
}
__global__ void k_zero_u32(unsigned int *p) { *p = 0u; }
__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// out[e] += sum_r rep[r][e]; rep[r][e] = 0 (ready for the next launch)
__global__ void __launch_bounds__(256) k_fold(double *out, double *rep, uint64_t n, uint32_t R)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        double s = 0.0;
        for (uint32_t r = 0; r < R; ++r) {
            s += rep[r * n + e];
            rep[r * n + e] = 0.0;
        }
        out[e] += s;
    }
}

// Planar image replicas (I, Q and U planes of np pixel entries each)
// folded into the caller's image, a pixel's I, Q, U side by side:
// out[3p + k] += sum_r rep[r][k][p]; the replicas are left zero.
__global__ void __launch_bounds__(256) k_fold_planar(double *out, double *rep, uint64_t np,
                                                     uint32_t R)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += stride) {
        double s[3] = {0.0, 0.0, 0.0};
        for (uint32_t r = 0; r < R; ++r) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double *q = rep + ((uint64_t)r * 3 + k) * np + p;
                s[k] += *q;
                *q = 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
            out[3 * p + k] += s[k];
    }
}

// Band-major replicas folded into the caller's cell-major J:
// out[c*nb + b] += sum_r rep[r][b*nc + c]; the replicas are left zero.
// 64-cell x 16-band tiles transposed through LDS so that both the replica
// reads (along cells) and the output updates (along bands) are contiguous.
__global__ void __launch_bounds__(256) k_fold_t(double *out, double *rep, uint64_t nc, int nb,
                                                uint32_t R)
{
    __shared__ double tile[16][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const uint64_t n_ct = (nc + 63) / 64;
    const uint64_t n_bt = (uint64_t)(nb + 15) / 16;
    const uint64_t nrep = nc * (uint64_t)nb;
    for (uint64_t t = blockIdx.x; t < n_ct * n_bt; t += gridDim.x) {
        const uint64_t c0 = (t % n_ct) * 64;
        const int b0 = (int)(t / n_ct) * 16;
        for (int k = ty; k < 16; k += 4) {
            const int b = b0 + k;
            const uint64_t c = c0 + tx;
            double sum = 0.0;
            if (b < nb && c < nc) {
                for (uint32_t r = 0; r < R; ++r) {
                    double *q = rep + r * nrep + (uint64_t)b * nc + c;
                    sum += *q;
                    *q = 0.0;
                }
            }
            tile[k][tx] = sum;
        }
        __syncthreads();
        for (int q = threadIdx.x; q < 1024; q += 256) {
            const int bb = q & 15, cc = q >> 4;
            const int b = b0 + bb;
            const uint64_t c = c0 + cc;
            if (b < nb && c < nc)
                out[c * nb + b] += tile[bb][cc];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_fill0(double *p, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride)
        p[e] = 0.0;
}

// Emission-direction sort key of every slot of a launch: the band, then the
// top `bits` bits of the packet's event-0 Philox words 0 (mu = 2u-1) and 1
// (azimuth) -- equal-solid-angle cells of the (mu, phi) square -- interleaved
// in Morton order, so that neighbouring keys are neighbouring directions.
__global__ void __launch_bounds__(256) k_dir_keys(kp::kparams P, int bits, int32_t *keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < P.total; s += stride) {
        const int band = kp::band_of(P, s);
        const uint64_t pid = P.packet_offset + (s - P.band_start[band]);
        uint32_t rn[4] = {(uint32_t)pid, (uint32_t)(pid >> 32), 0u, 0u};
        kp_philox4x32_10(rn, P.seed, (uint32_t)band);
        const uint32_t mu = rn[0] >> (32 - bits), ph = rn[1] >> (32 - bits);
        uint32_t m = 0;
        for (int b = 0; b < bits; ++b)
            m |= (((ph >> b) & 1u) << (2 * b)) | (((mu >> b) & 1u) << (2 * b + 1));
        keys[s] = (int32_t)(((uint32_t)band << (2 * bits)) | m);
    }
}

// Direction bits per axis of the sort key: as many as fit 30 bits beside the
// band (15 for one band: ~2^30 cells, finer than the packets of a launch).
int dir_key_bits(int n_bands, int &total_bits)
{
    int band_bits = 0;
    while ((1 << band_bits) < n_bands)
        ++band_bits;
    const int b = std::min(15, (30 - band_bits) / 2);
    total_bits = band_bits + 2 * b;
    return b;
}

// Spherical-mesh streaming (photon_sph.h): persistent lanes or one thread per
// packet; the source cell's J is summed in LDS when the launch gives it.
template <bool PERSIST> __global__ void __launch_bounds__(256, 4) k_sph(kp::kparams P)
{
    const bool lds = P.lds_n > 0;
    if (lds) {
        for (int k = threadIdx.x; k < P.lds_n; k += blockDim.x)
            kp::kp_lds[k] = 0.0;
        __syncthreads();
    }
    kp::f64::counters_t C;
    if constexpr (PERSIST)
        kp::sph::persist(P, lds, C);
    else
        kp::sph::history(P, lds, C);
    flush_counters(P, C);
    if (lds) {
        __syncthreads();
        for (int b = threadIdx.x; b < P.n_bands; b += blockDim.x)
            if (P.J && kp::kp_lds[b] != 0.0)
                kp::atomic_add_f64(kp::j_at(P, (uint64_t)P.src_gc, b), kp::kp_lds[b]);
    }
}

template <class Pol, int WPS, bool ADJ, bool DEFER = false, bool SYNC = false>
__global__ void __launch_bounds__(256, WPS) k_persist(kp::kparams P)
{
    bool lds = false;
    if constexpr (!ADJ) {
        if (P.lds_n > 0) {
            for (int k = threadIdx.x; k < P.lds_n; k += blockDim.x)
                kp::kp_lds[k] = 0.0;
            __syncthreads();
            lds = true;
        }
    }
    kp::lane_counters C;
    double jsum = 0.0;
    kp::persist_body<Pol, ADJ, DEFER, SYNC>(P, P.work, C, jsum, lds);
    flush_counters(P, C);
    if constexpr (!ADJ) {
        if (P.lds_n > 0) {
            __syncthreads();
            kp::flush_lds_acc<Pol>(P);
        }
    }
    if constexpr (ADJ) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            jsum += __shfl_xor(jsum, off, 64);
        if ((threadIdx.x & 63) == 0 && P.jlin && jsum != 0.0)
            atomicAdd(P.jlin, jsum);
    }
}

// register-budget variants of the persistent kernel
using persist_fn = void (*)(kp::kparams);
template <bool ADJ> persist_fn persist_variant(int precision, int wps)
{
    if (precision == KP_F64) {
        if (wps == 3)
            return k_persist<kp::pol64, 3, ADJ>;
        if (wps == 4)
            return k_persist<kp::pol64, 4, ADJ>;
        return k_persist<kp::pol64, 1, ADJ>;
    }
    if (wps == 1)
        return k_persist<kp::pol32, 1, ADJ>;
    if (wps == 5)
        return k_persist<kp::pol32, 5, ADJ>;
    if (wps == 6)
        return k_persist<kp::pol32, 6, ADJ>;
    return k_persist<kp::pol32, 4, ADJ>; // 128 VGPRs, no spills
}

// mixed forward in emission-direction order: waves refilled together, J
// deposits summed over runs of lanes
persist_fn sync_variant() { return k_persist<kp::pol32, 4, false, false, true>; }

// one-pass replay (deferred deposits)
persist_fn defer_variant(int precision, int wps)
{
    if (precision == KP_F64)
        return k_persist<kp::pol64, 1, true, true>;
    if (wps == 1)
        return k_persist<kp::pol32, 1, true, true>;
    return k_persist<kp::pol32, 4, true, true>;
}


// ---------------------------------------------------------------- block queue
// Emission peel-offs of the queue bands: every packet of a band starts at the
// source with the same Stokes vector, so its emission peel-off rays are one
// ray per observer.  One thread per band marches them, adds n_packets times
// the packet's deposit at the source pixel and records the rays' crossings
// (each packet's peel crossing count starts with them).
template <class Pol>
__global__ void __launch_bounds__(64) k_bq_emit_peel(kp::kparams P, kp::bq::qparams Q)
{
    using real = typename Pol::real;
    const int b = Q.first_band + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= P.n_bands)
        return;
    const uint64_t n = P.band_start[b + 1] - P.band_start[b];
    uint64_t np = 0;
    if (n > 0 && P.n_incl > 0) {
        typename Pol::ray r;
        r.d[0] = 0;
        r.d[1] = 0;
        r.d[2] = 1;
        Pol::at_source(P, r);
        const real ks = (real)P.bands[b].kappa_scale;
        for (int o = 0; o < P.n_incl; ++o) {
            double v;
            long px;
            if constexpr (sizeof(real) == 8) {
                const double tau = kp::f64::tau_to_edge(P, r, P.no[o], ks, np);
                v = P.w0 * (kp_exp(-tau) * KP_INV_4PI);
                px = kp::f64::pixel(P, b, o, r.x);
            } else {
                float no[3];
                kp::load3(P.no[o], no);
                const double tau = kp::f32::tau_to_edge(P, r, no, ks, np);
                const float att = kp::f32::fast_exp(-(float)tau) * kp::f32::kInv4Pi;
                v = (double)((float)P.w0 * att);
                float gx[3];
                kp::f32::global_pos(P, r, gx);
                px = kp::f32::pixel(P, b, o, gx);
            }
            if (P.image && px >= 0)
                kp::atomic_add_f64(kp::img_at(P, px, 0), (double)n * v);
        }
    }
    Q.emit_np[b] = (uint32_t)np;
}

// Records of the queue bands' packets at emission, all in the source block.
template <class Pol>
__global__ void __launch_bounds__(256) k_bq_generate(kp::kparams P, kp::bq::qparams Q)
{
    using real = typename Pol::real;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < Q.n_slots;
         k += stride) {
        kp::bq::qlane<Pol> L;
        L.slot = Q.first_slot + k;
        L.band = kp::band_of(P, L.slot);
        L.pid = P.packet_offset + (L.slot - P.band_start[L.band]);
        uint32_t rn[4] = {(uint32_t)L.pid, (uint32_t)(L.pid >> 32), 0u, 0u};
        kp_philox4x32_10(rn, P.seed, (uint32_t)L.band);
        if constexpr (sizeof(real) == 8)
            kp::f64::emit(rn, L.r.d, L.e1, L.tau);
        else
            kp::f32::emit(rn, L.r.d, L.e1, L.tau);
        L.I = (real)P.w0;
        L.Q = 0;
        L.U = 0;
        Pol::at_source(P, L.r);
        L.steps = P.step_max;
        L.events = 0;
        L.ncross = 0;
        L.npeel = Q.emit_np[L.band];
        if (P.end_state) // peel crossings are added to it as the rays are marched
            P.end_state[L.slot * KP_ES_STRIDE + 14] = 0.0;
        kp::bq::store<Pol>(L, static_cast<kp::bq::rec_t<real> *>(Q.out)[k]);
        Q.out_key[k] = (uint32_t)(L.band - Q.first_band) * (uint32_t)Q.n_blocks + (uint32_t)P.src_ib;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *Q.n_out = Q.n_slots;
}

// seg[k] = first sorted position whose key is >= k, for k in [0, n_keys]
__global__ void __launch_bounds__(256) k_bq_seg(const uint32_t *skey,
                                                const unsigned long long *n_p, uint32_t n_keys,
                                                uint32_t *seg)
{
    const uint64_t n = *n_p;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= n_keys; k += stride) {
        uint64_t lo = 0, hi = n; // lower bound of k in skey[0, n)
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (skey[mid] < (uint32_t)k)
                lo = mid + 1;
            else
                hi = mid;
        }
        seg[k] = (uint32_t)lo;
    }
}

// Tasks of a pass: each key's run of sorted positions split into pieces of at
// most `chunk`, so a task lies in one block (one staging of its kappa_rho and
// J tile) and the runs of sparse late passes spread over the workgroups
// instead of one workgroup walking many blocks.  toff[k] = number of tasks of
// key k (k_bq_tcount), then its exclusive prefix sum (k_bq_tscan, one
// workgroup; toff[n_keys] and *n_tasks the total), then one thread per task
// finds its key by binary search over toff (k_bq_tfill).
__global__ void __launch_bounds__(256) k_bq_tcount(const uint32_t *seg, uint32_t n_keys,
                                                   uint32_t chunk, uint32_t *toff)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_keys; k += stride)
        toff[k] = (seg[k + 1] - seg[k] + chunk - 1) / chunk;
}

__global__ void __launch_bounds__(1024) k_bq_tscan(uint32_t *toff, uint32_t n_keys,
                                                   unsigned long long *n_tasks,
                                                   unsigned long long *task_next)
{
    __shared__ uint32_t part[1024];
    const uint32_t per = (n_keys + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = min(b + per, n_keys);
    uint32_t sum = 0;
    for (uint32_t k = b; k < e; ++k)
        sum += toff[k];
    part[threadIdx.x] = sum;
    __syncthreads();
    // inclusive scan of the 1024 partial sums (Hillis-Steele)
    for (int d = 1; d < 1024; d <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum; // exclusive
    for (uint32_t k = b; k < e; ++k) {
        const uint32_t c = toff[k];
        toff[k] = run;
        run += c;
    }
    if (threadIdx.x == 1023) {
        toff[n_keys] = part[1023];
        *n_tasks = part[1023];
        *task_next = 0;
    }
}

__global__ void __launch_bounds__(256) k_bq_tfill(const uint32_t *seg, const uint32_t *toff,
                                                  uint32_t n_keys, uint32_t chunk,
                                                  const unsigned long long *n_tasks,
                                                  uint2 *task)
{
    const uint64_t nt = *n_tasks;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
        uint32_t lo = 0, hi = n_keys - 1; // last key with toff[k] <= t
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (toff[mid] <= (uint32_t)t)
                lo = mid;
            else
                hi = mid - 1;
        }
        task[t] = make_uint2(seg[lo] + ((uint32_t)t - toff[lo]) * chunk, lo);
    }
}

// One pass of the block queue: tasks of `chunk` sorted records, which the
// workgroups claim one at a time; per task (one block) the block's
// kappa_rho and a J tile in LDS, each packet marched to
// the block's edge or its end (in the FP64 path by the wave loop pass_block,
// with the block's face-neighbour entries in LDS too; in the mixed path one
// record per lane at a time), the tile added to J.
template <class Pol, int WPS>
__global__ void __launch_bounds__(256, WPS) k_bq_pass(kp::kparams P, kp::bq::qparams Q)
{
    using real = typename Pol::real;
    using tile = kp::bq::tile_t<Pol>;
    tile *lj = reinterpret_cast<tile *>(kp::kp_lds);
    float *lk = reinterpret_cast<float *>(lj + P.n3);
    __shared__ uint32_t s_key;
    __shared__ uint64_t s_t, s_p, s_end;
    __shared__ uint32_t s_next;
    __shared__ kp::blk_nb_t s_nb;
    __shared__ unsigned long long s_cnt[KP_NCOUNTERS];
    kp::bq::wg_counters_zero(s_cnt);
    const uint64_t n = *Q.n_in;
    const auto *in = static_cast<const kp::bq::rec_t<real> *>(Q.in);
    kp::bq::wave_peel_open();
    const uint64_t n_tasks = *Q.n_tasks;
    // tasks are claimed one at a time from a counter, so the workgroups that
    // draw light tasks take more of them
    for (;;) {
        if (threadIdx.x == 0)
            s_t = atomicAdd(Q.task_next, 1ull);
        __syncthreads();
        const uint64_t t = s_t;
        if (t >= n_tasks)
            break;
        {
            if (threadIdx.x == 0) {
                const uint2 tk = Q.tasks[t];
                uint64_t e = Q.seg[tk.y + 1];
                if (e > (uint64_t)tk.x + Q.chunk)
                    e = (uint64_t)tk.x + Q.chunk;
                if (e > n)
                    e = n;
                s_key = tk.y;
                s_p = tk.x;
                s_end = e;
                s_next = 0;
            }
            __syncthreads();
            const uint32_t key = s_key;
            const uint64_t p = s_p;
            const uint64_t e = s_end;
            const int ib = (int)(key % (uint32_t)Q.n_blocks);
            const int band = Q.first_band + (int)(key / (uint32_t)Q.n_blocks);
            const float *kb = P.kappa + (uint64_t)ib * P.n3;
            for (int c = threadIdx.x; c < P.n3; c += blockDim.x) {
                lk[c] = kb[c];
                lj[c] = (tile)0;
            }
            if constexpr (kp::bq::kPassWaveLoop<Pol>)
                kp::stage_nb(P, ib, &s_nb);
            __syncthreads();
            if constexpr (kp::bq::kPassWaveLoop<Pol>) {
                kp::bq::pass_block<Pol>(P, Q, in, p, e, lk, lj, &s_nb, s_cnt, &s_next);
            } else {
                for (uint64_t q = p + threadIdx.x; q < e; q += blockDim.x) {
                    kp::bq::qlane<Pol> L;
                    kp::bq::load<Pol>(P, in[kp::bq::record_at(Q, q)], L);
                    int m = kp::bq::march_block<Pol>(P, Q, L, lk, lj);
                    if (m == kp::bq::MARCH_BLOCK)
                        m = kp::bq::leave_block<Pol>(P, L, kp::nb_global{P});
                    if (m == kp::bq::MARCH_OUT)
                        kp::bq::push_out<Pol>(P, Q, L);
                    else
                        kp::bq::end_packet<Pol>(P, L, s_cnt);
                }
            }
            __syncthreads();
            if (P.J) {
                const uint64_t g0 = (uint64_t)ib * P.n3;
                for (int c = threadIdx.x; c < P.n3; c += blockDim.x) {
                    const double v = (double)lj[c];
                    if (v != 0.0)
                        kp::atomic_add_f64(kp::j_at(P, g0 + c, band), v);
                }
            }
            __syncthreads();
        }
    }
    kp::bq::wave_peel_close<real>(Q);
    __syncthreads();
    kp::bq::wg_counters_flush(P, s_cnt);
}

// The peel-offs queued by a pass: one interaction per lane, in queue order,
// its rays toward the observers one after the other (the interactions a
// wave queued lie in one block, so the lanes' rays run parallel).
// CLAIM false: the rays in grid-stride order over a grid of many more
// workgroups than fit at once.  CLAIM true: each wave claims the next 64
// rays from a counter until none are left, so the waves end together
// however long their rays are; with several observers a lane's work (one
// ray per observer) is long and uneven, and that pays for the claims.
template <class Pol, int WPS, bool CLAIM>
__global__ void __launch_bounds__(256, WPS) k_bq_peel(kp::kparams P, kp::bq::qparams Q)
{
    using real = typename Pol::real;
    kp::lane_counters C;
    const uint64_t n = *Q.pr_count < Q.pr_cap ? *Q.pr_count : Q.pr_cap;
    const auto *rays = static_cast<const kp::bq::ray_rec_t<real> *>(Q.pr);
    if constexpr (CLAIM) {
        const int lane = threadIdx.x & 63;
        for (;;) {
            unsigned long long base = 0;
            if (lane == 0)
                base = atomicAdd(Q.peel_next, 64ull);
            base = __shfl(base, 0, 64);
            if (base >= n)
                break;
            const uint64_t k = base + (uint64_t)lane;
            if (k < n && rays[k].ib >= 0) // not the unused rest of a wave's chunk
                C.peel += kp::bq::peel_ray<Pol>(P, rays[k]);
        }
    } else {
        const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
        for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
            if (rays[k].ib >= 0) // not the unused rest of a wave's chunk
                C.peel += kp::bq::peel_ray<Pol>(P, rays[k]);
    }
    flush_counters(P, C);
}

// Packets still alive after the last pass: each runs to its end in one thread.
template <class Pol>
__global__ void __launch_bounds__(256) k_bq_finish(kp::kparams P, kp::bq::qparams Q)
{
    using real = typename Pol::real;
    __shared__ unsigned long long s_cnt[KP_NCOUNTERS];
    kp::bq::wg_counters_zero(s_cnt);
    __syncthreads();
    const uint64_t n = *Q.n_in;
    const auto *in = static_cast<const kp::bq::rec_t<real> *>(Q.in);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    kp::bq::wave_peel_open();
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        kp::bq::qlane<Pol> L;
        kp::bq::load<Pol>(P, in[k], L);
        (void)kp::bq::march<Pol, false>(P, Q, L, nullptr, nullptr);
        kp::bq::end_packet<Pol>(P, L, s_cnt);
    }
    kp::bq::wave_peel_close<real>(Q);
    __syncthreads();
    kp::bq::wg_counters_flush(P, s_cnt);
}

// Sort key of every queued peel-off ray: observer, then the Morton code of
// its start point's 2^9 x 2^9 image-plane cell, then its depth along the line
// of sight (2^9 steps); unused slots sort last.
template <class Pol>
__global__ void __launch_bounds__(256) k_peel_keys(kp::kparams P, int32_t *keys)
{
    using real = typename Pol::real;
    const kp::peel_rec_t<real> *pq = static_cast<const kp::peel_rec_t<real> *>(P.pq);
    const uint64_t n = *P.pq_count < P.pq_cap ? *P.pq_count : (uint64_t)P.pq_cap;
    double D = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        D += (P.xmax[a] - P.xmin[a]) * (P.xmax[a] - P.xmin[a]);
    D = 0.5 * sqrt(D);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const kp::peel_rec_t<real> rec = pq[k];
        if (rec.ib < 0) { // unused rest of a wave's chunk
            keys[k] = (int32_t)((1u << 30) - 1u);
            continue;
        }
        const int o = (int)(rec.cell_o >> 24);
        typename Pol::ray t;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            t.x[a] = rec.x[a];
        t.ib = rec.ib;
        real g[3];
        Pol::gpos(P, t, g);
        double uh = 0.0, uv = 0.0, dz = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double r = (double)g[a] - P.center[a];
            uh += r * P.eh[o][a];
            uv += r * P.ev[o][a];
            dz += r * P.no[o][a];
        }
        auto q9 = [](double f) { return (uint32_t)fmin(fmax(f * 512.0, 0.0), 511.0); };
        const uint32_t u = q9((uh + P.image_half) / (2.0 * P.image_half));
        const uint32_t v = q9((uv + P.image_half) / (2.0 * P.image_half));
        const uint32_t z = q9((dz + D) / (2.0 * D));
        uint32_t m = 0;
        for (int b = 0; b < 9; ++b)
            m |= (((u >> b) & 1u) << (2 * b)) | (((v >> b) & 1u) << (2 * b + 1));
        keys[k] = (int32_t)(((uint32_t)o << 27) | (m << 9) | z);
    }
}

// Gradient deposits of the queued peel-off rays, d/dlog(rho_c) += -dtau_c A_k
// for every crossing: the same march of the precision, optical depth and
// cut-off as the replay's own peel-off march, one thread per ray in
// image-plane order.  Blocks blk of nblk (grid-stride over the rays).
template <class Pol>
__device__ __forceinline__ void peel_grad_body(const kp::kparams &P, uint64_t blk, uint64_t nblk)
{
    using real = typename Pol::real;
    const kp::peel_rec_t<real> *pq = static_cast<const kp::peel_rec_t<real> *>(P.pq);
    const uint64_t n = *P.pq_count < P.pq_cap ? *P.pq_count : (uint64_t)P.pq_cap;
    const int lane = threadIdx.x & 63;
    const uint64_t stride = nblk * blockDim.x;
    double *gl = kp::g_base(P, P.g_logrho);
    for (uint64_t base = blk * blockDim.x + (threadIdx.x & ~63u); base < n;
         base += stride) {
        const uint64_t k = base + lane;
        int64_t idx = -1;
        bool used = false;
        if (k < n) {
            const uint32_t s = (uint32_t)P.pq_order[k];
            if ((uint64_t)s < n && pq[s].ib >= 0) {
                used = true;
                // a NaN sum marks a packet left to the two-pass kernel
                if (!(P.pk_atot && isnan(P.pk_atot[pq[s].slot])))
                    idx = s;
            }
        }
        // used slots sort first: a wave with none ends this thread's share
        if (!__ballot(used))
            break;
        typename Pol::ray r;
        real ks = 0;
        double A = 0.0, tau = 0.0;
        bool live = idx >= 0;
        if (live) {
            const kp::peel_rec_t<real> rec = pq[idx];
            const int o = (int)(rec.cell_o >> 24);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                r.x[a] = rec.x[a];
                r.d[a] = (real)P.no[o][a];
            }
            r.i[0] = (int)(rec.cell_o & 255u);
            r.i[1] = (int)((rec.cell_o >> 8) & 255u);
            r.i[2] = (int)((rec.cell_o >> 16) & 255u);
            r.ib = rec.ib;
            Pol::relocate(P, r);
            ks = (real)P.bands[kp::band_of(P, (uint64_t)rec.slot)].kappa_scale;
            A = rec.A;
        }
        while (__ballot(live)) {
            bool has = false;
            int64_t gc = 0;
            double v = 0.0;
            if (live) {
                real l;
                const int ap = Pol::dl(r, l);
                if (ap < 0) {
                    live = false;
                } else {
                    gc = (int64_t)r.ib * P.n3 + Pol::cell(P, r);
                    const real kap = (real)P.kappa[gc] * ks;
                    const real dtau = kap * l;
                    tau = tau + (double)dtau;
                    has = true;
                    v = -((double)dtau * A);
                    if (tau > kp::kTauCut || Pol::cross(P, r, ap, l) != kp::f64::ST_IN)
                        live = false;
                }
            }
            kp::wave_run_add(gl, has, (uint64_t)gc, v);
        }
    }
}

// Deferred deposits of the one-pass mixed adjoint, one main-path vertex per
// lane, B = Atot - C with the packet's Atot from pk_atot:
//   interaction: d/dlog(rho) += B_v, d/d(albedo) += B_v / albedo;
//   scatter:     d/dg += B_s dlog P_HG(mu)/dg, plus the peel-off phase terms;
//   segment:     d/dlog(rho_c) -= kappa_c l_c B_s for every crossing of the
//                segment that starts at the vertex, marched with the main
//                path's own FP32 arithmetic (same cells, same lengths, same
//                end: interaction, escape or step budget).
// B_v / B_s: the A_k still ahead before / after the vertex's peel-offs.
// The interaction's d/dlog(rho) term and the segment's first crossing (same
// cell) go out as one atomic.  Blocks blk of nblk.
template <class Pol>
__device__ __forceinline__ void vtx_grad_body(const kp::kparams &P, uint64_t blk, uint64_t nblk)
{
    using real = typename Pol::real;
    const kp::vtx_rec_t<real> *vq = static_cast<const kp::vtx_rec_t<real> *>(P.vq);
    const uint64_t n = *P.vq_count < P.vq_cap ? *P.vq_count : (uint64_t)P.vq_cap;
    const int lane = threadIdx.x & 63;
    const uint64_t stride = nblk * blockDim.x;
    for (uint64_t base = blk * blockDim.x + (threadIdx.x & ~63u); base < n;
         base += stride) {
        const uint64_t k = base + lane;
        kp::vtx_rec_t<real> rec;
        bool ok = false;
        double Atot = 0.0;
        if (k < n) {
            rec = vq[k];
            if (rec.ib >= 0) {
                Atot = P.pk_atot[rec.slot];
                ok = !isnan(Atot);
            }
        }
        typename Pol::ray r;
        real ks = 0;
        double Bs = 0.0, tau = 0.0;
        int steps = 0;
        bool live = false;
        double carry = 0.0; // interaction term waiting for the first crossing
        int64_t carry_gc = -1;
        if (ok) {
            const int band = kp::band_of(P, (uint64_t)rec.slot);
            const uint32_t fl = P.bands[band].flags;
            const double Bv = Atot - rec.Cv;
            Bs = Atot - rec.Cs;
            if (rec.gc >= 0) {
                if (Bv != 0.0) {
                    if (P.g_logrho) {
                        carry = Bv;
                        carry_gc = rec.gc;
                    }
                    if ((fl & KP_BAND_CELL_ALBEDO) && P.g_alb)
                        kp::atomic_add_f64(kp::g_base(P, P.g_alb) + rec.gc, Bv / (double)rec.alb);
                }
                if (P.model == KP_HG_RAYLEIGH && (fl & KP_BAND_CELL_G) && P.g_g) {
                    // scatter term and peel-off phase terms: one cell, one atomic
                    double v = rec.phs;
                    if (Bs != 0.0) {
                        const double gg = P.g ? (double)P.g[rec.gc] : 0.0;
                        v += Bs * kp::dlogp_hg(gg, (double)rec.mu);
                    }
                    if (v != 0.0)
                        kp::atomic_add_f64(kp::g_base(P, P.g_g) + rec.gc, v);
                }
            }
            live = P.g_logrho && Bs != 0.0 && rec.steps > 0;
            if (live) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    r.x[a] = rec.x[a];
                    r.d[a] = rec.d[a];
                }
                r.i[0] = (int)(rec.cell & 255u);
                r.i[1] = (int)((rec.cell >> 8) & 255u);
                r.i[2] = (int)((rec.cell >> 16) & 255u);
                r.ib = rec.ib;
                Pol::relocate(P, r);
                ks = (real)P.bands[band].kappa_scale;
                tau = rec.tau;
                steps = rec.steps;
            }
        }
        double *gl = P.g_logrho ? kp::g_base(P, P.g_logrho) : nullptr;
        while (__ballot(live)) {
            if (live) {
                real l;
                const int ap = steps > 0 ? Pol::dl(r, l) : -1;
                if (ap < 0) {
                    live = false;
                } else {
                    const int64_t gc = (int64_t)r.ib * P.n3 + Pol::cell(P, r);
                    const real kap = (real)P.kappa[gc] * ks;
                    const real dtau = kap * l;
                    steps -= 1;
                    // the main path's own segment arithmetic of the precision
                    bool hit;
                    real seg;
                    if constexpr (sizeof(real) == 8) {
                        hit = kap > 0.0 && dtau >= tau;
                        seg = l;
                        if (hit) {
                            const double li = tau / kap;
                            seg = li > l ? l : li;
                        }
                    } else {
                        hit = kap > 0.0f && (double)dtau >= tau;
                        seg = hit ? fminf((float)tau * kp::f32::fast_rcp(kap), l) : l;
                    }
                    double v = -((double)(kap * seg) * Bs);
                    if (gc == carry_gc) {
                        v += carry;
                        carry_gc = -1;
                    }
                    kp::atomic_add_f64(gl + gc, v);
                    if (hit) {
                        live = false;
                    } else {
                        tau = tau - (double)dtau;
                        if (Pol::cross(P, r, ap, l) != kp::f64::ST_IN)
                            live = false;
                    }
                }
            }
        }
        if (carry_gc >= 0)
            kp::atomic_add_f64(kp::g_base(P, P.g_logrho) + carry_gc, carry);
    }
}

template <class Pol> __global__ void __launch_bounds__(256) k_vtx_grad(kp::kparams P)
{
    vtx_grad_body<Pol>(P, blockIdx.x, gridDim.x);
}

// Both deferred deposits in one launch: the first n_peel blocks march the
// peel-off rays (VALU-bound, runs of lanes summed), the others the vertex
// records (bound by scattered atomics), so the two overlap on the device.
template <class Pol>
__global__ void __launch_bounds__(256) k_defer_grad(kp::kparams P, uint32_t n_peel)
{
    if (blockIdx.x < n_peel)
        peel_grad_body<Pol>(P, blockIdx.x, n_peel);
    else
        vtx_grad_body<Pol>(P, blockIdx.x - n_peel, gridDim.x - n_peel);
}

// One thread per packet, both replay passes in that thread (FP64).
__global__ void __launch_bounds__(256) k_adjoint(kp::kparams P, kp::adj::aparams Q)
{
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double jsum = 0.0;
    for (uint64_t slot = tid; slot < P.total; slot += Q.n_threads) {
        const int band = kp::band_of(P, slot);
        const uint64_t pid = P.packet_offset + (slot - P.band_start[band]);
        kp::adj::adjoint_packet(P, Q, band, pid, tid, jsum);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        jsum += __shfl_xor(jsum, off, 64);
    if ((threadIdx.x & 63) == 0 && Q.jlin && jsum != 0.0)
        atomicAdd(Q.jlin, jsum);
}

void free_bufs(fwd_bufs &b)
{
    dfree(b.work);
    dfree(b.j_rep);
    dfree(b.img_rep);
    b.j_cap = b.img_cap = 0;
    dfree(b.keys);
    dfree(b.order);
    dfree(b.ws);
    b.ord_cap = b.ws_cap = 0;
    for (int k = 0; k < 2; ++k) {
        dfree(b.bq_rec[k]);
        dfree(b.bq_key[k]);
    }
    dfree(b.bq_seg);
    dfree(b.bq_toff);
    dfree(b.bq_task);
    dfree(b.bq_np);
    dfree(b.bq_ws);
    dfree(b.bq_cnt);
    dfree(b.bq_rays);
    b.bq_cap = b.bq_rec_bytes = b.bq_seg_cap = b.bq_np_cap = b.bq_ws_cap = 0;
    b.bq_task_cap = 0;
    b.bq_rays_bytes = 0;
}

void free_all(kp_engine *e)
{
    dfree(e->xf0);
    dfree(e->dx0);
    dfree(e->vol0);
    dfree(e->nb_id);
    dfree(e->nb_shift);
    dfree(e->phybnd);
    dfree(e->d_lvl);
    dfree(e->h32);
    dfree(e->org32);
    dfree(e->vol32);
    dfree(e->sph);
    dfree(e->kappa);
    dfree(e->albedo);
    dfree(e->g);
    dfree(e->bands_d);
    dfree(e->band_start_d);
    free_bufs(e->fb);
    dfree(e->adj_scratch);
    e->adj_cap = 0;
    dfree(e->g_rep);
    e->g_rep_cap = 0;
    dfree(e->pq);
    dfree(e->pq_count);
    dfree(e->pq_keys);
    dfree(e->pq_order);
    dfree(e->pq_ws);
    e->pq_bytes = e->pq_kcap = e->pq_ws_cap = 0;
    dfree(e->vq);
    dfree(e->vq_count);
    e->vq_bytes = 0;
    dfree(e->pk_atot);
    dfree(e->fb_list);
    dfree(e->fb_count);
    e->fb_cap = 0;
    if (e->stage)
        (void)hipHostFree(e->stage);
    e->stage = nullptr;
}

// geo_t::idx semantics: leaf block with x in [lo, hi) on every axis
bool locate(const kp_engine *e, const double x[3], int &ib, int i[3])
{
    for (int b = 0; b < e->n_blocks; ++b) {
        bool in = true;
        for (int a = 0; a < 3 && in; ++a) {
            const double lo = e->xf0_h[3 * b + a];
            const double hi = lo + e->n_cell * e->dx0_h[3 * b + a];
            in = (x[a] >= lo && x[a] < hi);
        }
        if (!in)
            continue;
        ib = b;
        for (int a = 0; a < 3; ++a) {
            int c = (int)((x[a] - e->xf0_h[3 * b + a]) / e->dx0_h[3 * b + a]);
            i[a] = std::min(std::max(c, 0), e->n_cell - 1);
        }
        return true;
    }
    return false;
}

// Cell of a Cartesian point on a spherical mesh, with the march's own face
// tests (kp_sph.h): cell k of an axis when the point is at or beyond face k
// and not at or beyond face k+1.
bool sph_locate(const kp_engine *e, const double x[3], int &ib, int i[3])
{
    const int n = e->n_cell, m = n + 1;
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    auto beyond = [&](const double *T, int q, int k) {
        if (q == 0)
            return r2 >= T[k] * T[k];
        if (q == 1)
            return x[2] <= T[m + k] * std::sqrt(r2);
        return x[1] * T[3 * m + k] - x[0] * T[2 * m + k] >= 0.0;
    };
    for (int b = 0; b < e->n_blocks; ++b) {
        const double *T = e->sph_h.data() + (size_t)b * KP_SPH_TAB(n);
        int c[3];
        bool in = true;
        for (int q = 0; q < 3 && in; ++q) {
            c[q] = -1;
            for (int k = 0; k < n; ++k)
                if (beyond(T, q, k) && !beyond(T, q, k + 1)) {
                    c[q] = k;
                    break;
                }
            in = c[q] >= 0;
        }
        if (!in)
            continue;
        ib = b;
        for (int q = 0; q < 3; ++q)
            i[q] = c[q];
        return true;
    }
    return false;
}

void upload_bands(kp_engine *e, const kp_run *run, uint64_t &total, hipStream_t stream)
{
    const int nb = run->n_bands;
    std::vector<unsigned char> img(nb * sizeof(kp::dband) + (nb + 1) * sizeof(uint64_t));
    kp::dband *db = (kp::dband *)img.data();
    uint64_t *bs = (uint64_t *)(img.data() + nb * sizeof(kp::dband));
    bs[0] = 0;
    for (int b = 0; b < nb; ++b) {
        const kp_band &s = run->bands[b];
        if (!(s.kappa_scale >= 0.0))
            throw kp::error(KP_ERR_ARG, "kp_forward: band kappa_scale must be >= 0");
        if (!(s.flags & KP_BAND_CELL_ALBEDO) && !(s.albedo >= 0.0 && s.albedo <= 1.0))
            throw kp::error(KP_ERR_ARG, "kp_forward: band albedo must be in [0,1]");
        if (!(s.flags & KP_BAND_CELL_G) && !(s.g > -1.0 && s.g < 1.0))
            throw kp::error(KP_ERR_ARG, "kp_forward: band g must be in (-1,1)");
        db[b].kappa_scale = s.kappa_scale;
        db[b].albedo = s.albedo;
        db[b].g = s.g;
        db[b].flags = s.flags;
        db[b].pad = 0;
        bs[b + 1] = bs[b] + s.n_packets;
    }
    total = bs[nb];
    if (img == e->shadow)
        return;
    if (nb > e->bands_cap) {
        dfree(e->bands_d);
        dfree(e->band_start_d);
        if (e->stage)
            (void)hipHostFree(e->stage);
        e->stage = nullptr;
        e->bands_d = dmalloc<kp::dband>(nb);
        e->band_start_d = dmalloc<uint64_t>(nb + 1);
        KP_HIP(hipHostMalloc(&e->stage, img.size(), hipHostMallocDefault));
        e->bands_cap = nb;
    }
    // the staging buffer may still feed an earlier copy on this stream
    KP_HIP(hipStreamSynchronize(stream));
    std::memcpy(e->stage, img.data(), img.size());
    KP_HIP(hipMemcpyAsync(e->bands_d, e->stage, nb * sizeof(kp::dband), hipMemcpyHostToDevice,
                          stream));
    KP_HIP(hipMemcpyAsync(e->band_start_d, (unsigned char *)e->stage + nb * sizeof(kp::dband),
                          (nb + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    KP_HIP(hipStreamSynchronize(stream));
    e->shadow = img;
}

// Replica count for an array of n elements: the schedule's value, or as many
// as fit a byte budget (capped by the number of resident workgroups).  The
// default budgets keep the live part of the replicas inside the 256 MB
// Infinity Cache: replicas that spill to HBM turn every FP64 atomic into a
// line fetch + write-back.  n_live is the part of one replica that the
// packets in flight update (one band's slice when bands are laid out
// band-major: slots are claimed in band order); the whole set of replicas
// is capped at 32 GiB (of the 288 GB of HBM).
uint32_t pick_replicas(int requested, uint64_t n, uint64_t n_live, uint64_t blocks,
                       uint64_t budget_bytes)
{
    uint64_t R =
        requested > 0 ? (uint64_t)requested : budget_bytes / (n_live * sizeof(double) + 1);
    const uint64_t hard = (32ull << 30) / (n * sizeof(double) + 1);
    if (R > hard)
        R = hard;
    if (R > blocks)
        R = blocks;
    if (R > 256)
        R = 256;
    if (R < 1)
        R = 1;
    return (uint32_t)R;
}

// Make sure `buf` holds R*n zeroed doubles (fresh allocations are cleared once;
// afterwards k_fold leaves them zero).
void ensure_replicas(double *&buf, size_t &cap, uint64_t need, hipStream_t stream)
{
    if (cap >= need)
        return;
    dfree(buf);
    cap = 0;
    buf = dmalloc<double>(need);
    hipLaunchKernelGGL(k_fill0, dim3(1024), dim3(256), 0, stream, buf, need);
    KP_HIP(hipGetLastError());
    cap = need;
}

void fill_params(kp_engine *e, const kp_source *src, const kp_run *run, hipStream_t stream,
                 kp::kparams &P, uint64_t &total)
{
        std::memset(&P, 0, sizeof(P));
        P.xf0 = e->xf0;
        P.dx0 = e->dx0;
        P.vol0 = e->vol0;
        P.nb_id = e->nb_id;
        P.nb_shift = e->nb_shift;
        P.d_lvl = e->d_lvl;
        P.phybnd = e->phybnd;
        P.h32 = e->h32;
        P.org32 = e->org32;
        P.vol32 = e->vol32;
        P.sph = e->sph;
        P.n_cell = e->n_cell;
        P.n3 = e->n3;
        for (int a = 0; a < 3; ++a) {
            P.periodic[a] = e->periodic[a];
            P.xmin[a] = e->xmin[a];
            P.xmax[a] = e->xmax[a];
            P.center[a] = (e->xmin[a] + e->xmax[a]) * 0.5;
        }
        P.kappa = e->kappa;
        P.albedo = e->albedo;
        P.g = e->g;
        total = 0;
        upload_bands(e, run, total, stream);
        P.bands = e->bands_d;
        P.band_start = e->band_start_d;
        P.n_bands = run->n_bands;
        P.j_cs = (uint64_t)run->n_bands; // cell-major J unless a plan says otherwise
        P.j_bs = 1;
        P.total = total;
        P.packet_offset = run->packet_offset;
        int ib, ii[3];
        const bool sph_mesh = e->coord == KP_COORD_SPHERICAL;
        if (!(sph_mesh ? sph_locate(e, src->x, ib, ii) : locate(e, src->x, ib, ii)))
            throw kp::error(KP_ERR_ARG, "source outside the mesh");
        for (int a = 0; a < 3; ++a) {
            P.src_x[a] = src->x[a];
            P.src_i[a] = ii[a];
            P.src_xl[a] = sph_mesh ? 0.0f
                                   : (float)((src->x[a] - e->xf0_h[3 * ib + a]) /
                                             e->dx0_h[3 * ib + a]);
        }
        P.src_ib = ib;
        P.lds_n = 0;
        P.src_gc = (int64_t)ib * e->n3 + ii[0] + (int64_t)e->n_cell * (ii[1] + (int64_t)e->n_cell * ii[2]);
        P.w0 = src->weight;
        P.seed = run->seed;
        P.model = run->scatter_model;
        P.pmax = run->p_max;
        P.step_max = run->step_max;
        P.rr_w = run->rr_weight;
        P.rr_s = run->rr_survive;
        P.n_incl = run->n_incl;
        for (int o = 0; o < run->n_incl; ++o) {
            double s, c;
            kp_sincos2pi(run->incl_deg[o] / 360.0, &s, &c);
            P.no[o][0] = s;
            P.no[o][1] = 0.0;
            P.no[o][2] = c;
            P.ev[o][0] = -c;
            P.ev[o][1] = 0.0;
            P.ev[o][2] = s;
            P.eh[o][0] = 0.0;
            P.eh[o][1] = 1.0;
            P.eh[o][2] = 0.0;
        }
        P.image_n = run->image_n;
        P.image_half = run->image_half;
        P.pix_scale = run->image_n > 0 ? (double)run->image_n / (2.0 * run->image_half) : 0.0;
}

// Launch plan of a forward run: grid and replica counts; allocates replica
// buffers (never called inside a stream capture).
// bytes of LDS accumulators a persistent forward workgroup may take
constexpr uint64_t kMaxLdsAcc = 32768;
constexpr double kMaxQueueBytes = 48e9; // one-pass adjoint queues, per engine

struct fwd_plan {
    uint64_t blocks = 0;
    int T = 256;
    uint64_t nJ = 0, nI = 0;
    bool j_band_major = false; // J accumulated band-major in replicas, folded by k_fold_t
    bool coherent = false;     // packets marched in emission-direction order
    int dir_bits = 0, key_bits = 0;
    // block queue (KP_SCHED_BLOCKQ): bands [bq_first_band, n_bands), their
    // launch slots [bq_first_slot, total); the persistent kernel takes the rest
    int bq_first_band = 0;
    uint64_t bq_first_slot = 0, bq_n_slots = 0;
    uint32_t bq_n_keys = 0;
    int bq_key_bits = 1, bq_passes = 20, n_blocks = 0;
    uint32_t bq_chunk = 1024, bq_yield = 128;
    uint64_t bq_pass_blocks = 0, bq_finish_blocks = 0;
    uint64_t bq_ray_cap = 0; // peel-off records a pass can queue (0: marched in place)
};

// Block-queue buffers of a plan, grown to n records of rec_bytes each.
void ensure_bq(fwd_bufs &fb, uint64_t n, size_t rec_bytes, uint32_t n_keys, int n_bands,
               uint32_t chunk)
{
    if (fb.bq_cap < n || fb.bq_rec_bytes < rec_bytes) {
        for (int k = 0; k < 2; ++k) {
            dfree(fb.bq_rec[k]);
            dfree(fb.bq_key[k]);
        }
        fb.bq_cap = fb.bq_rec_bytes = 0;
        for (int k = 0; k < 2; ++k) {
            fb.bq_rec[k] = dmalloc<unsigned char>(n * rec_bytes);
            fb.bq_key[k] = dmalloc<uint32_t>(n);
        }
        fb.bq_cap = n;
        fb.bq_rec_bytes = rec_bytes;
    }
    if (fb.bq_seg_cap < (size_t)n_keys + 1) {
        dfree(fb.bq_seg);
        dfree(fb.bq_toff);
        fb.bq_seg_cap = 0;
        fb.bq_seg = dmalloc<uint32_t>((size_t)n_keys + 1);
        fb.bq_toff = dmalloc<uint32_t>((size_t)n_keys + 1);
        fb.bq_seg_cap = (size_t)n_keys + 1;
    }
    // a key of len records makes ceil(len / chunk) tasks: at most n / chunk
    // plus one per key
    const size_t nt = (size_t)(n / chunk) + n_keys + 1;
    if (fb.bq_task_cap < nt) {
        dfree(fb.bq_task);
        fb.bq_task_cap = 0;
        fb.bq_task = dmalloc<uint2>(nt);
        fb.bq_task_cap = nt;
    }
    if (fb.bq_np_cap < (size_t)n_bands) {
        dfree(fb.bq_np);
        fb.bq_np_cap = 0;
        fb.bq_np = dmalloc<uint32_t>(n_bands);
        fb.bq_np_cap = n_bands;
    }
    const size_t wsb = kp_compact_workspace_bytes((int64_t)n, 1);
    if (fb.bq_ws_cap < wsb) {
        dfree(fb.bq_ws);
        fb.bq_ws_cap = 0;
        KP_HIP(hipMalloc(&fb.bq_ws, wsb));
        fb.bq_ws_cap = wsb;
    }
    if (!fb.bq_cnt)
        fb.bq_cnt = dmalloc<unsigned long long>(6);
}

// Peel-off ray queue of the block queue: `rays` records of ray_bytes each.
void ensure_bq_rays(fwd_bufs &fb, uint64_t rays, size_t ray_bytes)
{
    if (fb.bq_rays_bytes >= rays * ray_bytes)
        return;
    dfree(fb.bq_rays);
    fb.bq_rays_bytes = 0;
    fb.bq_rays = dmalloc<unsigned char>(rays * ray_bytes);
    fb.bq_rays_bytes = rays * ray_bytes;
}

fwd_plan plan_forward(kp_engine *e, fwd_bufs &fb, kp::kparams &P, uint64_t total,
                      const kp_run *run, hipStream_t stream)
{
    const kp_schedule &S = e->sched;
    fwd_plan pl;
    if (S.kind == KP_SCHED_HISTORY) {
        pl.T = 256;
        pl.blocks = (total + pl.T - 1) / pl.T;
        if (pl.blocks > 0x7fffffffull)
            pl.blocks = 0x7fffffffull;
    } else {
        pl.T = S.block_threads;
        pl.blocks = (uint64_t)e->n_cu * S.blocks_per_cu;
        const uint64_t need = (total + pl.T - 1) / pl.T;
        if (pl.blocks > need)
            pl.blocks = need;
    }
    pl.nJ = (uint64_t)e->n_blocks * e->n3 * run->n_bands;
    pl.nI = (uint64_t)run->n_bands * run->n_incl * run->image_n * run->image_n * 3;
    P.j_nrep = P.img_nrep = 1;
    P.j_stride = pl.nJ;
    P.img_stride = pl.nI;
    const uint64_t nc = (uint64_t)e->n_blocks * e->n3;
    P.j_cs = (uint64_t)run->n_bands;
    P.j_bs = 1;
    if (P.J) {
        P.j_nrep = pick_replicas(S.j_replicas, pl.nJ, run->n_bands > 1 ? nc : pl.nJ, pl.blocks,
                                 160ull << 20);
        // several bands: the packets in flight belong to one or two bands
        // (slots are claimed in order), so a band-major accumulator keeps
        // the live part of J to one band's cells; cell-major output is made
        // by the transposing fold
        pl.j_band_major = run->n_bands > 1;
        if (pl.j_band_major) {
            P.j_cs = 1;
            P.j_bs = nc;
        }
        if (P.j_nrep > 1 || pl.j_band_major)
            ensure_replicas(fb.j_rep, fb.j_cap, (uint64_t)P.j_nrep * pl.nJ, stream);
    }
    if (P.image && pl.nI > 0) {
        P.img_nrep = pick_replicas(S.img_replicas, pl.nI, pl.nI / (uint64_t)run->n_bands,
                                   pl.blocks, 32ull << 20);
        // the kernels add into planar replicas (even just one), folded into
        // the caller's image after the launch
        ensure_replicas(fb.img_rep, fb.img_cap, (uint64_t)P.img_nrep * pl.nI, stream);
    }
    // block queue: the bands from queue_from_band on
    pl.n_blocks = e->n_blocks;
    if (S.kind == KP_SCHED_BLOCKQ) {
        pl.bq_first_band = std::min(std::max(S.queue_from_band, 0), run->n_bands);
        for (int b = 0; b < pl.bq_first_band; ++b)
            pl.bq_first_slot += run->bands[b].n_packets;
        pl.bq_n_slots = total - pl.bq_first_slot;
        pl.bq_passes = S.queue_passes > 0 ? S.queue_passes : 32;
        pl.bq_chunk = S.queue_chunk > 0 ? (uint32_t)S.queue_chunk : 2048u;
        pl.bq_yield = S.queue_yield > 0 ? (uint32_t)S.queue_yield : 64u;
        pl.bq_pass_blocks = (uint64_t)e->n_cu * S.blocks_per_cu;
        pl.bq_finish_blocks = (uint64_t)e->n_cu * 4;
        if (pl.bq_n_slots > 0) {
            if (e->n3 > kp::bq::kMaxCells)
                throw kp::error(KP_ERR_UNSUPPORTED, "block queue: blocks of more than 4096 cells");
            if (pl.bq_n_slots >= (1ull << 31))
                throw kp::error(KP_ERR_UNSUPPORTED, "block queue: 2^31 packets or more in one launch");
            const uint64_t nk = (uint64_t)(run->n_bands - pl.bq_first_band) * e->n_blocks;
            if (nk >= (1ull << 30))
                throw kp::error(KP_ERR_UNSUPPORTED, "block queue: 2^30 (band, block) keys or more");
            pl.bq_n_keys = (uint32_t)nk;
            pl.bq_key_bits = 1;
            while (pl.bq_key_bits < 30 && (nk - 1) >> pl.bq_key_bits)
                ++pl.bq_key_bits;
            const size_t rb = run->precision == KP_F64 ? sizeof(kp::bq::rec_t<double>)
                                                       : sizeof(kp::bq::rec_t<float>);
            ensure_bq(fb, pl.bq_n_slots, rb, pl.bq_n_keys, run->n_bands, pl.bq_chunk);
            // peel-off records of a pass (one per interaction, all observers):
            // two per packet, at least 2^20 and at most 2^27 records or 8 GB;
            // interactions beyond it march their peel-offs in place
            if (run->n_incl > 0) {
                const size_t rr = run->precision == KP_F64 ? sizeof(kp::bq::ray_rec_t<double>)
                                                           : sizeof(kp::bq::ray_rec_t<float>);
                uint64_t cap = 2 * pl.bq_n_slots;
                cap = std::max<uint64_t>(cap, 1ull << 20);
                cap = std::min<uint64_t>(cap, std::min<uint64_t>(1ull << 27, (8ull << 30) / rr));
                pl.bq_ray_cap = cap;
                ensure_bq_rays(fb, cap, rr);
            }
        }
    }
    // LDS accumulators of the source cell's J and the emission peel-offs
    P.lds_n = 0;
    if (S.kind == KP_SCHED_PERSISTENT || S.kind == KP_SCHED_BLOCKQ) {
        const uint64_t n = (uint64_t)run->n_bands * (1 + 3 * (uint64_t)run->n_incl);
        if (n * sizeof(double) <= kMaxLdsAcc)
            P.lds_n = (int)n;
    }
    // emission-direction order: key and permutation arrays and the sort
    // workspace, grown to this launch
    // slot order instead when the launch or the band count is too large for
    // the 31-bit sort keys (fewer than 4 direction bits per axis left)
    pl.coherent = S.coherent != 0 && total < (1ull << 31);
    if (pl.coherent)
        pl.dir_bits = dir_key_bits(run->n_bands, pl.key_bits);
    pl.coherent = pl.coherent && pl.dir_bits >= 4;
    if (pl.coherent) {
        if (fb.ord_cap < total) {
            dfree(fb.keys);
            dfree(fb.order);
            fb.ord_cap = 0;
            fb.keys = dmalloc<int32_t>(total);
            fb.order = dmalloc<int32_t>(total);
            fb.ord_cap = total;
        }
        const size_t wsb = kp_compact_workspace_bytes((int64_t)total, 1);
        if (fb.ws_cap < wsb) {
            dfree(fb.ws);
            fb.ws_cap = 0;
            KP_HIP(hipMalloc(&fb.ws, wsb));
            fb.ws_cap = wsb;
        }
    }
    if (!fb.work) {
        fb.work = dmalloc<unsigned long long>(1);
        KP_HIP(hipMemsetAsync(fb.work, 0, sizeof(unsigned long long), stream));
    }
    P.work = fb.work;
    P.chunk = (uint32_t)S.chunk;
    P.batch = (uint32_t)S.event_batch;
    return pl;
}

// Waves per SIMD the peel kernel is compiled for: the mixed kernel keeps 4
// waves per SIMD resident (up to 128 VGPRs, no spills: at 8 waves and 64
// VGPRs it spilled and ran 3 % slower on config 2); the FP64 one runs at its
// natural register count.
template <class Pol> constexpr int kPeelWaves = sizeof(typename Pol::real) == 4 ? 4 : 1;

// The block-queue part of a forward run (kernels only): emission peel-offs
// and records, queue_passes passes (sort by key, segment table, pass), then
// the packets still alive to their end.
template <class Pol>
void issue_blockq(const fwd_bufs &fb, const kp::kparams &P, const fwd_plan &pl,
                  hipStream_t stream)
{
    kp::bq::qparams Q{};
    Q.n_keys = pl.bq_n_keys;
    Q.chunk = pl.bq_chunk;
    Q.yield = pl.bq_yield;
    Q.first_band = pl.bq_first_band;
    Q.n_blocks = pl.n_blocks;
    Q.first_slot = pl.bq_first_slot;
    Q.n_slots = pl.bq_n_slots;
    Q.emit_np = fb.bq_np;
    Q.pr = pl.bq_ray_cap > 0 ? fb.bq_rays : nullptr;
    Q.pr_count = fb.bq_cnt + 2;
    Q.peel_next = fb.bq_cnt + 5;
    Q.pr_cap = pl.bq_ray_cap;
    // one observer: grid-stride order over 32 times the workgroups that fit
    // (config 2 mixed, 1e8 packets: 420 ms per step, 429 with claims);
    // several: claims over the workgroups that fit (config 5, three
    // observers: 1010 ms per 1e8-packet chunk, 1044 in grid-stride order)
    const bool claim = P.n_incl > 1;
    const unsigned peel_blocks = (unsigned)(pl.bq_finish_blocks * (claim ? 1 : 32));
    const auto peel = claim ? k_bq_peel<Pol, kPeelWaves<Pol>, true>
                            : k_bq_peel<Pol, kPeelWaves<Pol>, false>;
    if (Q.pr)
        hipLaunchKernelGGL(k_zero_u64x2, dim3(1), dim3(1), 0, stream, Q.pr_count, Q.peel_next);
    const int nqb = P.n_bands - pl.bq_first_band;
    hipLaunchKernelGGL(k_bq_emit_peel<Pol>, dim3((unsigned)((nqb + 63) / 64)), dim3(64), 0,
                       stream, P, Q);
    Q.out = fb.bq_rec[0];
    Q.out_key = fb.bq_key[0];
    Q.n_out = fb.bq_cnt + 0;
    const auto grid_n = [](uint64_t n) {
        return (unsigned)std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 4096);
    };
    hipLaunchKernelGGL(k_bq_generate<Pol>, dim3(grid_n(pl.bq_n_slots)), dim3(256), 0, stream, P,
                       Q);
    int cur = 0;
    for (int pass = 0; pass < pl.bq_passes; ++pass) {
        const int nxt = 1 - cur;
        // the sort leaves the permutation and the sorted keys in its
        // workspace, where this pass reads them.  The mixed path's first
        // pass goes without: k_bq_generate writes the records in slot order,
        // whose keys (band, source block) ascend.  (Without it, and with its
        // window read through the same null check, the FP64 step measured
        // 1256 against 1242 ms per 1e8 packets, so the FP64 path keeps it.)
        const int32_t *order = nullptr;
        const uint32_t *skey = fb.bq_key[cur];
        if (pass > 0 || sizeof(typename Pol::real) == 8)
            kp::sort_in_place(reinterpret_cast<const int32_t *>(fb.bq_key[cur]), pl.bq_n_slots,
                              pl.bq_key_bits, fb.bq_ws, fb.bq_ws_cap, stream, fb.bq_cnt + cur,
                              &order, &skey);
        hipLaunchKernelGGL(k_bq_seg, dim3(grid_n((uint64_t)pl.bq_n_keys + 1)), dim3(256), 0,
                           stream, skey, fb.bq_cnt + cur, pl.bq_n_keys, fb.bq_seg);
        hipLaunchKernelGGL(k_bq_tcount, dim3(grid_n(pl.bq_n_keys)), dim3(256), 0, stream,
                           fb.bq_seg, pl.bq_n_keys, pl.bq_chunk, fb.bq_toff);
        hipLaunchKernelGGL(k_bq_tscan, dim3(1), dim3(1024), 0, stream, fb.bq_toff,
                           pl.bq_n_keys, fb.bq_cnt + 3, fb.bq_cnt + 4);
        hipLaunchKernelGGL(k_bq_tfill, dim3(grid_n(fb.bq_task_cap)), dim3(256), 0, stream,
                           fb.bq_seg, fb.bq_toff, pl.bq_n_keys, pl.bq_chunk, fb.bq_cnt + 3,
                           fb.bq_task);
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, fb.bq_cnt + nxt);
        Q.in = fb.bq_rec[cur];
        Q.in_key = fb.bq_key[cur];
        Q.order = order;
        Q.seg = fb.bq_seg;
        Q.tasks = fb.bq_task;
        Q.n_tasks = fb.bq_cnt + 3;
        Q.task_next = fb.bq_cnt + 4;
        Q.n_in = fb.bq_cnt + cur;
        Q.out = fb.bq_rec[nxt];
        Q.out_key = fb.bq_key[nxt];
        Q.n_out = fb.bq_cnt + nxt;
        // the mixed pass kernel is held to 4 waves per SIMD (128 VGPRs), the
        // FP64 one to 2 (256)
        constexpr int wps = sizeof(typename Pol::real) == 4 ? 4 : 2;
        hipLaunchKernelGGL((k_bq_pass<Pol, wps>), dim3((unsigned)pl.bq_pass_blocks), dim3(256),
                           (size_t)P.n3 * (sizeof(kp::bq::tile_t<Pol>) + sizeof(float)),
                           stream, P, Q);
        if (Q.pr) {
            hipLaunchKernelGGL(peel, dim3(peel_blocks), dim3(256), 0, stream, P, Q);
            hipLaunchKernelGGL(k_zero_u64x2, dim3(1), dim3(1), 0, stream, Q.pr_count,
                               Q.peel_next);
        }
        cur = nxt;
    }
    Q.in = fb.bq_rec[cur];
    Q.in_key = fb.bq_key[cur];
    Q.n_in = fb.bq_cnt + cur;
    hipLaunchKernelGGL(k_bq_finish<Pol>, dim3((unsigned)pl.bq_finish_blocks), dim3(256), 0,
                       stream, P, Q);
    if (Q.pr)
        hipLaunchKernelGGL(peel, dim3(peel_blocks), dim3(256), 0, stream, P, Q);
}

// Launches of a planned forward run on `stream`: n_chunks transport kernels
// (chunk k draws packet ids offset by k * chunk_stride), then the replica
// folds into the caller's arrays.  Kernels only -- no memsets, no host
// synchronisation -- so the sequence can be captured into a hipGraph.
void issue_forward(const kp_schedule &S, const fwd_bufs &fb, kp::kparams P, const fwd_plan &pl,
                   const kp_run *run, hipStream_t stream, int n_chunks, uint64_t chunk_stride)
{
    double *J_out = P.J, *I_out = P.image;
    if (P.J && (P.j_nrep > 1 || pl.j_band_major))
        P.J = fb.j_rep;
    if (P.image && pl.nI > 0) {
        P.image = fb.img_rep;
        P.img_plane = pl.nI / 3;
    }
    const uint64_t off0 = P.packet_offset;
    for (int k = 0; k < n_chunks; ++k) {
        P.packet_offset = off0 + (uint64_t)k * chunk_stride;
        if (pl.coherent) {
            hipLaunchKernelGGL(k_dir_keys, dim3(2048), dim3(256), 0, stream, P, pl.dir_bits,
                               fb.keys);
            kp::sort_order(fb.keys, P.total, pl.key_bits, fb.order, fb.ws, fb.ws_cap, stream);
            P.order = fb.order;
        }
        if (P.sph) {
            if (S.kind == KP_SCHED_HISTORY) {
                hipLaunchKernelGGL(k_sph<false>, dim3((unsigned)pl.blocks), dim3(pl.T),
                                   (size_t)P.lds_n * sizeof(double), stream, P);
            } else {
                hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, fb.work);
                hipLaunchKernelGGL(k_sph<true>, dim3((unsigned)pl.blocks), dim3(pl.T),
                                   (size_t)P.lds_n * sizeof(double), stream, P);
            }
        } else if (S.kind == KP_SCHED_HISTORY) {
            if (run->precision == KP_F64)
                hipLaunchKernelGGL(k_forward_f64, dim3((unsigned)pl.blocks), dim3(pl.T), 0,
                                   stream, P);
            else
                hipLaunchKernelGGL(k_forward_f32, dim3((unsigned)pl.blocks), dim3(pl.T), 0,
                                   stream, P);
        } else {
            // the persistent kernel: every band, or those before the block queue
            kp::kparams Pp = P;
            if (S.kind == KP_SCHED_BLOCKQ)
                Pp.total = pl.bq_first_slot;
            if (Pp.total > 0) {
                // mixed precision in emission-direction order: the lockstep variant
                const persist_fn f =
                    (run->precision == KP_F32 && P.order)
                        ? sync_variant()
                        : persist_variant<false>(run->precision, S.waves_per_simd);
                hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, fb.work);
                hipLaunchKernelGGL(f, dim3((unsigned)pl.blocks), dim3(pl.T),
                                   (size_t)P.lds_n * sizeof(double), stream, Pp);
            }
            if (S.kind == KP_SCHED_BLOCKQ && pl.bq_n_slots > 0) {
                if (run->precision == KP_F64)
                    issue_blockq<kp::pol64>(fb, P, pl, stream);
                else
                    issue_blockq<kp::pol32>(fb, P, pl, stream);
            }
        }
    }
    if (J_out && pl.j_band_major)
        hipLaunchKernelGGL(k_fold_t, dim3(4096), dim3(256), 0, stream, J_out, fb.j_rep,
                           pl.nJ / (uint64_t)run->n_bands, run->n_bands, P.j_nrep);
    else if (J_out && P.j_nrep > 1)
        hipLaunchKernelGGL(k_fold, dim3(2048), dim3(256), 0, stream, J_out, fb.j_rep, pl.nJ,
                           P.j_nrep);
    if (I_out && pl.nI > 0)
        hipLaunchKernelGGL(k_fold_planar, dim3(512), dim3(256), 0, stream, I_out, fb.img_rep,
                           pl.nI / 3, P.img_nrep);
    KP_HIP(hipGetLastError());
}

// Validation shared by kp_forward and kp_graph_create.
void check_forward_args(const kp_engine *e, const kp_source *src, const kp_run *run,
                        const kp_tallies *out, const char *who)
{
    const std::string w(who);
    if (!e || !src || !run || !out)
        throw kp::error(KP_ERR_ARG, w + ": null argument");
    if (run->n_bands < 1 || !run->bands)
        throw kp::error(KP_ERR_ARG, w + ": need at least one band");
    if (e->coord == KP_COORD_SPHERICAL) {
        // the reference's spherical photon streams without a medium
        for (int b = 0; b < run->n_bands; ++b)
            if (run->bands[b].kappa_scale != 0.0)
                throw kp::error(KP_ERR_UNSUPPORTED,
                                w + ": spherical meshes stream only (kappa_scale must be 0)");
        if (run->n_incl != 0)
            throw kp::error(KP_ERR_UNSUPPORTED, w + ": no peel-off images on spherical meshes");
        if (run->precision != KP_F64)
            throw kp::error(KP_ERR_UNSUPPORTED, w + ": spherical meshes run in FP64");
        return;
    }
    if (!e->media_set)
        throw kp::error(KP_ERR_STATE, w + ": kp_set_media not called");
    if (run->n_incl < 0 || run->n_incl > kp::kMaxIncl || (run->n_incl > 0 && !run->incl_deg))
        throw kp::error(KP_ERR_ARG, w + ": n_incl must be in [0,8]");
    if (run->n_incl > 0 && (e->periodic[0] || e->periodic[1] || e->periodic[2]))
        throw kp::error(KP_ERR_UNSUPPORTED, w + ": peel-off images need a non-periodic domain");
    if (run->n_incl > 0 && out->image && (run->image_n < 1 || !(run->image_half > 0.0)))
        throw kp::error(KP_ERR_ARG, w + ": bad image geometry");
    if (run->scatter_model != KP_THOMSON && run->scatter_model != KP_HG_RAYLEIGH)
        throw kp::error(KP_ERR_ARG, w + ": unknown scatter_model");
    if (run->precision != KP_F64 && run->precision != KP_F32)
        throw kp::error(KP_ERR_ARG, w + ": unknown precision");
    if (!(run->rr_survive > 0.0 && run->rr_survive <= 1.0))
        throw kp::error(KP_ERR_ARG, w + ": rr_survive must be in (0,1]");
}

} // namespace

extern "C" {

kp_status kp_engine_create(const kp_mesh_desc *m, int32_t device, kp_engine **out)
{
    return kp::guard([&] {
        if (!m || !out)
            throw kp::error(KP_ERR_ARG, "kp_engine_create: null argument");
        if (m->n_blocks < 1 || m->n_cell < 2 || (m->n_cell & 1))
            throw kp::error(KP_ERR_ARG, "kp_engine_create: bad block table");
        if ((uint64_t)m->n_blocks * m->n_cell * m->n_cell * m->n_cell >= (1ull << 31))
            throw kp::error(KP_ERR_UNSUPPORTED, "kp_engine_create: more than 2^31 cells");
        int ndev = 0;
        KP_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev)
            throw kp::error(KP_ERR_ARG, "kp_engine_create: no such device");
        device_guard dg(device);
        auto *e = new kp_engine();
        try {
            e->device = device;
            e->n_blocks = m->n_blocks;
            e->n_cell = m->n_cell;
            e->n3 = m->n_cell * m->n_cell * m->n_cell;
            for (int a = 0; a < 3; ++a) {
                e->periodic[a] = m->periodic[a];
                e->xmin[a] = m->x_min[a];
                e->xmax[a] = m->x_max[a];
            }
            const int nb = m->n_blocks;
            e->xf0_h.assign(m->xf0, m->xf0 + 3 * nb);
            e->dx0_h.assign(m->dx0, m->dx0 + 3 * nb);
            e->xf0 = dmalloc<double>(3 * nb);
            e->dx0 = dmalloc<double>(3 * nb);
            e->vol0 = dmalloc<double>(nb);
            e->nb_id = dmalloc<int32_t>(24 * nb);
            e->nb_shift = dmalloc<uint8_t>(24 * nb);
            e->d_lvl = dmalloc<int8_t>(6 * nb);
            e->phybnd = dmalloc<uint8_t>(nb);
            e->h32 = dmalloc<float>(3 * nb);
            e->org32 = dmalloc<float>(3 * nb);
            e->vol32 = dmalloc<float>(nb);
            upload(e->xf0, m->xf0, 3 * nb);
            upload(e->dx0, m->dx0, 3 * nb);
            upload(e->vol0, m->vol0, nb);
            upload(e->nb_id, m->nb_id, 24 * nb);
            upload(e->nb_shift, m->nb_shift, 24 * nb);
            upload(e->d_lvl, m->d_lvl, 6 * nb);
            upload(e->phybnd, m->phybnd, nb);
            if (m->coord == KP_COORD_SPHERICAL) {
                if (!m->sph)
                    throw kp::error(KP_ERR_ARG, "kp_engine_create: spherical mesh without table");
                const size_t nt = (size_t)nb * KP_SPH_TAB(m->n_cell);
                e->coord = KP_COORD_SPHERICAL;
                // streaming histories run best in lockstep, in emission-direction order
                e->sched.kind = KP_SCHED_HISTORY;
                e->sched.coherent = 1;
                e->sph_h.assign(m->sph, m->sph + nt);
                e->sph = dmalloc<double>(nt);
                upload(e->sph, m->sph, nt);
            } else if (m->coord != KP_COORD_CARTESIAN) {
                throw kp::error(KP_ERR_ARG, "kp_engine_create: unknown coordinate system");
            }
            e->fb.work = dmalloc<unsigned long long>(1);
            KP_HIP(hipDeviceGetAttribute(&e->n_cu, hipDeviceAttributeMultiprocessorCount, device));
            std::vector<float> h(3 * nb), o(3 * nb), v(nb);
            for (int b = 0; b < nb; ++b) {
                for (int a = 0; a < 3; ++a) {
                    h[3 * b + a] = (float)m->dx0[3 * b + a];
                    o[3 * b + a] = (float)m->xf0[3 * b + a];
                }
                v[b] = (float)m->vol0[b];
            }
            upload(e->h32, h.data(), 3 * nb);
            upload(e->org32, o.data(), 3 * nb);
            upload(e->vol32, v.data(), nb);
        } catch (...) {
            free_all(e);
            delete e;
            throw;
        }
        *out = e;
    });
}

void kp_engine_destroy(kp_engine *eng)
{
    if (!eng)
        return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(eng->device);
    (void)hipDeviceSynchronize();
    free_all(eng);
    (void)hipSetDevice(prev);
    delete eng;
}

kp_status kp_set_media(kp_engine *e, const float *kappa_rho, const float *albedo, const float *g)
{
    return kp::guard([&] {
        if (!e || !kappa_rho || !albedo)
            throw kp::error(KP_ERR_ARG, "kp_set_media: null argument");
        device_guard dg(e->device);
        const size_t nc = (size_t)e->n_blocks * e->n3;
        for (size_t q = 0; q < nc; ++q)
            if (!(kappa_rho[q] >= 0.0f) || !(albedo[q] >= 0.0f && albedo[q] <= 1.0f))
                throw kp::error(KP_ERR_ARG, "kp_set_media: kappa_rho must be >= 0 and "
                                            "albedo in [0,1]");
        if (g)
            for (size_t q = 0; q < nc; ++q)
                if (!(g[q] > -1.0f && g[q] < 1.0f))
                    throw kp::error(KP_ERR_ARG, "kp_set_media: g must be in (-1,1)");
        // launches and graph replays already issued on any stream of the
        // device read the old medium: wait for them before overwriting it
        KP_HIP(hipDeviceSynchronize());
        if (!e->kappa) {
            e->kappa = dmalloc<float>(nc);
            e->albedo = dmalloc<float>(nc);
            e->g = dmalloc<float>(nc);
        }
        upload(e->kappa, kappa_rho, nc);
        upload(e->albedo, albedo, nc);
        if (g) {
            upload(e->g, g, nc);
        } else {
            KP_HIP(hipMemset(e->g, 0, nc * sizeof(float)));
        }
        // later launches on any stream see the whole new medium
        KP_HIP(hipDeviceSynchronize());
        e->media_set = true;
    });
}

kp_status kp_forward(kp_engine *e, const kp_source *src, const kp_run *run, kp_tallies *out,
                     void *stream_)
{
    return kp::guard([&] {
        check_forward_args(e, src, run, out, "kp_forward");
        device_guard dg(e->device);
        hipStream_t stream = (hipStream_t)stream_;
        kp::kparams P;
        uint64_t total = 0;
        fill_params(e, src, run, stream, P, total);
        P.J = out->J;
        P.image = (run->n_incl > 0) ? out->image : nullptr;
        P.counters = (unsigned long long *)out->counters;
        P.end_state = out->end_state;
        if (total == 0)
            return;
        const fwd_plan pl = plan_forward(e, e->fb, P, total, run, stream);
        issue_forward(e->sched, e->fb, P, pl, run, stream, 1, 0);
    });
}

kp_status kp_adjoint(kp_engine *e, const kp_source *src, const kp_run *run,
                     const double *dL_dimage, kp_grads *grads, void *stream_)
{
    return kp::guard([&] {
        if (!e || !src || !run || !dL_dimage || !grads)
            throw kp::error(KP_ERR_ARG, "kp_adjoint: null argument");
        if (e->coord == KP_COORD_SPHERICAL)
            throw kp::error(KP_ERR_UNSUPPORTED, "kp_adjoint: no adjoint on spherical meshes");
        if (!e->media_set)
            throw kp::error(KP_ERR_STATE, "kp_adjoint: kp_set_media not called");
        if (run->precision != KP_F64 && run->precision != KP_F32)
            throw kp::error(KP_ERR_ARG, "kp_adjoint: unknown precision");
        if (run->precision == KP_F32 && e->sched.kind == KP_SCHED_HISTORY)
            throw kp::error(KP_ERR_UNSUPPORTED, "kp_adjoint: the one-thread-per-packet "
                                                "schedule replays FP64 histories only");
        if (run->n_bands < 1 || !run->bands)
            throw kp::error(KP_ERR_ARG, "kp_adjoint: need at least one band");
        if (run->n_incl < 1 || run->n_incl > kp::kMaxIncl || !run->incl_deg)
            throw kp::error(KP_ERR_ARG, "kp_adjoint: need 1..8 observers");
        if (run->image_n < 1 || !(run->image_half > 0.0))
            throw kp::error(KP_ERR_ARG, "kp_adjoint: bad image geometry");
        if (e->periodic[0] || e->periodic[1] || e->periodic[2])
            throw kp::error(KP_ERR_UNSUPPORTED, "kp_adjoint: needs a non-periodic domain");
        if (run->scatter_model != KP_THOMSON && run->scatter_model != KP_HG_RAYLEIGH)
            throw kp::error(KP_ERR_ARG, "kp_adjoint: unknown scatter_model");
        if (!(run->rr_survive > 0.0 && run->rr_survive <= 1.0))
            throw kp::error(KP_ERR_ARG, "kp_adjoint: rr_survive must be in (0,1]");
        device_guard dg(e->device);
        hipStream_t stream = (hipStream_t)stream_;
        kp::kparams P;
        uint64_t total = 0;
        fill_params(e, src, run, stream, P, total);
        if (total == 0)
            return;
        const kp_schedule &S = e->sched;
        const int T = S.kind == KP_SCHED_HISTORY ? 256 : S.block_threads;
        uint64_t blocks = (uint64_t)e->n_cu * (S.kind == KP_SCHED_HISTORY ? 4 : S.blocks_per_cu);
        const uint64_t need = (total + T - 1) / T;
        if (blocks > need)
            blocks = need;
        const uint64_t lanes = blocks * T;
        const uint32_t cap = (uint32_t)(32 * run->n_incl);
        const size_t want = (size_t)cap * lanes;
        if (e->adj_cap < want) {
            dfree(e->adj_scratch);
            e->adj_scratch = dmalloc<double>(want);
            e->adj_cap = want;
        }
        if (S.kind == KP_SCHED_HISTORY) {
            kp::adj::aparams Q;
            Q.w = dL_dimage;
            Q.d_logrho = grads->d_logrho;
            Q.d_alb = grads->d_albedo;
            Q.d_g = grads->d_g;
            Q.jlin = grads->jlin;
            Q.n_threads = lanes;
            Q.cap = cap;
            Q.scratch = e->adj_scratch;
            hipLaunchKernelGGL(k_adjoint, dim3((unsigned)blocks), dim3(T), 0, stream, P, Q);
            KP_HIP(hipGetLastError());
            return;
        }
        // persistent path replay with replicated gradient accumulators
        const uint64_t nc = (uint64_t)e->n_blocks * e->n3;
        P.adj_w = dL_dimage;
        P.adj_rows = e->adj_scratch;
        P.adj_cap = cap;
        P.adj_lanes = lanes;
        P.jlin = grads->jlin;
        P.g_stride = nc;
        P.g_nrep = pick_replicas(S.j_replicas, 3 * nc, 3 * nc, blocks, 400ull << 20);
        double *gl = grads->d_logrho, *ga = grads->d_albedo, *gg = grads->d_g;
        if (P.g_nrep > 1) {
            ensure_replicas(e->g_rep, e->g_rep_cap, 3ull * P.g_nrep * nc, stream);
            P.g_logrho = gl ? e->g_rep : nullptr;
            P.g_alb = ga ? e->g_rep + (uint64_t)P.g_nrep * nc : nullptr;
            P.g_g = gg ? e->g_rep + 2ull * P.g_nrep * nc : nullptr;
        } else {
            P.g_logrho = gl;
            P.g_alb = ga;
            P.g_g = gg;
        }
        P.lds_n = 0;
        P.work = e->fb.work;
        P.chunk = (uint32_t)S.chunk;
        P.batch = (uint32_t)S.event_batch;
        // one-pass replay (either precision): each packet is marched once;
        // its main-path vertices and peel-off rays are queued with what their
        // deposits need and deposited by separate kernels once the packets'
        // A_k sums are known (peel-off rays marched in image-plane order).
        // The queues hold vertex_slots (default 32) vertices per packet and
        // one peel-off ray per vertex and observer; runs whose queues would
        // take more than half of the free device memory, or 48 GB, go in
        // chunks of packets.  Packets whose records still do not fit are replayed by
        // the two-pass kernel after the last chunk.
        const bool any_grad = grads->d_logrho || grads->d_albedo || grads->d_g;
        const bool defer = S.peel_queue && (!grads->d_logrho || e->n_cell < 256) &&
                           total < (1ull << 31);
        const bool f64 = run->precision == KP_F64;
        const size_t prec_bytes = f64 ? sizeof(kp::peel_rec_t<double>)
                                      : sizeof(kp::peel_rec_t<float>);
        const size_t vrec_bytes = f64 ? sizeof(kp::vtx_rec_t<double>)
                                      : sizeof(kp::vtx_rec_t<float>);
        const bool queue = defer && grads->d_logrho;
        const bool vqueue = defer && any_grad;
        uint64_t chunk = total;
        if (defer) {
            const uint64_t vs = S.vertex_slots > 0 ? (uint64_t)S.vertex_slots : 32;
            const uint64_t ps = vs * (uint64_t)run->n_incl;
            const auto q_bytes = [&](uint64_t q) {
                return q ? (double)q * (prec_bytes + 2 * sizeof(int32_t)) +
                               (double)kp_compact_workspace_bytes((int64_t)q, 1)
                         : 0.0;
            };
            const auto v_bytes = [&](uint64_t v) { return (double)v * vrec_bytes; };
            const auto bytes = [&](uint64_t pk) {
                return (queue ? q_bytes(ps * pk) : 0.0) + (vqueue ? v_bytes(vs * pk) : 0.0);
            };
            if (vqueue) {
                size_t free_b = 0, total_b = 0;
                KP_HIP(hipMemGetInfo(&free_b, &total_b));
                // memory already held by the queues counts as free; the
                // engine keeps the queues until it is destroyed, so at most
                // kMaxQueueBytes
                const double budget = std::min(
                    kMaxQueueBytes,
                    0.5 * ((double)free_b + (double)e->pq_bytes +
                           (double)e->pq_kcap * 2 * sizeof(int32_t) + (double)e->pq_ws_cap +
                           (double)e->vq_bytes));
                // the largest chunk whose queues fit (at least 2^16 packets),
                // then equal chunks of the run
                const auto fits = [&](uint64_t c) {
                    return bytes(c) <= budget && vs * c < (1ull << 31) && ps * c < (1ull << 31);
                };
                if (!fits(chunk)) {
                    const double per = bytes(1ull << 20) / (double)(1ull << 20);
                    uint64_t cmax = std::max<uint64_t>(1ull << 16, (uint64_t)(budget / per));
                    while (cmax > (1ull << 16) && !fits(cmax))
                        cmax = std::max<uint64_t>(1ull << 16, cmax - cmax / 16);
                    const uint64_t nch = (total + cmax - 1) / cmax;
                    chunk = (total + nch - 1) / nch;
                }
            }
            if (S.adjoint_chunk > 0)
                chunk = std::min<uint64_t>(chunk, (uint64_t)S.adjoint_chunk);
            const uint64_t want_q = queue ? ps * chunk : 0, want_v = vqueue ? vs * chunk : 0;
            if (queue) {
                const size_t wsb = kp_compact_workspace_bytes((int64_t)want_q, 1);
                if (e->pq_bytes < want_q * prec_bytes) {
                    dfree(e->pq);
                    e->pq_bytes = 0;
                    e->pq = dmalloc<char>(want_q * prec_bytes);
                    e->pq_bytes = want_q * prec_bytes;
                }
                if (e->pq_kcap < want_q) {
                    dfree(e->pq_keys);
                    dfree(e->pq_order);
                    e->pq_kcap = 0;
                    e->pq_keys = dmalloc<int32_t>(want_q);
                    e->pq_order = dmalloc<int32_t>(want_q);
                    e->pq_kcap = want_q;
                }
                if (e->pq_ws_cap < wsb) {
                    dfree(e->pq_ws);
                    e->pq_ws_cap = 0;
                    KP_HIP(hipMalloc(&e->pq_ws, wsb));
                    e->pq_ws_cap = wsb;
                }
                if (!e->pq_count)
                    e->pq_count = dmalloc<unsigned long long>(1);
                P.pq = e->pq;
                P.pq_count = e->pq_count;
                P.pq_cap = (uint32_t)want_q; // this run's share of a larger cached queue
            }
            if (vqueue) {
                if (e->vq_bytes < want_v * vrec_bytes) {
                    dfree(e->vq);
                    e->vq_bytes = 0;
                    e->vq = dmalloc<char>(want_v * vrec_bytes);
                    e->vq_bytes = want_v * vrec_bytes;
                }
                if (!e->vq_count)
                    e->vq_count = dmalloc<unsigned long long>(1);
                P.vq = e->vq;
                P.vq_count = e->vq_count;
                P.vq_cap = (uint32_t)want_v;
            }
            if (e->fb_cap < total) {
                dfree(e->fb_list);
                dfree(e->pk_atot);
                e->fb_cap = 0;
                e->fb_list = dmalloc<int32_t>(total);
                e->pk_atot = dmalloc<double>(total);
                e->fb_cap = total;
            }
            if (!e->fb_count)
                e->fb_count = dmalloc<unsigned int>(1);
            P.pk_atot = e->pk_atot;
            P.fb_list = e->fb_list;
            P.fb_count = e->fb_count;
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, e->fb_count);
        }
        if (!defer) {
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, e->fb.work);
            hipLaunchKernelGGL(persist_variant<true>(run->precision, S.waves_per_simd),
                               dim3((unsigned)blocks), dim3(T), 0, stream, P);
            KP_HIP(hipGetLastError());
        }
        for (uint64_t c0 = 0; defer && c0 < total; c0 += chunk) {
            // packets [c0, c1): the claim counter starts at c0
            kp::kparams Pc = P;
            Pc.total = std::min<uint64_t>(total, c0 + chunk);
            hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, stream, e->fb.work,
                               (unsigned long long)c0);
            if (queue)
                hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, e->pq_count);
            if (vqueue)
                hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, e->vq_count);
            hipLaunchKernelGGL(defer_variant(run->precision, S.waves_per_simd),
                               dim3((unsigned)blocks), dim3(T), 0, stream, Pc);
            KP_HIP(hipGetLastError());
            const dim3 gq((unsigned)(e->n_cu * 8)), tq(256);
            if (queue) {
                if (f64)
                    hipLaunchKernelGGL(k_peel_keys<kp::pol64>, dim3(2048), tq, 0, stream, Pc,
                                       e->pq_keys);
                else
                    hipLaunchKernelGGL(k_peel_keys<kp::pol32>, dim3(2048), tq, 0, stream, Pc,
                                       e->pq_keys);
                kp::sort_order(e->pq_keys, Pc.pq_cap, 30, e->pq_order, e->pq_ws, e->pq_ws_cap,
                               stream, e->pq_count);
                Pc.pq_order = e->pq_order;
            }
            if (queue && vqueue) {
                // peel-off rays and vertex records together
                const uint32_t np = (uint32_t)(e->n_cu * 8);
                const dim3 g2(2 * np);
                if (f64)
                    hipLaunchKernelGGL(k_defer_grad<kp::pol64>, g2, tq, 0, stream, Pc, np);
                else
                    hipLaunchKernelGGL(k_defer_grad<kp::pol32>, g2, tq, 0, stream, Pc, np);
                KP_HIP(hipGetLastError());
            } else if (vqueue) {
                if (f64)
                    hipLaunchKernelGGL(k_vtx_grad<kp::pol64>, gq, tq, 0, stream, Pc);
                else
                    hipLaunchKernelGGL(k_vtx_grad<kp::pol32>, gq, tq, 0, stream, Pc);
                KP_HIP(hipGetLastError());
            }
        }
        if (defer) {
            // the packets left over, two-pass, in the order they were listed
            kp::kparams F = P;
            F.pq = nullptr;
            F.vq = nullptr;
            F.order = e->fb_list;
            F.total_dev = e->fb_count;
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, stream, e->fb.work);
            hipLaunchKernelGGL(persist_variant<true>(run->precision, S.waves_per_simd),
                               dim3((unsigned)blocks), dim3(T), 0, stream, F);
            KP_HIP(hipGetLastError());
        }
        if (P.g_nrep > 1) {
            if (gl)
                hipLaunchKernelGGL(k_fold, dim3(1024), dim3(256), 0, stream, gl, P.g_logrho, nc,
                                   P.g_nrep);
            if (ga)
                hipLaunchKernelGGL(k_fold, dim3(1024), dim3(256), 0, stream, ga, P.g_alb, nc,
                                   P.g_nrep);
            if (gg)
                hipLaunchKernelGGL(k_fold, dim3(1024), dim3(256), 0, stream, gg, P.g_g, nc,
                                   P.g_nrep);
            KP_HIP(hipGetLastError());
        }
    });
}


kp_status kp_graph_create(kp_engine *e, const kp_source *src, const kp_run *run, kp_tallies *out,
                          int32_t n_chunks, kp_graph **g_out)
{
    return kp::guard([&] {
        check_forward_args(e, src, run, out, "kp_graph_create");
        if (!g_out || n_chunks < 1)
            throw kp::error(KP_ERR_ARG, "kp_graph_create: need graph and n_chunks >= 1");
        if (out->end_state && n_chunks > 1)
            throw kp::error(KP_ERR_ARG, "kp_graph_create: end_state needs n_chunks == 1 (chunks "
                                        "would overwrite each other's end states)");
        device_guard dg(e->device);
        hipStream_t cs = nullptr;
        KP_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        auto *g = new kp_graph();
        g->e = e;
        try {
            kp::kparams P;
            uint64_t total = 0;
            fill_params(e, src, run, cs, P, total);
            if (total == 0)
                throw kp::error(KP_ERR_ARG, "kp_graph_create: no packets");
            // private copy of the band table: later runs may upload another
            const int nb = run->n_bands;
            g->bands = dmalloc<kp::dband>(nb);
            g->band_start = dmalloc<uint64_t>(nb + 1);
            KP_HIP(hipMemcpyAsync(g->bands, e->bands_d, nb * sizeof(kp::dband),
                                  hipMemcpyDeviceToDevice, cs));
            KP_HIP(hipMemcpyAsync(g->band_start, e->band_start_d, (nb + 1) * sizeof(uint64_t),
                                  hipMemcpyDeviceToDevice, cs));
            P.bands = g->bands;
            P.band_start = g->band_start;
            P.J = out->J;
            P.image = (run->n_incl > 0) ? out->image : nullptr;
            P.counters = (unsigned long long *)out->counters;
            P.end_state = out->end_state;
            const fwd_plan pl = plan_forward(e, g->fb, P, total, run, cs);
            KP_HIP(hipStreamSynchronize(cs));
            uint64_t stride = 0;
            for (int b = 0; b < nb; ++b)
                stride = std::max<uint64_t>(stride, run->bands[b].n_packets);
            KP_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            try {
                issue_forward(e->sched, g->fb, P, pl, run, cs, n_chunks, stride);
            } catch (...) {
                hipGraph_t dead = nullptr;
                (void)hipStreamEndCapture(cs, &dead);
                if (dead)
                    (void)hipGraphDestroy(dead);
                throw;
            }
            KP_HIP(hipStreamEndCapture(cs, &g->graph));
            KP_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        } catch (...) {
            kp_graph_destroy(g);
            (void)hipStreamDestroy(cs);
            throw;
        }
        (void)hipStreamDestroy(cs);
        *g_out = g;
    });
}

kp_status kp_graph_launch(kp_graph *g, void *stream)
{
    return kp::guard([&] {
        if (!g || !g->exec)
            throw kp::error(KP_ERR_ARG, "kp_graph_launch: null graph");
        device_guard dg(g->e->device);
        KP_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    });
}

void kp_graph_destroy(kp_graph *g)
{
    if (!g)
        return;
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    dfree(g->bands);
    dfree(g->band_start);
    free_bufs(g->fb);
    delete g;
}

kp_status kp_autotune(kp_engine *e, const kp_source *src, const kp_run *run,
                      uint64_t sample_packets, int32_t force, kp_schedule *best, void *stream_)
{
    return kp::guard([&] {
        kp_tallies probe_shape{};
        probe_shape.image = nullptr;
        check_forward_args(e, src, run, &probe_shape, "kp_autotune");
        if (!best || sample_packets < 1)
            throw kp::error(KP_ERR_ARG, "kp_autotune: need best and sample_packets >= 1");
        // the sample runs always render the images
        if (run->n_incl > 0 && (run->image_n < 1 || !(run->image_half > 0.0)))
            throw kp::error(KP_ERR_ARG, "kp_autotune: bad image geometry");
        device_guard dg(e->device);
        hipStream_t stream = (hipStream_t)stream_;
        uint64_t total = 0;
        for (int b = 0; b < run->n_bands; ++b)
            total += run->bands[b].n_packets;
        // signature of the workload shape the schedule is tuned for
        std::ostringstream key;
        key << run->precision << '/' << run->n_bands << '/' << run->n_incl << '/'
            << run->image_n << '/' << run->scatter_model << '/' << e->n_blocks << '/'
            << (int)std::log2((double)std::max<uint64_t>(total, 1));
        auto hit = e->tuned.find(key.str());
        if (!force && hit != e->tuned.end()) {
            e->sched = hit->second;
            *best = hit->second;
            return;
        }
        // sample run: every band scaled to about sample_packets in total
        std::vector<kp_band> bands(run->bands, run->bands + run->n_bands);
        if (total == 0)
            throw kp::error(KP_ERR_ARG, "kp_autotune: no packets");
        for (auto &b : bands)
            if (b.n_packets)
                b.n_packets = std::max<uint64_t>(
                    1, (uint64_t)((double)b.n_packets * (double)sample_packets / (double)total));
        kp_run r = *run;
        r.bands = bands.data();
        const uint64_t nJ = (uint64_t)e->n_blocks * e->n3 * run->n_bands;
        const uint64_t nI = (uint64_t)run->n_bands * run->n_incl * run->image_n * run->image_n * 3;
        kp_tallies t{};
        t.J = dmalloc<double>(nJ);
        t.image = nI ? dmalloc<double>(nI) : nullptr;
        t.counters = dmalloc<uint64_t>(KP_NCOUNTERS);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            dfree(t.J);
            dfree(t.image);
            dfree(t.counters);
            if (e0)
                (void)hipEventDestroy(e0);
            if (e1)
                (void)hipEventDestroy(e1);
        };
        const kp_schedule before = e->sched;
        try {
            KP_HIP(hipEventCreate(&e0));
            KP_HIP(hipEventCreate(&e1));
            auto time_it = [&](const kp_schedule &sc) -> float {
                e->sched = sc;
                kp::kparams P;
                uint64_t tot = 0;
                fill_params(e, src, &r, stream, P, tot);
                P.J = t.J;
                P.image = (r.n_incl > 0) ? t.image : nullptr;
                P.counters = (unsigned long long *)t.counters;
                P.end_state = nullptr;
                const fwd_plan pl = plan_forward(e, e->fb, P, tot, &r, stream);
                issue_forward(e->sched, e->fb, P, pl, &r, stream, 1, 0); // warm-up
                KP_HIP(hipEventRecord(e0, stream));
                issue_forward(e->sched, e->fb, P, pl, &r, stream, 2, 0);
                KP_HIP(hipEventRecord(e1, stream));
                KP_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                KP_HIP(hipEventElapsedTime(&ms, e0, e1));
                return ms;
            };
            // start from the persistent kind in slot order, whatever order an
            // earlier tuning (of another precision, say) left behind
            kp_schedule cur = e->sched;
            cur.kind = KP_SCHED_PERSISTENT;
            cur.coherent = 0;
            float best_ms = time_it(cur);
            // coordinate descent, one knob at a time
            auto try_knob = [&](int kp_schedule::*f, std::initializer_list<int> vals) {
                for (int v : vals) {
                    if (cur.*f == v)
                        continue;
                    kp_schedule c = cur;
                    c.*f = v;
                    const float ms = time_it(c);
                    if (ms < best_ms) {
                        best_ms = ms;
                        cur = c;
                    }
                }
            };
            const bool sph_mesh = e->coord == KP_COORD_SPHERICAL;
            // schedule kind and packet order together: lockstep histories
            // pay off in emission-direction order.  FP64 persistent lanes in
            // that order collide on the same J addresses, more so the denser
            // the packets: a sample run can favour it where the full run
            // loses, so that pair is not a candidate (callers may still set
            // it).  The mixed persistent kernel in that order refills whole
            // waves and sums runs of lanes (wins on optically thin bands).
            // With observers the one-thread-per-packet kernels march every
            // peel-off ray inside the history, unbatched: not a candidate.
            for (int kind : {KP_SCHED_PERSISTENT, KP_SCHED_HISTORY})
                for (int coh : {0, 1}) {
                    const bool lockstep_persist = run->precision == KP_F32 && !sph_mesh;
                    if ((cur.kind == kind && cur.coherent == coh) ||
                        (coh && (total >= (1ull << 31) ||
                                 (kind == KP_SCHED_PERSISTENT && !lockstep_persist))) ||
                        (kind == KP_SCHED_HISTORY && run->n_incl > 0))
                        continue;
                    kp_schedule c = cur;
                    c.kind = kind;
                    c.coherent = coh;
                    const float ms = time_it(c);
                    if (ms < best_ms) {
                        best_ms = ms;
                        cur = c;
                    }
                }
            try_knob(&kp_schedule::j_replicas, {0, 1, 2, 3, 4, 8});
            if (!sph_mesh) // the spherical kernel has no images and one register budget
                try_knob(&kp_schedule::img_replicas, {0, 16, 64});
            try_knob(&kp_schedule::event_batch, {4, 8, 16});
            if (sph_mesh)
                ;
            else if (run->precision == KP_F64)
                try_knob(&kp_schedule::waves_per_simd, {0, 3, 4});
            else
                try_knob(&kp_schedule::waves_per_simd, {0, 1, 5});
            try_knob(&kp_schedule::blocks_per_cu, {2, 4, 8});
            try_knob(&kp_schedule::block_threads, {128, 256});
            // block queue for the optically thinnest bands: the longest tail
            // of bands whose extinction scale is at most t, for a few t
            // (every band at t = infinity); the bands before it stay on the
            // persistent kernel
            if (!sph_mesh && e->n3 <= kp::bq::kMaxCells && total < (1ull << 31)) {
                const kp_schedule base = cur;
                float queue_best = best_ms;
                kp_schedule queue_pick = cur;
                std::vector<int> firsts;
                for (double t : {1e300, 1.0, 0.5, 0.25, 0.1}) {
                    int k = run->n_bands;
                    while (k > 0 && run->bands[k - 1].kappa_scale <= t)
                        --k;
                    if (k < run->n_bands &&
                        std::find(firsts.begin(), firsts.end(), k) == firsts.end())
                        firsts.push_back(k);
                }
                for (int k : firsts) {
                    kp_schedule c = base;
                    c.kind = KP_SCHED_BLOCKQ;
                    c.queue_from_band = k;
                    // the queue's own defaults: all its pass workgroups
                    // resident (a pass hands out tasks statically)
                    c.blocks_per_cu = 4;
                    c.queue_passes = 0;
                    c.queue_chunk = 0;
                    c.queue_yield = 0;
                    const float ms = time_it(c);
                    if (ms < queue_best) {
                        queue_best = ms;
                        queue_pick = c;
                    }
                }
                if (queue_best < best_ms) {
                    best_ms = queue_best;
                    cur = queue_pick;
                    // passes and yield together: a packet's share of a pass
                    // and the number of passes trade against each other
                    const kp_schedule q0 = cur;
                    for (const auto &py : {std::pair<int, int>{20, 64}, {48, 64}, {64, 64},
                                           {96, 64}, {32, 128}, {64, 128}}) {
                        kp_schedule c = q0;
                        c.queue_passes = py.first;
                        c.queue_yield = py.second;
                        const float ms = time_it(c);
                        if (ms < best_ms) {
                            best_ms = ms;
                            cur = c;
                        }
                    }
                    try_knob(&kp_schedule::queue_chunk, {1024, 4096});
                    try_knob(&kp_schedule::blocks_per_cu, {8});
                }
            }
            e->sched = cur;
            e->tuned[key.str()] = cur;
            *best = cur;
        } catch (...) {
            e->sched = before;
            cleanup();
            throw;
        }
        cleanup();
    });
}

kp_status kp_set_schedule(kp_engine *e, const kp_schedule *s)
{
    return kp::guard([&] {
        if (!e || !s)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: null argument");
        if (s->kind != KP_SCHED_HISTORY && s->kind != KP_SCHED_PERSISTENT &&
            s->kind != KP_SCHED_BLOCKQ)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: unknown kind");
        if (s->kind == KP_SCHED_BLOCKQ && e->coord == KP_COORD_SPHERICAL)
            throw kp::error(KP_ERR_UNSUPPORTED,
                            "kp_set_schedule: no block queue on spherical meshes");
        if (s->kind == KP_SCHED_BLOCKQ && e->n3 > kp::bq::kMaxCells)
            throw kp::error(KP_ERR_UNSUPPORTED,
                            "kp_set_schedule: block queue needs n_cell^3 <= 4096");
        if (s->queue_from_band < 0 || s->queue_passes < 0 || s->queue_passes > 4096 ||
            s->queue_chunk < 0 || s->queue_chunk > (1 << 20) || s->queue_yield < 0)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: queue_from_band >= 0, queue_passes in "
                                        "[0,4096], queue_chunk in [0,2^20], queue_yield >= 0");
        if (s->block_threads != 64 && s->block_threads != 128 && s->block_threads != 256)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: block_threads must be 64, 128 or 256");
        if (s->blocks_per_cu < 1 || s->blocks_per_cu > 64)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: blocks_per_cu must be in [1,64]");
        if (s->chunk < 1 || s->chunk > 4096)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: chunk must be in [1,4096]");
        if (s->waves_per_simd < 0 || s->waves_per_simd > 6 || s->waves_per_simd == 2)
            throw kp::error(KP_ERR_ARG,
                            "kp_set_schedule: waves_per_simd must be 0, 1, 3, 4, 5 or 6");
        if (s->event_batch < 1 || s->event_batch > 64)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: event_batch must be in [1,64]");
        if (s->j_replicas < 0 || s->j_replicas > 256 || s->img_replicas < 0 ||
            s->img_replicas > 256)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: replicas must be in [0,256]");
        if (s->coherent != 0 && s->coherent != 1)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: coherent must be 0 or 1");
        if (s->peel_queue != 0 && s->peel_queue != 1)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: peel_queue must be 0 or 1");
        if (s->vertex_slots < 0 || s->vertex_slots > 4096)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: vertex_slots must be in [0, 4096]");
        if (s->adjoint_chunk != 0 && s->adjoint_chunk < 1024)
            throw kp::error(KP_ERR_ARG, "kp_set_schedule: adjoint_chunk must be 0 or >= 1024");
        e->sched = *s;
    });
}

kp_status kp_get_schedule(const kp_engine *e, kp_schedule *s)
{
    return kp::guard([&] {
        if (!e || !s)
            throw kp::error(KP_ERR_ARG, "kp_get_schedule: null argument");
        *s = e->sched;
    });
}

} // extern "C"
