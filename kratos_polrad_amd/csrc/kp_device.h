// kp_device.h -- device-side data views shared by the transport kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kp_math.h"

namespace kp {

constexpr int kMaxIncl = 8;
constexpr double kTauCut = 40.0; // peel-off rays stop once tau exceeds this

// One band (wavelength / source set) as the kernels see it.
struct dband {
    double kappa_scale;
    double albedo;
    double g;
    uint32_t flags;
    uint32_t pad;
};

// A main-path vertex of a one-pass adjoint replay (emission or scatter),
// with what its deferred gradient deposits need once the packet's sum of
// peel-off contributions Atot is known: B = Atot - C, C being the sum of the
// packet's A_k before the point.  Written when the vertex's scatter is done;
// the segment that follows starts here.  R is the transport precision
// (FP32: block-local position; FP64: global position).
template <class R> struct vtx_rec_t {
    R x[3];          // segment start: position, block, cell
    int32_t ib;
    R d[3];          // direction of the segment
    uint32_t cell;   // i0 | i1 << 8 | i2 << 16
    double tau;      // free-flight optical depth of the segment
    double Cv;       // A_k sum before the vertex's peel-offs (interaction terms)
    double Cs;       // A_k sum after them (scatter term and the segment)
    double phs;      // sum over observers of A_k dlog P_HG(mu_obs)/dg
    int32_t slot;    // packet slot
    int32_t gc;      // interaction cell, -1 for the emission vertex
    int32_t steps;   // main-path crossings left (0: the packet ended here)
    R alb;           // albedo at the interaction
    R mu;            // cosine of the scattering angle
};

// A peel-off ray of an adjoint replay whose gradient deposits are marched
// by a separate kernel: start point (position as in vtx_rec_t), block, cell,
// observer, packet slot and its A_k.
template <class R> struct peel_rec_t {
    R x[3];
    int32_t ib;
    uint32_t cell_o; // i0 | i1 << 8 | i2 << 16 | observer << 24
    int32_t slot;
    double A;
};

// Everything a transport launch needs, passed by value.
struct kparams {
    // mesh (block table, FP64)
    const double *xf0;
    const double *dx0;
    const double *vol0;
    const int32_t *nb_id;
    const uint8_t *nb_shift;
    const int8_t *d_lvl;
    const uint8_t *phybnd;
    // FP32 block-local view: per block h[3] (cell width), origin[3], volume
    const float *h32;
    const float *org32;
    const float *vol32;
    // spherical mesh: per-block face/volume table (KP_SPH_TAB doubles), else null
    const double *sph;
    int n_cell;
    int n3;
    int periodic[3];
    double xmin[3], xmax[3];
    // medium
    const float *kappa;
    const float *albedo;
    const float *g;
    // bands
    const dband *bands;
    const uint64_t *band_start; // [n_bands+1]
    int n_bands;
    uint64_t total;
    uint64_t packet_offset;
    // source
    double src_x[3];
    double w0;
    int src_ib;
    int src_i[3];
    int64_t src_gc;  // global index of the source cell
    int lds_n;       // doubles of LDS accumulators per workgroup (0: none)
    float src_xl[3]; // source position in block-local cell units (FP32 path)
    // run
    uint32_t seed;
    int model;
    double pmax;
    int step_max;
    double rr_w;
    double rr_s;
    int n_incl;
    double no[kMaxIncl][3], ev[kMaxIncl][3], eh[kMaxIncl][3];
    double center[3];
    int image_n;
    double image_half;
    double pix_scale;
    // outputs
    double *J;
    double *image;
    unsigned long long *counters;
    double *end_state;
    // persistent schedule
    unsigned long long *work; // slot counter, zeroed before the launch
    const int32_t *order;     // claim index -> slot (emission-direction order), or null
    uint32_t chunk;           // slots a wave claims per refill
    uint32_t batch;           // waiting lanes that trigger an action phase
    // J and image are replicated R times; workgroup b accumulates into
    // replica b % R, spreading FP64 atomics on hot cells/pixels
    uint32_t j_nrep, img_nrep;
    uint64_t j_stride, img_stride; // elements per replica
    // 0: a pixel's I, Q, U side by side (the caller's layout); otherwise the
    // replicas hold three planes of img_plane pixel entries each
    uint64_t img_plane;
    uint64_t j_cs, j_bs;           // J element of (cell gc, band b): gc*j_cs + b*j_bs
    // adjoint (path replay) launches
    const double *adj_w;           // dL/dimage, image layout
    double *g_logrho, *g_alb, *g_g; // gradient replica bases (may be null)
    uint32_t g_nrep;
    uint64_t g_stride;             // cells per gradient replica
    double *adj_rows;              // [adj_cap][adj_lanes] peel-off contributions A_k
    uint32_t adj_cap;
    uint64_t adj_lanes;
    double *jlin;                  // device scalar
    void *pq;                      // queue of peel_rec_t of the precision (or null)
    unsigned long long *pq_count;
    uint32_t pq_cap;
    const int32_t *pq_order;       // queue slots in image-plane order
    void *vq;                      // vtx_rec_t of the one-pass replay (or null)
    unsigned long long *vq_count;
    uint32_t vq_cap;
    double *pk_atot;               // per slot: Atot, or NaN for a packet left over
    int32_t *fb_list;              // slots of the packets left to the two-pass kernel
    unsigned int *fb_count;
    const unsigned int *total_dev; // device-side claim bound (or null)
};

// The face-neighbour entries of a block that a block change reads (the
// proxy's neighbors / d_lvl / physical-boundary bits): from the engine's
// block table (nb_global), or from a copy of one block's entries in LDS
// (nb_local) where every lane is in that block.
struct nb_global {
    const kparams &P;
    __device__ uint32_t phy(int ib) const { return P.phybnd[ib]; }
    __device__ int dl(int ib, int k) const { return P.d_lvl[6 * ib + k]; }
    __device__ int id(int ib, int k) const { return P.nb_id[24 * ib + k]; }
    __device__ int shift(int ib, int k) const { return P.nb_shift[24 * ib + k]; }
};
struct blk_nb_t {
    int32_t id[24];
    uint8_t shift[24];
    int8_t dl[6];
    uint8_t phy;
};
struct nb_local {
    const blk_nb_t *t;
    __device__ uint32_t phy(int) const { return t->phy; }
    __device__ int dl(int, int k) const { return t->dl[k]; }
    __device__ int id(int, int k) const { return t->id[k]; }
    __device__ int shift(int, int k) const { return t->shift[k]; }
};
// Copies block ib's entries into t (threads 0..23 of the workgroup; the
// caller synchronises before use).
__device__ __forceinline__ void stage_nb(const kparams &P, int ib, blk_nb_t *t)
{
    const int k = threadIdx.x;
    if (k < 24) {
        t->id[k] = P.nb_id[24 * ib + k];
        t->shift[k] = P.nb_shift[24 * ib + k];
    }
    if (k < 6)
        t->dl[k] = P.d_lvl[6 * ib + k];
    if (k == 0)
        t->phy = P.phybnd[ib];
}

__device__ __forceinline__ double *g_base(const kparams &P, double *base)
{
    return base + (uint64_t)(blockIdx.x % P.g_nrep) * P.g_stride;
}

__device__ __forceinline__ double *j_base(const kparams &P)
{
    return P.J + (uint64_t)(blockIdx.x % P.j_nrep) * P.j_stride;
}
// address of the J tally of global cell gc in band b (this workgroup's replica);
// cell ids and strides are below 2^32, so the products are 32x32->64 bit
__device__ __forceinline__ double *j_at(const kparams &P, uint64_t gc, int b)
{
    return j_base(P) + (uint64_t)(uint32_t)gc * (uint32_t)P.j_cs +
           (uint64_t)(uint32_t)b * (uint32_t)P.j_bs;
}
__device__ __forceinline__ double *img_base(const kparams &P)
{
    return P.image + (uint64_t)(blockIdx.x % P.img_nrep) * P.img_stride;
}

// Stokes component k (0 I, 1 Q, 2 U) of pixel entry px (band, observer and
// pixel as pixel() numbers them) in the workgroup's image.  In the planar
// replicas the three atomic adds of a peel-off go to three cache lines; side
// by side in one line, each waited in the L2 for the one before.
__device__ __forceinline__ double *img_at(const kparams &P, long px, int k)
{
    double *b = img_base(P);
    return P.img_plane ? b + (uint64_t)k * P.img_plane + (uint64_t)px : b + 3 * px + k;
}

__device__ __forceinline__ void atomic_add_f64(double *p, double v)
{
    unsafeAtomicAdd(p, v); // native global_atomic_add_f64 (coarse-grained memory)
}

// Slot of the k-th claimed packet of a launch.
__device__ __forceinline__ uint64_t slot_at(const kparams &P, uint64_t k)
{
    return P.order ? (uint64_t)(uint32_t)P.order[k] : k;
}

// Band of a global packet slot: the last b with band_start[b] <= slot
// (binary search; band_start is small and cache resident).  Empty bands
// (equal starts) are skipped the same way a linear scan would skip them.
__device__ __forceinline__ int band_of(const kparams &P, uint64_t slot)
{
    int lo = 0, hi = P.n_bands - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (slot >= P.band_start[mid])
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// base[idx] += v for every lane with `has`, called by all 64 lanes of the
// wave together.  Runs of neighbouring lanes with the same idx are summed by
// a segmented scan across the wave first and the last lane of each run
// issues one atomic: FP64 atomics on one address serialise in the L2, and
// lanes that march neighbouring rays in lockstep share cells.
__device__ __forceinline__ void wave_run_add(double *base, bool has, uint64_t idx, double v)
{
    const int lane = threadIdx.x & 63;
    const uint64_t key = has ? idx : ~(uint64_t)lane; // lanes without a value: own keys
    const uint64_t kprev = __shfl_up(key, 1, 64);
    const uint64_t knext = __shfl_down(key, 1, 64);
    int head = (lane == 0 || kprev != key) ? 1 : 0;
    double sum = has ? v : 0.0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double so = __shfl_up(sum, d, 64);
        const int ho = __shfl_up(head, d, 64);
        if (lane >= d && !head) {
            sum += so;
            head = ho;
        }
    }
    if (has && (lane == 63 || knext != key))
        atomic_add_f64(base + key, sum);
}

// Dynamic LDS of persistent forward launches: the workgroup's accumulators,
// [n_bands] J of the source cell, then [n_bands][n_incl][3] I,Q,U of the
// emission peel-offs.  Every packet starts in the source cell and its
// emission peel-off lands on the source pixel, so those deposits would
// otherwise serialise on a few global addresses.
extern __shared__ double kp_lds[];

// J deposit of v into cell gc of band b (the source cell to LDS when `lds`).
__device__ __forceinline__ void j_deposit(const kparams &P, bool lds, int64_t gc, int b,
                                          double v)
{
    if (lds && gc == P.src_gc)
        atomicAdd(&kp_lds[b], v);
    else
        atomic_add_f64(j_at(P, (uint64_t)gc, b), v);
}

} // namespace kp
