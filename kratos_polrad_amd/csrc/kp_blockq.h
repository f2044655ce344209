// kp_blockq.h -- block-bucketed packet queues (schedule KP_SCHED_BLOCKQ).
//
// The reference moves packets between per-block pools: a packet that leaves
// its block is marked for relocation and compacted out of the pool
// (types/particle_pool.h:191-248), one that leaves the rank is registered and
// gathered per destination (boundary/comm.h:118-127, 226-239).  Here the same
// idea runs per block, on one GPU:
//   - every packet lives in a record (AoS, 144 B in FP64, 96 B in FP32:
//     position, direction, Stokes frame and vector, optical depth left, block,
//     cell, counts), keyed by (band, block);
//   - a pass sorts the records by key (stable LSD radix sort) and gives each
//     workgroup a run of sorted positions; per block of that run the
//     workgroup stages the block's kappa_rho (n_cell^3 floats) and a J tile
//     (FP32 in the mixed path, FP64 in the FP64 path) in LDS, marches each of
//     its packets until the packet leaves the block or ends, deposits J into
//     the LDS tile, and flushes the tile once with one FP64 atomic per
//     touched cell;
//   - an interaction queues what its peel-offs are made of; the peel kernel
//     computes and marches them after the pass (kappa_rho from HBM);
//   - a packet that crosses into another block is written to the next pass's
//     records at a slot claimed once per wave (ballot / prefix rank), with
//     its new key; one that ends is counted and its end state written.
// After queue_passes passes the packets still alive run to their end in one
// thread each, with HBM kappa_rho and global J atomics.
// Every packet performs the per-packet arithmetic of its precision in the
// same order as the other schedules (FP64 histories are bit-identical to the
// CPU oracle); only the J summation order differs.
#pragma once

#include "kp_persist.h"

namespace kp {
namespace bq {

// Largest block (n_cell^3 cells) whose kappa_rho and J tile fit the LDS of a
// workgroup: 16^3 cells, 48 KB.
constexpr int kMaxCells = 4096;

template <class R> struct rec_t;
template <> struct __align__(16) rec_t<double> {
    double x[3], d[3], e1[3], s[3]; // position (global), direction, frame, Stokes I Q U
    double tau;                      // optical depth left of the free flight
    uint32_t slot;                   // launch slot (band and packet id)
    int32_t ib;                      // block
    uint32_t cell;                   // i0 | i1 << 8 | i2 << 16
    int32_t steps;                   // march steps left
    uint32_t events, ncross, npeel, pad;
};
template <> struct __align__(16) rec_t<float> {
    float x[3], d[3], e1[3], s[3];   // position (block-local cell units), ...
    double tau;
    uint32_t slot;
    int32_t ib;
    uint32_t cell;
    int32_t steps;
    uint32_t events, ncross, npeel, pad;
};
static_assert(sizeof(rec_t<double>) == 144 && sizeof(rec_t<float>) == 96, "record sizes");

// The peel-offs of one interaction of a queued packet, computed and marched
// by the peel kernel after the pass: the interaction point (position as in
// rec_t, block, cell), the packet slot, and what the peel-off Stokes vectors
// toward the observers are made of -- the incoming direction and Stokes
// frame, the Stokes vector after the albedo weighting, and the cell's g.
// 64 B in FP32, 128 B in FP64.
template <class R> struct __align__(16) ray_rec_t {
    R x[3];
    int32_t ib;
    uint32_t cell;                   // i0 | i1 << 8 | i2 << 16
    uint32_t slot;
    R d[3], e1[3], s[3];
    R g;
};
static_assert(sizeof(ray_rec_t<float>) == 64 && sizeof(ray_rec_t<double>) == 128,
              "peel-off record sizes");

// Launch parameters of the queue kernels (besides kparams).
struct qparams {
    const void *in;                   // records of this pass
    const uint32_t *in_key;           // their keys
    const int32_t *order;             // sorted position -> record (mixed path: null = the same)
    const uint32_t *seg;              // [n_keys + 1] first sorted position of each key
    const uint2 *tasks;               // (first sorted position, key) of each task
    const unsigned long long *n_tasks; // tasks of this pass
    unsigned long long *task_next;     // next task to claim (zeroed with the task table)
    const unsigned long long *n_in;   // records of this pass
    void *out;                        // records that crossed into another block
    uint32_t *out_key;
    unsigned long long *n_out;        // claimed out records (zeroed before the pass)
    uint32_t n_keys;                  // queue bands x blocks
    uint32_t chunk;                   // sorted positions per workgroup task
    uint32_t yield;                   // steps + interactions of a packet per pass
    int first_band;                   // key = (band - first_band) * n_blocks + block
    int n_blocks;
    uint64_t first_slot;              // first launch slot of the queue bands
    uint64_t n_slots;                 // slots of the queue bands
    uint32_t *emit_np;                // [n_bands] peel-off crossings of the emission peel-offs
    // peel-off rays of this pass (null: every lane marches its own)
    void *pr;
    unsigned long long *pr_count;
    unsigned long long *peel_next;    // next ray the peel kernel hands out (zeroed with pr_count)
    uint64_t pr_cap;
};

// The record at sorted position q.
__device__ __forceinline__ int32_t record_at(const qparams &Q, uint64_t q)
{
    return Q.order ? Q.order[q] : (int32_t)q;
}

// A packet between march steps.
template <class Pol> struct qlane {
    using real = typename Pol::real;
    typename Pol::ray r;
    real e1[3];
    real I, Q, U;
    real ks;
    double tau;
    uint64_t slot, pid;
    int band, events, steps, status;
    int blk;       // MARCH_BLOCK: exit axis + 4 * (exit through the upper face)
    int ret;       // pass kernel: how the lane's march in the block ended (MARCH_*)
    int cell;      // pass kernel: cell of the pending interaction
    uint32_t work; // pass kernel: steps and interactions in this pass
    uint32_t ncross, npeel;
};

template <class Pol> __device__ __forceinline__ void unpack_cell(typename Pol::ray &r, uint32_t c)
{
    r.i[0] = (int)(c & 255u);
    r.i[1] = (int)((c >> 8) & 255u);
    r.i[2] = (int)((c >> 16) & 255u);
}

template <class Pol>
__device__ __forceinline__ void load(const kparams &P, const rec_t<typename Pol::real> &q,
                                     qlane<Pol> &L)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        L.r.x[a] = q.x[a];
        L.r.d[a] = q.d[a];
        L.e1[a] = q.e1[a];
    }
    L.I = q.s[0];
    L.Q = q.s[1];
    L.U = q.s[2];
    L.tau = q.tau;
    L.slot = q.slot;
    L.r.ib = q.ib;
    unpack_cell<Pol>(L.r, q.cell);
    L.steps = q.steps;
    L.events = (int)q.events;
    L.ncross = q.ncross;
    L.npeel = q.npeel;
    L.band = band_of(P, L.slot);
    L.pid = P.packet_offset + (L.slot - P.band_start[L.band]);
    L.ks = (typename Pol::real)P.bands[L.band].kappa_scale;
    Pol::relocate(P, L.r); // FP64: face bounds; FP32: per-ray scales of the block
}

template <class Pol>
__device__ __forceinline__ void store(const qlane<Pol> &L, rec_t<typename Pol::real> &q)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        q.x[a] = L.r.x[a];
        q.d[a] = L.r.d[a];
        q.e1[a] = L.e1[a];
    }
    q.s[0] = L.I;
    q.s[1] = L.Q;
    q.s[2] = L.U;
    q.tau = L.tau;
    q.slot = (uint32_t)L.slot;
    q.ib = L.r.ib;
    q.cell = (uint32_t)L.r.i[0] | ((uint32_t)L.r.i[1] << 8) | ((uint32_t)L.r.i[2] << 16);
    q.steps = L.steps;
    q.events = (uint32_t)L.events;
    q.ncross = L.ncross;
    q.npeel = L.npeel;
    q.pad = 0;
}

// Interaction at the packet's point (in cell gc): Philox words of the event,
// absorption weighting, the peel-offs toward every observer (rays marched to
// the domain edge through HBM kappa_rho), the scattering and Russian
// roulette.  The operations and their order are those of the other
// schedules.  Returns false when roulette ends the packet.
// One queue slot for each calling lane (called by those lanes together): one
// atomic per wave, lanes ranked by ballot.
__device__ __forceinline__ uint64_t claim(unsigned long long *count)
{
    const int lane = threadIdx.x & 63;
    const uint64_t m = __ballot(1);
    const int leader = __ffsll((long long)m) - 1;
    const uint32_t rank = (uint32_t)__popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
    unsigned long long base = 0;
    if (lane == leader)
        base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + rank;
}

// Queue the peel-offs of the lane's interaction (cell g gg; L holds the
// incoming direction and frame and the albedo-weighted Stokes vector) for
// the peel kernel; false when the queue is full.  Called by the interacting
// lanes together.  Slots come from the wave's chunk of the queue (one
// counter atomic per chunk of kQChunk records; wave_peel_open first,
// wave_peel_close when the wave is done).
template <class Pol>
__device__ __forceinline__ bool push_peel(const qparams &Q, const qlane<Pol> &L,
                                          typename Pol::real gg)
{
    using real = typename Pol::real;
    const uint64_t k = claim_q(Q.pr_count, kp_pq_base, kp_pq_end);
    if (k >= Q.pr_cap)
        return false;
    ray_rec_t<real> r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.x[a] = L.r.x[a];
        r.d[a] = L.r.d[a];
        r.e1[a] = L.e1[a];
    }
    r.ib = L.r.ib;
    r.cell = (uint32_t)L.r.i[0] | ((uint32_t)L.r.i[1] << 8) | ((uint32_t)L.r.i[2] << 16);
    r.slot = (uint32_t)L.slot;
    r.s[0] = L.I;
    r.s[1] = L.Q;
    r.s[2] = L.U;
    r.g = gg;
    static_cast<ray_rec_t<real> *>(Q.pr)[k] = r;
    return true;
}

// The calling wave's peel-off chunk: none yet.
__device__ __forceinline__ void wave_peel_open()
{
    if ((threadIdx.x & 63) == 0) {
        kp_pq_base[threadIdx.x >> 6] = 0;
        kp_pq_end[threadIdx.x >> 6] = 0;
    }
}

// Marks the unused rest of the calling wave's peel-off chunk (ib = -1; the
// peel kernel skips such records).  Called by the whole wave when it is
// done.
template <class R> __device__ __forceinline__ void wave_peel_close(const qparams &Q)
{
    if (!Q.pr)
        return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t e = kp_pq_end[w] < Q.pr_cap ? kp_pq_end[w] : Q.pr_cap;
    for (uint64_t k = kp_pq_base[w] + lane; k < e; k += 64)
        static_cast<ray_rec_t<R> *>(Q.pr)[k].ib = -1;
}

template <class Pol>
__device__ bool interact(const kparams &P, const qparams &Qp, qlane<Pol> &L, int64_t gc)
{
    using real = typename Pol::real;
    L.events += 1;
    uint32_t u[4] = {(uint32_t)L.pid, (uint32_t)(L.pid >> 32), (uint32_t)L.events, 0u};
    kp_philox4x32_10(u, P.seed, (uint32_t)L.band);
    const dband &bd = P.bands[L.band];
    uint64_t np = 0;
    if constexpr (sizeof(real) == 8) {
        const double alb = (bd.flags & KP_BAND_CELL_ALBEDO) ? (double)P.albedo[gc] : bd.albedo;
        const double gg = (bd.flags & KP_BAND_CELL_G) ? (P.g ? (double)P.g[gc] : 0.0) : bd.g;
        L.I = L.I * alb;
        L.Q = L.Q * alb;
        L.U = L.U * alb;
        double e2[3];
        f64::cross3(L.r.d, L.e1, e2);
        if (P.n_incl > 0 && !(Qp.pr && push_peel<Pol>(Qp, L, gg))) {
            for (int o = 0; o < P.n_incl; ++o) {
                double Ip, Qi, Ui;
                f64::peel_stokes(P.model, gg, P.pmax, L.r.d, L.e1, e2, L.I, L.Q, L.U, P.no[o],
                                 P.ev[o], Ip, Qi, Ui);
                const double tau = f64::tau_to_edge(P, L.r, P.no[o], L.ks, np);
                const double att = kp_exp(-tau) * KP_INV_4PI;
                f64::deposit(P, L.band, o, L.r.x, Ip * att, Qi * att, Ui * att);
            }
        }
        const double mu = (P.model == KP_THOMSON) ? kp_thomson_mu(kp_u01(u[0]))
                                                  : kp_hg_mu(gg, kp_u01(u[0]));
        f64::scatter_stokes(P.model, P.pmax, mu, kp_u01(u[1]), L.r.d, L.e1, e2, L.I, L.Q, L.U);
        L.tau = -kp_log(kp_u01(u[2]));
        L.npeel += (uint32_t)np;
        if (L.I < P.rr_w * P.w0) {
            if (kp_u01(u[3]) < P.rr_s) {
                L.I = L.I / P.rr_s;
                L.Q = L.Q / P.rr_s;
                L.U = L.U / P.rr_s;
            } else {
                return false;
            }
        }
    } else {
        const float alb = (bd.flags & KP_BAND_CELL_ALBEDO) ? P.albedo[gc] : (float)bd.albedo;
        const float gg = (bd.flags & KP_BAND_CELL_G) ? (P.g ? P.g[gc] : 0.0f) : (float)bd.g;
        L.I *= alb;
        L.Q *= alb;
        L.U *= alb;
        float e2[3];
        f32::cross3(L.r.d, L.e1, e2);
        if (P.n_incl > 0 && !(Qp.pr && push_peel<Pol>(Qp, L, gg))) {
            float gx[3];
            f32::global_pos(P, L.r, gx);
            for (int o = 0; o < P.n_incl; ++o) {
                float no[3], ev[3];
                load3(P.no[o], no);
                load3(P.ev[o], ev);
                float Ip, Qi, Ui;
                f32::peel_stokes(P.model, gg, (float)P.pmax, L.r.d, L.e1, e2, L.I, L.Q, L.U, no,
                                 ev, Ip, Qi, Ui);
                const double tau = f32::tau_to_edge(P, L.r, no, L.ks, np);
                const float att = f32::fast_exp(-(float)tau) * f32::kInv4Pi;
                f32::deposit(P, L.band, o, gx, Ip * att, Qi * att, Ui * att);
            }
        }
        const float xi0 = f32::u01(u[0]);
        const float mu = (P.model == KP_THOMSON) ? f32::thomson_mu(xi0) : f32::hg_mu(gg, xi0);
        f32::scatter_stokes(P.model, (float)P.pmax, mu, f32::u01(u[1]), L.r.d, L.e1, e2, L.I,
                            L.Q, L.U);
        Pol::set_dir(P, L.r, L.r.d);
        L.tau = -(double)f32::fast_log(f32::u01(u[2]));
        L.npeel += (uint32_t)np;
        if (L.I < (float)P.rr_w * (float)P.w0) {
            if (f32::u01(u[3]) < (float)P.rr_s) {
                const float inv = (float)(1.0 / P.rr_s);
                L.I *= inv;
                L.Q *= inv;
                L.U *= inv;
            } else {
                return false;
            }
        }
    }
    return true;
}

enum : int { MARCH_END = 0, MARCH_OUT = 1, MARCH_BLOCK = 2 };

// The packet's march from its current point.  LOCAL: kappa_rho from the LDS
// tile lk of the packet's block and J into the LDS tile lj; returns
// MARCH_BLOCK as soon as the ray has reached its block's edge (leave_block
// then takes it into the next block: a wave's lanes do that together once
// all are done with this block), or MARCH_OUT after Q.yield steps and
// interactions in this block (it goes back to the queue where it stands).  Otherwise kappa_rho from HBM and J by global
// atomics, to the packet's end.  MARCH_END: L.status holds the end status.
// J tile of a block in LDS: FP64 for the FP64 path; FP32 for the mixed path,
// whose per-task sums (a few hundred contributions per cell at most) keep
// FP32's 7 digits, far below the Monte Carlo noise, before the FP64 atomic
// flush.  Half the LDS lets more workgroups share a CU.
template <class Pol> using tile_t = typename Pol::real;

// march<LOCAL> leaves the block change to leave_block, after the wave's
// lanes are done with the block: the block-change code then runs once per
// wave and march instead of in nearly every step.
template <class Pol> constexpr bool kQueueDeferBlock = true;

// march<true> written with one exit from its step loop: each step ends in an
// event code (none, yield, step budget, lost, interaction point, block edge)
// and the loop runs while the code is none.  The steps and their order are
// those of march; with one exit the compiler keeps each piece of lane state
// in one register, where the five exits of march's step loop made it copy
// the state at every step (41 of 111 VALU instructions of the mixed step).
enum : int { EV_NONE = 0, EV_YIELD, EV_STEPS, EV_LOST, EV_HIT, EV_BLOCK };

// The mixed path's march_block.  Within its step loop the ray keeps,
// instead of the integer cell indices, the face ahead on each axis as a
// float (the value f32::get_dl converts the index to) and the linear cell
// index, moved by +-1, +-n or +-n^2; the indices are rebuilt from the faces
// when the loop ends.  Lengths, positions and optical depths are computed as
// in get_dl and cross_in_block, so they are the same numbers.
template <class Pol>
__device__ int march_block_f32(const kparams &P, const qparams &Q, qlane<Pol> &L,
                               const float *lk, tile_t<Pol> *lj)
{
    static_assert(sizeof(typename Pol::real) == 4, "mixed path only");
    const int n = P.n_cell;
    const float nf = (float)n;
    uint32_t work = 0; // steps and interactions in this pass
    for (;;) {
        // per direction (set again after each scattering)
        const bool nz0 = L.r.d[0] != 0.0f, nz1 = L.r.d[1] != 0.0f, nz2 = L.r.d[2] != 0.0f;
        const f32::face_steps S(L.r.d, n);
        float ff[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            ff[a] = (float)(L.r.i[a] + S.up(a));
        int c = f32::cell_index(P, L.r);
        int ev = EV_NONE;
        int cell = 0;
        do {
            if (work >= Q.yield) {
                ev = EV_YIELD;
            } else {
                ++work;
                if (L.steps <= 0) {
                    ev = EV_STEPS;
                } else {
                    // get_dl: the nearest face ahead, ties to the later axis
                    float t[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        t[a] = fmaxf((ff[a] - L.r.x[a]) * L.r.lpc[a], 0.0f);
                    int ap = -1;
                    float l = f32::kInf;
                    if (t[0] <= l && nz0) {
                        ap = 0;
                        l = t[0];
                    }
                    if (t[1] <= l && nz1) {
                        ap = 1;
                        l = t[1];
                    }
                    if (t[2] <= l && nz2) {
                        ap = 2;
                        l = t[2];
                    }
                    if (ap < 0) {
                        ev = EV_LOST;
                    } else {
                        cell = c;
                        const float kap = lk[cell] * L.ks;
                        const float dtau = kap * l;
                        L.ncross += 1;
                        L.steps -= 1;
                        const bool hit = kap > 0.0f && (double)dtau >= L.tau;
                        const float seg =
                            hit ? fminf((float)L.tau * f32::fast_rcp(kap), l) : l;
                        if (P.J)
                            atomicAdd(&lj[cell], L.I * seg * L.r.iv);
                        if (hit) {
                            Pol::advance(L.r, seg);
                            ev = EV_HIT;
                        } else {
                            L.tau = L.tau - (double)dtau;
                            // cross_in_block: the exit axis onto its face, the
                            // others along the ray
#pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                if (a == ap) {
                                    L.r.x[a] = ff[a];
                                    ff[a] += S.sf(a);
                                    c += S.sc(a);
                                } else {
                                    L.r.x[a] = fmaf(l, L.r.cpl[a], L.r.x[a]);
                                }
                            }
                            const float fhi = fmaxf(fmaxf(ff[0], ff[1]), ff[2]);
                            const float flo = fminf(fminf(ff[0], ff[1]), ff[2]);
                            if (fhi > nf) {
                                L.blk = ap + 4;
                                ev = EV_BLOCK;
                            } else if (flo < 0.0f) {
                                L.blk = ap;
                                ev = EV_BLOCK;
                            }
                        }
                    }
                }
            }
        } while (ev == EV_NONE);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            L.r.i[a] = (int)ff[a] - S.up(a);
        if (ev == EV_YIELD)
            return MARCH_OUT;
        if (ev == EV_BLOCK)
            return MARCH_BLOCK;
        if (ev != EV_HIT) {
            L.status = (ev == EV_STEPS) ? KP_END_STEPS : KP_END_LOST;
            return MARCH_END;
        }
        if (!interact<Pol>(P, Q, L, (int64_t)L.r.ib * P.n3 + cell)) {
            L.status = KP_END_RR;
            return MARCH_END;
        }
        ++work;
    }
}

template <class Pol>
__device__ int march_block(const kparams &P, const qparams &Q, qlane<Pol> &L, const float *lk,
                           tile_t<Pol> *lj)
{
    using real = typename Pol::real;
    if constexpr (sizeof(real) == 4)
        return march_block_f32<Pol>(P, Q, L, lk, lj);
    uint32_t work = 0; // steps and interactions in this pass
    for (;;) {
        int ev = EV_NONE;
        int cell = 0;
        do {
            if (work >= Q.yield) {
                ev = EV_YIELD;
            } else {
                ++work;
                if (L.steps <= 0) {
                    ev = EV_STEPS;
                } else {
                    real l;
                    const int ap = Pol::dl(L.r, l);
                    if (ap < 0) {
                        ev = EV_LOST;
                    } else {
                        cell = Pol::cell(P, L.r);
                        const real kap = (real)lk[cell] * L.ks;
                        const real dtau = kap * l;
                        L.ncross += 1;
                        L.steps -= 1;
                        bool hit;
                        real seg;
                        if constexpr (sizeof(real) == 8) {
                            hit = kap > 0.0 && dtau >= L.tau;
                            seg = l;
                            if (hit) {
                                const double li = L.tau / kap;
                                seg = li > l ? l : li;
                            }
                        } else {
                            hit = kap > 0.0f && (double)dtau >= L.tau;
                            seg = hit ? fminf((float)L.tau * f32::fast_rcp(kap), l) : l;
                        }
                        if (P.J) {
                            real v;
                            if constexpr (sizeof(real) == 8)
                                v = L.I * (seg / P.vol0[L.r.ib]);
                            else
                                v = L.I * seg * L.r.iv;
                            atomicAdd(&lj[cell], v);
                        }
                        if (hit) {
                            Pol::advance(L.r, seg);
                            ev = EV_HIT;
                        } else {
                            L.tau = L.tau - (double)dtau;
                            const int st = Pol::cross_step(P, L.r, ap, l);
                            if (st != f32::ST_IN) {
                                L.blk = ap + (st == f32::ST_BLOCK_HI ? 4 : 0);
                                ev = EV_BLOCK;
                            }
                        }
                    }
                }
            }
        } while (ev == EV_NONE);
        if (ev == EV_YIELD)
            return MARCH_OUT;
        if (ev == EV_BLOCK)
            return MARCH_BLOCK;
        if (ev != EV_HIT) {
            L.status = (ev == EV_STEPS) ? KP_END_STEPS : KP_END_LOST;
            return MARCH_END;
        }
        if (!interact<Pol>(P, Q, L, (int64_t)L.r.ib * P.n3 + cell)) {
            L.status = KP_END_RR;
            return MARCH_END;
        }
        ++work;
    }
}

template <class Pol, bool LOCAL>
__device__ int march(const kparams &P, const qparams &Q, qlane<Pol> &L, const float *lk,
                     tile_t<Pol> *lj)
{
    using real = typename Pol::real;
    uint32_t work = 0; // steps and interactions in this pass (LOCAL)
    for (;;) {
        int64_t gc = 0;
        for (;;) {
            if constexpr (LOCAL) {
                // back to the queue, still in this block: every packet's
                // share of a pass stays bounded
                if (work >= Q.yield)
                    return MARCH_OUT;
                ++work;
            }
            if (L.steps <= 0) {
                L.status = KP_END_STEPS;
                return MARCH_END;
            }
            real l;
            const int ap = Pol::dl(L.r, l);
            if (ap < 0) {
                L.status = KP_END_LOST;
                return MARCH_END;
            }
            const int cell = Pol::cell(P, L.r);
            gc = (int64_t)L.r.ib * P.n3 + cell;
            const float kraw = LOCAL ? lk[cell] : P.kappa[gc];
            const real kap = (real)kraw * L.ks;
            const real dtau = kap * l;
            L.ncross += 1;
            L.steps -= 1;
            bool hit;
            real seg;
            if constexpr (sizeof(real) == 8) {
                hit = kap > 0.0 && dtau >= L.tau;
                seg = l;
                if (hit) {
                    const double li = L.tau / kap;
                    seg = li > l ? l : li;
                }
            } else {
                hit = kap > 0.0f && (double)dtau >= L.tau;
                seg = hit ? fminf((float)L.tau * f32::fast_rcp(kap), l) : l;
            }
            if (P.J) {
                real v;
                if constexpr (sizeof(real) == 8)
                    v = L.I * (seg / P.vol0[L.r.ib]);
                else
                    v = L.I * seg * L.r.iv;
                if constexpr (LOCAL)
                    atomicAdd(&lj[cell], v);
                else
                    atomic_add_f64(j_at(P, (uint64_t)gc, L.band), (double)v);
            }
            if (hit) {
                Pol::advance(L.r, seg);
                break;
            }
            L.tau = L.tau - (double)dtau;
            int st = Pol::cross_step(P, L.r, ap, l);
            if (st == f32::ST_IN)
                continue;
            if constexpr (LOCAL && kQueueDeferBlock<Pol>) {
                L.blk = ap + (st == f32::ST_BLOCK_HI ? 4 : 0);
                return MARCH_BLOCK;
            }
            st = Pol::change_block(P, L.r, ap, st == f32::ST_BLOCK_HI);
            if (st == f32::ST_ESCAPED) {
                L.status = KP_END_ESCAPED;
                return MARCH_END;
            }
            if (st != f32::ST_IN) {
                L.status = KP_END_LOST;
                return MARCH_END;
            }
            if constexpr (LOCAL)
                return MARCH_OUT;
        }
        if (!interact<Pol>(P, Q, L, gc)) {
            L.status = KP_END_RR;
            return MARCH_END;
        }
        if constexpr (LOCAL)
            ++work;
    }
}

// Modes of a lane in the pass kernel's wave loop (pass_block).
enum : int { Q_IDLE = 0, Q_MARCH = 1, Q_INT = 2, Q_RETIRE = 3 };
// A wave runs the interactions of its lanes once this many wait (or no lane
// marches), and retires and refills lanes once this many are done.
constexpr int kPassIntBatch = 16;
constexpr int kPassRetireBatch = 32;
// Whether the pass kernel runs its records in the wave loop (pass_block) or
// one record per lane at a time (march<true>, each wave waiting for its
// slowest lane).  The FP64 step is costly enough that keeping lanes busy
// wins (config 2, 2e7 packets: 213 against 280 ms); the mixed step is cheap,
// and the wave loop's phases cost more than the idle lanes (121-141 against
// 108 ms over batch sizes 16-64).
template <class Pol> constexpr bool kPassWaveLoop = sizeof(typename Pol::real) == 8;

// One step of a marching lane in its block (kappa_rho from the LDS tile lk,
// J into the LDS tile lj): the step of march<true>, ending in a mode instead
// of a loop exit.  Q_INT: the interaction point is reached (cell in L.cell);
// Q_RETIRE: the lane's march in this pass ends (L.ret).
template <class Pol>
__device__ __forceinline__ int pass_step(const kparams &P, const qparams &Q, qlane<Pol> &L,
                                         const float *lk, tile_t<Pol> *lj)
{
    using real = typename Pol::real;
    if (L.work >= Q.yield) {
        L.ret = MARCH_OUT;
        return Q_RETIRE;
    }
    ++L.work;
    if (L.steps <= 0) {
        L.status = KP_END_STEPS;
        L.ret = MARCH_END;
        return Q_RETIRE;
    }
    real l;
    const int ap = Pol::dl(L.r, l);
    if (ap < 0) {
        L.status = KP_END_LOST;
        L.ret = MARCH_END;
        return Q_RETIRE;
    }
    const int cell = Pol::cell(P, L.r);
    const real kap = (real)lk[cell] * L.ks;
    const real dtau = kap * l;
    L.ncross += 1;
    L.steps -= 1;
    bool hit;
    real seg;
    if constexpr (sizeof(real) == 8) {
        hit = kap > 0.0 && dtau >= L.tau;
        seg = l;
        if (hit) {
            const double li = L.tau / kap;
            seg = li > l ? l : li;
        }
    } else {
        hit = kap > 0.0f && (double)dtau >= L.tau;
        seg = hit ? fminf((float)L.tau * f32::fast_rcp(kap), l) : l;
    }
    if (P.J) {
        real v;
        if constexpr (sizeof(real) == 8)
            v = L.I * (seg / P.vol0[L.r.ib]);
        else
            v = L.I * seg * L.r.iv;
        atomicAdd(&lj[cell], v);
    }
    if (hit) {
        Pol::advance(L.r, seg);
        L.cell = cell;
        return Q_INT;
    }
    L.tau = L.tau - (double)dtau;
    const int st = Pol::cross_step(P, L.r, ap, l);
    if (st == f32::ST_IN)
        return Q_MARCH;
    L.blk = ap + (st == f32::ST_BLOCK_HI ? 4 : 0);
    L.ret = MARCH_BLOCK;
    return Q_RETIRE;
}

// The block change of a lane that march<LOCAL> left at its block's edge:
// MARCH_OUT with the packet in the next block, or MARCH_END (escaped through
// a domain face, or lost).
template <class Pol, class NB = nb_global>
__device__ __forceinline__ int leave_block(const kparams &P, qlane<Pol> &L, const NB &T)
{
    const int st = Pol::change_block(P, L.r, L.blk & 3, L.blk >> 2, T);
    if (st == f32::ST_IN)
        return MARCH_OUT;
    L.status = (st == f32::ST_ESCAPED) ? KP_END_ESCAPED : KP_END_LOST;
    return MARCH_END;
}

// Counters of a workgroup in LDS (KP_C_* order): a packet's end adds to them
// with LDS atomics, and the workgroup adds them to the launch's counters
// once (wg_counters_flush).  Kept out of the lanes' registers, which the
// march needs.
__device__ __forceinline__ void wg_counters_zero(unsigned long long *sc)
{
    if (threadIdx.x < KP_NCOUNTERS)
        sc[threadIdx.x] = 0;
}
__device__ __forceinline__ void wg_counters_flush(const kparams &P, const unsigned long long *sc)
{
    if (P.counters && threadIdx.x < KP_NCOUNTERS && sc[threadIdx.x] != 0)
        atomicAdd(P.counters + threadIdx.x, sc[threadIdx.x]);
}

// Counters and end state of a packet that ended.
template <class Pol>
__device__ __forceinline__ void end_packet(const kparams &P, const qlane<Pol> &L,
                                           unsigned long long *sc)
{
    atomicAdd(&sc[KP_C_CROSS_MAIN], (unsigned long long)L.ncross);
    atomicAdd(&sc[KP_C_CROSS_PEEL], (unsigned long long)L.npeel);
    atomicAdd(&sc[KP_C_EVENTS], (unsigned long long)L.events);
    atomicAdd(&sc[KP_C_PACKETS], 1ull);
    const int k = L.status == KP_END_ESCAPED ? KP_C_ESCAPED
                  : L.status == KP_END_RR    ? KP_C_KILLED_RR
                  : L.status == KP_END_STEPS ? KP_C_KILLED_STEPS
                                             : KP_C_LOST;
    atomicAdd(&sc[k], 1ull);
    if (P.end_state) {
        double *es = P.end_state + L.slot * KP_ES_STRIDE;
        double g[3];
        Pol::gpos_d(P, L.r, g);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            es[KP_ES_X + a] = g[a];
            es[KP_ES_D + a] = (double)L.r.d[a];
        }
        es[KP_ES_I] = (double)L.I;
        es[KP_ES_Q] = (double)L.Q;
        es[KP_ES_U] = (double)L.U;
        es[KP_ES_EVENTS] = (double)L.events;
        es[KP_ES_CROSS] = (double)L.ncross;
        es[KP_ES_STATUS] = (double)L.status;
        es[KP_ES_IB] = (double)L.r.ib;
        es[KP_ES_CELL] = (double)Pol::cell(P, L.r);
        // the peel kernel adds the crossings of the rays it marched for this
        // packet, before or after this (the generator cleared the entry)
        atomicAdd(&es[14], (double)L.npeel);
        es[15] = 0.0;
    }
}

// The queued peel-offs of one interaction: per observer the peel-off Stokes
// vector (the interaction's own arithmetic, on the recorded direction,
// frame, Stokes vector and g), the ray marched to the domain edge
// (kappa_rho from HBM), the attenuated vector deposited at the pixel of the
// interaction point.  Returns the rays' crossings.
template <class Pol>
__device__ __forceinline__ uint32_t peel_ray(const kparams &P, const ray_rec_t<typename Pol::real> &q)
{
    using real = typename Pol::real;
    typename Pol::ray t;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        t.x[a] = q.x[a];
        t.d[a] = 0;
    }
    t.ib = q.ib;
    unpack_cell<Pol>(t, q.cell);
    const int band = band_of(P, q.slot);
    const real ks = (real)P.bands[band].kappa_scale;
    real e2[3];
    uint64_t np = 0;
    if constexpr (sizeof(real) == 8) {
        f64::cross3(q.d, q.e1, e2);
        f64::gl_load(P, t);
        for (int o = 0; o < P.n_incl; ++o) {
            double Ip, Qi, Ui;
            f64::peel_stokes(P.model, q.g, P.pmax, q.d, q.e1, e2, q.s[0], q.s[1], q.s[2],
                             P.no[o], P.ev[o], Ip, Qi, Ui);
            const double tau = f64::tau_to_edge_sync(P, t, P.no[o], ks, np);
            const double att = kp_exp(-tau) * KP_INV_4PI;
            f64::deposit(P, band, o, t.x, Ip * att, Qi * att, Ui * att);
        }
    } else {
        f32::cross3(q.d, q.e1, e2);
        float gx[3];
        f32::global_pos(P, t, gx);
        for (int o = 0; o < P.n_incl; ++o) {
            float no[3], ev[3];
            load3(P.no[o], no);
            load3(P.ev[o], ev);
            float Ip, Qi, Ui;
            f32::peel_stokes(P.model, q.g, (float)P.pmax, q.d, q.e1, e2, q.s[0], q.s[1],
                             q.s[2], no, ev, Ip, Qi, Ui);
            const double tau = f32::tau_to_edge_sync(P, t, no, ks, np);
            const float att = f32::fast_exp(-(float)tau) * f32::kInv4Pi;
            f32::deposit(P, band, o, gx, Ip * att, Qi * att, Ui * att);
        }
    }
    if (P.end_state)
        atomicAdd(&P.end_state[(uint64_t)q.slot * KP_ES_STRIDE + 14], (double)np);
    return (uint32_t)np;
}

// The records of one block segment of a pass, sorted positions p .. e-1,
// marched by the workgroup's waves with the block's kappa_rho in lk, its J
// tile in lj and its face-neighbour entries in nbt.  Each lane holds one packet at a time; per loop iteration every
// marching lane takes one step, and the wave runs
//   - the interactions of its waiting lanes together once kPassIntBatch
//     lanes wait (or none marches),
//   - the ends of its lanes' marches together (block change and record of
//     the next pass, or the packet's end), each lane then taking the next
//     record of the segment, once kPassRetireBatch lanes are done (or none
//     marches); a wave claims the segment's sorted positions 64 at a time
//     (*s_next: the next unclaimed one, shared by the workgroup's waves) and
//     holds their record indices one per lane, so a refill waits for one
//     load (the record), not two,
// so lanes do not wait for the wave's slowest packet.  Every packet runs the
// steps and interactions of march<true>, in the same order.
template <class Pol>
__device__ void pass_block(const kparams &P, const qparams &Q,
                           const rec_t<typename Pol::real> *in, uint64_t p, uint64_t e,
                           const float *lk, tile_t<Pol> *lj, const blk_nb_t *nbt,
                           unsigned long long *sc, uint32_t *s_next)
{
    const uint32_t n = (uint32_t)(e - p);
    const int lane = threadIdx.x & 63;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    qlane<Pol> L;
    int mode = Q_IDLE;
    bool more = true; // wave-uniform: the segment may still hold records
    // the wave's window of claimed positions (wave-uniform): wlen of them
    // from wpos, wused taken; lane k holds the record index of wpos + k
    uint32_t wpos = 0, wlen = 0, wused = 0;
    int32_t word = 0;
    for (;;) {
        if (mode == Q_MARCH)
            mode = pass_step<Pol>(P, Q, L, lk, lj);
        const int nm = __popcll(__ballot(mode == Q_MARCH));
        const int ni = __popcll(__ballot(mode == Q_INT));
        if (ni >= kPassIntBatch || (ni > 0 && nm == 0)) {
            if (mode == Q_INT) {
                if (interact<Pol>(P, Q, L, (int64_t)L.r.ib * P.n3 + L.cell)) {
                    ++L.work;
                    mode = Q_MARCH;
                } else {
                    L.status = KP_END_RR;
                    L.ret = MARCH_END;
                    mode = Q_RETIRE;
                }
            }
            continue;
        }
        const int nr = __popcll(__ballot(mode == Q_RETIRE));
        const int nfree = 64 - nm - ni; // retiring and idle lanes
        if (nr >= kPassRetireBatch || (more && nfree >= kPassRetireBatch) || nm == 0) {
            if (mode == Q_RETIRE) {
                int m = L.ret;
                if (m == MARCH_BLOCK)
                    m = leave_block<Pol>(P, L, nb_local{nbt});
                if (m == MARCH_OUT)
                    push_out<Pol>(P, Q, L);
                else
                    end_packet<Pol>(P, L, sc);
                mode = Q_IDLE;
            }
            for (uint64_t idle = __ballot(mode == Q_IDLE); more && idle;
                 idle = __ballot(mode == Q_IDLE)) {
                if (wused == wlen) { // a new window of positions
                    uint32_t b = 0;
                    if (lane == 0)
                        b = atomicAdd(s_next, 64u);
                    b = __shfl(b, 0, 64);
                    if (b >= n) {
                        more = false;
                        break;
                    }
                    wpos = b;
                    wlen = n - b < 64u ? n - b : 64u;
                    wused = 0;
                    word = (uint32_t)lane < wlen ? Q.order[p + b + lane] : 0;
                }
                const uint32_t nidle = (uint32_t)__popcll(idle);
                const uint32_t take = nidle < wlen - wused ? nidle : wlen - wused;
                const uint32_t rk = (uint32_t)__popcll(idle & below);
                const int src = (int)(wused + (rk < take ? rk : take - 1));
                const int32_t o = __shfl(word, src, 64);
                if (mode == Q_IDLE && rk < take) {
                    load<Pol>(P, in[o], L);
                    L.work = 0;
                    mode = Q_MARCH;
                }
                wused += take;
                if (wused == wlen && wpos + wlen >= n)
                    more = false;
            }
            if (!more && __ballot(mode != Q_IDLE) == 0)
                break;
        }
    }
}

// Record of the next pass for each calling lane that crossed into another
// block: one slot claim per wave (ballot, then the lane's rank among the
// writers), the record and its key written.  Called by the writing lanes
// together.
template <class Pol>
__device__ __forceinline__ void push_out(const kparams &P, const qparams &Q, const qlane<Pol> &L)
{
    using real = typename Pol::real;
    const uint64_t k = claim(Q.n_out);
    store<Pol>(L, static_cast<rec_t<real> *>(Q.out)[k]);
    Q.out_key[k] = (uint32_t)(L.band - Q.first_band) * (uint32_t)Q.n_blocks + (uint32_t)L.r.ib;
}

} // namespace bq
} // namespace kp
