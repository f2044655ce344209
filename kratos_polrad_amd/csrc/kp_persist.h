// kp_persist.h -- persistent transport kernel: one march step per iteration.
//
// The reference runs each packet's whole history in one thread of one
// launch (integrator.h:80-106 -> photon.h:104-125).  Here every lane runs a
// small state machine instead:
//   MAIN  - one crossing of the packet's own path (tally, optical depth);
//   PEEL  - one crossing of a peel-off ray toward an observer;
//   PEND  - waiting for the wave's next action phase with a pending action:
//           an interaction, the end of a peel-off ray, or the end of the
//           packet;
//   NEED  - waiting for a new packet (the wave claims at least `chunk` slots
//           per global atomic and hands them out by ballot rank);
//   DONE  - no work left.
// Main-path and peel-off crossings execute the same face-distance and
// block-change code, so the wave stays converged on the crossing loop.  The
// long, divergent code (Philox, scattering, Stokes rotation, peel-off set-up,
// deposits, emission) runs in action phases, entered when at least
// `event_batch` lanes (or half the live lanes, or all marching lanes) are
// waiting, so it executes once for many lanes instead of once per lane.
// Each history performs exactly the per-packet arithmetic of the precision
// policy in the same order (bit-identical to the CPU oracle in FP64).
//
// Adjoint launches replay each history for its gradients.  The two-pass
// replay (ADJ) marches a packet twice: pass 1 sums its peel-off
// contributions A_k, pass 2 re-marches it and deposits the gradients with
// B = the A_k still ahead.  The one-pass replay (ADJ + DEFER, mixed
// precision) marches the packet once and defers every deposit: each
// main-path vertex becomes a record holding C = the sum of the packet's A_k
// so far (so B = Atot - C) and the start of the segment that follows, each
// peel-off ray a record with its A_k; the packet's Atot goes to a per-slot
// array.  Separate kernels then deposit the vertex terms and march the
// segments and the peel-off rays.  A packet whose records do not fit is
// marked (Atot = NaN, its records are skipped) and left to the two-pass
// kernel.
#pragma once

#include "kp_transport32.h"
#include "kp_transport64.h"

namespace kp {

// the march code reads the statuses of both precisions by one set of names
static_assert(f32::ST_IN == f64::ST_IN && f32::ST_ESCAPED == f64::ST_ESCAPED &&
                  f32::ST_BLOCK_LO == f64::ST_BLOCK_LO && f32::ST_BLOCK_HI == f64::ST_BLOCK_HI &&
                  f32::ST_LOST == f64::ST_LOST,
              "FP32 and FP64 crossing statuses differ");
enum : int { M_NEED = 0, M_MAIN = 1, M_PEEL = 2, M_DONE = 3, M_PEND = 4 };
enum : int { PA_NONE = 0, PA_INTERACT = 1, PA_PEEL_END = 2, PA_SCATTER = 3, PA_FINISH = 4,
             PA_BLOCK = 5 };
enum : int { EV_EMIT = 0, EV_SCATTER = 1 };

// ---------------------------------------------------------------- policies
struct pol64 {
    using real = double;
    using ray = f64::ray_t;
    __device__ static void at_source(const kparams &P, ray &r)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            r.x[a] = P.src_x[a];
            r.i[a] = P.src_i[a];
        }
        r.ib = P.src_ib;
        f64::gl_load(P, r);
    }
    __device__ static void relocate(const kparams &P, ray &r) { f64::gl_load(P, r); }
    __device__ static void set_dir(const kparams &, ray &, const real *) {}
    // block changes wait for a block phase of the wave (block_stage)
    static constexpr bool kDeferBlock = true;
    __device__ static int dl(const ray &r, real &l) { return f64::get_dl(r, l); }
    __device__ static int cross(const kparams &P, ray &r, int ap, real l)
    {
        bool rl;
        const int st = f64::cross_face(P, r, ap, l, rl);
        if (st == f64::ST_IN && rl)
            f64::gl_load(P, r);
        return st;
    }
    __device__ static int cross_step(const kparams &P, ray &r, int ap, real l)
    {
        bool rl;
        const int st = f64::cross_in_block(P, r, ap, l, rl);
        if (st == f64::ST_IN && rl)
            f64::gl_load(P, r);
        return st;
    }
    template <class NB>
    __device__ static int change_block(const kparams &P, ray &r, int a, int side, const NB &T)
    {
        bool rl = false;
        const int st = f64::change_block(P, r, a, side, rl, T);
        if (st == f64::ST_IN && rl)
            f64::gl_load(P, r);
        return st;
    }
    __device__ static int change_block(const kparams &P, ray &r, int a, int side)
    {
        return change_block(P, r, a, side, nb_global{P});
    }
    __device__ static void advance(ray &r, real l)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            r.x[a] = r.x[a] + l * r.d[a];
    }
    __device__ static void gpos(const kparams &, const ray &r, real *g)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            g[a] = r.x[a];
    }
    __device__ static void gpos_d(const kparams &, const ray &r, double *g)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            g[a] = r.x[a];
    }
    __device__ static int cell(const kparams &P, const ray &r) { return f64::cell_index(P, r); }
};

struct pol32 {
    using real = float;
    using ray = f32::ray_t;
    __device__ static void at_source(const kparams &P, ray &r)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            r.x[a] = P.src_xl[a];
            r.i[a] = P.src_i[a];
        }
        r.ib = P.src_ib;
        f32::set_scales(P, r);
    }
    __device__ static void relocate(const kparams &P, ray &r) { f32::set_scales(P, r); }
    __device__ static void set_dir(const kparams &P, ray &r, const real *) { f32::set_scales(P, r); }
    // block changes wait for a block phase of the wave (block_stage)
    static constexpr bool kDeferBlock = true;
    __device__ static int dl(const ray &r, real &l) { return f32::get_dl(r, l); }
    __device__ static int cross(const kparams &P, ray &r, int ap, real l)
    {
        return f32::cross_face(P, r, ap, l);
    }
    __device__ static int cross_step(const kparams &P, ray &r, int ap, real l)
    {
        return f32::cross_in_block(P, r, ap, l);
    }
    template <class NB>
    __device__ static int change_block(const kparams &P, ray &r, int a, int side, const NB &T)
    {
        return f32::change_block(P, r, a, side, T);
    }
    __device__ static int change_block(const kparams &P, ray &r, int a, int side)
    {
        return f32::change_block(P, r, a, side);
    }
    __device__ static void advance(ray &r, real l)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            r.x[a] = fmaf(l, r.cpl[a], r.x[a]);
    }
    __device__ static void gpos(const kparams &P, const ray &r, real *g)
    {
        f32::global_pos(P, r, g);
    }
    __device__ static void gpos_d(const kparams &P, const ray &r, double *g)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            g[a] = (double)P.org32[3 * r.ib + a] + (double)r.x[a] * (double)P.h32[3 * r.ib + a];
    }
    __device__ static int cell(const kparams &P, const ray &r) { return f32::cell_index(P, r); }
};

// FP32 observers: loaded per use from the launch parameters
template <class R> __device__ __forceinline__ void load3(const double *s, R *d)
{
    d[0] = (R)s[0];
    d[1] = (R)s[1];
    d[2] = (R)s[2];
}

// ---------------------------------------------------------------- kernel
template <class Pol> struct lane_t {
    using real = typename Pol::real;
    using ray = typename Pol::ray;
    ray cur;          // ray being marched (main path or peel-off)
    real sx[3];       // main path saved while peel-off rays run
    real sd[3];
    int si[3];
    int sib;
    real e1[3];
    real I, Q, U;     // packet Stokes vector
    real Pi, Pq, Pu;  // Stokes vector of the current peel-off ray
    double tau;       // MAIN: optical depth left; PEEL: accumulated
    double tau_main;
    real gg;          // HG asymmetry at the current interaction
    real alb;         // albedo at the current interaction
    uint32_t u[4];    // Philox words of the current interaction
    uint64_t slot, pid;
    int band;
    int mode, ev_kind, o;
    int pa, status;   // pending action and, for PA_FINISH, the end status
    int blk;          // PA_BLOCK: exit axis + 4 * face side + 8 * the march mode to resume
    int64_t gc;       // cell of the pending / current interaction
    int64_t kgc;      // cell whose kappa_rho is in kcache (issued one step ahead)
    float kcache;
    int events, steps;
    uint32_t ncross, npeel;
    real ks;            // extinction scale of the packet's band
    f32::tally_acc acc; // FP32: Kahan register accumulator of the current cell
    // path replay (adjoint launches only)
    int pass;         // 1: collect peel-off contributions A_k; 2: gradients
    int peel_sub;     // pass 2: 0 = re-measuring an A_k beyond the row, 1 = depositing
    uint32_t kslot;   // next A_k slot of this packet
    double Atot;      // sum of the packet's A_k
    double B;         // pass 2: sum of the A_k still ahead
    double Acur;      // A_k of the current peel-off ray
    double jsum;      // lane's sum of A_k over its packets
    // one-pass replay (deferred deposits)
    double Cv;        // A_k sum before the current vertex's peel-offs
    double phs;       // the vertex's peel-off phase terms
    int tover;        // a queue overflowed: the two-pass kernel redoes the packet
};

struct lane_counters {
    uint32_t cross = 0, peel = 0, events = 0, escaped = 0, rr = 0, steps = 0, lost = 0,
             packets = 0;
};

__device__ __forceinline__ double dlogp_hg(double gg, double mu)
{
    const double D = (1.0 + gg * gg) - 2.0 * gg * mu;
    return -2.0 * gg / (1.0 - gg * gg) - 3.0 * (gg - mu) / D;
}

template <class Pol, bool ADJ, bool DEFER> __device__ __forceinline__ void emit_packet(const kparams &P, lane_t<Pol> &L);

template <class Pol, bool ADJ, bool DEFER>
__device__ __forceinline__ void finish(const kparams &P, lane_t<Pol> &L, lane_counters &C,
                                       int status, bool lds)
{
    if constexpr (!ADJ && sizeof(typename Pol::real) == 4)
        L.acc.flush(P, L.band, lds);
    C.cross += L.ncross;
    C.peel += L.npeel;
    C.events += (uint32_t)L.events;
    if constexpr (ADJ && DEFER) {
        if (L.tover) {
            P.pk_atot[L.slot] = __builtin_nan("");
            const unsigned int f = atomicAdd(P.fb_count, 1u);
            P.fb_list[f] = (int32_t)L.slot;
        } else {
            P.pk_atot[L.slot] = L.Atot;
            L.jsum += L.Atot;
        }
    } else if constexpr (ADJ) {
        if (L.pass == 1) {
            // replay the same history for the gradients
            L.pass = 2;
            emit_packet<Pol, ADJ, DEFER>(P, L);
            return;
        }
        L.jsum += L.Atot;
        L.pass = 1;
    }
    C.packets += 1;
    if (status == KP_END_ESCAPED)
        C.escaped += 1;
    else if (status == KP_END_RR)
        C.rr += 1;
    else if (status == KP_END_STEPS)
        C.steps += 1;
    else
        C.lost += 1;
    if (!ADJ && P.end_state) {
        double *es = P.end_state + L.slot * KP_ES_STRIDE;
        double g[3];
        Pol::gpos_d(P, L.cur, g);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            es[KP_ES_X + a] = g[a];
            es[KP_ES_D + a] = (double)L.cur.d[a];
        }
        es[KP_ES_I] = (double)L.I;
        es[KP_ES_Q] = (double)L.Q;
        es[KP_ES_U] = (double)L.U;
        es[KP_ES_EVENTS] = (double)L.events;
        es[KP_ES_CROSS] = (double)L.ncross;
        es[KP_ES_STATUS] = (double)status;
        es[KP_ES_IB] = (double)L.cur.ib;
        es[KP_ES_CELL] = (double)Pol::cell(P, L.cur);
        es[14] = (double)L.npeel;
        es[15] = 0.0;
    }
    L.mode = M_NEED;
}

// Wave-private chunks [base, end) of the peel-off and vertex queues, in LDS
// (one entry per wave of the workgroup): a wave claims kQChunk slots with one
// atomic and fills them itself, so a queue's counter sees one atomic per
// chunk, and lanes that write together write neighbouring records.  The
// counters and bounds are 64-bit: waves keep claiming chunks past a queue's
// capacity (those records are not written), so a 32-bit counter could wrap
// back below it.
constexpr uint32_t kQChunk = 256;
__shared__ uint64_t kp_pq_base[4], kp_pq_end[4], kp_vq_base[4], kp_vq_end[4];

// One queue slot for each calling lane (called by those lanes together).
__device__ __forceinline__ uint64_t claim_q(unsigned long long *count, uint64_t *qbase,
                                            uint64_t *qend)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(1);
    const int leader = __ffsll((long long)m) - 1;
    const uint32_t n = (uint32_t)__popcll(m);
    const uint32_t rank = (uint32_t)__popcll((lane == 0) ? 0ull : (m & (~0ull >> (64 - lane))));
    const uint64_t b0 = qbase[w], e0 = qend[w];
    const uint32_t avail = (uint32_t)(e0 - b0);
    uint64_t k = b0 + rank;
    if (n > avail) {
        // the rest of the chunk, then a new one
        unsigned long long nb = 0;
        if (lane == leader)
            nb = atomicAdd(count, (unsigned long long)kQChunk);
        nb = __shfl(nb, leader, 64);
        if (rank >= avail)
            k = (uint64_t)nb + (rank - avail);
        if (lane == leader) {
            qbase[w] = nb + (n - avail);
            qend[w] = nb + kQChunk;
        }
    } else if (lane == leader) {
        qbase[w] = b0 + n;
    }
    return k;
}

// Queue the finished peel-off ray (start point L.sx / L.si / L.sib,
// observer L.o, A_k = L.Acur) for the gradient-march kernel.  Called by the
// lanes that queue, together.  Returns false beyond the queue's capacity.
template <class Pol>
__device__ __forceinline__ bool queue_peel(const kparams &P, const lane_t<Pol> &L)
{
    using real = typename Pol::real;
    const uint64_t k = claim_q(P.pq_count, kp_pq_base, kp_pq_end);
    if (k >= P.pq_cap)
        return false;
    peel_rec_t<real> r;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        r.x[a] = L.sx[a];
    r.ib = L.sib;
    r.cell_o = (uint32_t)L.si[0] | ((uint32_t)L.si[1] << 8) | ((uint32_t)L.si[2] << 16) |
               ((uint32_t)L.o << 24);
    r.slot = (int32_t)L.slot;
    r.A = L.Acur;
    static_cast<peel_rec_t<real> *>(P.pq)[k] = r;
    return true;
}

// Queue the current main-path vertex: the ray L.cur starts the next segment
// (free flight L.tau, L.steps crossings left, or none when `ended`).  Called
// by the lanes that queue, together.  Returns false beyond the capacity.
template <class Pol>
__device__ __forceinline__ bool queue_vtx(const kparams &P, const lane_t<Pol> &L, int32_t gc,
                                          typename Pol::real mu, bool ended)
{
    using real = typename Pol::real;
    const uint64_t k = claim_q(P.vq_count, kp_vq_base, kp_vq_end);
    if (k >= P.vq_cap)
        return false;
    vtx_rec_t<real> r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.x[a] = L.cur.x[a];
        r.d[a] = L.cur.d[a];
    }
    r.ib = L.cur.ib;
    r.cell = (uint32_t)L.cur.i[0] | ((uint32_t)L.cur.i[1] << 8) | ((uint32_t)L.cur.i[2] << 16);
    r.tau = L.tau;
    r.Cv = L.Cv;
    r.Cs = L.Atot;
    r.phs = L.phs;
    r.slot = (int32_t)L.slot;
    r.gc = gc;
    r.steps = ended ? 0 : L.steps;
    r.alb = L.alb;
    r.mu = mu;
    static_cast<vtx_rec_t<real> *>(P.vq)[k] = r;
    return true;
}

// Marks the unused rest of the wave's queue chunks (ib = -1) when the wave
// is done; the queue kernels skip such records.
template <class R> __device__ __forceinline__ void close_q_chunks(const kparams &P)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (P.pq) {
        const uint64_t e = kp_pq_end[w] < P.pq_cap ? kp_pq_end[w] : P.pq_cap;
        for (uint64_t k = kp_pq_base[w] + lane; k < e; k += 64)
            static_cast<peel_rec_t<R> *>(P.pq)[k].ib = -1;
    }
    if (P.vq) {
        const uint64_t e = kp_vq_end[w] < P.vq_cap ? kp_vq_end[w] : P.vq_cap;
        for (uint64_t k = kp_vq_base[w] + lane; k < e; k += 64)
            static_cast<vtx_rec_t<R> *>(P.vq)[k].ib = -1;
    }
}

// Start the peel-off ray toward observer L.o from the saved main position.
template <class Pol, bool ADJ, bool DEFER>
__device__ __forceinline__ void start_peel(const kparams &P, lane_t<Pol> &L, bool restart = false)
{
    using real = typename Pol::real;
    real no[3], ev[3];
    load3(P.no[L.o], no);
    load3(P.ev[L.o], ev);
    if (!restart) {
        if (L.ev_kind == EV_EMIT) {
            L.Pi = L.I;
            L.Pq = 0;
            L.Pu = 0;
        } else {
            real e2[3];
            if constexpr (sizeof(real) == 8) {
                f64::cross3(L.sd, L.e1, e2);
                f64::peel_stokes(P.model, L.gg, P.pmax, L.sd, L.e1, e2, L.I, L.Q, L.U, no, ev,
                                 L.Pi, L.Pq, L.Pu);
            } else {
                f32::cross3(L.sd, L.e1, e2);
                f32::peel_stokes(P.model, L.gg, (float)P.pmax, L.sd, L.e1, e2, L.I, L.Q, L.U, no,
                                 ev, L.Pi, L.Pq, L.Pu);
            }
        }
        if constexpr (ADJ) {
            if (!DEFER && L.pass == 2) {
                const uint64_t lane_id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
                if (L.kslot < P.adj_cap) {
                    L.Acur = P.adj_rows[(uint64_t)L.kslot * P.adj_lanes + lane_id];
                    L.peel_sub = 1;
                } else {
                    L.peel_sub = 0; // beyond the row: measure A_k first
                }
            } else {
                L.peel_sub = 0;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        L.cur.x[a] = L.sx[a];
        L.cur.i[a] = L.si[a];
        L.cur.d[a] = no[a];
    }
    L.cur.ib = L.sib;
    Pol::relocate(P, L.cur);
    L.tau = 0.0;
    L.kgc = -1;
    L.mode = M_PEEL;
}

template <class Pol> __device__ __forceinline__ void save_main(lane_t<Pol> &L)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        L.sx[a] = L.cur.x[a];
        L.sd[a] = L.cur.d[a];
        L.si[a] = L.cur.i[a];
    }
    L.sib = L.cur.ib;
    L.tau_main = L.tau;
}

// Scatter, new free flight, Russian roulette (after the interaction's
// peel-offs).  Leaves the lane in MAIN, or pending PA_FINISH when killed.
template <class Pol, bool ADJ, bool DEFER>
__device__ __forceinline__ void scatter_stage(const kparams &P, lane_t<Pol> &L)
{
    using real = typename Pol::real;
    real e2[3];
    bool killed = false;
    const bool hg = P.model == KP_HG_RAYLEIGH && (P.bands[L.band].flags & KP_BAND_CELL_G);
    if constexpr (sizeof(real) == 8) {
        f64::cross3(L.cur.d, L.e1, e2);
        const double mu = (P.model == KP_THOMSON) ? kp_thomson_mu(kp_u01(L.u[0]))
                                                  : kp_hg_mu(L.gg, kp_u01(L.u[0]));
        if constexpr (ADJ && !DEFER) {
            if (L.pass == 2 && hg && P.g_g && L.B != 0.0)
                atomic_add_f64(g_base(P, P.g_g) + L.gc, L.B * dlogp_hg(L.gg, mu));
        }
        f64::scatter_stokes(P.model, P.pmax, mu, kp_u01(L.u[1]), L.cur.d, L.e1, e2, L.I, L.Q,
                            L.U);
        L.tau = -kp_log(kp_u01(L.u[2]));
        if (L.I < P.rr_w * P.w0) {
            if (kp_u01(L.u[3]) < P.rr_s) {
                L.I = L.I / P.rr_s;
                L.Q = L.Q / P.rr_s;
                L.U = L.U / P.rr_s;
            } else {
                killed = true;
            }
        }
        if constexpr (ADJ && DEFER) {
            if (P.vq && !L.tover)
                if (!queue_vtx<Pol>(P, L, (int32_t)L.gc, mu, killed))
                    L.tover = 1;
        }
    } else {
        f32::cross3(L.cur.d, L.e1, e2);
        const float xi0 = f32::u01(L.u[0]);
        const float mu = (P.model == KP_THOMSON) ? f32::thomson_mu(xi0) : f32::hg_mu(L.gg, xi0);
        if constexpr (ADJ && !DEFER) {
            if (L.pass == 2 && hg && P.g_g && L.B != 0.0)
                atomic_add_f64(g_base(P, P.g_g) + L.gc,
                               L.B * dlogp_hg((double)L.gg, (double)mu));
        }
        f32::scatter_stokes(P.model, (float)P.pmax, mu, f32::u01(L.u[1]), L.cur.d, L.e1, e2, L.I,
                            L.Q, L.U);
        Pol::set_dir(P, L.cur, L.cur.d);
        L.tau = -(double)f32::fast_log(f32::u01(L.u[2]));
        if (L.I < (float)P.rr_w * (float)P.w0) {
            if (f32::u01(L.u[3]) < (float)P.rr_s) {
                const float inv = (float)(1.0 / P.rr_s);
                L.I *= inv;
                L.Q *= inv;
                L.U *= inv;
            } else {
                killed = true;
            }
        }
        if constexpr (ADJ && DEFER) {
            if (P.vq && !L.tover)
                if (!queue_vtx<Pol>(P, L, (int32_t)L.gc, mu, killed))
                    L.tover = 1;
        }
    }
    L.kgc = -1;
    if (killed) {
        L.pa = PA_FINISH;
        L.status = KP_END_RR;
    } else {
        L.mode = M_MAIN;
        L.pa = PA_NONE;
    }
}

// A_k of the finished peel-off ray: dL/dimage at its pixel . S exp(-tau)/(4 pi)
template <class Pol> __device__ __forceinline__ double peel_A(const kparams &P, const lane_t<Pol> &L)
{
    using real = typename Pol::real;
    long px;
    double Si, Sq, Su;
    if constexpr (sizeof(real) == 8) {
        const double att = kp_exp(-L.tau) * KP_INV_4PI;
        px = f64::pixel(P, L.band, L.o, L.sx);
        Si = L.Pi * att;
        Sq = L.Pq * att;
        Su = L.Pu * att;
    } else {
        f32::ray_t tmp;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            tmp.x[a] = L.sx[a];
        tmp.ib = L.sib;
        float gx[3];
        f32::global_pos(P, tmp, gx);
        const float att = f32::fast_exp(-(float)L.tau) * f32::kInv4Pi;
        px = f32::pixel(P, L.band, L.o, gx);
        Si = (double)(L.Pi * att);
        Sq = (double)(L.Pq * att);
        Su = (double)(L.Pu * att);
    }
    if (px < 0)
        return 0.0;
    const double *w = P.adj_w + 3 * px;
    return w[0] * Si + w[1] * Sq + w[2] * Su;
}

// End of a peel-off ray: deposit (forward) or A_k bookkeeping (adjoint),
// then the next observer, or resume the main path (EMIT: march on;
// SCATTER: pending scatter).
template <class Pol, bool ADJ, bool DEFER>
__device__ __forceinline__ void peel_end_stage(const kparams &P, lane_t<Pol> &L, bool lds)
{
    using real = typename Pol::real;
    if constexpr (ADJ && DEFER) {
        const double A = peel_A<Pol>(P, L);
        L.Atot += A;
        if (L.ev_kind == EV_SCATTER && P.model == KP_HG_RAYLEIGH &&
            (P.bands[L.band].flags & KP_BAND_CELL_G) && P.g_g && A != 0.0) {
            real no[3];
            load3(P.no[L.o], no);
            double mo;
            if constexpr (sizeof(real) == 8)
                mo = f64::dot3(L.sd, no);
            else
                mo = (double)f32::dot3(L.sd, no);
            mo = fmin(fmax(mo, -1.0), 1.0);
            L.phs += A * dlogp_hg((double)L.gg, mo);
        }
        if (P.pq && P.g_logrho && A != 0.0 && !L.tover) {
            L.Acur = A;
            if (!queue_peel<Pol>(P, L))
                L.tover = 1;
        }
    } else if constexpr (ADJ) {
        const uint64_t lane_id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (L.pass == 1) {
            const double A = peel_A<Pol>(P, L);
            if (L.kslot < P.adj_cap)
                P.adj_rows[(uint64_t)L.kslot * P.adj_lanes + lane_id] = A;
            L.Atot += A;
        } else {
            if (L.peel_sub == 0) {
                // measured an A_k that did not fit the row: march it again
                L.Acur = peel_A<Pol>(P, L);
                L.peel_sub = 1;
                start_peel<Pol, ADJ, DEFER>(P, L, true);
                L.pa = PA_NONE;
                return;
            }
            if (L.ev_kind == EV_SCATTER && P.model == KP_HG_RAYLEIGH &&
                (P.bands[L.band].flags & KP_BAND_CELL_G) && P.g_g && L.Acur != 0.0) {
                real no[3];
                load3(P.no[L.o], no);
                double mo;
                if constexpr (sizeof(real) == 8)
                    mo = f64::dot3(L.sd, no);
                else
                    mo = (double)f32::dot3(L.sd, no);
                mo = fmin(fmax(mo, -1.0), 1.0);
                atomic_add_f64(g_base(P, P.g_g) + L.gc, L.Acur * dlogp_hg((double)L.gg, mo));
            }
            L.B -= L.Acur;
        }
        L.kslot += 1;
    } else if (P.image) {
        real v[3];
        long px;
        if constexpr (sizeof(real) == 8) {
            const double att = kp_exp(-L.tau) * KP_INV_4PI;
            px = f64::pixel(P, L.band, L.o, L.sx);
            v[0] = L.Pi * att;
            v[1] = L.Pq * att;
            v[2] = L.Pu * att;
        } else {
            f32::ray_t tmp;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                tmp.x[a] = L.sx[a];
            tmp.ib = L.sib;
            float gx[3];
            f32::global_pos(P, tmp, gx);
            const float att = f32::fast_exp(-(float)L.tau) * f32::kInv4Pi;
            px = f32::pixel(P, L.band, L.o, gx);
            v[0] = L.Pi * att;
            v[1] = L.Pq * att;
            v[2] = L.Pu * att;
        }
        if (px >= 0) {
            // emission peel-offs of one band all land on the source pixel
            // with equal values: summed in LDS, added once per workgroup
            if (lds && L.ev_kind == EV_EMIT) {
                const int k = P.n_bands + 3 * (L.band * P.n_incl + L.o);
                atomicAdd(&kp_lds[k + 0], (double)v[0]);
                atomicAdd(&kp_lds[k + 1], (double)v[1]);
                atomicAdd(&kp_lds[k + 2], (double)v[2]);
            } else {
                atomic_add_f64(img_at(P, px, 0), (double)v[0]);
                atomic_add_f64(img_at(P, px, 1), (double)v[1]);
                atomic_add_f64(img_at(P, px, 2), (double)v[2]);
            }
        }
    }
    L.o += 1;
    if (L.o < P.n_incl) {
        start_peel<Pol, ADJ, DEFER>(P, L);
        L.pa = PA_NONE;
        return;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        L.cur.x[a] = L.sx[a];
        L.cur.d[a] = L.sd[a];
        L.cur.i[a] = L.si[a];
    }
    L.cur.ib = L.sib;
    Pol::relocate(P, L.cur);
    L.tau = L.tau_main;
    L.kgc = -1;
    if (L.ev_kind == EV_EMIT) {
        L.mode = M_MAIN;
        L.pa = PA_NONE;
        if constexpr (ADJ && DEFER) {
            // the emission vertex: the first segment, no interaction terms
            if (P.vq && !L.tover) {
                L.Cv = L.Atot;
                if (!queue_vtx<Pol>(P, L, -1, 0.0f, false))
                    L.tover = 1;
            }
        }
    } else {
        L.pa = PA_SCATTER;
    }
}

template <class Pol, bool ADJ, bool DEFER>
__device__ __forceinline__ void emit_packet(const kparams &P, lane_t<Pol> &L)
{
    L.band = band_of(P, L.slot);
    L.ks = (typename Pol::real)P.bands[L.band].kappa_scale;
    L.pid = P.packet_offset + (L.slot - P.band_start[L.band]);
    uint32_t rn[4] = {(uint32_t)L.pid, (uint32_t)(L.pid >> 32), 0u, 0u};
    kp_philox4x32_10(rn, P.seed, (uint32_t)L.band);
    if constexpr (sizeof(typename Pol::real) == 8)
        f64::emit(rn, L.cur.d, L.e1, L.tau);
    else
        f32::emit(rn, L.cur.d, L.e1, L.tau);
    L.I = (typename Pol::real)P.w0;
    L.Q = 0;
    L.U = 0;
    Pol::at_source(P, L.cur);
    L.kgc = -1;
    L.events = 0;
    L.steps = P.step_max;
    L.ncross = 0;
    L.npeel = 0;
    L.pa = PA_NONE;
    if constexpr (ADJ) {
        L.kslot = 0;
        if (L.pass == 1)
            L.Atot = 0.0;
        else
            L.B = L.Atot;
        if constexpr (DEFER) {
            L.tover = 0;
            L.Cv = 0.0;
            L.phs = 0.0;
            L.alb = (typename Pol::real)1;
        }
    }
    if (P.n_incl > 0) {
        save_main(L);
        L.ev_kind = EV_EMIT;
        L.o = 0;
        start_peel<Pol, ADJ, DEFER>(P, L);
    } else {
        L.mode = M_MAIN;
    }
}

// Interaction at the current point: Philox words of the event, absorption,
// then the peel-offs (or the scatter directly when there is no observer).
template <class Pol, bool ADJ, bool DEFER>
__device__ __forceinline__ void interact_stage(const kparams &P, lane_t<Pol> &L)
{
    using real = typename Pol::real;
    const int64_t gc = L.gc;
    L.events += 1;
    L.u[0] = (uint32_t)L.pid;
    L.u[1] = (uint32_t)(L.pid >> 32);
    L.u[2] = (uint32_t)L.events;
    L.u[3] = 0u;
    kp_philox4x32_10(L.u, P.seed, (uint32_t)L.band);
    const dband &bd = P.bands[L.band];
    real alb, gg;
    if constexpr (sizeof(real) == 8) {
        alb = (bd.flags & KP_BAND_CELL_ALBEDO) ? (double)P.albedo[gc] : bd.albedo;
        gg = (bd.flags & KP_BAND_CELL_G) ? (P.g ? (double)P.g[gc] : 0.0) : bd.g;
        L.I = L.I * alb;
        L.Q = L.Q * alb;
        L.U = L.U * alb;
    } else {
        alb = (bd.flags & KP_BAND_CELL_ALBEDO) ? P.albedo[gc] : (float)bd.albedo;
        gg = (bd.flags & KP_BAND_CELL_G) ? (P.g ? P.g[gc] : 0.0f) : (float)bd.g;
        L.I *= alb;
        L.Q *= alb;
        L.U *= alb;
    }
    L.gg = gg;
    L.alb = alb;
    if constexpr (ADJ && DEFER) {
        L.Cv = L.Atot;
        L.phs = 0.0;
    } else if constexpr (ADJ) {
        if (L.pass == 2 && L.B != 0.0) {
            if (P.g_logrho)
                atomic_add_f64(g_base(P, P.g_logrho) + gc, L.B);
            if ((bd.flags & KP_BAND_CELL_ALBEDO) && P.g_alb)
                atomic_add_f64(g_base(P, P.g_alb) + gc, L.B / (double)alb);
        }
    }
    if (P.n_incl > 0) {
        save_main(L);
        L.ev_kind = EV_SCATTER;
        L.o = 0;
        start_peel<Pol, ADJ, DEFER>(P, L);
        L.pa = PA_NONE;
    } else {
        L.pa = PA_SCATTER;
    }
}

// Adds the workgroup's LDS accumulators (kp_lds) to J and the image.
// Called by every thread of the workgroup after a barrier.
template <class Pol> __device__ void flush_lds_acc(const kparams &P)
{
    const double *sh = kp_lds;
    using real = typename Pol::real;
    for (int b = threadIdx.x; b < P.n_bands; b += blockDim.x)
        if (P.J && sh[b] != 0.0)
            atomic_add_f64(j_at(P, (uint64_t)P.src_gc, b), sh[b]);
    const int n = P.n_bands * P.n_incl;
    if (!P.image)
        return;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double *v = sh + P.n_bands + 3 * k;
        if (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0)
            continue;
        const int b = k / P.n_incl, o = k - b * P.n_incl;
        // the emission point exactly as the packets saw it
        typename Pol::ray r;
        Pol::at_source(P, r);
        real g[3];
        Pol::gpos(P, r, g);
        long px;
        if constexpr (sizeof(real) == 8)
            px = f64::pixel(P, b, o, g);
        else
            px = f32::pixel(P, b, o, g);
        if (px < 0)
            continue;
        atomic_add_f64(img_at(P, px, 0), v[0]);
        atomic_add_f64(img_at(P, px, 1), v[1]);
        atomic_add_f64(img_at(P, px, 2), v[2]);
    }
}

// Block change of a lane that left its block (PA_BLOCK), run for the
// waiting lanes of a wave together: the neighbour-table lookups and the
// level rescale are issued once per block phase instead of in nearly every
// march step.  The lane resumes the march it left, with the next cell's
// kappa_rho load issued, or ends that march as the march step would have.
template <class Pol> __device__ __forceinline__ void block_stage(const kparams &P, lane_t<Pol> &L)
{
    const int mode = L.blk >> 3;
    const int st = Pol::change_block(P, L.cur, L.blk & 3, (L.blk >> 2) & 1);
    if (st == f32::ST_IN) {
        L.mode = mode;
        L.pa = PA_NONE;
        L.kgc = (int64_t)L.cur.ib * P.n3 + Pol::cell(P, L.cur);
        L.kcache = P.kappa[L.kgc];
    } else if (mode == M_MAIN) {
        L.pa = PA_FINISH;
        L.status = (st == f32::ST_ESCAPED) ? KP_END_ESCAPED : KP_END_LOST;
    } else {
        L.pa = PA_PEEL_END;
    }
}

// Waiting block changes that start a block phase of their own.
constexpr uint32_t kBlockBatch = 16;

// lds: the launch gives each workgroup the kp_lds accumulators.
// SYNC (mixed forward in emission-direction order): a wave takes new packets
// only when all its lanes are idle, so lanes that start together fly
// neighbouring rays in step; each main-path J deposit is handed to the next
// loop iteration, where the wave sums runs of lanes in one cell before one
// atomic (no per-lane register accumulator).
template <class Pol, bool ADJ, bool DEFER, bool SYNC = false>
__device__ void persist_body(const kparams &P, unsigned long long *work, lane_counters &C,
                             double &jsum, bool lds)
{
    using real = typename Pol::real;
    // forward launches defer block changes; the replays keep them in the
    // march step (their larger lane state would spill), and so does the
    // lockstep variant (its waves would wait on each parked lane's
    // un-prefetched kappa_rho load)
    constexpr bool kDB = Pol::kDeferBlock && !ADJ && !SYNC;
    lane_t<Pol> L;
    L.mode = M_NEED;
    L.pa = PA_NONE;
    L.blk = 0;
    L.kgc = -1;
    L.kcache = 0.0f;
    L.pass = 1;
    L.jsum = 0.0;
    L.Atot = 0.0;
    L.B = 0.0;
    L.Acur = 0.0;
    L.peel_sub = 0;
    L.tover = 0;
    L.Cv = 0.0;
    L.phs = 0.0;
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t wbase = 0, wend = 0; // wave-uniform slot range
    if constexpr (DEFER) {
        if (lane == 0) {
            kp_pq_base[threadIdx.x >> 6] = 0;
            kp_pq_end[threadIdx.x >> 6] = 0;
            kp_vq_base[threadIdx.x >> 6] = 0;
            kp_vq_end[threadIdx.x >> 6] = 0;
        }
    }

    bool jh = false; // SYNC: a J deposit waits for the wave (element jidx, value jv)
    uint64_t jidx = 0;
    double jv = 0.0;

    for (;;) {
        if constexpr (SYNC) {
            if (__ballot(jh)) {
                const uint64_t key = jh ? jidx : ~(uint64_t)lane;
                const uint64_t kprev = __shfl_up(key, 1, 64);
                if (__ballot(jh && lane > 0 && kprev == key))
                    wave_run_add(j_base(P), jh, jidx, jv);
                else if (jh)
                    atomic_add_f64(j_base(P) + jidx, jv);
            }
            jh = false;
        }
        const uint64_t live = __ballot(L.mode != M_DONE);
        if (!live)
            break;
        // lanes waiting for a block change do not count towards an action
        // phase; they get a block phase when enough of them wait, when no
        // lane marches, or with an action phase
        const bool bwait = kDB && L.mode == M_PEND && L.pa == PA_BLOCK;
        const uint64_t wait = __ballot((L.mode == M_PEND && !bwait) || L.mode == M_NEED);
        const uint64_t march = __ballot(L.mode == M_MAIN || L.mode == M_PEEL);
        const uint32_t nw = __popcll(wait);
        const bool phase = wait && (nw >= P.batch || 2 * nw >= (uint32_t)__popcll(live) || !march);
        if constexpr (kDB) {
            const uint64_t bw = __ballot(bwait);
            if (bw && (phase || !march || (uint32_t)__popcll(bw) >= kBlockBatch)) {
                if (bwait)
                    block_stage<Pol>(P, L);
            }
        }
        if (phase) {
            // ---------------- action phase: every pending lane together
            if (L.mode == M_PEND && L.pa == PA_PEEL_END)
                peel_end_stage<Pol, ADJ, DEFER>(P, L, lds);
            if (L.mode == M_PEND && L.pa == PA_INTERACT)
                interact_stage<Pol, ADJ, DEFER>(P, L);
            if (L.mode == M_PEND && L.pa == PA_SCATTER)
                scatter_stage<Pol, ADJ, DEFER>(P, L);
            if (L.mode == M_PEND && L.pa == PA_FINISH)
                finish<Pol, ADJ, DEFER>(P, L, C, L.status, lds);
            // refill idle lanes from the wave's slot range (SYNC: only once
            // every lane is idle)
            const uint64_t need = __ballot(L.mode == M_NEED);
            if (need && (!SYNC || !__ballot(L.mode != M_NEED && L.mode != M_DONE))) {
                const uint32_t n = __popcll(need);
                const uint32_t rank = __popcll(need & lt_mask);
                const uint64_t avail = wend - wbase;
                uint64_t my = wbase + rank;
                if (n > avail) {
                    // claim enough for every idle lane, at least one chunk
                    const uint64_t take = (n - avail) > P.chunk ? (n - avail) : P.chunk;
                    uint64_t nb = 0;
                    if (lane == __ffsll((long long)need) - 1)
                        nb = atomicAdd(work, (unsigned long long)take);
                    nb = __shfl(nb, __ffsll((long long)need) - 1, 64);
                    if (rank >= avail)
                        my = nb + (rank - avail);
                    wbase = nb + (n - avail);
                    wend = nb + take;
                } else {
                    wbase += n;
                }
                if (L.mode == M_NEED) {
                    if (my < P.total && (!P.total_dev || my < *P.total_dev)) {
                        L.slot = slot_at(P, my);
                        emit_packet<Pol, ADJ, DEFER>(P, L);
                    } else {
                        L.mode = M_DONE;
                    }
                }
            }
        }
        if (L.mode != M_MAIN && L.mode != M_PEEL)
            continue;

        // ---------------- one march step of the current ray
        const bool is_main = (L.mode == M_MAIN);
        if (is_main && L.steps <= 0) {
            L.mode = M_PEND;
            L.pa = PA_FINISH;
            L.status = KP_END_STEPS;
            continue;
        }
        real l;
        const int ap = Pol::dl(L.cur, l);
        if (ap < 0) {
            L.mode = M_PEND;
            L.pa = is_main ? PA_FINISH : PA_PEEL_END;
            L.status = KP_END_LOST;
            continue;
        }
        // the cell and its kappa_rho were fetched by the previous crossing
        // unless the ray was (re)started since
        if (L.kgc < 0) {
            L.kgc = (int64_t)L.cur.ib * P.n3 + Pol::cell(P, L.cur);
            L.kcache = P.kappa[L.kgc];
        }
        const int64_t gc = L.kgc;
        const float kraw = L.kcache;
        const real kap = (real)kraw * L.ks;
        const real dtau = kap * l;
        if (is_main) {
            L.ncross += 1;
            L.steps -= 1;
            // the segment ends at the exit face, or at the interaction point
            // when the remaining optical depth runs out inside the cell
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
            if constexpr (ADJ && DEFER) {
                // deposited by the vertex kernel
            } else if constexpr (ADJ) {
                if (L.pass == 2 && P.g_logrho && L.B != 0.0)
                    atomic_add_f64(g_base(P, P.g_logrho) + gc, -((double)(kap * seg) * L.B));
            } else if constexpr (sizeof(real) == 8) {
                if (P.J)
                    j_deposit(P, lds, gc, L.band, L.I * (seg / P.vol0[L.cur.ib]));
            } else if constexpr (SYNC) {
                if (P.J) {
                    const double v = (double)(L.I * seg * L.cur.iv);
                    if (lds && gc == P.src_gc) {
                        atomicAdd(&kp_lds[L.band], v);
                    } else {
                        jh = true;
                        jidx = (uint64_t)(uint32_t)gc * (uint32_t)P.j_cs +
                               (uint64_t)(uint32_t)L.band * (uint32_t)P.j_bs;
                        jv = v;
                    }
                }
            } else {
                L.acc.add(P, L.band, gc, L.I * seg * L.cur.iv, lds);
            }
            if (hit) {
                Pol::advance(L.cur, seg);
                L.gc = gc;
                L.mode = M_PEND;
                L.pa = PA_INTERACT;
                continue;
            }
            L.tau = L.tau - (double)dtau;
        } else {
            L.tau = L.tau + (double)dtau;
            L.npeel += 1;
            if constexpr (ADJ && !DEFER) {
                if (L.pass == 2 && L.peel_sub == 1 && P.g_logrho && L.Acur != 0.0)
                    atomic_add_f64(g_base(P, P.g_logrho) + gc, -((double)dtau * L.Acur));
            }
            if (L.tau > kTauCut) {
                L.mode = M_PEND;
                L.pa = PA_PEEL_END;
                continue;
            }
        }
        int st;
        if constexpr (kDB) {
            st = Pol::cross_step(P, L.cur, ap, l);
            if (st != f32::ST_IN) {
                L.blk = ap + 4 * (st == f32::ST_BLOCK_HI) + 8 * L.mode;
                L.mode = M_PEND;
                L.pa = PA_BLOCK;
                continue;
            }
        } else {
            st = Pol::cross(P, L.cur, ap, l);
        }
        if (st == f64::ST_IN) {
            // issue the next cell's kappa_rho load now; it is consumed next step
            L.kgc = (int64_t)L.cur.ib * P.n3 + Pol::cell(P, L.cur);
            L.kcache = P.kappa[L.kgc];
        } else {
            L.mode = M_PEND;
            if (is_main) {
                L.pa = PA_FINISH;
                L.status = (st == f64::ST_ESCAPED) ? KP_END_ESCAPED : KP_END_LOST;
            } else {
                L.pa = PA_PEEL_END;
            }
        }
    }
    jsum = L.jsum;
    if constexpr (DEFER)
        close_q_chunks<real>(P);
}

} // namespace kp
