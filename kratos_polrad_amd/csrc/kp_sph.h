// kp_sph.h -- FP64 streaming march on spherical (r, theta, phi) block meshes.
//
// The reference's spherical photon (radiation/photon_sph.h:10-266) under the
// step loop of photon.h:104-125, with shift_blk (types/proxy.h:64-121), the
// path-length tally of photon.h:73-83 and the cell volumes of
// meshgen/spherical.h:31-42.  Positions stay Cartesian.  Each block's face
// trigonometry (r_f, cos th_f, sin ph_f, cos ph_f) and volume factors are a
// table made once on the host, and the position tests against theta and phi
// faces are algebraic (z against cos(th_f)|x|, the sign of
// y cos(ph_f) - x sin(ph_f)), so a step needs no acos/atan2/sin/cos: only the
// square roots and divisions of the sphere, cone and plane intersections.
// The departures from the reference are those of the CPU restatement: exact
// polar angle in the theta tests, every axis that left the block resolved in
// the same step, the fine tangential cell chosen by position, a packet with
// no positive face distance ended as lost, and the equatorial face taken as
// the plane z = 0 with a double-root intersection.  Operation order is fixed (no
// contraction), so histories equal the CPU restatement's bit for bit.
#pragma once

#include "../../include/kpolrad.h"
#include "kp_device.h"

namespace kp {
namespace sph {

enum { ST_IN = 0, ST_ESCAPED = 1, ST_LOST = 4 };

struct ray_t {
    double x[3], d[3];
    int ib, i0, i1, i2;
};

// the current cell's faces: r, cos(theta), sin/cos(phi) at the lower and
// upper face of each axis, and its volume
struct faces_t {
    double rl, rr, cl, cr, sl, sr, ql, qr, vol;
};

__device__ __forceinline__ const double *tab(const kparams &P, int ib)
{
    const int n = P.n_cell;
    return P.sph + (size_t)ib * (size_t)(4 * (n + 1) + 3 * n);
}

__device__ __forceinline__ void load_faces(const kparams &P, const ray_t &r, faces_t &f)
{
    const int m = P.n_cell + 1;
    const double *T = tab(P, r.ib);
    f.rl = T[r.i0];
    f.rr = T[r.i0 + 1];
    f.cl = T[m + r.i1];
    f.cr = T[m + r.i1 + 1];
    f.sl = T[2 * m + r.i2];
    f.sr = T[2 * m + r.i2 + 1];
    f.ql = T[3 * m + r.i2];
    f.qr = T[3 * m + r.i2 + 1];
    const int n = P.n_cell;
    f.vol = (T[4 * m + r.i0] * T[4 * m + n + r.i1]) * T[4 * m + 2 * n + r.i2];
}

// sph_t::volume (spherical.h:31-42) from the per-axis factors
__device__ __forceinline__ double volume(const kparams &P, const ray_t &r)
{
    const int n = P.n_cell, m = n + 1;
    const double *T = tab(P, r.ib);
    return (T[4 * m + r.i0] * T[4 * m + n + r.i1]) * T[4 * m + 2 * n + r.i2];
}

__device__ __forceinline__ int cell_index(const kparams &P, const ray_t &r)
{
    return r.i0 + P.n_cell * (r.i1 + P.n_cell * r.i2);
}

__device__ __forceinline__ int sgn(double v) { return v > 0.0 ? 1 : (v < 0.0 ? -1 : 0); }

// sph_t::root (photon_sph.h:24-40): both roots of a t^2 + b t + c, ordered;
// `plane`: the equatorial face z = 0, a double root (discriminant exactly 0)
__device__ __forceinline__ bool root(double a, double b, double c, bool plane, double &x1,
                                     double &x2)
{
    const double det = plane ? 0.0 : b * b - 4.0 * a * c;
    if (det < 0.0)
        return false;
    const double nom = copysign(sqrt(det), b) + b;
    const double x0 = nom / (-2.0 * a);
    const double y0 = (-2.0 * c) / nom;
    x1 = x0 < y0 ? x0 : y0;
    x2 = x0 > y0 ? x0 : y0;
    return true;
}

// sph_t::get_dl (photon_sph.h:42-180): exit axis (-1 if no face distance is
// positive), its distance, and the index direction per axis after the
// backward-face fallbacks.
__device__ __forceinline__ int get_dl(const ray_t &r, const faces_t &f, double &lmin, int &s0,
                                      int &s1, int &s2)
{
    const double x0 = r.x[0], x1 = r.x[1], x2 = r.x[2];
    const double d0 = r.d[0], d1 = r.d[1], d2 = r.d[2];
    const double xd = x0 * d0 + x1 * d1 + x2 * d2;
    const double r2 = x0 * x0 + x1 * x1 + x2 * x2;
    const double rho2 = x0 * x0 + x1 * x1;
    const double dt = x2 * (x0 * d0 + x1 * d1) - rho2 * d2;
    const double df = x0 * d1 - x1 * d0;
    s0 = sgn(xd);
    s1 = sgn(dt);
    s2 = sgn(df);
    double l0, l1, l2, t1, t2;
    {
        // r: the sphere ahead, else (a ray missing the inner sphere) the
        // far side of the outer one
        const double b = 2.0 * xd;
        const double fw = s0 > 0 ? f.rr : f.rl;
        if (root(1.0, b, r2 - fw * fw, false, t1, t2)) {
            l0 = t1 > 0.0 ? t1 : t2;
        } else {
            const double bk = s0 < 0 ? f.rr : f.rl;
            l0 = root(1.0, b, r2 - bk * bk, false, t1, t2) ? t2 : -1.0;
            s0 = -s0;
        }
    }
    {
        // theta: the cone z^2 = cos^2(th_f) |x|^2
        const double cf = s1 > 0 ? f.cr : f.cl;
        double mu2 = cf * cf;
        double a = d2 * d2 - mu2;
        double b = 2.0 * (x2 * d2 - mu2 * xd);
        double c = x2 * x2 - mu2 * r2;
        if (a != 0.0) {
            if (root(a, b, c, mu2 == 0.0, t1, t2)) {
                l1 = t1 > 0.0 ? t1 : t2;
            } else {
                const double cb = s1 < 0 ? f.cr : f.cl;
                mu2 = cb * cb;
                a = d2 * d2 - mu2;
                b = 2.0 * (x2 * d2 - mu2 * xd);
                c = x2 * x2 - mu2 * r2;
                if (a != 0.0)
                    l1 = root(a, b, c, mu2 == 0.0, t1, t2) ? t2 : -1.0;
                else
                    l1 = -c / b;
                s1 = -s1;
            }
        } else {
            l1 = -c / b;
        }
    }
    {
        // phi: the plane through the polar axis
        const double sf = s2 > 0 ? f.sr : f.sl, qf = s2 > 0 ? f.qr : f.ql;
        double a = d0 * sf - d1 * qf;
        double b = x1 * qf - x0 * sf;
        l2 = a != 0.0 ? b / a : -1.0;
        if (l2 < 0.0) {
            const double sb = s2 < 0 ? f.sr : f.sl, qb = s2 < 0 ? f.qr : f.ql;
            a = d0 * sb - d1 * qb;
            b = x1 * qb - x0 * sb;
            l2 = a != 0.0 ? b / a : -1.0;
            s2 = -s2;
        }
    }
    int ap = -1;
    double best = 1.7976931348623157e308;
    if (l0 > 0.0 && l0 < best) {
        ap = 0;
        best = l0;
    }
    if (l1 > 0.0 && l1 < best) {
        ap = 1;
        best = l1;
    }
    if (l2 > 0.0 && l2 < best) {
        ap = 2;
        best = l2;
    }
    lmin = best;
    return ap;
}

// position at or beyond face k of axis Q of block table T
template <int Q>
__device__ __forceinline__ bool beyond(const double *T, int n, int k, const double *x)
{
    const int m = n + 1;
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    if constexpr (Q == 0)
        return r2 >= T[k] * T[k];
    else if constexpr (Q == 1)
        return x[2] <= T[m + k] * sqrt(r2);
    else
        return x[1] * T[3 * m + k] - x[0] * T[2 * m + k] >= 0.0;
}

template <int Q> __device__ __forceinline__ int &idx(ray_t &r)
{
    if constexpr (Q == 0)
        return r.i0;
    else if constexpr (Q == 1)
        return r.i1;
    else
        return r.i2;
}

// tangential index of axis Q after a level jump into block nb
template <int Q>
__device__ __forceinline__ void remap(const kparams &P, ray_t &r, int dlv, int sh, int nb)
{
    const int n = P.n_cell;
    int &iq = idx<Q>(r);
    const int is = iq;
    if (dlv < 0) {
        const int h = is >= 0 ? is / 2 : -((1 - is) / 2);
        iq = h + sh * (n / 2);
    } else if (dlv > 0) {
        int fi = (is < n / 2 ? is : is - n / 2) * 2;
        if (is >= 0 && is < n && beyond<Q>(tab(P, nb), n, fi + 1, r.x))
            fi += 1;
        iq = fi;
    }
}

// shift_blk (proxy.h:64-121) for axis A
template <int A> __device__ __forceinline__ int shift_axis(const kparams &P, ray_t &r)
{
    constexpr int Q0 = (A + 1) % 3, Q1 = (A + 2) % 3;
    const int n = P.n_cell;
    int &ia = idx<A>(r);
    if (ia >= 0 && ia < n)
        return ST_IN;
    const int side = ia > 0;
    if (P.phybnd[r.ib] & (1u << (2 * A + side)))
        return ST_ESCAPED;
    const int dlv = P.d_lvl[6 * r.ib + 2 * A + side];
    int j0 = 0, j1 = 0;
    if (dlv > 0) {
        j0 = idx<Q0>(r) >= n / 2;
        j1 = idx<Q1>(r) >= n / 2;
    }
    const int slot = 24 * r.ib + ((A * 2 + side) * 2 + j0) * 2 + j1;
    const int nb = P.nb_id[slot];
    if (nb < 0)
        return ST_LOST;
    const int shift = P.nb_shift[slot];
    r.ib = nb;
    ia = side ? 0 : n - 1;
    remap<Q0>(P, r, dlv, shift & 1, nb);
    remap<Q1>(P, r, dlv, (shift >> 1) & 1, nb);
    return ST_IN;
}

// sph_t::proc_geo (photon_sph.h:222-243) + shift_blk of every axis
__device__ __forceinline__ int cross(const kparams &P, ray_t &r, const faces_t &f, int ap,
                                     double l, int s0, int s1, int s2)
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
        r.x[a] = r.x[a] + (l * r.d[a]) * 1.001;
    if (ap == 0)
        r.i0 += s0;
    else if (ap == 1)
        r.i1 += s1;
    else
        r.i2 += s2;
    const double x0 = r.x[0], x1 = r.x[1], x2 = r.x[2];
    const double r2 = x0 * x0 + x1 * x1 + x2 * x2;
    if (ap != 0) {
        if (r2 >= f.rr * f.rr)
            r.i0 += 1;
        else if (r2 <= f.rl * f.rl)
            r.i0 -= 1;
    }
    if (ap != 1) {
        const double rn = sqrt(r2);
        if (x2 <= f.cr * rn)
            r.i1 += 1;
        else if (x2 >= f.cl * rn)
            r.i1 -= 1;
    }
    if (ap != 2) {
        if (x1 * f.qr - x0 * f.sr >= 0.0)
            r.i2 += 1;
        else if (x1 * f.ql - x0 * f.sl <= 0.0)
            r.i2 -= 1;
    }
    int st = shift_axis<0>(P, r);
    if (st == ST_IN)
        st = shift_axis<1>(P, r);
    if (st == ST_IN)
        st = shift_axis<2>(P, r);
    return st;
}

// Emission: isotropic direction from the packet's event-0 Philox words (the
// same draw as the Cartesian path's emission).
__device__ __forceinline__ void start(const kparams &P, ray_t &r, int band, uint64_t pid)
{
    uint32_t rn[4] = {(uint32_t)pid, (uint32_t)(pid >> 32), 0u, 0u};
    kp_philox4x32_10(rn, P.seed, (uint32_t)band);
    const double mu = fma(2.0, kp_u01(rn[0]), -1.0);
    double sf, cf;
    kp_sincos2pi(kp_u01(rn[1]), &sf, &cf);
    double s2 = fma(-mu, mu, 1.0);
    if (s2 < 0.0)
        s2 = 0.0;
    const double st = sqrt(s2);
    r.d[0] = st * cf;
    r.d[1] = st * sf;
    r.d[2] = mu;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        r.x[a] = P.src_x[a];
    r.ib = P.src_ib;
    r.i0 = P.src_i[0];
    r.i1 = P.src_i[1];
    r.i2 = P.src_i[2];
}

// One march step.  Returns 0 while the packet goes on, else its KP_END_*
// status; the step's tally (cell gc, value v) is handed back for deposit()
// (gc = -1: none).
__device__ __forceinline__ int step(const kparams &P, ray_t &r, int &steps, uint64_t &ncross,
                                    int64_t &gc, double &v)
{
    gc = -1;
    v = 0.0;
    if (steps <= 0)
        return KP_END_STEPS;
    faces_t f;
    load_faces(P, r, f);
    double l;
    int s0, s1, s2;
    const int ap = get_dl(r, f, l, s0, s1, s2);
    if (ap < 0)
        return KP_END_LOST;
    gc = (int64_t)r.ib * P.n3 + cell_index(P, r);
    v = P.w0 * (l / f.vol);
    ncross += 1;
    steps -= 1;
    const int st = cross(P, r, f, ap, l, s0, s1, s2);
    if (st == ST_ESCAPED)
        return KP_END_ESCAPED;
    if (st == ST_LOST)
        return KP_END_LOST;
    return 0;
}

// J deposit of every lane of the wave (called by all 64 lanes together;
// `has` marks lanes with a tally).  The source cell goes to the workgroup's
// LDS accumulator.  With `runs`, lanes whose J element equals their lower
// neighbour's are summed first by a segmented scan across the wave and the
// last lane of each run issues one atomic: in emission-direction order the
// lanes of a wave fly neighbouring rays, and FP64 atomics on one address
// serialise in the L2.
__device__ __forceinline__ void deposit(const kparams &P, bool lds, bool runs, bool has,
                                        int64_t gc, int band, double v)
{
    if (!P.J)
        return;
    if (has && lds && gc == P.src_gc) {
        atomicAdd(&kp_lds[band], v);
        has = false;
    }
    if (!runs) {
        if (has)
            atomic_add_f64(j_at(P, (uint64_t)gc, band), v);
        return;
    }
    // the J element (j_at without the base) as the run key
    const uint64_t e = (uint64_t)(uint32_t)gc * (uint32_t)P.j_cs +
                       (uint64_t)(uint32_t)band * (uint32_t)P.j_bs;
    wave_run_add(j_base(P), has, e, v);
}

template <class C>
__device__ __forceinline__ void finish(const kparams &P, const ray_t &r, uint64_t slot,
                                       uint64_t ncross, int status, C &c)
{
    c.cross += ncross;
    c.packets += 1;
    if (status == KP_END_ESCAPED)
        c.escaped += 1;
    else if (status == KP_END_STEPS)
        c.steps += 1;
    else
        c.lost += 1;
    if (P.end_state) {
        double *es = P.end_state + slot * KP_ES_STRIDE;
#pragma unroll
        for (int k = 0; k < KP_ES_STRIDE; ++k)
            es[k] = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            es[KP_ES_X + a] = r.x[a];
            es[KP_ES_D + a] = r.d[a];
        }
        es[KP_ES_I] = P.w0;
        es[KP_ES_CROSS] = (double)ncross;
        es[KP_ES_STATUS] = (double)status;
        es[KP_ES_IB] = (double)r.ib;
        es[KP_ES_CELL] = (double)cell_index(P, r);
    }
}

// Whole histories, one thread per packet (the integrator's launch shape);
// the wave steps its lanes together so deposit() sees all 64.
template <class C> __device__ void history(const kparams &P, bool lds, C &c)
{
    const int lane = threadIdx.x & 63;
    const bool runs = P.order != nullptr;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
         base < P.total; base += stride) {
        const uint64_t k = base + lane;
        ray_t r;
        int band = 0, steps = P.step_max, status = -1;
        uint64_t slot = 0, ncross = 0;
        if (k < P.total) {
            slot = slot_at(P, k);
            band = band_of(P, slot);
            start(P, r, band, P.packet_offset + (slot - P.band_start[band]));
            status = 0;
        }
        while (__ballot(status == 0)) {
            int64_t gc = -1;
            double v = 0.0;
            if (status == 0)
                status = step(P, r, steps, ncross, gc, v);
            deposit(P, lds, runs, gc >= 0, gc, band, v);
            if (status > 0) {
                finish(P, r, slot, ncross, status, c);
                status = -1;
            }
        }
    }
}

// Persistent lanes: one step per iteration, finished lanes refilled from a
// wave-wide slot range claimed with one atomic per `chunk` packets, once
// `batch` lanes wait (or half the live lanes, or none is marching).
template <class C> __device__ void persist(const kparams &P, bool lds, C &c)
{
    enum { M_NEED = 0, M_RUN = 1, M_DONE = 2 };
    ray_t r;
    int mode = M_NEED, band = 0, steps = 0;
    uint64_t slot = 0, ncross = 0;
    const int lane = threadIdx.x & 63;
    const bool runs = P.order != nullptr;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t wbase = 0, wend = 0; // wave-uniform slot range
    for (;;) {
        const uint64_t live = __ballot(mode != M_DONE);
        if (!live)
            break;
        const uint64_t need = __ballot(mode == M_NEED);
        const uint64_t run = __ballot(mode == M_RUN);
        const uint32_t nw = __popcll(need);
        if (need && (nw >= P.batch || 2 * nw >= (uint32_t)__popcll(live) || !run)) {
            const uint32_t rank = __popcll(need & lt_mask);
            const uint64_t avail = wend - wbase;
            uint64_t my = wbase + rank;
            if (nw > avail) {
                const uint64_t take = (nw - avail) > P.chunk ? (nw - avail) : P.chunk;
                const int leader = __ffsll((long long)need) - 1;
                uint64_t nb = 0;
                if (lane == leader)
                    nb = atomicAdd(P.work, (unsigned long long)take);
                nb = __shfl(nb, leader, 64);
                if (rank >= avail)
                    my = nb + (rank - avail);
                wbase = nb + (nw - avail);
                wend = nb + take;
            } else {
                wbase += nw;
            }
            if (mode == M_NEED) {
                if (my < P.total) {
                    slot = slot_at(P, my);
                    band = band_of(P, slot);
                    start(P, r, band, P.packet_offset + (slot - P.band_start[band]));
                    steps = P.step_max;
                    ncross = 0;
                    mode = M_RUN;
                } else {
                    mode = M_DONE;
                }
            }
        }
        int status = 0;
        int64_t gc = -1;
        double v = 0.0;
        if (mode == M_RUN)
            status = step(P, r, steps, ncross, gc, v);
        deposit(P, lds, runs, gc >= 0, gc, band, v);
        if (mode == M_RUN && status != 0) {
            finish(P, r, slot, ncross, status, c);
            mode = M_NEED;
        }
    }
}

} // namespace sph
} // namespace kp
