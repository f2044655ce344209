// kp_transport64.h -- FP64 transport, global coordinates.
//
// The plain method: the reference march of photon.h:20-125 / proxy.h:64-121
// in FP64 global coordinates (the PRECISION=2 build of types.h:12-27), with
// Philox-keyed sampling, polarized scattering, absorption weighting and
// peel-off images on top.  Operation order is fixed (no contraction;
// explicit fma only) so that results match a CPU run of the same algorithm
// bit for bit on the same packet ids.
#pragma once

#include "kp_device.h"

namespace kp {
namespace f64 {

struct ray_t {
    double x[3], d[3];
    double xl[3], xr[3];
    int i[3];
    int ib;
};

// ST_BLOCK_LO / ST_BLOCK_HI: the index left [0,n) through the lower / upper
// face of the exit axis (cross_in_block); change_block takes the ray on.
enum { ST_IN = 0, ST_ESCAPED = 1, ST_BLOCK_LO = 2, ST_BLOCK_HI = 3, ST_LOST = 4 };

__device__ __forceinline__ double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// geo_t::x_fc (geometry.h:80-81)
__device__ __forceinline__ double x_fc(const kparams &P, int ib, int a, int i)
{
    return P.xf0[3 * ib + a] + (double)i * P.dx0[3 * ib + a];
}

// geo_loc_t::load (particle_base.h:91-100)
__device__ __forceinline__ void gl_load(const kparams &P, ray_t &r)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.xl[a] = x_fc(P, r.ib, a, r.i[a]);
        r.xr[a] = x_fc(P, r.ib, a, r.i[a] + 1);
    }
}

// cart_t::get_dl (photon.h:20-44): ties go to the highest axis.
__device__ __forceinline__ int get_dl(const ray_t &r, double &lmin)
{
    double dl[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (r.d[a] == 0.0) {
            dl[a] = -1.0;
        } else {
            const double xf = r.d[a] > 0.0 ? r.xr[a] : r.xl[a];
            double v = (xf - r.x[a]) / r.d[a];
            if (v < 0.0)
                v = 0.0;
            dl[a] = v;
        }
    }
    int ap = -1;
    double best = 1.7976931348623157e308;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (dl[a] >= 0.0 && dl[a] <= best) {
            ap = a;
            best = dl[a];
        }
    lmin = best;
    return ap;
}

// proc_geo (photon.h:54-71) + shift_blk (proxy.h:64-121) + regulate
// (particle_base.h:37-56), in three parts that keep the reference's order:
// the exit-axis step (cross_in_block), the block change when the index left
// [0,n) (change_block), then the periodic wrap (regulate).
__device__ __forceinline__ void regulate(const kparams &P, ray_t &r, int a, bool &reload)
{
    // periodic wrap (par_base_t::regulate); a ray that just left through the
    // lower domain face sits exactly on it and belongs to the upper image
#pragma unroll
    for (int q = 0; q < 3; ++q)
        if (P.periodic[q]) {
            const double lbox = P.xmax[q] - P.xmin[q];
            const bool down_exit = (q == a) && !(r.d[q] > 0.0);
            if (r.x[q] < P.xmin[q] || (down_exit && r.x[q] == P.xmin[q])) {
                r.x[q] += lbox;
                reload = true;
            } else if (r.x[q] >= P.xmax[q]) {
                r.x[q] -= lbox;
                reload = true;
            }
        }
}

// Step to the exit face of axis ap after physical length l: ST_IN (wrapped
// periodically), or ST_BLOCK_LO / ST_BLOCK_HI.
__device__ __forceinline__ int cross_in_block(const kparams &P, ray_t &r, int ap, double l,
                                              bool &reload)
{
    reload = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == ap) {
            const bool rhs = r.d[a] > 0.0;
            r.x[a] = rhs ? r.xr[a] : r.xl[a];
            r.i[a] += rhs ? 1 : -1;
            if (rhs) {
                r.xl[a] = r.xr[a];
                r.xr[a] = x_fc(P, r.ib, a, r.i[a] + 1);
            } else {
                r.xr[a] = r.xl[a];
                r.xl[a] = x_fc(P, r.ib, a, r.i[a]);
            }
        } else {
            r.x[a] = r.x[a] + l * r.d[a];
        }
    }
    const int n = P.n_cell;
    // only the exit axis' index moved, so it is the one that can have left
    // [0,n): tested on all three axes, without selecting the axis (a select
    // chain on ap compiles to branches)
    const bool hi = (r.i[0] >= n) | (r.i[1] >= n) | (r.i[2] >= n);
    const bool lo = (r.i[0] < 0) | (r.i[1] < 0) | (r.i[2] < 0);
    if (hi | lo)
        return hi ? ST_BLOCK_HI : ST_BLOCK_LO;
    regulate(P, r, ap, reload);
    return ST_IN;
}

// Block change of a ray whose index left [0,n) through the exit face of axis
// a (side 1: the upper face), then the periodic wrap; reload is set.
template <class NB>
__device__ __forceinline__ int change_block(const kparams &P, ray_t &r, int a, int side,
                                            bool &reload, const NB &T)
{
    const int n = P.n_cell;
    if (T.phy(r.ib) & (1u << (2 * a + side)))
        return ST_ESCAPED;
    const int ap0 = a == 2 ? 0 : a + 1;
    const int dl = T.dl(r.ib, 2 * a + side);
    int j0 = 0, j1 = 0;
    if (dl > 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q == a)
                continue;
            const int jj = r.i[q] >= n / 2;
            if (q == ap0)
                j0 = jj;
            else
                j1 = jj;
        }
    }
    const int slot = ((a * 2 + side) * 2 + j0) * 2 + j1;
    const int nb = T.id(r.ib, slot);
    if (nb < 0)
        return ST_LOST;
    const int shift = T.shift(r.ib, slot);
    r.ib = nb;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q == a) {
            r.i[q] = side ? 0 : n - 1;
            continue;
        }
        const int j = (q == ap0) ? 0 : 1;
        if (dl < 0) {
            r.i[q] = r.i[q] / 2 + ((shift >> j) & 1) * (n / 2);
        } else if (dl > 0) {
            int f = (r.i[q] < n / 2 ? r.i[q] : r.i[q] - n / 2) * 2;
            if (r.x[q] >= x_fc(P, nb, q, f + 1))
                f += 1;
            r.i[q] = f;
        }
    }
    reload = true;
    regulate(P, r, a, reload);
    return ST_IN;
}

__device__ __forceinline__ int change_block(const kparams &P, ray_t &r, int a, int side,
                                            bool &reload)
{
    return change_block(P, r, a, side, reload, nb_global{P});
}

__device__ __forceinline__ int cross_face(const kparams &P, ray_t &r, int ap, double l,
                                          bool &reload)
{
    const int st = cross_in_block(P, r, ap, l, reload);
    return st == ST_IN ? ST_IN : change_block(P, r, ap, st == ST_BLOCK_HI, reload);
}

__device__ __forceinline__ int cell_index(const kparams &P, const ray_t &r)
{
    return r.i[0] + P.n_cell * (r.i[1] + P.n_cell * r.i[2]);
}

// cell_index with the indices clamped into the block (see f32).
__device__ __forceinline__ int cell_index_clamped(const kparams &P, const ray_t &r)
{
    const int m = P.n_cell - 1;
    return min(max(r.i[0], 0), m) +
           P.n_cell * (min(max(r.i[1], 0), m) + P.n_cell * min(max(r.i[2], 0), m));
}

__device__ double tau_to_edge(const kparams &P, const ray_t &from, const double *dir, double ks,
                              uint64_t &ncross)
{
    ray_t r = from;
    r.d[0] = dir[0];
    r.d[1] = dir[1];
    r.d[2] = dir[2];
    double tau = 0.0;
    // kappa_rho of the next cell is loaded one step ahead: the crossing is
    // taken before the current cell's optical depth is added (the ray is a
    // copy, so the crossing after the last cell is simply dropped)
    float kc = P.kappa[(size_t)r.ib * P.n3 + cell_index(P, r)];
    for (;;) {
        double l;
        const int ap = get_dl(r, l);
        if (ap < 0)
            break;
        const double kap = (double)kc * ks;
        bool reload;
        const int st = cross_face(P, r, ap, l, reload);
        if (st == ST_IN) {
            if (reload)
                gl_load(P, r);
            kc = P.kappa[(size_t)r.ib * P.n3 + cell_index(P, r)];
        }
        tau = tau + kap * l;
        ncross += 1;
        if (tau > kTauCut || st != ST_IN)
            break;
    }
    return tau;
}

// tau_to_edge with the block changes of a wave's rays taken together (see
// f32::tau_to_edge_sync); same operations in the same order per ray.
__device__ double tau_to_edge_sync(const kparams &P, const ray_t &from, const double *dir,
                                   double ks, uint64_t &ncross)
{
    ray_t r = from;
    r.d[0] = dir[0];
    r.d[1] = dir[1];
    r.d[2] = dir[2];
    double tau = 0.0;
    float kc = P.kappa[(size_t)r.ib * P.n3 + cell_index(P, r)];
    for (;;) {
        int ap, st;
        for (;;) {
            double l;
            ap = get_dl(r, l);
            if (ap < 0) {
                st = ST_LOST;
                break;
            }
            const double kap = (double)kc * ks;
            bool reload;
            st = cross_in_block(P, r, ap, l, reload);
            // issued whether or not the ray is still in the block (see f32)
            kc = P.kappa[(size_t)r.ib * P.n3 + cell_index_clamped(P, r)];
            if (st == ST_IN && reload)
                gl_load(P, r);
            tau = tau + kap * l;
            ncross += 1;
            if (tau > kTauCut) {
                st = ST_LOST;
                break;
            }
            if (st != ST_IN)
                break;
        }
        if (st != ST_BLOCK_LO && st != ST_BLOCK_HI)
            break;
        bool reload = false;
        if (change_block(P, r, ap, st == ST_BLOCK_HI, reload) != ST_IN)
            break;
        if (reload)
            gl_load(P, r);
        kc = P.kappa[(size_t)r.ib * P.n3 + cell_index(P, r)];
    }
    return tau;
}

// Image pixel (linear index into [band][incl][N][N]) of point x, or -1.
__device__ __forceinline__ long pixel(const kparams &P, int band, int o, const double *x)
{
    const int N = P.image_n;
    double rel[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        rel[a] = x[a] - P.center[a];
    const double uh = dot3(rel, P.eh[o]);
    const double uv = dot3(rel, P.ev[o]);
    const double fx = floor((uh + P.image_half) * P.pix_scale);
    const double fy = floor((uv + P.image_half) * P.pix_scale);
    // written so that a NaN coordinate is outside too
    if (!(fx >= 0.0 && fy >= 0.0 && fx < (double)N && fy < (double)N))
        return -1;
    return (long)((((size_t)band * P.n_incl + o) * N + (size_t)fy) * N + (size_t)fx);
}

__device__ __forceinline__ void deposit(const kparams &P, int band, int o, const double *x,
                                        double I, double Q, double U)
{
    if (!P.image)
        return;
    const long pix = pixel(P, band, o, x);
    if (pix < 0)
        return;
    atomic_add_f64(img_at(P, pix, 0), I);
    atomic_add_f64(img_at(P, pix, 1), Q);
    atomic_add_f64(img_at(P, pix, 2), U);
}

__device__ __forceinline__ void phase(int model, double gg, double pmax, double mu, double &z11,
                                      double &z12, double &z33)
{
    const double m2 = mu * mu;
    if (model == KP_THOMSON) {
        z11 = 0.75 * (1.0 + m2);
        z12 = 0.75 * (m2 - 1.0);
        z33 = 1.5 * mu;
    } else {
        const double D = (1.0 + gg * gg) - 2.0 * gg * mu;
        const double ph = (1.0 - gg * gg) / (D * sqrt(D));
        const double pol = pmax * (1.0 - m2) / (1.0 + m2);
        z11 = ph;
        z12 = -pol * ph;
        z33 = ph * (2.0 * mu / (1.0 + m2));
    }
}

// Peel-off Stokes vector toward no in the image frame (Q>0 along ev), without
// the 1/(4 pi) and attenuation factors.  Same operation order as the oracle.
__device__ __forceinline__ void peel_stokes(int model, double gg, double pmax, const double *d,
                                            const double *e1, const double *e2, double I,
                                            double Q, double U, const double *no,
                                            const double *ev, double &Io, double &Qo,
                                            double &Uo)
{
    double mo = dot3(d, no);
    if (mo > 1.0)
        mo = 1.0;
    if (mo < -1.0)
        mo = -1.0;
    double p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        p[a] = no[a] - mo * d[a];
    const double pn2 = dot3(p, p);
    double cf = 1.0, sf = 0.0, st = 0.0;
    if (pn2 > 1e-24) {
        const double ip = 1.0 / sqrt(pn2);
        cf = dot3(p, e1) * ip;
        sf = dot3(p, e2) * ip;
        st = pn2 * ip;
    }
    const double c2 = cf * cf - sf * sf;
    const double s2 = 2.0 * cf * sf;
    const double Qr = Q * c2 + U * s2;
    const double Ur = U * c2 - Q * s2;
    double z11, z12, z33;
    phase(model, gg, pmax, mo, z11, z12, z33);
    const double Ip = z11 * I + z12 * Qr;
    const double Qp = z12 * I + z11 * Qr;
    const double Up = z33 * Ur;
    double e1o[3], e2o[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double e1r = cf * e1[a] + sf * e2[a];
        e1o[a] = mo * e1r - st * d[a];
    }
    cross3(no, e1o, e2o);
    const double cps = dot3(ev, e1o);
    const double sps = dot3(ev, e2o);
    const double nrm = cps * cps + sps * sps;
    const double c2p = (cps * cps - sps * sps) / nrm;
    const double s2p = 2.0 * cps * sps / nrm;
    Io = Ip;
    Qo = Qp * c2p + Up * s2p;
    Uo = Up * c2p - Qp * s2p;
}

// One scattering: new direction d and frame e1 in place, Stokes weight
// Z S / Z11.  Same operation order as the oracle.
__device__ __forceinline__ void scatter_stokes(int model, double pmax, double mu, double phi_u,
                                               double *d, double *e1, const double *e2,
                                               double &I, double &Q, double &U)
{
    double s2m = fma(-mu, mu, 1.0);
    if (s2m < 0.0)
        s2m = 0.0;
    const double st = sqrt(s2m);
    double sf, cf;
    kp_sincos2pi(phi_u, &sf, &cf);
    const double c2 = cf * cf - sf * sf;
    const double s2 = 2.0 * cf * sf;
    const double Qr = Q * c2 + U * s2;
    const double Ur = U * c2 - Q * s2;
    const double m2 = mu * mu;
    const double pl = (model == KP_THOMSON) ? 1.0 : pmax;
    const double a12 = -pl * (1.0 - m2) / (1.0 + m2);
    const double a33 = 2.0 * mu / (1.0 + m2);
    const double In = I + a12 * Qr;
    const double Qn = a12 * I + Qr;
    const double Un = a33 * Ur;
    I = In;
    Q = Qn;
    U = Un;
    double dn[3], en[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double e1r = cf * e1[a] + sf * e2[a];
        dn[a] = mu * d[a] + st * e1r;
        en[a] = mu * e1r - st * d[a];
    }
    const double id = 1.0 / sqrt(dot3(dn, dn));
#pragma unroll
    for (int a = 0; a < 3; ++a)
        dn[a] = dn[a] * id;
    const double pr = dot3(en, dn);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        en[a] = en[a] - pr * dn[a];
    const double ie = 1.0 / sqrt(dot3(en, en));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d[a] = dn[a];
        e1[a] = en[a] * ie;
    }
}

// Emission (event 0): isotropic direction, theta-hat reference axis, first
// free flight.
__device__ __forceinline__ void emit(const uint32_t rn[4], double *d, double *e1, double &tau)
{
    const double mu = fma(2.0, kp_u01(rn[0]), -1.0);
    double sf, cf;
    kp_sincos2pi(kp_u01(rn[1]), &sf, &cf);
    double s2 = fma(-mu, mu, 1.0);
    if (s2 < 0.0)
        s2 = 0.0;
    const double st = sqrt(s2);
    d[0] = st * cf;
    d[1] = st * sf;
    d[2] = mu;
    e1[0] = mu * cf;
    e1[1] = mu * sf;
    e1[2] = -st;
    tau = -kp_log(kp_u01(rn[2]));
}

struct counters_t {
    uint64_t cross = 0, peel = 0, events = 0, escaped = 0, rr = 0, steps = 0, lost = 0,
             packets = 0;
};

// One packet history: emission, march, interactions, peel-offs.
__device__ void packet(const kparams &P, int band, uint64_t pid, uint64_t slot, counters_t &C)
{
    const dband bd = P.bands[band];
    const double ks = bd.kappa_scale;
    const double w0 = P.w0;
    const uint32_t k0 = P.seed, k1 = (uint32_t)band;
    const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32);
    uint64_t ncross = 0, npeel = 0;
    int events = 0, status = 0;

    uint32_t rn[4] = {plo, phi, 0u, 0u};
    kp_philox4x32_10(rn, k0, k1);
    ray_t r;
    double e1[3];
    double tau_rem;
    emit(rn, r.d, e1, tau_rem);
    double I = w0, Q = 0.0, U = 0.0;
    r.x[0] = P.src_x[0];
    r.x[1] = P.src_x[1];
    r.x[2] = P.src_x[2];
    r.ib = P.src_ib;
    r.i[0] = P.src_i[0];
    r.i[1] = P.src_i[1];
    r.i[2] = P.src_i[2];
    gl_load(P, r);
    for (int o = 0; o < P.n_incl; ++o) {
        const double tau = tau_to_edge(P, r, P.no[o], ks, npeel);
        const double att = kp_exp(-tau) * KP_INV_4PI;
        deposit(P, band, o, r.x, I * att, 0.0, 0.0);
    }

    int steps = P.step_max;
    int cell = 0;
    for (;;) {
        bool interact = false;
        for (;;) {
            if (steps <= 0) {
                status = KP_END_STEPS;
                break;
            }
            double l;
            const int ap = get_dl(r, l);
            if (ap < 0) {
                status = KP_END_LOST;
                break;
            }
            cell = cell_index(P, r);
            const size_t gc = (size_t)r.ib * P.n3 + cell;
            const double kap = (double)P.kappa[gc] * ks;
            const double dtau = kap * l;
            const double V = P.vol0[r.ib];
            ncross += 1;
            steps -= 1;
            if (kap > 0.0 && dtau >= tau_rem) {
                double li = tau_rem / kap;
                if (li > l)
                    li = l;
                if (P.J)
                    atomic_add_f64(j_at(P, gc, band), I * (li / V));
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    r.x[a] = r.x[a] + li * r.d[a];
                interact = true;
                break;
            }
            tau_rem = tau_rem - dtau;
            if (P.J)
                atomic_add_f64(j_at(P, gc, band), I * (l / V));
            bool reload;
            const int st = cross_face(P, r, ap, l, reload);
            if (st == ST_ESCAPED) {
                status = KP_END_ESCAPED;
                break;
            }
            if (st == ST_LOST) {
                status = KP_END_LOST;
                break;
            }
            if (reload)
                gl_load(P, r);
        }
        if (!interact)
            break;

        events += 1;
        uint32_t u[4] = {plo, phi, (uint32_t)events, 0u};
        kp_philox4x32_10(u, k0, k1);
        const size_t gc = (size_t)r.ib * P.n3 + cell;
        const double alb = (bd.flags & KP_BAND_CELL_ALBEDO) ? (double)P.albedo[gc] : bd.albedo;
        const double gg =
            (bd.flags & KP_BAND_CELL_G) ? (P.g ? (double)P.g[gc] : 0.0) : bd.g;
        I = I * alb;
        Q = Q * alb;
        U = U * alb;
        double e2[3];
        cross3(r.d, e1, e2);

        for (int o = 0; o < P.n_incl; ++o) {
            double Ip, Qi, Ui;
            peel_stokes(P.model, gg, P.pmax, r.d, e1, e2, I, Q, U, P.no[o], P.ev[o], Ip, Qi, Ui);
            const double tau = tau_to_edge(P, r, P.no[o], ks, npeel);
            const double att = kp_exp(-tau) * KP_INV_4PI;
            deposit(P, band, o, r.x, Ip * att, Qi * att, Ui * att);
        }

        {
            const double mu = (P.model == KP_THOMSON) ? kp_thomson_mu(kp_u01(u[0]))
                                                      : kp_hg_mu(gg, kp_u01(u[0]));
            scatter_stokes(P.model, P.pmax, mu, kp_u01(u[1]), r.d, e1, e2, I, Q, U);
        }
        tau_rem = -kp_log(kp_u01(u[2]));
        if (I < P.rr_w * w0) {
            if (kp_u01(u[3]) < P.rr_s) {
                I = I / P.rr_s;
                Q = Q / P.rr_s;
                U = U / P.rr_s;
            } else {
                status = KP_END_RR;
                break;
            }
        }
    }

    C.cross += ncross;
    C.peel += npeel;
    C.events += (uint64_t)events;
    C.packets += 1;
    if (status == KP_END_ESCAPED)
        C.escaped += 1;
    else if (status == KP_END_RR)
        C.rr += 1;
    else if (status == KP_END_STEPS)
        C.steps += 1;
    else
        C.lost += 1;
    if (P.end_state) {
        double *es = P.end_state + slot * KP_ES_STRIDE;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            es[KP_ES_X + a] = r.x[a];
            es[KP_ES_D + a] = r.d[a];
        }
        es[KP_ES_I] = I;
        es[KP_ES_Q] = Q;
        es[KP_ES_U] = U;
        es[KP_ES_EVENTS] = (double)events;
        es[KP_ES_CROSS] = (double)ncross;
        es[KP_ES_STATUS] = (double)status;
        es[KP_ES_IB] = (double)r.ib;
        es[KP_ES_CELL] = (double)cell_index(P, r);
        es[14] = (double)npeel;
        es[15] = 0.0;
    }
}

} // namespace f64
} // namespace kp
