// kp_adjoint.h -- gradients of a linear image functional by path replay.
//
// Jlin = sum_pix w_pix . S_pix with w = dL/dimage.  For each packet the
// history is replayed twice from its Philox key (no history is stored):
//  pass 1 computes the adjoint-weighted peel-off contribution A_k of every
//         peel-off (emission k=0, interactions k>=1);
//  pass 2 replays again carrying B = sum of the A_k still ahead, adding the
//         score-function terms of each path segment (-kappa l B to
//         dJ/dlog rho), vertex (+B to dJ/dlog rho, +B/a to dJ/da),
//         scattering (B dlogP_HG(mu)/dg), peel-off phase (A_k dlogP_HG/dg)
//         and peel-off ray (-kappa l A_k).
// Path replay backpropagation: Vicini, Speierer & Jakob, ACM TOG 40(4), 2021.
// The A_k of a packet are kept in a per-thread global scratch row; packets
// with more peel-offs than the row holds recompute the overflow A_k by
// marching the peel-off ray once more (identical FP64 arithmetic).
// FP64 transport (the same per-packet arithmetic as the forward kernel) and
// FP64 gradient accumulation.
#pragma once

#include "kp_transport64.h"

namespace kp {
namespace adj {

struct aparams {
    const double *w;   // dL/dimage, image layout
    double *d_logrho;  // per cell (may be null)
    double *d_alb;
    double *d_g;
    double *jlin;      // device scalar, += sum of A_k
    double *scratch;   // [cap][n_threads]
    uint32_t cap;      // A_k slots per thread
    uint64_t n_threads;
};

__device__ __forceinline__ double dlogp_hg_h(double gg, double mu)
{
    const double D = (1.0 + gg * gg) - 2.0 * gg * mu;
    return -2.0 * gg / (1.0 - gg * gg) - 3.0 * (gg - mu) / D;
}

__device__ void peel_grad(const kparams &P, const f64::ray_t &from, const double *dir, double ks,
                          double A, double *d_logrho)
{
    f64::ray_t r = from;
    r.d[0] = dir[0];
    r.d[1] = dir[1];
    r.d[2] = dir[2];
    double tau = 0.0;
    for (;;) {
        double l;
        const int ap = f64::get_dl(r, l);
        if (ap < 0)
            break;
        const size_t gc = (size_t)r.ib * P.n3 + f64::cell_index(P, r);
        const double kap = (double)P.kappa[gc] * ks;
        tau = tau + kap * l;
        if (d_logrho && A != 0.0)
            atomic_add_f64(d_logrho + gc, -((kap * l) * A));
        if (tau > kTauCut)
            break;
        bool reload;
        const int st = f64::cross_face(P, r, ap, l, reload);
        if (st != f64::ST_IN)
            break;
        if (reload)
            f64::gl_load(P, r);
    }
}

// A of one peel-off: w(pixel) . S * exp(-tau) / (4 pi)
__device__ __forceinline__ double peel_A(const kparams &P, const aparams &Q, const f64::ray_t &r,
                                         int band, int o, double ks, double I, double Qs,
                                         double U)
{
    uint64_t dummy = 0;
    const double tau = f64::tau_to_edge(P, r, P.no[o], ks, dummy);
    const double att = kp_exp(-tau) * KP_INV_4PI;
    const long px = f64::pixel(P, band, o, r.x);
    if (px < 0)
        return 0.0;
    return Q.w[3 * px] * (I * att) + Q.w[3 * px + 1] * (Qs * att) + Q.w[3 * px + 2] * (U * att);
}

__device__ void adjoint_packet(const kparams &P, const aparams &Q, int band, uint64_t pid,
                               uint64_t tid, double &jsum)
{
    const dband bd = P.bands[band];
    const double ks = bd.kappa_scale;
    const double w0 = P.w0;
    const uint32_t k0 = P.seed, k1 = (uint32_t)band;
    const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32);
    const int no_ = P.n_incl;
    const bool hg = P.model == KP_HG_RAYLEIGH && (bd.flags & KP_BAND_CELL_G);
    const bool calb = (bd.flags & KP_BAND_CELL_ALBEDO) != 0;
    double A_tot = 0.0;
    auto slot = [&](int ev, int o) -> double * {
        const uint32_t k = (uint32_t)ev * no_ + o;
        return k < Q.cap ? Q.scratch + (uint64_t)k * Q.n_threads + tid : nullptr;
    };

    for (int pass = 1; pass <= 2; ++pass) {
        uint32_t rn[4] = {plo, phi, 0u, 0u};
        kp_philox4x32_10(rn, k0, k1);
        f64::ray_t r;
        double e1[3];
        double tau_rem;
        f64::emit(rn, r.d, e1, tau_rem);
        double I = w0, Qs = 0.0, U = 0.0;
        r.x[0] = P.src_x[0];
        r.x[1] = P.src_x[1];
        r.x[2] = P.src_x[2];
        r.ib = P.src_ib;
        r.i[0] = P.src_i[0];
        r.i[1] = P.src_i[1];
        r.i[2] = P.src_i[2];
        f64::gl_load(P, r);
        double B = 0.0;
        for (int o = 0; o < no_; ++o) {
            double *sA = slot(0, o);
            if (pass == 1) {
                const double a = peel_A(P, Q, r, band, o, ks, I, 0.0, 0.0);
                if (sA)
                    *sA = a;
                A_tot += a;
            } else {
                const double a = sA ? *sA : peel_A(P, Q, r, band, o, ks, I, 0.0, 0.0);
                peel_grad(P, r, P.no[o], ks, a, Q.d_logrho);
            }
        }
        if (pass == 2) {
            B = A_tot;
            for (int o = 0; o < no_; ++o) {
                double *sA = slot(0, o);
                B -= sA ? *sA : peel_A(P, Q, r, band, o, ks, I, 0.0, 0.0);
            }
        }
        int steps = P.step_max;
        int events = 0;
        for (;;) {
            bool interact = false;
            size_t gc = 0;
            for (;;) {
                if (steps <= 0)
                    break;
                double l;
                const int ap = f64::get_dl(r, l);
                if (ap < 0)
                    break;
                gc = (size_t)r.ib * P.n3 + f64::cell_index(P, r);
                const double kap = (double)P.kappa[gc] * ks;
                const double dtau = kap * l;
                steps -= 1;
                if (kap > 0.0 && dtau >= tau_rem) {
                    double li = tau_rem / kap;
                    if (li > l)
                        li = l;
                    if (pass == 2 && Q.d_logrho && B != 0.0)
                        atomic_add_f64(Q.d_logrho + gc, -((kap * li) * B));
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        r.x[a] = r.x[a] + li * r.d[a];
                    interact = true;
                    break;
                }
                tau_rem = tau_rem - dtau;
                if (pass == 2 && Q.d_logrho && B != 0.0)
                    atomic_add_f64(Q.d_logrho + gc, -(dtau * B));
                bool reload;
                const int st = f64::cross_face(P, r, ap, l, reload);
                if (st != f64::ST_IN)
                    break;
                if (reload)
                    f64::gl_load(P, r);
            }
            if (!interact)
                break;
            events += 1;
            uint32_t u[4] = {plo, phi, (uint32_t)events, 0u};
            kp_philox4x32_10(u, k0, k1);
            const double alb = calb ? (double)P.albedo[gc] : bd.albedo;
            const double gg =
                (bd.flags & KP_BAND_CELL_G) ? (P.g ? (double)P.g[gc] : 0.0) : bd.g;
            I = I * alb;
            Qs = Qs * alb;
            U = U * alb;
            if (pass == 2 && B != 0.0) {
                if (Q.d_logrho)
                    atomic_add_f64(Q.d_logrho + gc, B);
                if (calb && Q.d_alb)
                    atomic_add_f64(Q.d_alb + gc, B / alb);
            }
            double e2[3];
            f64::cross3(r.d, e1, e2);
            for (int o = 0; o < no_; ++o) {
                double Ip, Qi, Ui;
                f64::peel_stokes(P.model, gg, P.pmax, r.d, e1, e2, I, Qs, U, P.no[o], P.ev[o], Ip,
                                 Qi, Ui);
                double *sA = slot(events, o);
                if (pass == 1) {
                    const double a = peel_A(P, Q, r, band, o, ks, Ip, Qi, Ui);
                    if (sA)
                        *sA = a;
                    A_tot += a;
                } else {
                    const double a = sA ? *sA : peel_A(P, Q, r, band, o, ks, Ip, Qi, Ui);
                    peel_grad(P, r, P.no[o], ks, a, Q.d_logrho);
                    if (hg && Q.d_g && a != 0.0) {
                        double mo = f64::dot3(r.d, P.no[o]);
                        mo = fmin(fmax(mo, -1.0), 1.0);
                        atomic_add_f64(Q.d_g + gc, a * dlogp_hg_h(gg, mo));
                    }
                    B -= a;
                }
            }
            const double mu = (P.model == KP_THOMSON) ? kp_thomson_mu(kp_u01(u[0]))
                                                      : kp_hg_mu(gg, kp_u01(u[0]));
            if (pass == 2 && hg && Q.d_g && B != 0.0)
                atomic_add_f64(Q.d_g + gc, B * dlogp_hg_h(gg, mu));
            f64::scatter_stokes(P.model, P.pmax, mu, kp_u01(u[1]), r.d, e1, e2, I, Qs, U);
            tau_rem = -kp_log(kp_u01(u[2]));
            if (I < P.rr_w * w0) {
                if (kp_u01(u[3]) < P.rr_s) {
                    I = I / P.rr_s;
                    Qs = Qs / P.rr_s;
                    U = U / P.rr_s;
                } else {
                    break;
                }
            }
        }
    }
    jsum += A_tot;
}

} // namespace adj
} // namespace kp
