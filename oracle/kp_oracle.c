/*
 * kp_oracle.c -- FP64 CPU oracle of the photon-packet transport path.
 *
 * TEST INFRASTRUCTURE ONLY (see kp_oracle.h).  Built with
 * -ffp-contract=off so every multiply and add rounds on its own; fused
 * operations are the explicit fma() calls.  The leaf arithmetic (Philox,
 * log, exp, sincos, scattering-angle inversion) comes from the shared
 * deterministic header and is pinned separately against libm and the
 * Random123 known-answer vectors.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

#include "../kratos_polrad_amd/csrc/kp_math.h"
#include "kp_oracle.h"

#define TAU_CUT 40.0

/* Planted mistakes: built with -DKP_PLANT=k (k > 0), the oracle carries the
 * k-th deliberate error below, for the test that checks every one of them
 * fails some pin.  The default build (KP_PLANT = 0) has none. */
#ifndef KP_PLANT
#define KP_PLANT 0
#endif
#define PLANT(k) (KP_PLANT == (k))

/* ------------------------------------------------------------------ */
/* Leaf wrappers                                                       */

void or_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    kp_philox4x32_10(c, k0, k1);
    memcpy(out, c, sizeof(c));
}
double or_log(double x) { return kp_log(x); }
double or_exp(double x) { return kp_exp(x); }
void or_sincos2pi(double u, double *s, double *c) { kp_sincos2pi(u, s, c); }
double or_thomson_mu(double xi) { return kp_thomson_mu(xi); }
double or_hg_mu(double g, double xi) { return kp_hg_mu(g, xi); }

/* ------------------------------------------------------------------ */
/* Ray state and the reference march                                   */

typedef struct {
    double x[3], d[3];
    double xl[3], xr[3]; /* geo_loc_t face bounds of the current cell */
    int i[3];
    int ib;
} ray_t;

enum { ST_IN = 0, ST_ESCAPED = 1, ST_LOST = 4 };

static double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* geo_t::x_fc for uniform axes (geometry.h:80-81) */
static double x_fc(const kp_mesh_desc *m, int ib, int a, int i)
{
    return m->xf0[3 * ib + a] + (double)i * m->dx0[3 * ib + a];
}

/* geo_loc_t::load (particle_base.h:91-100) */
static void gl_load(const kp_mesh_desc *m, ray_t *r)
{
    for (int a = 0; a < 3; ++a) {
        r->xl[a] = x_fc(m, r->ib, a, r->i[a]);
        r->xr[a] = x_fc(m, r->ib, a, r->i[a] + 1);
    }
}

/* cart_t::get_dl (photon.h:20-44): face distance per axis, argmin with ties
 * going to the highest axis.  Returns the axis, or -1 if no axis moves. */
static int get_dl(const ray_t *r, double *lmin)
{
    double dl[3];
    for (int a = 0; a < 3; ++a) {
        if (r->d[a] == 0.0) {
            dl[a] = -1.0;
            continue;
        }
        const double xf = r->d[a] > 0.0 ? r->xr[a] : r->xl[a];
        dl[a] = (xf - r->x[a]) / r->d[a];
        if (dl[a] < 0.0)
            dl[a] = 0.0; /* deviation (2) */
    }
    int ap = -1;
    double best = DBL_MAX;
    for (int a = 0; a < 3; ++a)
        if (dl[a] >= 0.0 && (PLANT(10) ? dl[a] < best : dl[a] <= best)) {
            ap = a;
            best = dl[a];
        }
    *lmin = best;
    return ap;
}

/* proc_geo (photon.h:54-71) + shift_blk (proxy.h:64-121) + regulate
 * (particle_base.h:37-56).  *reload is set when the cell bounds must be
 * recomputed from the (new) block. */
static int cross_face(const kp_mesh_desc *m, ray_t *r, int ap, double l, int *reload)
{
    *reload = 0;
    for (int a = 0; a < 3; ++a) {
        if (a == ap) {
            const int rhs = r->d[a] > 0.0;
            r->x[a] = rhs ? r->xr[a] : r->xl[a];
            r->i[a] += rhs ? 1 : -1;
            if (rhs) {
                r->xl[a] = r->xr[a];
                r->xr[a] = x_fc(m, r->ib, a, r->i[a] + 1);
            } else {
                r->xr[a] = r->xl[a];
                r->xl[a] = x_fc(m, r->ib, a, r->i[a]);
            }
        } else {
            r->x[a] = r->x[a] + l * r->d[a];
        }
    }
    const int n = m->n_cell;
    for (int a = 0; a < 3; ++a) {
        if (r->i[a] >= 0 && r->i[a] < n)
            continue;
        const int side = r->i[a] > 0;
        if (m->phybnd[r->ib] & (1u << (2 * a + side)))
            return ST_ESCAPED;
        const int ap0 = (a + 1) % 3, ap1 = (a + 2) % 3;
        const int dl = m->d_lvl[6 * r->ib + 2 * a + side];
        int j0 = 0, j1 = 0;
        if (dl > 0) {
            j0 = r->i[ap0] >= n / 2;
            j1 = r->i[ap1] >= n / 2;
        }
        const int slot = 24 * r->ib + ((a * 2 + side) * 2 + j0) * 2 + j1;
        const int nb = m->nb_id[slot];
        if (nb < 0)
            return ST_LOST;
        const int shift = m->nb_shift[slot];
        r->ib = nb;
        r->i[a] = side ? 0 : n - 1;
        const int apj[2] = {ap0, ap1};
        if (dl < 0) {
            for (int j = 0; j <= 1; ++j) {
                const int is = r->i[apj[j]];
                r->i[apj[j]] = is / 2 + ((shift >> j) & 1) * (n / 2);
            }
        } else if (dl > 0) {
            for (int j = 0; j <= 1; ++j) {
                const int b = apj[j];
                const int is = r->i[b];
                int f = (is < n / 2 ? is : is - n / 2) * 2;
                if (r->x[b] >= x_fc(m, nb, b, f + 1)) /* deviation (1) */
                    f += 1;
                r->i[b] = f;
            }
        }
        *reload = 1;
        break;
    }
    /* periodic wrap (particle_base.h:37-56); deviation (3): a ray that just
     * left through the lower domain face sits exactly on it and is moved to
     * the upper image (the reference's strict x < x_min leaves it at x_min) */
    for (int a = 0; a < 3; ++a)
        if (m->periodic[a]) {
            const double lbox = m->x_max[a] - m->x_min[a];
            const int down_exit = (a == ap) && !(r->d[a] > 0.0);
            if (r->x[a] < m->x_min[a] || (down_exit && r->x[a] == m->x_min[a])) {
                r->x[a] += lbox;
                *reload = 1;
            } else if (r->x[a] >= m->x_max[a]) {
                r->x[a] -= lbox;
                *reload = 1;
            }
        }
    return ST_IN;
}

static int cell_index(const kp_mesh_desc *m, const ray_t *r)
{
    const int n = m->n_cell;
    return r->i[0] + n * (r->i[1] + n * r->i[2]);
}

int or_locate(const kp_mesh_desc *m, const double x[3], int32_t *ib, int32_t i[3])
{
    const int n = m->n_cell;
    for (int b = 0; b < m->n_blocks; ++b) {
        int in = 1;
        for (int a = 0; a < 3 && in; ++a) {
            const double lo = m->xf0[3 * b + a];
            const double hi = lo + n * m->dx0[3 * b + a];
            in = (x[a] >= lo && x[a] < hi);
        }
        if (!in)
            continue;
        *ib = b;
        for (int a = 0; a < 3; ++a) {
            int c = (int)((x[a] - m->xf0[3 * b + a]) / m->dx0[3 * b + a]);
            if (c < 0) c = 0;
            if (c > n - 1) c = n - 1;
            i[a] = c;
        }
        return 0;
    }
    *ib = -1;
    return -1;
}

int or_trace(const kp_mesh_desc *m, const double x[3], const double d[3], int max_steps,
             int32_t *out_ib, int32_t *out_cell, double *out_len, double x_end[3],
             int *status)
{
    ray_t r;
    memcpy(r.x, x, sizeof(r.x));
    memcpy(r.d, d, sizeof(r.d));
    *status = ST_LOST;
    if (or_locate(m, x, &r.ib, r.i) != 0)
        return 0;
    gl_load(m, &r);
    int k = 0;
    *status = 0;
    while (k < max_steps) {
        double l;
        const int ap = get_dl(&r, &l);
        if (ap < 0) {
            *status = ST_LOST;
            break;
        }
        out_ib[k] = r.ib;
        out_cell[k] = cell_index(m, &r);
        out_len[k] = l;
        ++k;
        int reload;
        const int st = cross_face(m, &r, ap, l, &reload);
        if (st != ST_IN) {
            *status = st;
            break;
        }
        if (reload)
            gl_load(m, &r);
    }
    memcpy(x_end, r.x, sizeof(r.x));
    return k;
}

/* ------------------------------------------------------------------ */
/* Forward transport                                                   */

typedef struct {
    const kp_mesh_desc *m;
    const float *kappa, *albedo, *g;
    const kp_source *src;
    const kp_run *run;
    int32_t src_ib, src_i[3];
    int n_incl;
    double no[8][3], ev[8][3], eh[8][3];
    double center[3];
    double pix_scale;
    uint64_t total;
    uint64_t band_start[1024];
} fctx_t;

typedef struct {
    const fctx_t *c;
    uint64_t k_lo, k_hi; /* global packet slots */
    double *J, *image;
    uint64_t cnt[KP_NCOUNTERS];
    double *end_state;
} fthr_t;

/* optical depth from r to the domain edge along d (peel-off ray) */
static double tau_to_edge(const fctx_t *c, const ray_t *from, const double *dir, double ks,
                          uint64_t *ncross)
{
    const kp_mesh_desc *m = c->m;
    const int n3 = m->n_cell * m->n_cell * m->n_cell;
    ray_t r = *from;
    memcpy(r.d, dir, sizeof(r.d));
    double tau = 0.0;
    for (;;) {
        double l;
        const int ap = get_dl(&r, &l);
        if (ap < 0)
            break;
        const double kap = (double)c->kappa[(size_t)r.ib * n3 + cell_index(m, &r)] * ks;
        tau = tau + kap * l;
        *ncross += 1;
        if (tau > TAU_CUT)
            break;
        int reload;
        const int st = cross_face(m, &r, ap, l, &reload);
        if (st != ST_IN)
            break;
        if (reload)
            gl_load(m, &r);
    }
    return tau;
}

static void deposit(const fctx_t *c, fthr_t *t, int band, int o, const double *x, double I,
                    double Q, double U)
{
    if (!t->image)
        return;
    const int N = c->run->image_n;
    double rel[3];
    for (int a = 0; a < 3; ++a)
        rel[a] = x[a] - c->center[a];
    const double uh = dot3(rel, PLANT(8) ? c->ev[o] : c->eh[o]);
    const double uv = dot3(rel, PLANT(8) ? c->eh[o] : c->ev[o]);
    const double fx = floor((uh + c->run->image_half) * c->pix_scale);
    const double fy = floor((uv + c->run->image_half) * c->pix_scale);
    if (!(fx >= 0.0 && fy >= 0.0 && fx < (double)N && fy < (double)N))
        return;
    const size_t pix = (((size_t)band * c->n_incl + o) * N + (size_t)fy) * N + (size_t)fx;
    t->image[3 * pix + 0] += I;
    t->image[3 * pix + 1] += Q;
    t->image[3 * pix + 2] += U;
}

/* normalized phase-matrix elements at scattering cosine mu */
static void phase(int model, double gg, double pmax, double mu, double *z11, double *z12,
                  double *z33)
{
    const double m2 = mu * mu;
    if (model == KP_THOMSON) {
        *z11 = 0.75 * (1.0 + m2);
        *z12 = PLANT(13) ? 0.75 * (1.0 - m2) : 0.75 * (m2 - 1.0);
        *z33 = 1.5 * mu;
    } else {
        const double D = (1.0 + gg * gg) - 2.0 * gg * mu;
        const double ph = PLANT(6) ? (1.0 - gg * gg) / D : (1.0 - gg * gg) / (D * sqrt(D));
        const double pol = PLANT(1) ? pmax * (1.0 - m2) : pmax * (1.0 - m2) / (1.0 + m2);
        *z11 = ph;
        *z12 = -pol * ph;
        *z33 = PLANT(2) ? ph * (2.0 * mu) : PLANT(3) ? ph * mu : ph * (2.0 * mu / (1.0 + m2));
    }
}

/* Stokes vector (I,Q,U) referenced to e1 (Q>0: polarized along e1), e2 = d x e1.
 * Rotation of the reference axis by phi: Q' = Q cos2phi + U sin2phi,
 * U' = U cos2phi - Q sin2phi (Chandrasekhar 1960, Radiative Transfer, ch. I). */

/* Peel-off toward observer direction no with image vertical axis ev: the
 * scattered Stokes vector Z(mu_o) R(phi_o) S in the image frame (Q>0 along
 * ev), without the 1/(4 pi) and attenuation factors. */
static void peel_stokes(int model, double gg, double pmax, const double *d, const double *e1,
                        const double *e2, const double *S, const double *no, const double *ev,
                        double *Si)
{
    double mo = dot3(d, no);
    if (mo > 1.0) mo = 1.0;
    if (mo < -1.0) mo = -1.0;
    double p[3];
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
    const double Qr = S[1] * c2 + S[2] * s2;
    const double Ur = S[2] * c2 - S[1] * s2;
    double z11, z12, z33;
    phase(model, gg, pmax, mo, &z11, &z12, &z33);
    const double Ip = z11 * S[0] + z12 * Qr;
    const double Qp = z12 * S[0] + z11 * Qr;
    const double Up = z33 * Ur;
    double e1o[3], e2o[3];
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
    Si[0] = Ip;
    Si[1] = Qp * c2p + Up * s2p;
    Si[2] = PLANT(7) ? Qp * s2p - Up * c2p : Up * c2p - Qp * s2p;
}

/* Scattering by cos(theta)=mu and azimuth 2 pi phi_u measured from e1
 * toward e2: new direction and frame in d/e1, Stokes weight Z S / Z11. */
static void scatter_stokes(int model, double pmax, double mu, double phi_u, double *d,
                           double *e1, const double *e2, double *S)
{
    double s2m = fma(-mu, mu, 1.0);
    if (s2m < 0.0) s2m = 0.0;
    const double st = sqrt(s2m);
    double sf, cf;
    kp_sincos2pi(phi_u, &sf, &cf);
    const double c2 = cf * cf - sf * sf;
    const double s2 = 2.0 * cf * sf;
    const double Qr = PLANT(9) ? S[1] * c2 - S[2] * s2 : S[1] * c2 + S[2] * s2;
    const double Ur = PLANT(9) ? S[2] * c2 + S[1] * s2 : S[2] * c2 - S[1] * s2;
    const double m2 = mu * mu;
    const double pl = (model == KP_THOMSON) ? 1.0 : pmax;
    const double a12 = (PLANT(5) ? pl : -pl) * (1.0 - m2) / (1.0 + m2);
    const double a33 = (PLANT(4) ? pl : 1.0) * (2.0 * mu / (1.0 + m2));
    const double In = S[0] + a12 * Qr;
    const double Qn = a12 * S[0] + Qr;
    const double Un = a33 * Ur;
    S[0] = In;
    S[1] = Qn;
    S[2] = Un;
    double dn[3], en[3];
    for (int a = 0; a < 3; ++a) {
        const double e1r = cf * e1[a] + sf * e2[a];
        dn[a] = mu * d[a] + st * e1r;
        en[a] = mu * e1r - st * d[a];
    }
    const double id = 1.0 / sqrt(dot3(dn, dn));
    for (int a = 0; a < 3; ++a)
        dn[a] = dn[a] * id;
    const double pr = dot3(en, dn);
    for (int a = 0; a < 3; ++a)
        en[a] = en[a] - pr * dn[a];
    const double ie = 1.0 / sqrt(dot3(en, en));
    for (int a = 0; a < 3; ++a) {
        d[a] = dn[a];
        e1[a] = en[a] * ie;
    }
}

void or_peel_stokes(int model, double g, double pmax, const double d[3], const double e1[3],
                    const double S[3], const double no[3], const double ev[3], double Si[3])
{
    double e2[3];
    cross3(d, e1, e2);
    peel_stokes(model, g, pmax, d, e1, e2, S, no, ev, Si);
}

void or_scatter_stokes(int model, double pmax, double mu, double phi_u, double d[3],
                       double e1[3], double S[3])
{
    double e2[3];
    cross3(d, e1, e2);
    scatter_stokes(model, pmax, mu, phi_u, d, e1, e2, S);
}

static void packet(const fctx_t *c, fthr_t *t, int band, uint64_t pid, uint64_t slot)
{
    const kp_mesh_desc *m = c->m;
    const kp_run *run = c->run;
    const kp_band *bd = &run->bands[band];
    const int n3 = m->n_cell * m->n_cell * m->n_cell;
    const int nbands = run->n_bands;
    const double ks = bd->kappa_scale;
    const double w0 = c->src->weight;
    const uint32_t k0 = run->seed, k1 = (uint32_t)band;
    const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32);

    uint64_t ncross = 0, npeel = 0;
    int events = 0, status = 0;

    /* emission (event 0) */
    uint32_t rn[4] = {plo, phi, 0u, 0u};
    kp_philox4x32_10(rn, k0, k1);
    ray_t r;
    double e1[3];
    {
        const double mu = fma(2.0, kp_u01(rn[0]), -1.0);
        double sf, cf;
        kp_sincos2pi(kp_u01(rn[1]), &sf, &cf);
        double s2 = fma(-mu, mu, 1.0);
        if (s2 < 0.0) s2 = 0.0;
        const double st = sqrt(s2);
        r.d[0] = st * cf;
        r.d[1] = st * sf;
        r.d[2] = mu;
        e1[0] = mu * cf;
        e1[1] = mu * sf;
        e1[2] = -st;
    }
    double tau_rem = -kp_log(kp_u01(rn[2]));
    double I = w0, Q = 0.0, U = 0.0;
    memcpy(r.x, c->src->x, sizeof(r.x));
    r.ib = c->src_ib;
    memcpy(r.i, c->src_i, sizeof(r.i));
    gl_load(m, &r);
    for (int o = 0; o < c->n_incl; ++o) {
        const double tau = tau_to_edge(c, &r, c->no[o], ks, &npeel);
        const double att = kp_exp(-tau) * KP_INV_4PI;
        deposit(c, t, band, o, r.x, I * att, 0.0, 0.0);
    }

    int steps = run->step_max;
    int cell = 0;
    for (;;) {
        /* march to the next interaction or out of the domain */
        int interact = 0;
        for (;;) {
            if (steps <= 0) {
                status = KP_END_STEPS;
                break;
            }
            double l;
            const int ap = get_dl(&r, &l);
            if (ap < 0) {
                status = KP_END_LOST;
                break;
            }
            cell = cell_index(m, &r);
            const size_t gc = (size_t)r.ib * n3 + cell;
            const double kap = (double)c->kappa[gc] * ks;
            const double dtau = kap * l;
            const double V = m->vol0[r.ib];
            ncross += 1;
            steps -= 1;
            if (kap > 0.0 && dtau >= tau_rem) {
                double li = tau_rem / kap;
                if (li > l)
                    li = l;
                if (t->J)
                    t->J[gc * nbands + band] += I * (PLANT(11) ? li : li / V);
                for (int a = 0; a < 3; ++a)
                    r.x[a] = r.x[a] + li * r.d[a];
                interact = 1;
                break;
            }
            tau_rem = tau_rem - dtau;
            if (t->J)
                t->J[gc * nbands + band] += I * (PLANT(11) ? l : l / V);
            int reload;
            const int st = cross_face(m, &r, ap, l, &reload);
            if (st == ST_ESCAPED) {
                status = KP_END_ESCAPED;
                break;
            }
            if (st == ST_LOST) {
                status = KP_END_LOST;
                break;
            }
            if (reload)
                gl_load(m, &r);
        }
        if (!interact)
            break;

        /* interaction */
        events += 1;
        uint32_t u[4] = {plo, phi, (uint32_t)events, 0u};
        kp_philox4x32_10(u, k0, k1);
        const size_t gc = (size_t)r.ib * n3 + cell;
        const double alb = PLANT(12) ? 1.0
                         : (bd->flags & KP_BAND_CELL_ALBEDO) ? (double)c->albedo[gc] : bd->albedo;
        const double gg = (bd->flags & KP_BAND_CELL_G) ? (c->g ? (double)c->g[gc] : 0.0) : bd->g;
        I = I * alb;
        Q = Q * alb;
        U = U * alb;

        double e2[3];
        cross3(r.d, e1, e2);

        /* peel-off toward every observer */
        for (int o = 0; o < c->n_incl; ++o) {
            double S[3] = {I, Q, U}, Si[3];
            peel_stokes(run->scatter_model, gg, run->p_max, r.d, e1, e2, S, c->no[o], c->ev[o],
                        Si);
            const double tau = tau_to_edge(c, &r, c->no[o], ks, &npeel);
            const double att = PLANT(14) ? KP_INV_4PI : kp_exp(-tau) * KP_INV_4PI;
            deposit(c, t, band, o, r.x, Si[0] * att, Si[1] * att, Si[2] * att);
        }

        /* scatter: sample from the unpolarized phase function, carry the
         * polarization in the Stokes weight */
        {
            const double mu = (run->scatter_model == KP_THOMSON) ? kp_thomson_mu(kp_u01(u[0]))
                                                                  : kp_hg_mu(gg, kp_u01(u[0]));
            double S[3] = {I, Q, U};
            scatter_stokes(run->scatter_model, run->p_max, mu, kp_u01(u[1]), r.d, e1, e2, S);
            I = S[0];
            Q = S[1];
            U = S[2];
        }
        tau_rem = -kp_log(kp_u01(u[2]));
        if (I < run->rr_weight * w0) {
            if (kp_u01(u[3]) < run->rr_survive) {
                I = I / run->rr_survive;
                Q = Q / run->rr_survive;
                U = U / run->rr_survive;
            } else {
                status = KP_END_RR;
                break;
            }
        }
    }

    t->cnt[KP_C_CROSS_MAIN] += ncross;
    t->cnt[KP_C_CROSS_PEEL] += npeel;
    t->cnt[KP_C_EVENTS] += (uint64_t)events;
    t->cnt[KP_C_PACKETS] += 1;
    if (status == KP_END_ESCAPED) t->cnt[KP_C_ESCAPED] += 1;
    else if (status == KP_END_RR) t->cnt[KP_C_KILLED_RR] += 1;
    else if (status == KP_END_STEPS) t->cnt[KP_C_KILLED_STEPS] += 1;
    else t->cnt[KP_C_LOST] += 1;
    if (t->end_state) {
        double *es = t->end_state + slot * KP_ES_STRIDE;
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
        es[KP_ES_CELL] = (double)cell_index(m, &r);
        es[14] = (double)npeel;
        es[15] = 0.0;
    }
}

static void *worker(void *arg)
{
    fthr_t *t = (fthr_t *)arg;
    const fctx_t *c = t->c;
    int band = 0;
    for (uint64_t k = t->k_lo; k < t->k_hi; ++k) {
        while (k >= c->band_start[band + 1])
            ++band;
        const uint64_t pid = c->run->packet_offset + (k - c->band_start[band]);
        packet(c, t, band, pid, k);
    }
    return NULL;
}

/* wall time of the last or_forward's transport phase (threads started to
 * threads joined: no tally allocation or per-thread reduction) */
static double g_transport_seconds = 0.0;
double or_last_transport_seconds(void) { return g_transport_seconds; }

int or_forward(const kp_mesh_desc *m, const float *kappa, const float *albedo, const float *g,
               const kp_source *src, const kp_run *run, double *J, double *image,
               uint64_t *counters, double *end_state, int n_threads)
{
    if (!m || !kappa || !src || !run || run->n_bands < 1 || run->n_bands > 1023)
        return -1;
    if (run->n_incl < 0 || run->n_incl > 8)
        return -1;
    fctx_t *c = (fctx_t *)calloc(1, sizeof(fctx_t));
    c->m = m;
    c->kappa = kappa;
    c->albedo = albedo;
    c->g = g;
    c->src = src;
    c->run = run;
    if (or_locate(m, src->x, &c->src_ib, c->src_i) != 0) {
        free(c);
        return -2;
    }
    c->n_incl = run->n_incl;
    for (int o = 0; o < run->n_incl; ++o) {
        double s, co;
        kp_sincos2pi(run->incl_deg[o] / 360.0, &s, &co);
        c->no[o][0] = s;  c->no[o][1] = 0.0; c->no[o][2] = co;
        c->ev[o][0] = -co; c->ev[o][1] = 0.0; c->ev[o][2] = s;
        c->eh[o][0] = 0.0; c->eh[o][1] = 1.0; c->eh[o][2] = 0.0;
    }
    for (int a = 0; a < 3; ++a)
        c->center[a] = (m->x_min[a] + m->x_max[a]) * 0.5;
    c->pix_scale = run->image_n > 0 ? (double)run->image_n / (2.0 * run->image_half) : 0.0;
    c->band_start[0] = 0;
    for (int b = 0; b < run->n_bands; ++b)
        c->band_start[b + 1] = c->band_start[b] + run->bands[b].n_packets;
    c->total = c->band_start[run->n_bands];
    for (int b = run->n_bands + 1; b < 1024; ++b)
        c->band_start[b] = c->total;

    if (n_threads < 1)
        n_threads = 1;
    if ((uint64_t)n_threads > c->total && c->total > 0)
        n_threads = (int)c->total;
    const size_t ncell = (size_t)m->n_blocks * m->n_cell * m->n_cell * m->n_cell;
    const size_t nJ = ncell * run->n_bands;
    const size_t nimg = (size_t)run->n_bands * run->n_incl * run->image_n * run->image_n * 3;
    fthr_t *th = (fthr_t *)calloc((size_t)n_threads, sizeof(fthr_t));
    pthread_t *tid = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    int rc = 0;
    for (int k = 0; k < n_threads; ++k) {
        th[k].c = c;
        th[k].k_lo = c->total * k / n_threads;
        th[k].k_hi = c->total * (k + 1) / n_threads;
        th[k].end_state = end_state;
        if (J) {
            th[k].J = (k == 0) ? J : (double *)calloc(nJ, sizeof(double));
            if (!th[k].J) rc = -3;
        }
        if (image && nimg) {
            th[k].image = (k == 0) ? image : (double *)calloc(nimg, sizeof(double));
            if (!th[k].image) rc = -3;
        }
    }
    if (rc == 0) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (int k = 1; k < n_threads; ++k)
            pthread_create(&tid[k], NULL, worker, &th[k]);
        worker(&th[0]);
        for (int k = 1; k < n_threads; ++k)
            pthread_join(tid[k], NULL);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        g_transport_seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        for (int k = 1; k < n_threads; ++k) {
            if (J)
                for (size_t q = 0; q < nJ; ++q)
                    J[q] += th[k].J[q];
            if (image && nimg)
                for (size_t q = 0; q < nimg; ++q)
                    image[q] += th[k].image[q];
        }
        if (counters)
            for (int k = 0; k < n_threads; ++k)
                for (int q = 0; q < KP_NCOUNTERS; ++q)
                    counters[q] += th[k].cnt[q];
    }
    for (int k = 1; k < n_threads; ++k) {
        if (J) free(th[k].J);
        if (image && nimg) free(th[k].image);
    }
    free(th);
    free(tid);
    free(c);
    return rc;
}


/* ------------------------------------------------------------------ */
/* Adjoint by path replay                                              */
/*
 * Gradient of the linear image functional Jlin = sum_pix w_pix . S_pix
 * (w = dL/dimage, Stokes I,Q,U per pixel) with respect to per-cell
 * log-density (kappa_rho scales with rho), albedo and HG g.  The image
 * estimator is the path integral f/p over sampled histories, so
 * dE[Jlin] = E[(f/p) dlog f]: for the peel-off contribution A_k of event k,
 *   dlog f/dlog rho_c = (#vertices j<=k in c) - tau_c(main path up to x_k)
 *                       - tau_c(peel-off ray k),
 *   dlog f/d a_c      = (#vertices j<=k in c) / a_c,
 *   dlog f/d g_c      = sum_{scatterings j<k in c} dlogP_HG(mu_j)/dg
 *                       + [x_k in c] dlogP_HG(mu_obs,k)/dg.
 * Path replay backpropagation (Vicini, Speierer & Jakob 2021): pass 1 replays
 * the history to get every A_k; pass 2 replays it again carrying the suffix
 * sum B_j = sum_{k>=j} A_k, so each segment, vertex and scattering adds its
 * term once.  Russian roulette adds no term (its probability is not
 * differentiated; the estimator stays unbiased).
 */

typedef struct {
    const double *w;          /* dL/dimage, image layout */
    double *d_logrho, *d_alb, *d_g;
    double jlin;
    double *A;                /* A per (event, observer) of the current packet */
    size_t A_cap;
} athr_t;

static double dlogp_hg(double gg, double mu)
{
    const double D = (1.0 + gg * gg) - 2.0 * gg * mu;
    return -2.0 * gg / (1.0 - gg * gg) - 3.0 * (gg - mu) / D;
}

/* pixel index of x for observer o, or -1 */
static long pix_index(const fctx_t *c, int band, int o, const double *x)
{
    const int N = c->run->image_n;
    double rel[3];
    for (int a = 0; a < 3; ++a)
        rel[a] = x[a] - c->center[a];
    const double uh = dot3(rel, PLANT(8) ? c->ev[o] : c->eh[o]);
    const double uv = dot3(rel, PLANT(8) ? c->eh[o] : c->ev[o]);
    const double fx = floor((uh + c->run->image_half) * c->pix_scale);
    const double fy = floor((uv + c->run->image_half) * c->pix_scale);
    if (!(fx >= 0.0 && fy >= 0.0 && fx < (double)N && fy < (double)N))
        return -1;
    return (long)((((size_t)band * c->n_incl + o) * N + (size_t)fy) * N + (size_t)fx);
}

/* peel-off ray gradient: d_logrho[c] -= kappa_c l_c A along the ray, with the
 * same march and the same early stop as tau_to_edge */
static void peel_grad(const fctx_t *c, const ray_t *from, const double *dir, double ks, double A,
                      double *d_logrho)
{
    const kp_mesh_desc *m = c->m;
    const int n3 = m->n_cell * m->n_cell * m->n_cell;
    ray_t r = *from;
    memcpy(r.d, dir, sizeof(r.d));
    double tau = 0.0;
    for (;;) {
        double l;
        const int ap = get_dl(&r, &l);
        if (ap < 0)
            break;
        const size_t gc = (size_t)r.ib * n3 + cell_index(m, &r);
        const double kap = (double)c->kappa[gc] * ks;
        tau = tau + kap * l;
        d_logrho[gc] -= (kap * l) * A;
        if (tau > TAU_CUT)
            break;
        int reload;
        const int st = cross_face(m, &r, ap, l, &reload);
        if (st != ST_IN)
            break;
        if (reload)
            gl_load(m, &r);
    }
}

static void adjoint_packet(const fctx_t *c, athr_t *t, int band, uint64_t pid)
{
    const kp_mesh_desc *m = c->m;
    const kp_run *run = c->run;
    const kp_band *bd = &run->bands[band];
    const int n3 = m->n_cell * m->n_cell * m->n_cell;
    const double ks = bd->kappa_scale;
    const double w0 = c->src->weight;
    const uint32_t k0 = run->seed, k1 = (uint32_t)band;
    const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32);
    const int no_ = c->n_incl;
    const int hg = run->scatter_model == KP_HG_RAYLEIGH && (bd->flags & KP_BAND_CELL_G);
    const int calb = (bd->flags & KP_BAND_CELL_ALBEDO) != 0;
    double A_tot = 0.0;
    uint64_t dummy = 0;

    for (int pass = 1; pass <= 2; ++pass) {
        uint32_t rn[4] = {plo, phi, 0u, 0u};
        kp_philox4x32_10(rn, k0, k1);
        ray_t r;
        double e1[3];
        {
            const double mu = fma(2.0, kp_u01(rn[0]), -1.0);
            double sf, cf;
            kp_sincos2pi(kp_u01(rn[1]), &sf, &cf);
            double s2 = fma(-mu, mu, 1.0);
            if (s2 < 0.0) s2 = 0.0;
            const double st = sqrt(s2);
            r.d[0] = st * cf;
            r.d[1] = st * sf;
            r.d[2] = mu;
            e1[0] = mu * cf;
            e1[1] = mu * sf;
            e1[2] = -st;
        }
        double tau_rem = -kp_log(kp_u01(rn[2]));
        double I = w0, Q = 0.0, U = 0.0;
        memcpy(r.x, c->src->x, sizeof(r.x));
        r.ib = c->src_ib;
        memcpy(r.i, c->src_i, sizeof(r.i));
        gl_load(m, &r);
        double B = 0.0;
        /* event 0: emission peel-offs */
        for (int o = 0; o < no_; ++o) {
            if (pass == 1) {
                const double tau = tau_to_edge(c, &r, c->no[o], ks, &dummy);
                const double att = kp_exp(-tau) * KP_INV_4PI;
                const long px = pix_index(c, band, o, r.x);
                double Ak = 0.0;
                if (px >= 0)
                    Ak = t->w[3 * px] * (I * att);
                t->A[o] = Ak;
                A_tot += Ak;
            } else {
                peel_grad(c, &r, c->no[o], ks, t->A[o], t->d_logrho);
            }
        }
        if (pass == 2) {
            B = A_tot;
            for (int o = 0; o < no_; ++o)
                B -= t->A[o];
        }
        int steps = run->step_max;
        int events = 0;
        for (;;) {
            int interact = 0;
            size_t gc = 0;
            for (;;) {
                if (steps <= 0)
                    break;
                double l;
                const int ap = get_dl(&r, &l);
                if (ap < 0)
                    break;
                gc = (size_t)r.ib * n3 + cell_index(m, &r);
                const double kap = (double)c->kappa[gc] * ks;
                const double dtau = kap * l;
                steps -= 1;
                if (kap > 0.0 && dtau >= tau_rem) {
                    double li = tau_rem / kap;
                    if (li > l)
                        li = l;
                    if (pass == 2)
                        t->d_logrho[gc] -= (kap * li) * B;
                    for (int a = 0; a < 3; ++a)
                        r.x[a] = r.x[a] + li * r.d[a];
                    interact = 1;
                    break;
                }
                tau_rem = tau_rem - dtau;
                if (pass == 2)
                    t->d_logrho[gc] -= dtau * B;
                int reload;
                const int st = cross_face(m, &r, ap, l, &reload);
                if (st != ST_IN)
                    break;
                if (reload)
                    gl_load(m, &r);
            }
            if (!interact)
                break;
            events += 1;
            uint32_t u[4] = {plo, phi, (uint32_t)events, 0u};
            kp_philox4x32_10(u, k0, k1);
            const double alb = calb ? (double)c->albedo[gc] : bd->albedo;
            const double gg = (bd->flags & KP_BAND_CELL_G) ? (c->g ? (double)c->g[gc] : 0.0) : bd->g;
            I = I * alb;
            Q = Q * alb;
            U = U * alb;
            if (pass == 2) {
                t->d_logrho[gc] += B;
                if (calb && t->d_alb)
                    t->d_alb[gc] += B / alb;
            }
            double e2[3];
            cross3(r.d, e1, e2);
            if ((size_t)(events + 1) * no_ > t->A_cap) {
                t->A_cap = 2 * (size_t)(events + 1) * no_ + 16;
                t->A = (double *)realloc(t->A, t->A_cap * sizeof(double));
            }
            for (int o = 0; o < no_; ++o) {
                double S[3] = {I, Q, U}, Si[3];
                peel_stokes(run->scatter_model, gg, run->p_max, r.d, e1, e2, S, c->no[o],
                            c->ev[o], Si);
                double *Ak = &t->A[(size_t)events * no_ + o];
                if (pass == 1) {
                    const double tau = tau_to_edge(c, &r, c->no[o], ks, &dummy);
                    const double att = kp_exp(-tau) * KP_INV_4PI;
                    const long px = pix_index(c, band, o, r.x);
                    double a = 0.0;
                    if (px >= 0)
                        a = t->w[3 * px] * (Si[0] * att) + t->w[3 * px + 1] * (Si[1] * att) +
                            t->w[3 * px + 2] * (Si[2] * att);
                    *Ak = a;
                    A_tot += a;
                } else {
                    peel_grad(c, &r, c->no[o], ks, *Ak, t->d_logrho);
                    if (hg && t->d_g)
                        t->d_g[gc] += *Ak * dlogp_hg(gg, fmin(fmax(dot3(r.d, c->no[o]), -1.0), 1.0));
                    B -= *Ak;
                }
            }
            const double mu = (run->scatter_model == KP_THOMSON) ? kp_thomson_mu(kp_u01(u[0]))
                                                                  : kp_hg_mu(gg, kp_u01(u[0]));
            if (pass == 2 && hg && t->d_g)
                t->d_g[gc] += B * dlogp_hg(gg, mu);
            double S[3] = {I, Q, U};
            scatter_stokes(run->scatter_model, run->p_max, mu, kp_u01(u[1]), r.d, e1, e2, S);
            I = S[0];
            Q = S[1];
            U = S[2];
            tau_rem = -kp_log(kp_u01(u[2]));
            if (I < run->rr_weight * w0) {
                if (kp_u01(u[3]) < run->rr_survive) {
                    I = I / run->rr_survive;
                    Q = Q / run->rr_survive;
                    U = U / run->rr_survive;
                } else {
                    break;
                }
            }
        }
    }
    t->jlin += A_tot;
}

typedef struct {
    const fctx_t *c;
    uint64_t k_lo, k_hi;
    athr_t a;
} awork_t;

static void *aworker(void *arg)
{
    awork_t *w = (awork_t *)arg;
    const fctx_t *c = w->c;
    int band = 0;
    for (uint64_t k = w->k_lo; k < w->k_hi; ++k) {
        while (k >= c->band_start[band + 1])
            ++band;
        const uint64_t pid = c->run->packet_offset + (k - c->band_start[band]);
        adjoint_packet(c, &w->a, band, pid);
    }
    return NULL;
}

int or_adjoint(const kp_mesh_desc *m, const float *kappa, const float *albedo, const float *g,
               const kp_source *src, const kp_run *run, const double *dL_dimage,
               double *d_logrho, double *d_albedo, double *d_g, double *jlin, int n_threads)
{
    if (!m || !kappa || !src || !run || !dL_dimage || !d_logrho || run->n_bands < 1 ||
        run->n_bands > 1023 || run->n_incl < 1 || run->n_incl > 8)
        return -1;
    fctx_t *c = (fctx_t *)calloc(1, sizeof(fctx_t));
    c->m = m;
    c->kappa = kappa;
    c->albedo = albedo;
    c->g = g;
    c->src = src;
    c->run = run;
    if (or_locate(m, src->x, &c->src_ib, c->src_i) != 0) {
        free(c);
        return -2;
    }
    c->n_incl = run->n_incl;
    for (int o = 0; o < run->n_incl; ++o) {
        double s, co;
        kp_sincos2pi(run->incl_deg[o] / 360.0, &s, &co);
        c->no[o][0] = s;  c->no[o][1] = 0.0; c->no[o][2] = co;
        c->ev[o][0] = -co; c->ev[o][1] = 0.0; c->ev[o][2] = s;
        c->eh[o][0] = 0.0; c->eh[o][1] = 1.0; c->eh[o][2] = 0.0;
    }
    for (int a = 0; a < 3; ++a)
        c->center[a] = (m->x_min[a] + m->x_max[a]) * 0.5;
    c->pix_scale = (double)run->image_n / (2.0 * run->image_half);
    c->band_start[0] = 0;
    for (int b = 0; b < run->n_bands; ++b)
        c->band_start[b + 1] = c->band_start[b] + run->bands[b].n_packets;
    c->total = c->band_start[run->n_bands];
    for (int b = run->n_bands + 1; b < 1024; ++b)
        c->band_start[b] = c->total;
    if (n_threads < 1)
        n_threads = 1;
    if ((uint64_t)n_threads > c->total && c->total > 0)
        n_threads = (int)c->total;
    const size_t ncell = (size_t)m->n_blocks * m->n_cell * m->n_cell * m->n_cell;
    awork_t *w = (awork_t *)calloc((size_t)n_threads, sizeof(awork_t));
    pthread_t *tid = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    for (int k = 0; k < n_threads; ++k) {
        w[k].c = c;
        w[k].k_lo = c->total * k / n_threads;
        w[k].k_hi = c->total * (k + 1) / n_threads;
        w[k].a.w = dL_dimage;
        w[k].a.d_logrho = k == 0 ? d_logrho : (double *)calloc(ncell, sizeof(double));
        w[k].a.d_alb = d_albedo ? (k == 0 ? d_albedo : (double *)calloc(ncell, sizeof(double))) : NULL;
        w[k].a.d_g = d_g ? (k == 0 ? d_g : (double *)calloc(ncell, sizeof(double))) : NULL;
        w[k].a.A_cap = 64;
        w[k].a.A = (double *)malloc(64 * sizeof(double));
    }
    for (int k = 1; k < n_threads; ++k)
        pthread_create(&tid[k], NULL, aworker, &w[k]);
    aworker(&w[0]);
    for (int k = 1; k < n_threads; ++k)
        pthread_join(tid[k], NULL);
    double jl = 0.0;
    for (int k = 0; k < n_threads; ++k) {
        jl += w[k].a.jlin;
        if (k > 0) {
            for (size_t q = 0; q < ncell; ++q) {
                d_logrho[q] += w[k].a.d_logrho[q];
                if (d_albedo)
                    d_albedo[q] += w[k].a.d_alb[q];
                if (d_g)
                    d_g[q] += w[k].a.d_g[q];
            }
            free(w[k].a.d_logrho);
            if (d_albedo)
                free(w[k].a.d_alb);
            if (d_g)
                free(w[k].a.d_g);
        }
        free(w[k].a.A);
    }
    if (jlin)
        *jlin += jl;
    free(w);
    free(tid);
    free(c);
    return 0;
}
