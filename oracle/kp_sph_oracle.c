/*
 * kp_sph_oracle.c -- FP64 CPU restatement of the reference's spherical
 * photon march.
 *
 * TEST INFRASTRUCTURE ONLY (see kp_oracle.h).  Built with -ffp-contract=off.
 *
 * Follows (paths relative to the reference src/):
 *   sph_t::get_dl ............ modules/particle/radiation/photon_sph.h:42-180
 *                              (sphere and cone faces by the quadratic root of
 *                              root(), 24-40; the phi face as a plane through
 *                              the polar axis; argmin over strictly positive
 *                              distances, ties to the lower axis)
 *   sph_t::proc_geo .......... photon_sph.h:222-243 (move by 1.001 times the
 *                              distance, step the exit axis by its direction,
 *                              the other axes by position against the old cell)
 *   cart_t::proc_phys ........ photon.h:73-83 (tally proper*dl/V, V from
 *                              sph_t::volume, mesh/meshgen/spherical.h:31-42)
 *   cart_t::proc / proc_step . photon.h:104-125, photon_sph.h:245-263
 *   proxy_base_t::shift_blk .. types/proxy.h:64-121
 *
 * Arithmetic form.  Positions and directions are Cartesian; the face
 * trigonometry comes from the mesh's per-block table (kp_mesh_desc::sph).
 * Where the reference converts the position to (r, theta, phi) with
 * sqrt/acos/atan2 and compares with face coordinates, this file compares in
 * Cartesian form, equal up to rounding:
 *   r >= r_f      <=>  x.x >= r_f*r_f
 *   theta >= th_f <=>  z <= cos(th_f)*|x|
 *   phi >= ph_f   <=>  y*cos(ph_f) - x*sin(ph_f) >= 0   (cells narrower than pi)
 * and the signs of dtheta/dl and dphi/dl are those of
 * z(x dx + y dy) - (x^2+y^2) dz and x dy - y dx (the reference's
 * theta_hat.dir and df_dl).  The periodic phi wrap of regulate() has no
 * counterpart: Cartesian positions need none.
 *
 * Deviations from the reference:
 *  (S1) theta comparisons use the exact polar angle.  The reference's to_sph
 *       (photon_sph.h:193-208) sets theta = 1.570796 - z/r for every
 *       z/r < 0.01, which is acos(z/r) only near the equator and is off by
 *       up to 0.43 rad in the southern hemisphere.
 *  (S2) every axis whose index left the block is resolved in the same step,
 *       in axis order (the reference resolves the first and keeps the other
 *       index out of range for one step, whose tally then lands in another
 *       cell of the array); tangential halving uses floor division.
 *  (S3) after a coarse -> fine crossing the tangential fine cell is the one
 *       containing the position (the reference takes the lower one).
 *  (S4) a step with no strictly positive face distance ends the packet as
 *       lost (the reference would move along axis 0 by a non-positive
 *       distance).
 *  (S5) a theta face at pi/2 is the plane z = 0 (its table cosine is 0), and
 *       its intersection is taken as the double root it is: discriminant 0.
 *       In the reference b^2 - 4ac of that face is rounding noise, so about
 *       half of the equator crossings are missed and the step is tallied in
 *       the wrong hemisphere's cell.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../kratos_polrad_amd/csrc/kp_math.h"
#include "kp_oracle.h"

enum { S_IN = 0, S_ESCAPED = 1, S_LOST = 4 };

typedef struct {
    double x[3], d[3];
    int ib, i[3];
} sray_t;

static const double *tab(const kp_mesh_desc *m, int ib)
{
    return m->sph + (size_t)ib * KP_SPH_TAB(m->n_cell);
}

static int sgn(double v) { return v > 0.0 ? 1 : (v < 0.0 ? -1 : 0); }

/* sph_t::root (photon_sph.h:24-40); `plane`: the equatorial face z = 0,
 * whose intersection is a double root (discriminant exactly 0) */
static int sroot(double a, double b, double c, int plane, double *x1, double *x2)
{
    const double det = plane ? 0.0 : b * b - 4.0 * a * c;
    if (det < 0.0)
        return 0;
    const double nom = copysign(sqrt(det), b) + b;
    const double x0 = nom / (-2.0 * a);
    const double y0 = (-2.0 * c) / nom;
    *x1 = x0 < y0 ? x0 : y0;
    *x2 = x0 > y0 ? x0 : y0;
    return 1;
}

/* "position at or beyond face k of axis q" in block table T */
static int beyond(const double *T, int n, int q, int k, const double x[3])
{
    const int m = n + 1;
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    if (q == 0)
        return r2 >= T[k] * T[k];
    if (q == 1)
        return x[2] <= T[m + k] * sqrt(r2);
    return x[1] * T[3 * m + k] - x[0] * T[2 * m + k] >= 0.0;
}

/* sph_t::get_dl: returns the exit axis (or -1), its distance and the
 * per-axis index directions after the backward-face fallbacks. */
static int sph_get_dl(const kp_mesh_desc *m, const sray_t *r, double *lmin, int s[3])
{
    const int n = m->n_cell, mm = n + 1;
    const double *T = tab(m, r->ib);
    const double rl = T[r->i[0]], rr = T[r->i[0] + 1];
    const double cl = T[mm + r->i[1]], cr = T[mm + r->i[1] + 1];
    const double sl = T[2 * mm + r->i[2]], sr = T[2 * mm + r->i[2] + 1];
    const double ql = T[3 * mm + r->i[2]], qr = T[3 * mm + r->i[2] + 1];
    const double *x = r->x, *d = r->d;
    const double xd = x[0] * d[0] + x[1] * d[1] + x[2] * d[2];
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const double rho2 = x[0] * x[0] + x[1] * x[1];
    const double dt = x[2] * (x[0] * d[0] + x[1] * d[1]) - rho2 * d[2];
    const double df = x[0] * d[1] - x[1] * d[0];
    s[0] = sgn(xd);
    s[1] = sgn(dt);
    s[2] = sgn(df);
    double dl[3], x1, x2;

    /* r: the sphere ahead; a ray that misses the inner sphere leaves
     * through the outer one, from inside it */
    {
        const double b = 2.0 * xd;
        const double f = s[0] > 0 ? rr : rl;
        if (sroot(1.0, b, r2 - f * f, 0, &x1, &x2)) {
            dl[0] = x1 > 0.0 ? x1 : x2;
        } else {
            const double k = s[0] < 0 ? rr : rl;
            dl[0] = sroot(1.0, b, r2 - k * k, 0, &x1, &x2) ? x2 : -1.0;
            s[0] = -s[0];
        }
    }
    /* theta: the cone z^2 = cos^2(th_f) |x|^2 */
    {
        const double z = x[2], dz = d[2];
        const double cf = s[1] > 0 ? cr : cl;
        double mu2 = cf * cf;
        double a = dz * dz - mu2;
        double b = 2.0 * (z * dz - mu2 * xd);
        double c = z * z - mu2 * r2;
        if (a != 0.0) {
            if (sroot(a, b, c, mu2 == 0.0, &x1, &x2)) {
                dl[1] = x1 > 0.0 ? x1 : x2;
            } else {
                const double cb = s[1] < 0 ? cr : cl;
                mu2 = cb * cb;
                a = dz * dz - mu2;
                b = 2.0 * (z * dz - mu2 * xd);
                c = z * z - mu2 * r2;
                if (a != 0.0)
                    dl[1] = sroot(a, b, c, mu2 == 0.0, &x1, &x2) ? x2 : -1.0;
                else
                    dl[1] = -c / b;
                s[1] = -s[1];
            }
        } else {
            dl[1] = -c / b;
        }
    }
    /* phi: the plane through the polar axis at ph_f */
    {
        const double sf = s[2] > 0 ? sr : sl, qf = s[2] > 0 ? qr : ql;
        double a = d[0] * sf - d[1] * qf;
        double b = x[1] * qf - x[0] * sf;
        dl[2] = a != 0.0 ? b / a : -1.0;
        if (dl[2] < 0.0) {
            const double sb = s[2] < 0 ? sr : sl, qb = s[2] < 0 ? qr : ql;
            a = d[0] * sb - d[1] * qb;
            b = x[1] * qb - x[0] * sb;
            dl[2] = a != 0.0 ? b / a : -1.0;
            s[2] = -s[2];
        }
    }
    int ap = -1;
    double best = DBL_MAX;
    for (int a = 0; a < 3; ++a)
        if (dl[a] > 0.0 && dl[a] < best) {
            ap = a;
            best = dl[a];
        }
    *lmin = best;
    return ap;
}

/* proc_geo + shift_blk of one step of length l through axis ap */
static int sph_cross(const kp_mesh_desc *m, sray_t *r, int ap, double l, const int s[3])
{
    const int n = m->n_cell, mm = n + 1;
    const double *T = tab(m, r->ib);
    const double rl = T[r->i[0]], rr = T[r->i[0] + 1];
    const double cl = T[mm + r->i[1]], cr = T[mm + r->i[1] + 1];
    const double sl = T[2 * mm + r->i[2]], sr = T[2 * mm + r->i[2] + 1];
    const double ql = T[3 * mm + r->i[2]], qr = T[3 * mm + r->i[2] + 1];
    for (int a = 0; a < 3; ++a)
        r->x[a] = r->x[a] + (l * r->d[a]) * 1.001;
    r->i[ap] += s[ap];
    const double *x = r->x;
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    if (ap != 0) {
        if (r2 >= rr * rr)
            r->i[0] += 1;
        else if (r2 <= rl * rl)
            r->i[0] -= 1;
    }
    if (ap != 1) {
        const double rn = sqrt(r2);
        if (x[2] <= cr * rn)
            r->i[1] += 1;
        else if (x[2] >= cl * rn)
            r->i[1] -= 1;
    }
    if (ap != 2) {
        if (x[1] * qr - x[0] * sr >= 0.0)
            r->i[2] += 1;
        else if (x[1] * ql - x[0] * sl <= 0.0)
            r->i[2] -= 1;
    }
    for (int a = 0; a < 3; ++a) {
        if (r->i[a] >= 0 && r->i[a] < n)
            continue;
        const int side = r->i[a] > 0;
        if (m->phybnd[r->ib] & (1u << (2 * a + side)))
            return S_ESCAPED;
        const int apj[2] = {(a + 1) % 3, (a + 2) % 3};
        const int dlv = m->d_lvl[6 * r->ib + 2 * a + side];
        int j0 = 0, j1 = 0;
        if (dlv > 0) {
            j0 = r->i[apj[0]] >= n / 2;
            j1 = r->i[apj[1]] >= n / 2;
        }
        const int slot = 24 * r->ib + ((a * 2 + side) * 2 + j0) * 2 + j1;
        const int nb = m->nb_id[slot];
        if (nb < 0)
            return S_LOST;
        const int shift = m->nb_shift[slot];
        r->ib = nb;
        r->i[a] = side ? 0 : n - 1;
        for (int j = 0; j <= 1; ++j) {
            const int q = apj[j];
            const int is = r->i[q];
            if (dlv < 0) {
                const int h = is >= 0 ? is / 2 : -((1 - is) / 2);
                r->i[q] = h + ((shift >> j) & 1) * (n / 2);
            } else if (dlv > 0) {
                int f = (is < n / 2 ? is : is - n / 2) * 2;
                if (is >= 0 && is < n && beyond(tab(m, nb), n, q, f + 1, r->x))
                    f += 1;
                r->i[q] = f;
            }
        }
    }
    return S_IN;
}

int or_sph_locate(const kp_mesh_desc *m, const double x[3], int32_t *ib, int32_t i[3])
{
    const int n = m->n_cell;
    *ib = -1;
    if (!m->sph)
        return -1;
    for (int b = 0; b < m->n_blocks; ++b) {
        const double *T = tab(m, b);
        int c[3], in = 1;
        for (int q = 0; q < 3 && in; ++q) {
            c[q] = -1;
            for (int k = 0; k < n; ++k)
                if (beyond(T, n, q, k, x) && !beyond(T, n, q, k + 1, x)) {
                    c[q] = k;
                    break;
                }
            in = c[q] >= 0;
        }
        if (!in)
            continue;
        *ib = b;
        for (int q = 0; q < 3; ++q)
            i[q] = c[q];
        return 0;
    }
    return -1;
}

static double cell_volume(const kp_mesh_desc *m, const sray_t *r)
{
    const int n = m->n_cell, mm = n + 1;
    const double *T = tab(m, r->ib);
    return (T[4 * mm + r->i[0]] * T[4 * mm + n + r->i[1]]) * T[4 * mm + 2 * n + r->i[2]];
}

int or_sph_trace(const kp_mesh_desc *m, const double x[3], const double d[3], int max_steps,
                 int32_t *out_ib, int32_t *out_cell, double *out_len, double x_end[3],
                 int *status)
{
    sray_t r;
    memcpy(r.x, x, sizeof(r.x));
    memcpy(r.d, d, sizeof(r.d));
    *status = S_LOST;
    if (or_sph_locate(m, x, &r.ib, r.i) != 0)
        return 0;
    const int n = m->n_cell;
    int k = 0;
    *status = 0;
    while (k < max_steps) {
        double l;
        int s[3];
        const int ap = sph_get_dl(m, &r, &l, s);
        if (ap < 0) {
            *status = S_LOST;
            break;
        }
        out_ib[k] = r.ib;
        out_cell[k] = r.i[0] + n * (r.i[1] + n * r.i[2]);
        out_len[k] = l;
        ++k;
        const int st = sph_cross(m, &r, ap, l, s);
        if (st != S_IN) {
            *status = st;
            break;
        }
    }
    memcpy(x_end, r.x, sizeof(r.x));
    return k;
}

/* ------------------------------------------------------------------ */
/* Streaming forward run                                               */

typedef struct {
    const kp_mesh_desc *m;
    const kp_source *src;
    const kp_run *run;
    int32_t src_ib, src_i[3];
    uint64_t band_start[1025];
    uint64_t total;
} sctx_t;

typedef struct {
    const sctx_t *c;
    uint64_t k_lo, k_hi;
    double *J, *end_state;
    uint64_t cnt[KP_NCOUNTERS];
} sthr_t;

static void sph_packet(const sctx_t *c, sthr_t *t, int band, uint64_t pid, uint64_t slot)
{
    const kp_mesh_desc *m = c->m;
    const int n = m->n_cell;
    const size_t n3 = (size_t)n * n * n;
    const int nbands = c->run->n_bands;
    uint32_t rn[4] = {(uint32_t)pid, (uint32_t)(pid >> 32), 0u, 0u};
    kp_philox4x32_10(rn, c->run->seed, (uint32_t)band);
    sray_t r;
    {
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
    }
    for (int a = 0; a < 3; ++a) {
        r.x[a] = c->src->x[a];
        r.i[a] = c->src_i[a];
    }
    r.ib = c->src_ib;
    const double I = c->src->weight;
    int steps = c->run->step_max, status = 0;
    uint64_t ncross = 0;
    for (;;) {
        if (steps <= 0) {
            status = KP_END_STEPS;
            break;
        }
        double l;
        int s[3];
        const int ap = sph_get_dl(m, &r, &l, s);
        if (ap < 0) {
            status = KP_END_LOST;
            break;
        }
        const size_t gc = (size_t)r.ib * n3 + r.i[0] + (size_t)n * (r.i[1] + (size_t)n * r.i[2]);
        if (t->J)
            t->J[gc * nbands + band] += I * (l / cell_volume(m, &r));
        ncross += 1;
        steps -= 1;
        const int st = sph_cross(m, &r, ap, l, s);
        if (st == S_ESCAPED) {
            status = KP_END_ESCAPED;
            break;
        }
        if (st == S_LOST) {
            status = KP_END_LOST;
            break;
        }
    }
    t->cnt[KP_C_CROSS_MAIN] += ncross;
    t->cnt[KP_C_PACKETS] += 1;
    if (status == KP_END_ESCAPED)
        t->cnt[KP_C_ESCAPED] += 1;
    else if (status == KP_END_STEPS)
        t->cnt[KP_C_KILLED_STEPS] += 1;
    else
        t->cnt[KP_C_LOST] += 1;
    if (t->end_state) {
        double *es = t->end_state + slot * KP_ES_STRIDE;
        memset(es, 0, KP_ES_STRIDE * sizeof(double));
        for (int a = 0; a < 3; ++a) {
            es[KP_ES_X + a] = r.x[a];
            es[KP_ES_D + a] = r.d[a];
        }
        es[KP_ES_I] = I;
        es[KP_ES_CROSS] = (double)ncross;
        es[KP_ES_STATUS] = (double)status;
        es[KP_ES_IB] = (double)r.ib;
        es[KP_ES_CELL] = (double)(r.i[0] + n * (r.i[1] + n * r.i[2]));
    }
}

static void *sph_worker(void *arg)
{
    sthr_t *t = (sthr_t *)arg;
    const sctx_t *c = t->c;
    int band = 0;
    for (uint64_t k = t->k_lo; k < t->k_hi; ++k) {
        while (band + 1 < c->run->n_bands && k >= c->band_start[band + 1])
            ++band;
        const uint64_t pid = c->run->packet_offset + (k - c->band_start[band]);
        sph_packet(c, t, band, pid, k);
    }
    return NULL;
}

int or_sph_forward(const kp_mesh_desc *m, const kp_source *src, const kp_run *run, double *J,
                   uint64_t *counters, double *end_state, int n_threads)
{
    if (!m || !m->sph || !src || !run || run->n_bands < 1 || run->n_bands > 1024)
        return -1;
    sctx_t *c = (sctx_t *)calloc(1, sizeof(sctx_t));
    c->m = m;
    c->src = src;
    c->run = run;
    if (or_sph_locate(m, src->x, &c->src_ib, c->src_i) != 0) {
        free(c);
        return -2;
    }
    c->band_start[0] = 0;
    for (int b = 0; b < run->n_bands; ++b)
        c->band_start[b + 1] = c->band_start[b] + run->bands[b].n_packets;
    c->total = c->band_start[run->n_bands];
    if (n_threads < 1)
        n_threads = 1;
    if ((uint64_t)n_threads > c->total && c->total > 0)
        n_threads = (int)c->total;
    const size_t nJ = (size_t)m->n_blocks * m->n_cell * m->n_cell * m->n_cell * run->n_bands;
    sthr_t *th = (sthr_t *)calloc((size_t)n_threads, sizeof(sthr_t));
    pthread_t *tid = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
    int rc = 0;
    for (int k = 0; k < n_threads; ++k) {
        th[k].c = c;
        th[k].k_lo = c->total * k / n_threads;
        th[k].k_hi = c->total * (k + 1) / n_threads;
        th[k].end_state = end_state;
        if (J) {
            th[k].J = (k == 0) ? J : (double *)calloc(nJ, sizeof(double));
            if (!th[k].J)
                rc = -3;
        }
    }
    if (rc == 0) {
        for (int k = 1; k < n_threads; ++k)
            pthread_create(&tid[k], NULL, sph_worker, &th[k]);
        sph_worker(&th[0]);
        for (int k = 1; k < n_threads; ++k)
            pthread_join(tid[k], NULL);
        if (J)
            for (int k = 1; k < n_threads; ++k)
                for (size_t q = 0; q < nJ; ++q)
                    J[q] += th[k].J[q];
        if (counters)
            for (int k = 0; k < n_threads; ++k)
                for (int q = 0; q < KP_NCOUNTERS; ++q)
                    counters[q] += th[k].cnt[q];
    }
    for (int k = 1; k < n_threads; ++k)
        if (J)
            free(th[k].J);
    free(th);
    free(tid);
    free(c);
    return rc;
}
