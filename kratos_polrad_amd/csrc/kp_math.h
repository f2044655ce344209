/*
 * kp_math.h -- deterministic leaf arithmetic shared by host and device.
 *
 * Everything here is written so that gcc (host, -ffp-contract=off) and hipcc
 * (gfx950, -ffp-contract=off) execute the same IEEE-754 double operations in
 * the same order: every fused multiply-add is an explicit fma(), every other
 * operation rounds on its own.  Division and square root are the correctly
 * rounded IEEE operations on both targets.  This is what lets the FP64
 * transport kernel reproduce a CPU run bit for bit on the same packet ids.
 *
 *  - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror & Shaw,
 *    "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Multipliers and
 *    Weyl constants are the published ones.
 *  - kp_log / kp_exp: Cody & Waite argument reduction with the fdlibm split
 *    of ln 2, then an atanh series (log) or a Taylor series (exp).
 *  - kp_sincos2pi: exact quarter-turn reduction of a turn fraction, then
 *    Taylor series on |x| <= pi/4.
 *  - kp_thomson_mu: Newton iteration on the cubic CDF of (1+mu^2).
 *  - kp_hg_mu: inversion of the Henyey & Greenstein (1941) CDF.
 *
 * The file is valid C99 and HIP C++.
 */
#ifndef KP_MATH_H
#define KP_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define KP_HD __host__ __device__ static inline
#else
#include <math.h>
#define KP_HD static inline
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                       */

#define KP_PHILOX_M0 0xD2511F53u
#define KP_PHILOX_M1 0xCD9E8D57u
#define KP_PHILOX_W0 0x9E3779B9u
#define KP_PHILOX_W1 0xBB67AE85u

/* In place: c[0..3] is the counter on entry and the output on exit. */
KP_HD void kp_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)KP_PHILOX_M0 * (uint64_t)c[0];
        const uint64_t p1 = (uint64_t)KP_PHILOX_M1 * (uint64_t)c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += KP_PHILOX_W0;
        k1 += KP_PHILOX_W1;
    }
}

/* Uniform on the open interval (0,1) from 32 random bits, exact in double. */
KP_HD double kp_u01(uint32_t x)
{
    return ((double)x + 0.5) * 2.3283064365386962890625e-10; /* 2^-32 */
}

/* ------------------------------------------------------------------ */
/* Bit casts                                                           */

KP_HD uint64_t kp_dbits(double x)
{
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}
KP_HD double kp_dfrom(uint64_t u)
{
    union { double d; uint64_t u; } v;
    v.u = u;
    return v.d;
}

/* fdlibm's split of ln 2: LN2_HI has 32 trailing zero bits so k*LN2_HI is
 * exact for |k| < 2^11. */
#define KP_LN2_HI 6.93147180369123816490e-01
#define KP_LN2_LO 1.90821492927058770002e-10
#define KP_INV_LN2 1.44269504088896338700e+00
#define KP_TWO_PI 6.28318530717958647692e+00
#define KP_INV_4PI 7.95774715459476678844e-02

/* Natural log for positive normal x. */
KP_HD double kp_log(double x)
{
    const uint64_t b = kp_dbits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = kp_dfrom((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double f = m - 1.0; /* exact for m in [0.70, 1.42] */
    const double s = f / (m + 1.0);
    const double z = s * s;
    /* log(m) = 2 atanh(s) = 2s + s z sum_{k>=1} 2/(2k+1) z^(k-1) */
    double q = 2.0 / 25.0;
    q = fma(q, z, 2.0 / 23.0);
    q = fma(q, z, 2.0 / 21.0);
    q = fma(q, z, 2.0 / 19.0);
    q = fma(q, z, 2.0 / 17.0);
    q = fma(q, z, 2.0 / 15.0);
    q = fma(q, z, 2.0 / 13.0);
    q = fma(q, z, 2.0 / 11.0);
    q = fma(q, z, 2.0 / 9.0);
    q = fma(q, z, 2.0 / 7.0);
    q = fma(q, z, 2.0 / 5.0);
    q = fma(q, z, 2.0 / 3.0);
    const double r = fma(s * z, q, 2.0 * s);
    const double ed = (double)e;
    return fma(ed, KP_LN2_HI, fma(ed, KP_LN2_LO, r));
}

/* exp(x) for x <= ~700; returns 0 below -708 (attenuation use only). */
KP_HD double kp_exp(double x)
{
    if (x < -708.0)
        return 0.0;
    const double kd = floor(fma(x, KP_INV_LN2, 0.5));
    double r = fma(-kd, KP_LN2_HI, x);
    r = fma(-kd, KP_LN2_LO, r);
    double p = 1.0 / 6227020800.0; /* 1/13! */
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    const int64_t k = (int64_t)kd;
    return p * kp_dfrom((uint64_t)(k + 1023) << 52);
}

/* sin(2 pi u), cos(2 pi u) for a turn fraction u (|u| < 2^20). */
KP_HD void kp_sincos2pi(double u, double *s_out, double *c_out)
{
    const double q = floor(fma(u, 4.0, 0.5));
    const double r = u - q * 0.25; /* exact, |r| <= 1/8 */
    const double x = r * KP_TWO_PI;
    const double z = x * x;
    double sp = 1.0 / 355687428096000.0;  /* +1/17! */
    sp = fma(sp, z, -1.0 / 1307674368000.0); /* -1/15! */
    sp = fma(sp, z, 1.0 / 6227020800.0);     /* +1/13! */
    sp = fma(sp, z, -1.0 / 39916800.0);      /* -1/11! */
    sp = fma(sp, z, 1.0 / 362880.0);         /* +1/9!  */
    sp = fma(sp, z, -1.0 / 5040.0);          /* -1/7!  */
    sp = fma(sp, z, 1.0 / 120.0);            /* +1/5!  */
    sp = fma(sp, z, -1.0 / 6.0);             /* -1/3!  */
    const double sn = fma(x * z, sp, x);
    double cp = -1.0 / 6402373705728000.0;   /* -1/18! */
    cp = fma(cp, z, 1.0 / 20922789888000.0);  /* +1/16! */
    cp = fma(cp, z, -1.0 / 87178291200.0);    /* -1/14! */
    cp = fma(cp, z, 1.0 / 479001600.0);       /* +1/12! */
    cp = fma(cp, z, -1.0 / 3628800.0);        /* -1/10! */
    cp = fma(cp, z, 1.0 / 40320.0);           /* +1/8!  */
    cp = fma(cp, z, -1.0 / 720.0);            /* -1/6!  */
    cp = fma(cp, z, 1.0 / 24.0);              /* +1/4!  */
    cp = fma(cp, z, -0.5);                    /* -1/2!  */
    const double cs = fma(z, cp, 1.0);
    const int qi = ((int)q) & 3;
    double s, c;
    if (qi == 0) { s = sn;  c = cs;  }
    else if (qi == 1) { s = cs;  c = -sn; }
    else if (qi == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
    *s_out = s;
    *c_out = c;
}

/* Scattering cosine for Thomson (Rayleigh) scattering, density (3/8)(1+mu^2).
 * CDF = xi  <=>  mu^3 + 3 mu + 4 - 8 xi = 0; six Newton steps from mu = c/4. */
KP_HD double kp_thomson_mu(double xi)
{
    const double c = fma(8.0, xi, -4.0);
    double mu = 0.25 * c;
    for (int it = 0; it < 6; ++it) {
        const double mu2 = mu * mu;
        const double f = fma(mu2, mu, fma(3.0, mu, -c));
        const double fp = fma(3.0, mu2, 3.0);
        mu = mu - f / fp;
    }
    if (mu > 1.0) mu = 1.0;
    if (mu < -1.0) mu = -1.0;
    return mu;
}

/* Scattering cosine for the Henyey-Greenstein phase function. */
KP_HD double kp_hg_mu(double g, double xi)
{
    double mu;
    if (g == 0.0) {
        mu = fma(2.0, xi, -1.0);
    } else {
        const double g2 = g * g;
        const double t = (1.0 - g2) / fma(2.0 * g, xi, 1.0 - g);
        mu = (1.0 + g2 - t * t) / (2.0 * g);
    }
    if (mu > 1.0) mu = 1.0;
    if (mu < -1.0) mu = -1.0;
    return mu;
}

#endif /* KP_MATH_H */
