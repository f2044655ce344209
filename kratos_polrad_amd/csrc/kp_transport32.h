// kp_transport32.h -- mixed-precision transport: FP32 block-local march,
// FP64 optical depth, FP64 tallies.
//
// Positions are kept in cell units of the current block, so every face sits
// at an integer and every block change (same level, 2:1 coarser, 2:1 finer)
// is an exact shift or power-of-two scale in FP32.  Face distances use a
// per-ray "length per cell" reciprocal instead of three divisions per step.
// Optical depth is accumulated in FP64 (tiny per-cell increments at deep
// refinement levels do not vanish), tallies and images are FP64 atomics, and
// consecutive contributions of one packet to one cell are summed with Kahan
// compensation in registers before the FP64 atomic.
// Transcendentals are the gfx950 hardware ops (v_log_f32, v_exp_f32,
// v_sin/cos_f32 in revolutions, v_rsq/rcp_f32).
#pragma once

#include "kp_device.h"

namespace kp {
namespace f32 {

constexpr float kInf = __builtin_huge_valf();
constexpr float kLn2 = 0.693147180559945309f;
constexpr float kLog2e = 1.442695040888963407f;
constexpr float kInv4Pi = 0.0795774715459476679f;

__device__ __forceinline__ float u01(uint32_t x)
{
    return ((float)(x >> 9) + 0.5f) * 1.1920928955078125e-07f; // exact, (0,1)
}
__device__ __forceinline__ float fast_log(float x) { return __builtin_amdgcn_logf(x) * kLn2; }
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * kLog2e); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float dot3(const float *a, const float *b)
{
    return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}
__device__ __forceinline__ void cross3(const float *a, const float *b, float *c)
{
    c[0] = fmaf(a[1], b[2], -a[2] * b[1]);
    c[1] = fmaf(a[2], b[0], -a[0] * b[2]);
    c[2] = fmaf(a[0], b[1], -a[1] * b[0]);
}

struct ray_t {
    float x[3];   // position in cell units of block ib, in [0, n]
    float d[3];   // unit direction
    float lpc[3]; // physical length per cell unit along axis a (h/d)
    float cpl[3]; // cell units per physical length along a (d/h)
    int i[3];
    int ib;
    float iv;     // 1 / cell volume of block ib
};

__device__ __forceinline__ void set_scales(const kparams &P, ray_t &r)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float h = P.h32[3 * r.ib + a];
        r.cpl[a] = r.d[a] * fast_rcp(h);
        r.lpc[a] = (r.d[a] != 0.0f) ? h * fast_rcp(r.d[a]) : kInf;
    }
    r.iv = fast_rcp(P.vol32[r.ib]);
}

__device__ __forceinline__ int get_dl(const ray_t &r, float &lmin)
{
    float t[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // computed for every axis and selected, so the wave does not branch
        const float face = (float)(r.i[a] + (r.d[a] > 0.0f ? 1 : 0));
        const float ta = fmaxf((face - r.x[a]) * r.lpc[a], 0.0f);
        t[a] = (r.d[a] == 0.0f) ? kInf : ta;
    }
    int ap = -1;
    float best = kInf;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (t[a] <= best && r.d[a] != 0.0f) {
            ap = a;
            best = t[a];
        }
    lmin = best;
    return ap;
}

// ST_BLOCK_LO / ST_BLOCK_HI: the index left [0,n) through the lower / upper
// face of the exit axis; the ray changes block (change_block) before it
// marches on.
enum { ST_IN = 0, ST_ESCAPED = 1, ST_BLOCK_LO = 2, ST_BLOCK_HI = 3, ST_LOST = 4 };

// Step to the exit face of axis ap after physical length l.
__device__ __forceinline__ int cross_in_block(const kparams &P, ray_t &r, int ap, float l)
{
    const int n = P.n_cell;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == ap) {
            const bool rhs = r.d[a] > 0.0f;
            r.x[a] = (float)(rhs ? r.i[a] + 1 : r.i[a]);
            r.i[a] += rhs ? 1 : -1;
        } else {
            r.x[a] = fmaf(l, r.cpl[a], r.x[a]);
        }
    }
    // only the exit axis' index moved, so it is the one that can have left
    // [0,n): tested on all three axes, without selecting the axis (a select
    // chain on ap compiles to branches)
    const bool hi = (r.i[0] >= n) | (r.i[1] >= n) | (r.i[2] >= n);
    const bool lo = (r.i[0] < 0) | (r.i[1] < 0) | (r.i[2] < 0);
    return hi ? ST_BLOCK_HI : (lo ? ST_BLOCK_LO : ST_IN);
}

// Block change of a ray whose index left [0,n) through the exit face of axis
// a (side 1: the upper face): through the face-neighbour table to the
// neighbour block, with the position rescaled across a level jump.
template <class NB>
__device__ __forceinline__ int change_block(const kparams &P, ray_t &r, int a, int side,
                                            const NB &T)
{
    const int n = P.n_cell;
    const float nf = (float)n;
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
    const float half = (float)(n / 2);
    const float sc = dl < 0 ? 2.0f : (dl > 0 ? 0.5f : 1.0f);
    const float isc = dl < 0 ? 0.5f : (dl > 0 ? 2.0f : 1.0f); // 1 / sc, exact
    // cell volumes of neighbouring levels differ by exactly 8
    r.iv *= dl < 0 ? 0.125f : (dl > 0 ? 8.0f : 1.0f);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        r.lpc[q] *= sc;
        r.cpl[q] *= isc;
        if (q == a) {
            r.i[q] = side ? 0 : n - 1;
            r.x[q] = side ? 0.0f : nf;
            continue;
        }
        const int j = (q == ap0) ? 0 : 1;
        if (dl < 0)
            r.x[q] = fmaf(r.x[q], 0.5f, (float)((shift >> j) & 1) * half);
        else if (dl > 0)
            r.x[q] = (r.x[q] - (float)(j == 0 ? j0 : j1) * half) * 2.0f;
        if (dl != 0)
            r.i[q] = min(max((int)floorf(r.x[q]), 0), n - 1);
    }
    return ST_IN;
}

__device__ __forceinline__ int change_block(const kparams &P, ray_t &r, int a, int side)
{
    return change_block(P, r, a, side, nb_global{P});
}

__device__ __forceinline__ int cross_face(const kparams &P, ray_t &r, int ap, float l)
{
    const int st = cross_in_block(P, r, ap, l);
    return st == ST_IN ? ST_IN : change_block(P, r, ap, st == ST_BLOCK_HI);
}

__device__ __forceinline__ int cell_index(const kparams &P, const ray_t &r)
{
    return r.i[0] + P.n_cell * (r.i[1] + P.n_cell * r.i[2]);
}

// cell_index with the indices clamped into the block: the cell itself while
// the ray is in its block, some cell of the block after it stepped out (an
// index of -1 maps to n-1: one unsigned minimum per axis).
__device__ __forceinline__ int cell_index_clamped(const kparams &P, const ray_t &r)
{
    const uint32_t m = (uint32_t)P.n_cell - 1u;
    // 24-bit products: cell indices of a block stay far below 2^24
    return (int)(min((uint32_t)r.i[0], m) +
                 __umul24((uint32_t)P.n_cell,
                          min((uint32_t)r.i[1], m) +
                              __umul24((uint32_t)P.n_cell, min((uint32_t)r.i[2], m))));
}

__device__ double tau_to_edge(const kparams &P, const ray_t &from, const float *dir, float ks,
                              uint64_t &ncross)
{
    ray_t r = from;
    r.d[0] = dir[0];
    r.d[1] = dir[1];
    r.d[2] = dir[2];
    set_scales(P, r);
    double tau = 0.0;
    // kappa_rho of the next cell is loaded one step ahead: the crossing is
    // taken before the current cell's optical depth is added (the ray is a
    // copy, so the crossing after the last cell is simply dropped)
    float kc = P.kappa[(size_t)r.ib * P.n3 + cell_index(P, r)];
    for (;;) {
        float l;
        const int ap = get_dl(r, l);
        if (ap < 0)
            break;
        const float kap = kc * ks;
        const int st = cross_face(P, r, ap, l);
        if (st == ST_IN)
            kc = P.kappa[(size_t)r.ib * P.n3 + cell_index(P, r)];
        tau += (double)(kap * l);
        ncross += 1;
        if (tau > kTauCut || st != ST_IN)
            break;
    }
    return tau;
}

// Per axis of a direction: whether the face ahead is the upper one, and the
// moves of the face ahead (+-1) and of the linear cell index (+-1, +-n,
// +-n^2) when the ray crosses it.  Held as named registers and picked by
// selects (the axis is a constant wherever the loops are unrolled).
struct face_steps {
    int u0, u1, u2, c0, c1, c2;
    __device__ face_steps(const float *d, int n)
    {
        u0 = d[0] > 0.0f ? 1 : 0;
        u1 = d[1] > 0.0f ? 1 : 0;
        u2 = d[2] > 0.0f ? 1 : 0;
        c0 = u0 ? 1 : -1;
        c1 = u1 ? n : -n;
        c2 = u2 ? n * n : -(n * n);
    }
    __device__ int up(int a) const { return a == 0 ? u0 : (a == 1 ? u1 : u2); }
    __device__ float sf(int a) const { return up(a) ? 1.0f : -1.0f; }
    __device__ int sc(int a) const { return a == 0 ? c0 : (a == 1 ? c1 : c2); }
};

// tau_to_edge with the block changes of a wave's rays taken together: each
// lane marches to the edge of its block (or its end), and the rays that left
// their block change it once every ray of the wave has left its block or
// ended.  Parallel rays of one wave that start in one block then enter the
// following blocks together, and the block-change code runs once per block
// instead of in nearly every step.  Per ray the operations and their order
// are those of tau_to_edge, so the result is the same.
//
// Within a block the step keeps, instead of the integer cell indices, the
// face ahead on each axis as a float (the value get_dl converts the index
// to) and the linear cell index, moved by +-1, +-n or +-n^2.  The ray has
// left the block when a face ahead passed n or 0 (the indices are rebuilt
// from the faces for change_block).  The lengths, positions and optical
// depths are computed as in get_dl and cross_in_block, so they are the same
// numbers.
__device__ double tau_to_edge_sync(const kparams &P, const ray_t &from, const float *dir,
                                   float ks, uint64_t &ncross)
{
    ray_t r = from;
    r.d[0] = dir[0];
    r.d[1] = dir[1];
    r.d[2] = dir[2];
    set_scales(P, r);
    double tau = 0.0;
    // an axis with d = 0 never has the nearest face; with none left the ray
    // has no face ahead at all (get_dl returns -1)
    const bool nz0 = r.d[0] != 0.0f, nz1 = r.d[1] != 0.0f, nz2 = r.d[2] != 0.0f;
    if (!(nz0 | nz1 | nz2))
        return tau;
    const int n = P.n_cell;
    const float nf = (float)n;
    const face_steps S(r.d, n);
    const uint32_t cmax = (uint32_t)P.n3 - 1u;
    const float *kb = P.kappa + (size_t)r.ib * P.n3; // the block's kappa_rho
    int c = cell_index(P, r);
    float kc = kb[c];
    float ff[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        ff[a] = (float)(r.i[a] + S.up(a));
    for (;;) {
        int ap, st;
        for (;;) {
            // get_dl: the nearest face ahead, ties to the later axis
            float t[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                t[a] = fmaxf((ff[a] - r.x[a]) * r.lpc[a], 0.0f);
            ap = -1;
            float l = kInf;
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
            if (ap < 0) { // no face ahead (get_dl's -1), as in tau_to_edge
                st = ST_LOST;
                break;
            }
            const float kap = kc * ks;
            // cross_in_block: the exit axis onto its face, the others along
            // the ray
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (a == ap) {
                    r.x[a] = ff[a];
                    ff[a] += S.sf(a);
                    c += S.sc(a);
                } else {
                    r.x[a] = fmaf(l, r.cpl[a], r.x[a]);
                }
            }
            const float fhi = fmaxf(fmaxf(ff[0], ff[1]), ff[2]);
            const float flo = fminf(fminf(ff[0], ff[1]), ff[2]);
            st = fhi > nf ? ST_BLOCK_HI : (flo < 0.0f ? ST_BLOCK_LO : ST_IN);
            // issued whether or not the ray is still in the block (then the
            // value is not used), so kc is one register loaded one step
            // ahead: a conditional load here made the compiler copy it into
            // place, waiting for the load within the step
            kc = kb[min((uint32_t)c, cmax)];
            tau += (double)(kap * l);
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
#pragma unroll
        for (int a = 0; a < 3; ++a)
            r.i[a] = (int)ff[a] - S.up(a);
        if (change_block(P, r, ap, st == ST_BLOCK_HI) != ST_IN)
            break;
        kb = P.kappa + (size_t)r.ib * P.n3;
        c = cell_index(P, r);
        kc = kb[c];
#pragma unroll
        for (int a = 0; a < 3; ++a)
            ff[a] = (float)(r.i[a] + S.up(a));
    }
    return tau;
}

// Position in global coordinates.
__device__ __forceinline__ void global_pos(const kparams &P, const ray_t &r, float *gx)
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
        gx[a] = fmaf(r.x[a], P.h32[3 * r.ib + a], P.org32[3 * r.ib + a]);
}

// Image pixel (linear index into [band][incl][N][N]) of global point gx, or -1.
__device__ __forceinline__ long pixel(const kparams &P, int band, int o, const float *gx)
{
    const int N = P.image_n;
    float rel[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        rel[a] = gx[a] - (float)P.center[a];
    const float eh[3] = {(float)P.eh[o][0], (float)P.eh[o][1], (float)P.eh[o][2]};
    const float ev[3] = {(float)P.ev[o][0], (float)P.ev[o][1], (float)P.ev[o][2]};
    const float uh = dot3(rel, eh);
    const float uv = dot3(rel, ev);
    const float half = (float)P.image_half, sc = (float)P.pix_scale;
    const float fx = floorf((uh + half) * sc);
    const float fy = floorf((uv + half) * sc);
    // written so that a NaN coordinate is outside too
    if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)N && fy < (float)N))
        return -1;
    return (long)((((size_t)band * P.n_incl + o) * N + (size_t)fy) * N + (size_t)fx);
}

__device__ __forceinline__ void deposit(const kparams &P, int band, int o, const float *gx,
                                        float I, float Q, float U)
{
    if (!P.image)
        return;
    const long pix = pixel(P, band, o, gx);
    if (pix < 0)
        return;
    atomic_add_f64(img_at(P, pix, 0), (double)I);
    atomic_add_f64(img_at(P, pix, 1), (double)Q);
    atomic_add_f64(img_at(P, pix, 2), (double)U);
}

__device__ __forceinline__ void phase(int model, float gg, float pmax, float mu, float &z11,
                                      float &z12, float &z33)
{
    const float m2 = mu * mu;
    if (model == KP_THOMSON) {
        z11 = 0.75f * (1.0f + m2);
        z12 = 0.75f * (m2 - 1.0f);
        z33 = 1.5f * mu;
    } else {
        const float D = fmaf(-2.0f * gg, mu, fmaf(gg, gg, 1.0f));
        const float ph = (1.0f - gg * gg) * fast_rsq(D) * fast_rcp(D);
        const float r1 = fast_rcp(1.0f + m2);
        z11 = ph;
        z12 = -pmax * (1.0f - m2) * r1 * ph;
        z33 = ph * (2.0f * mu * r1);
    }
}

__device__ __forceinline__ float thomson_mu(float xi)
{
    const float c = fmaf(8.0f, xi, -4.0f);
    float mu = 0.25f * c;
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const float mu2 = mu * mu;
        const float f = fmaf(mu2, mu, fmaf(3.0f, mu, -c));
        const float fp = fmaf(3.0f, mu2, 3.0f);
        mu = mu - f * fast_rcp(fp);
    }
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

__device__ __forceinline__ float hg_mu(float g, float xi)
{
    float mu;
    if (g == 0.0f) {
        mu = fmaf(2.0f, xi, -1.0f);
    } else {
        const float g2 = g * g;
        const float t = (1.0f - g2) * fast_rcp(fmaf(2.0f * g, xi, 1.0f - g));
        mu = (1.0f + g2 - t * t) * fast_rcp(2.0f * g);
    }
    return fminf(fmaxf(mu, -1.0f), 1.0f);
}

// Peel-off Stokes vector toward no in the image frame (Q>0 along ev).
__device__ __forceinline__ void peel_stokes(int model, float gg, float pmax, const float *d,
                                            const float *e1, const float *e2, float I, float Q,
                                            float U, const float *no, const float *ev,
                                            float &Io, float &Qo, float &Uo)
{
    const float mo = fminf(fmaxf(dot3(d, no), -1.0f), 1.0f);
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        p[a] = fmaf(-mo, d[a], no[a]);
    const float pn2 = dot3(p, p);
    float cf = 1.0f, sf = 0.0f, st = 0.0f;
    if (pn2 > 1e-12f) {
        const float ip = fast_rsq(pn2);
        cf = dot3(p, e1) * ip;
        sf = dot3(p, e2) * ip;
        st = pn2 * ip;
    }
    const float c2 = fmaf(cf, cf, -sf * sf);
    const float s2 = 2.0f * cf * sf;
    const float Qr = fmaf(Q, c2, U * s2);
    const float Ur = fmaf(U, c2, -Q * s2);
    float z11, z12, z33;
    phase(model, gg, pmax, mo, z11, z12, z33);
    const float Ip = fmaf(z11, I, z12 * Qr);
    const float Qp = fmaf(z12, I, z11 * Qr);
    const float Up = z33 * Ur;
    float e1o[3], e2o[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float e1r = fmaf(cf, e1[a], sf * e2[a]);
        e1o[a] = fmaf(mo, e1r, -st * d[a]);
    }
    cross3(no, e1o, e2o);
    const float cps = dot3(ev, e1o);
    const float sps = dot3(ev, e2o);
    const float inrm = fast_rcp(fmaf(cps, cps, sps * sps));
    const float c2p = fmaf(cps, cps, -sps * sps) * inrm;
    const float s2p = 2.0f * cps * sps * inrm;
    Io = Ip;
    Qo = fmaf(Qp, c2p, Up * s2p);
    Uo = fmaf(Up, c2p, -Qp * s2p);
}

// One scattering by (mu, azimuth 2 pi phi_u from e1): new d, e1, Stokes weight.
__device__ __forceinline__ void scatter_stokes(int model, float pmax, float mu, float phi_u,
                                               float *d, float *e1, const float *e2, float &I,
                                               float &Q, float &U)
{
    const float st = fast_sqrt(fmaxf(fmaf(-mu, mu, 1.0f), 0.0f));
    const float sf = __builtin_amdgcn_sinf(phi_u), cf = __builtin_amdgcn_cosf(phi_u);
    const float c2 = fmaf(cf, cf, -sf * sf);
    const float s2 = 2.0f * cf * sf;
    const float Qr = fmaf(Q, c2, U * s2);
    const float Ur = fmaf(U, c2, -Q * s2);
    const float m2 = mu * mu;
    const float pl = (model == KP_THOMSON) ? 1.0f : pmax;
    const float r1 = fast_rcp(1.0f + m2);
    const float a12 = -pl * (1.0f - m2) * r1;
    const float a33 = 2.0f * mu * r1;
    const float In = fmaf(a12, Qr, I);
    const float Qn = fmaf(a12, I, Qr);
    const float Un = a33 * Ur;
    I = In;
    Q = Qn;
    U = Un;
    float dn[3], en[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float e1r = fmaf(cf, e1[a], sf * e2[a]);
        dn[a] = fmaf(mu, d[a], st * e1r);
        en[a] = fmaf(mu, e1r, -st * d[a]);
    }
    const float id = fast_rsq(dot3(dn, dn));
#pragma unroll
    for (int a = 0; a < 3; ++a)
        dn[a] *= id;
    const float pr = dot3(en, dn);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        en[a] = fmaf(-pr, dn[a], en[a]);
    const float ie = fast_rsq(dot3(en, en));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d[a] = dn[a];
        e1[a] = en[a] * ie;
    }
}

__device__ __forceinline__ void emit(const uint32_t rn[4], float *d, float *e1, double &tau)
{
    const float mu = fmaf(2.0f, u01(rn[0]), -1.0f);
    const float ph = u01(rn[1]);
    const float sf = __builtin_amdgcn_sinf(ph), cf = __builtin_amdgcn_cosf(ph);
    const float st = fast_sqrt(fmaxf(fmaf(-mu, mu, 1.0f), 0.0f));
    d[0] = st * cf;
    d[1] = st * sf;
    d[2] = mu;
    e1[0] = mu * cf;
    e1[1] = mu * sf;
    e1[2] = -st;
    tau = -(double)fast_log(u01(rn[2]));
}

// Kahan-compensated register accumulator for one cell of the J tally.
// The source cell's sums go to the LDS accumulators when `lds`.
struct tally_acc {
    int64_t gc = -1;
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void flush(const kparams &P, int band, bool lds = false)
    {
        if (gc >= 0 && P.J)
            j_deposit(P, lds, gc, band, (double)s - (double)c);
        gc = -1;
    }
    __device__ __forceinline__ void add(const kparams &P, int band, int64_t cell, float v,
                                        bool lds = false)
    {
        if (cell == gc) {
            const float y = v - c;
            const float t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            flush(P, band, lds);
            gc = cell;
            s = v;
            c = 0.0f;
        }
    }
};

struct counters_t {
    uint64_t cross = 0, peel = 0, events = 0, escaped = 0, rr = 0, steps = 0, lost = 0,
             packets = 0;
};

__device__ void packet(const kparams &P, int band, uint64_t pid, uint64_t slot, counters_t &C)
{
    const dband bd = P.bands[band];
    const float ks = (float)bd.kappa_scale;
    const float w0 = (float)P.w0;
    const uint32_t k0 = P.seed, k1 = (uint32_t)band;
    const uint32_t plo = (uint32_t)pid, phi = (uint32_t)(pid >> 32);
    const float pmax = (float)P.pmax;
    uint64_t ncross = 0, npeel = 0;
    int events = 0, status = 0;
    tally_acc acc;

    uint32_t rn[4] = {plo, phi, 0u, 0u};
    kp_philox4x32_10(rn, k0, k1);
    ray_t r;
    float e1[3];
    {
        const float mu = fmaf(2.0f, u01(rn[0]), -1.0f);
        const float ph = u01(rn[1]);
        const float sf = __builtin_amdgcn_sinf(ph), cf = __builtin_amdgcn_cosf(ph);
        const float st = fast_sqrt(fmaxf(fmaf(-mu, mu, 1.0f), 0.0f));
        r.d[0] = st * cf;
        r.d[1] = st * sf;
        r.d[2] = mu;
        e1[0] = mu * cf;
        e1[1] = mu * sf;
        e1[2] = -st;
    }
    double tau_rem = -(double)fast_log(u01(rn[2]));
    float I = w0, Q = 0.0f, U = 0.0f;
    r.ib = P.src_ib;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.x[a] = P.src_xl[a];
        r.i[a] = P.src_i[a];
    }
    set_scales(P, r);
    float nof[kMaxIncl][3];
    for (int o = 0; o < P.n_incl; ++o)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            nof[o][a] = (float)P.no[o][a];
    if (P.n_incl > 0) {
        float gx[3];
        global_pos(P, r, gx);
        for (int o = 0; o < P.n_incl; ++o) {
            const double tau = tau_to_edge(P, r, nof[o], ks, npeel);
            const float att = fast_exp(-(float)tau) * kInv4Pi;
            deposit(P, band, o, gx, I * att, 0.0f, 0.0f);
        }
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
            float l;
            const int ap = get_dl(r, l);
            if (ap < 0) {
                status = KP_END_LOST;
                break;
            }
            cell = cell_index(P, r);
            const int64_t gc = (int64_t)r.ib * P.n3 + cell;
            const float kap = P.kappa[gc] * ks;
            const float dtau = kap * l;
            const float iV = r.iv;
            ncross += 1;
            steps -= 1;
            if (kap > 0.0f && (double)dtau >= tau_rem) {
                const float li = fminf((float)tau_rem * fast_rcp(kap), l);
                acc.add(P, band, gc, I * li * iV);
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    r.x[a] = fmaf(li, r.cpl[a], r.x[a]);
                interact = true;
                break;
            }
            tau_rem -= (double)dtau;
            acc.add(P, band, gc, I * l * iV);
            const int st = cross_face(P, r, ap, l);
            if (st == ST_ESCAPED) {
                status = KP_END_ESCAPED;
                break;
            }
            if (st == ST_LOST) {
                status = KP_END_LOST;
                break;
            }
        }
        if (!interact)
            break;

        events += 1;
        uint32_t u[4] = {plo, phi, (uint32_t)events, 0u};
        kp_philox4x32_10(u, k0, k1);
        const int64_t gc = (int64_t)r.ib * P.n3 + cell;
        const float alb = (bd.flags & KP_BAND_CELL_ALBEDO) ? P.albedo[gc] : (float)bd.albedo;
        const float gg = (bd.flags & KP_BAND_CELL_G) ? (P.g ? P.g[gc] : 0.0f) : (float)bd.g;
        I *= alb;
        Q *= alb;
        U *= alb;
        float e2[3];
        cross3(r.d, e1, e2);

        if (P.n_incl > 0) {
            float gx[3];
            global_pos(P, r, gx);
            for (int o = 0; o < P.n_incl; ++o) {
                const float *no = nof[o];
                const float mo = fminf(fmaxf(dot3(r.d, no), -1.0f), 1.0f);
                float p[3];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    p[a] = fmaf(-mo, r.d[a], no[a]);
                const float pn2 = dot3(p, p);
                float cf = 1.0f, sf = 0.0f, st = 0.0f;
                if (pn2 > 1e-12f) {
                    const float ip = fast_rsq(pn2);
                    cf = dot3(p, e1) * ip;
                    sf = dot3(p, e2) * ip;
                    st = pn2 * ip;
                }
                const float c2 = fmaf(cf, cf, -sf * sf);
                const float s2 = 2.0f * cf * sf;
                const float Qr = fmaf(Q, c2, U * s2);
                const float Ur = fmaf(U, c2, -Q * s2);
                float z11, z12, z33;
                phase(P.model, gg, pmax, mo, z11, z12, z33);
                const float Ip = fmaf(z11, I, z12 * Qr);
                const float Qp = fmaf(z12, I, z11 * Qr);
                const float Up = z33 * Ur;
                float e1o[3], e2o[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float e1r = fmaf(cf, e1[a], sf * e2[a]);
                    e1o[a] = fmaf(mo, e1r, -st * r.d[a]);
                }
                cross3(no, e1o, e2o);
                const float ev[3] = {(float)P.ev[o][0], (float)P.ev[o][1], (float)P.ev[o][2]};
                const float cps = dot3(ev, e1o);
                const float sps = dot3(ev, e2o);
                const float inrm = fast_rcp(fmaf(cps, cps, sps * sps));
                const float c2p = fmaf(cps, cps, -sps * sps) * inrm;
                const float s2p = 2.0f * cps * sps * inrm;
                const float Qi = fmaf(Qp, c2p, Up * s2p);
                const float Ui = fmaf(Up, c2p, -Qp * s2p);
                const double tau = tau_to_edge(P, r, no, ks, npeel);
                const float att = fast_exp(-(float)tau) * kInv4Pi;
                deposit(P, band, o, gx, Ip * att, Qi * att, Ui * att);
            }
        }

        {
            const float xi0 = u01(u[0]);
            const float mu = (P.model == KP_THOMSON) ? thomson_mu(xi0) : hg_mu(gg, xi0);
            const float st = fast_sqrt(fmaxf(fmaf(-mu, mu, 1.0f), 0.0f));
            const float ph = u01(u[1]);
            const float sf = __builtin_amdgcn_sinf(ph), cf = __builtin_amdgcn_cosf(ph);
            const float c2 = fmaf(cf, cf, -sf * sf);
            const float s2 = 2.0f * cf * sf;
            const float Qr = fmaf(Q, c2, U * s2);
            const float Ur = fmaf(U, c2, -Q * s2);
            const float m2 = mu * mu;
            const float pl = (P.model == KP_THOMSON) ? 1.0f : pmax;
            const float r1 = fast_rcp(1.0f + m2);
            const float a12 = -pl * (1.0f - m2) * r1;
            const float a33 = 2.0f * mu * r1;
            const float In = fmaf(a12, Qr, I);
            const float Qn = fmaf(a12, I, Qr);
            const float Un = a33 * Ur;
            I = In;
            Q = Qn;
            U = Un;
            float dn[3], en[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float e1r = fmaf(cf, e1[a], sf * e2[a]);
                dn[a] = fmaf(mu, r.d[a], st * e1r);
                en[a] = fmaf(mu, e1r, -st * r.d[a]);
            }
            const float id = fast_rsq(dot3(dn, dn));
#pragma unroll
            for (int a = 0; a < 3; ++a)
                dn[a] *= id;
            const float pr = dot3(en, dn);
#pragma unroll
            for (int a = 0; a < 3; ++a)
                en[a] = fmaf(-pr, dn[a], en[a]);
            const float ie = fast_rsq(dot3(en, en));
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                r.d[a] = dn[a];
                e1[a] = en[a] * ie;
            }
            set_scales(P, r);
        }
        tau_rem = -(double)fast_log(u01(u[2]));
        if (I < (float)P.rr_w * w0) {
            if (u01(u[3]) < (float)P.rr_s) {
                const float inv = (float)(1.0 / P.rr_s);
                I *= inv;
                Q *= inv;
                U *= inv;
            } else {
                status = KP_END_RR;
                break;
            }
        }
    }
    acc.flush(P, band);

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
            es[KP_ES_X + a] = (double)P.org32[3 * r.ib + a] +
                              (double)r.x[a] * (double)P.h32[3 * r.ib + a];
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

} // namespace f32
} // namespace kp
