/*
 * kp_oracle.h -- FP64 CPU oracle of the photon-packet transport path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker or the CPU baseline; the product path never calls it.
 *
 * It restates, in plain C with pthreads over packet-id ranges:
 *  - the reference march: get_dl (photon.h:20-44), proc_phys tally
 *    (photon.h:73-83), proc_geo snap (photon.h:54-71), the step budget and
 *    block reload of proc (photon.h:104-125), shift_blk level jumps
 *    (proxy.h:64-121), x_fc (geometry.h:68-83), periodic regulate
 *    (particle_base.h:37-56);
 *  - the physics the reference has no code for, from published methods:
 *    Philox4x32-10 keyed by (seed, band) with counter (packet id, event),
 *    tau = -ln xi free flights, survival-biased absorption with Russian
 *    roulette, Thomson and Henyey-Greenstein/Rayleigh-form (White 1979)
 *    Mueller scattering with Stokes frame rotation (Chandrasekhar 1960),
 *    peel-off images (Yusef-Zadeh, Morris & White 1984) and the path-length
 *    mean-intensity tally (Lucy 1999).
 * Deviations from the reference march (the first two inert on the dyadic
 * meshes the reference produces for pure streaming):
 *  (1) after a coarse->fine face crossing the tangential fine cell is the one
 *      containing the position (the reference always takes the lower one,
 *      proxy.h:383-389, and geo_loc_t::fix, particle_base.h:119-127, does
 *      not update the index);
 *  (2) a negative face distance on a moving axis (position one ulp past a
 *      face after an interaction) is clamped to zero;
 *  (3) a ray leaving through the lower face of a periodic axis is moved to
 *      the upper periodic image at once (the reference's strict
 *      x < x_min test in regulate leaves it on x_min, inside the upper
 *      block's index range but outside its face bounds, for one step).
 */
#ifndef KP_ORACLE_H
#define KP_ORACLE_H

#include "../include/kpolrad.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Full forward run.  Outputs are zeroed by the caller and accumulated.
 * Any of J, image, counters, end_state may be NULL. */
int or_forward(const kp_mesh_desc *mesh, const float *kappa, const float *albedo,
               const float *g, const kp_source *src, const kp_run *run, double *J,
               double *image, uint64_t *counters, double *end_state, int n_threads);

/* Wall time (seconds) of the transport phase of the last or_forward call:
 * from starting its threads to joining them, without the allocation and the
 * reduction of the per-thread tallies. */
double or_last_transport_seconds(void);

/* Adjoint by path replay of the linear image functional
 * Jlin = sum_pix dL_dimage[pix] . image[pix] (image layout, I,Q,U):
 * accumulates dJlin/dlog(rho_c), dJlin/d(albedo_c), dJlin/d(g_c) per cell
 * (d_albedo / d_g may be NULL) and adds the primal Jlin to *jlin.  FP64. */
int or_adjoint(const kp_mesh_desc *mesh, const float *kappa, const float *albedo,
               const float *g, const kp_source *src, const kp_run *run,
               const double *dL_dimage, double *d_logrho, double *d_albedo, double *d_g,
               double *jlin, int n_threads);

/* Pure-streaming trace of one ray (the reference photon.h:104-125 march with
 * no medium): records every step's block, in-block cell index and length.
 * Returns the number of steps written (<= max_steps); *status is 1 when
 * the ray left through a physical boundary, 4 when lost, 0 when truncated. */
int or_trace(const kp_mesh_desc *mesh, const double x[3], const double d[3],
             int max_steps, int32_t *out_ib, int32_t *out_cell, double *out_len,
             double x_end[3], int *status);

/* Spherical meshes (kp_sph_oracle.c): the reference's spherical photon march
 * (photon_sph.h:10-266) as pure streaming with the path-length tally.  The
 * source point and traced rays are Cartesian; or_sph_locate finds the cell
 * containing a Cartesian point with the march's own face comparisons. */
int or_sph_forward(const kp_mesh_desc *mesh, const kp_source *src, const kp_run *run,
                   double *J, uint64_t *counters, double *end_state, int n_threads);
int or_sph_trace(const kp_mesh_desc *mesh, const double x[3], const double d[3],
                 int max_steps, int32_t *out_ib, int32_t *out_cell, double *out_len,
                 double x_end[3], int *status);
int or_sph_locate(const kp_mesh_desc *mesh, const double x[3], int32_t *ib, int32_t i[3]);

/* Source cell by brute-force point location (geo_t::idx, geometry.h:120-140). */
int or_locate(const kp_mesh_desc *mesh, const double x[3], int32_t *ib, int32_t i[3]);

/* Stokes algebra of one scattering (new d, e1, S in place; azimuth 2 pi phi_u
 * from e1 toward e2 = d x e1) and of one peel-off toward no with image
 * vertical axis ev (Si in the image frame, no 1/4pi, no attenuation). */
void or_scatter_stokes(int model, double pmax, double mu, double phi_u, double d[3],
                       double e1[3], double S[3]);
void or_peel_stokes(int model, double g, double pmax, const double d[3], const double e1[3],
                    const double S[3], const double no[3], const double ev[3], double Si[3]);

/* Leaf arithmetic exposed for pinning against libm / published vectors. */
void or_philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]);

/* Pool compaction and per-destination binning (kp_compact_oracle.c). */
int64_t or_pool_keep(void *pool, int64_t n, int64_t unit, const uint8_t *keep);
int64_t or_bin_by_key(const int32_t *key, int64_t n, int32_t n_bins, int64_t *offsets,
                      int32_t *order, int32_t *slot);
void or_gather(const void *src, int64_t unit, const int32_t *order, int64_t n, void *dst);
double or_log(double x);
double or_exp(double x);
void or_sincos2pi(double u, double *s, double *c);
double or_thomson_mu(double xi);
double or_hg_mu(double g, double xi);

#ifdef __cplusplus
}
#endif

#endif
