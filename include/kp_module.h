/*
 * kp_module.h -- C ABI of the photon module: the host-side counterpart of
 * the reference's particle::radiation::base_t (src/modules/particle/
 * radiation/radiation.h:19-55 over particle::base_t, driver.h:15-102,
 * driver.cpp:1-73), built on the engine ABI of kpolrad.h.
 *
 *   read   -> kp_module_create   (INI text, reference key names)
 *   init   -> kp_module_init     (mesh from [mesh]/[refine_region*], engine)
 *   step   -> kp_module_step     (n_cycle_stp integrate cycles)
 *   save   -> kp_module_save     (binary_io container, io/binary/binary_io.h)
 *
 * INI keys read (reference keys first, then the photon keys this engine
 * adds; "section|key" as in io/args/input.cpp):
 *   [mesh] x_min, x_max (3 reals), n_cell_global, n_cell_block (3 ints;
 *          cubic blocks) ......................... meshgen.cpp:13-40
 *          coord (cartesian|spherical: meshgen::sph_t axes r, theta,
 *          phi, spherical.h:9-140), r_ratio (1) .... cylindrical.h:18-32
 *   [boundary] kinds (6 words, "per..." = periodic) meshgen.cpp:47-52
 *   [refine_region*] level, x_min/x_max or loc ... meshgen.cpp:111-142
 *   [particle] n_hv (1) ........................... radiation.h:26
 *              n_cycle_stp (1) .................... driver.cpp:18
 *              n_th (64) .......................... driver_base.h:30
 *              n_par (1.0, packets per band per cycle), step_max (100)
 *                                                   generate_par.h:37-41
 *   [photon] seed (42), model (thomson|hg_rayleigh), p_max (1), kappa
 *            (uniform kappa_rho, 0), albedo, g (uniform, per band lists
 *            allowed), kappa_scale (n_hv values, 1), source (3 reals,
 *            centre of the domain), weight (1), incl (degrees, none),
 *            image_n (128), image_half (0.5), precision (f64|f32),
 *            rr_weight (0.01), rr_survive (0.1), schedule (persistent|history|
 *            block_queue),
 *            order (slot|direction: packets of a launch in emission-direction
 *            order, kp_schedule::coherent; the engine's default if absent)
 *   On a spherical mesh the photon streams (photon_sph.h): kappa_scale
 *   defaults to 0, the source (a Cartesian point) to the centre of the
 *   (r, theta, phi) box, precision must be f64, incl empty, and the engine's
 *   one-thread-per-packet schedule in emission-direction order is kept unless
 *   photon|schedule is given.
 */
#ifndef KP_MODULE_H
#define KP_MODULE_H

#include "kpolrad.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kp_module kp_module;

typedef struct kp_module_info {
    int64_t n_blocks;
    int64_t n_cells;        /* n_blocks * n_cell^3                          */
    int32_t n_cell;
    int32_t n_bands;        /* [particle] n_hv                              */
    int32_t n_incl;
    int32_t image_n;
    int64_t n_par;          /* packets per band per cycle                   */
    int32_t n_cycle;        /* [particle] n_cycle_stp                       */
    int32_t initialised;    /* kp_module_init done                          */
    int64_t packets_done;   /* packets per band run so far                  */
} kp_module_info;

/* particle::radiation::base_t::read (radiation.h:24-29, driver.cpp:18,
 * driver_base.h:30, generate_par.h:37-41): parses INI text; host only. */
kp_status kp_module_create(const char *ini_text, kp_module **out);
void kp_module_destroy(kp_module *m);

/* particle::base_t::init (driver.cpp:7-24) with meshgen::base_t::init and
 * load_refinement (meshgen.cpp:42-142): builds the block mesh and, when
 * device >= 0, the engine on that device with the [photon] medium. */
kp_status kp_module_init(kp_module *m, int32_t device);

/* Per-cell medium (host arrays in the mesh cell order), replacing the
 * uniform [photon] kappa / albedo / g (block data of rad_data.h:10-42). */
kp_status kp_module_set_media(kp_module *m, const float *kappa_rho, const float *albedo,
                              const float *g);

/* particle::base_t::step (driver.cpp:55-60): n_cycle_stp cycles, each one
 * kp_forward of n_par packets per band with fresh packet ids; tallies
 * accumulate.  Synchronises `stream` at the end like the reference. */
kp_status kp_module_step(kp_module *m, void *stream);

/* particle::base_t::save (driver.cpp:33-38) / rad_t::write
 * (rad_data.cpp:45-49): a binary_io container (format of
 * io/binary/binary_io.h:6-32) holding input_args, and per block
 * block_<id>|i_logic, |level, |rad_flx_flux_{n_int,n_step,n_cell,n_gh,field}
 * (FP64 J, [cell][band]) and |rad_flx_mfpi_* (kappa_rho), plus
 * photon|image (FP64 [band][incl][y][x][IQU]) and photon|counters. */
kp_status kp_module_save(const kp_module *m, const char *path);

kp_status kp_module_get_info(const kp_module *m, kp_module_info *info);

/* Host copies of the accumulated tallies (any pointer may be NULL):
 * J [n_cells][n_bands], image [n_bands][n_incl][image_n][image_n][3],
 * counters [KP_NCOUNTERS]. */
kp_status kp_module_tallies(const kp_module *m, double *J, double *image, uint64_t *counters);

/* The module's mesh (owned by the module; valid after kp_module_init). */
kp_status kp_module_mesh(const kp_module *m, const kp_mesh **mesh);

#ifdef __cplusplus
}
#endif

#endif /* KP_MODULE_H */
