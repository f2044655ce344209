/*
 * kpolrad.h -- C ABI of the MI355X photon-packet transport engine.
 *
 * Drop-in for the photon module of kratos_polrad
 * (src/modules/particle/radiation/).  Every entry point takes plain pointers
 * and sizes, never throws, and returns a kp_status; kp_last_error() gives the
 * message of the last failure on the calling thread.  Device pointers are
 * hipMalloc'ed memory owned by the caller unless stated otherwise; the
 * stream argument is a hipStream_t passed as void* (NULL = default stream).
 *
 * Reference interfaces replaced (paths relative to the reference src/):
 *   mesh tree + face-neighbour table ... mesh/tree/tree.cpp:177-218,
 *                                        modules/particle/types/proxy.h:22-62,
 *                                        modules/particle/types/block_map.h:53-103
 *   block geometry ..................... mesh/meshgen/meshgen.cpp:286-351,378-432
 *   per-cell radiation storage ......... modules/particle/radiation/rad_data.cpp:13-25
 *   one particle step (generate +
 *     integrate + comm round) .......... modules/particle/driver.h:59-79,
 *                                        evolve/generate_par.h:51-62,
 *                                        evolve/integrator.h:47-106,
 *                                        radiation/photon.h:20-125,
 *                                        types/proxy.h:64-121
 *   pool compaction / send gather ...... types/particle_pool.h:191-248,
 *                                        boundary/comm.h:118-127,226-239
 *   spherical coordinates (mesh and
 *     photon march) .................... mesh/meshgen/spherical.h:9-140,
 *                                        mesh/meshgen/cylindrical.h:18-43,
 *                                        radiation/photon_sph.h:10-266
 */
#ifndef KPOLRAD_H
#define KPOLRAD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kp_status {
    KP_OK = 0,
    KP_ERR_ARG = 1,         /* invalid argument (std::invalid_argument)      */
    KP_ERR_HIP = 2,         /* HIP runtime error (reference hip_api throws)   */
    KP_ERR_ALLOC = 3,       /* allocation failure                             */
    KP_ERR_STATE = 4,       /* call order violated (e.g. no media set)        */
    KP_ERR_UNSUPPORTED = 5  /* valid request the engine does not implement   */
} kp_status;

/* Message of the last failure on this thread ("" if none). */
const char *kp_last_error(void);

/* ================================================================== */
/* Host mesh: block tree with 2:1 balance, ids, geometry, face table.  */
/* ================================================================== */

typedef struct kp_mesh kp_mesh;

/* Coordinate system of a mesh: the meshgen class the reference's problem
 * code instantiates (meshgen::base_t or meshgen::sph_t). */
enum { KP_COORD_CARTESIAN = 0, KP_COORD_SPHERICAL = 1 };

/* Per-block geometry table of a spherical mesh (doubles), with m = n_cell+1:
 *   [0,m)         r of the r faces            (geo_t::x_fc axis 0)
 *   [m,2m)        cos(theta) of the theta faces
 *   [2m,3m)       sin(phi) of the phi faces
 *   [3m,4m)       cos(phi) of the phi faces
 *   [4m,4m+n)     (r_{k+1}^3 - r_k^3)/3       (sph_t::volume factors,
 *   [4m+n,4m+2n)  |cos th_k - cos th_{k+1}|    spherical.h:31-42; a cell's
 *   [4m+2n,4m+3n) phi_{k+1} - phi_k            volume is (f_r*f_th)*f_phi)  */
#define KP_SPH_TAB(n_cell) (4 * ((n_cell) + 1) + 3 * (n_cell))

/* Block-table view of a mesh (arrays owned by the kp_mesh).  Block b has
 * cells i in [0,n_cell)^3; its cell (i0,i1,i2) has linear index
 *   b*n_cell^3 + i0 + n_cell*(i1 + n_cell*i2)
 * (the dat_3d_t order of mesh/types/dat_3d.h:145-147 without ghosts).
 * nb_id[b][a][s][j0][j1] is the face neighbour across side s of axis a;
 * j0/j1 pick the sub-face along axes (a+1)%3/(a+2)%3 when the neighbour is
 * finer (proxy.h:11-20 neighbors[3][2][2][2]).  nb_shift bit j is shift[j]
 * (which half of a coarser neighbour's face this block touches). */
typedef struct kp_mesh_desc {
    int32_t n_blocks;
    int32_t n_cell;        /* cells per block edge, even (n_cell_block)    */
    int32_t periodic[3];
    double x_min[3];
    double x_max[3];
    const double *xf0;     /* [n_blocks][3] first face coordinate          */
    const double *dx0;     /* [n_blocks][3] cell width                     */
    const double *vol0;    /* [n_blocks]    cell volume                    */
    const int32_t *level;  /* [n_blocks]                                   */
    const int32_t *logic;  /* [n_blocks][3] logical block index at level   */
    const int32_t *nb_id;  /* [n_blocks][24] neighbour block id or -1      */
    const uint8_t *nb_shift; /* [n_blocks][24]                             */
    const int8_t *d_lvl;   /* [n_blocks][6]  neighbour level - own level   */
    const uint8_t *phybnd; /* [n_blocks] bit 2a+s: physical boundary       */
    int32_t coord;         /* KP_COORD_*; spherical axes are (r,theta,phi) */
    int32_t pad_;
    const double *xf;      /* [n_blocks][3][n_cell+1] face coordinates     */
    const double *sph;     /* [n_blocks][KP_SPH_TAB(n_cell)] or NULL       */
} kp_mesh_desc;

/* meshgen::base_t::read/init (meshgen.cpp:13-106): root grid of
 * root_blocks[a] blocks of n_cell^3 cells on [x_min,x_max]. */
kp_status kp_mesh_create(const int32_t root_blocks[3], int32_t n_cell,
                         const double x_min[3], const double x_max[3],
                         const int32_t periodic[3], kp_mesh **out);
void kp_mesh_destroy(kp_mesh *mesh);

/* Coordinate system, before kp_mesh_finalize.  KP_COORD_SPHERICAL makes the
 * axes (r, theta, phi) (meshgen::sph_t, spherical.h:9-140): x_min/x_max of
 * kp_mesh_create are the r, theta and phi limits; 0 < r_min, 0 <= theta_min
 * < theta_max <= pi, phi width <= 2 pi (periodic only when it is 2 pi) and
 * every phi cell narrower than pi.  r_ratio is [mesh] r_ratio of
 * cyl_t::read (cylindrical.h:18-32): 1 = uniform r, > 0 = ratio of
 * consecutive root-level r widths, < 0 = logarithmic spacing
 * (r_max/r_min)^(1/n_r). */
kp_status kp_mesh_set_coord(kp_mesh *mesh, int32_t coord, double r_ratio);

/* tree_t::insert (tree.cpp:202-218) of n logical regions at `level`
 * (idx is [n][3]); parents are inserted and masked, 2:1 balance kept. */
kp_status kp_mesh_refine(kp_mesh *mesh, int32_t level, const int32_t *idx,
                         int64_t n);

/* load_refinement SMR box (meshgen.cpp:111-142): every block at `level`
 * whose index range covers [x_lo, x_hi] (search_blk_idx, meshgen.cpp:353-373). */
kp_status kp_mesh_refine_region(kp_mesh *mesh, int32_t level, const double x_lo[3],
                                const double x_hi[3]);

/* fill_base_level + dist_default ids + neighbours + geometry
 * (meshgen.cpp:144-188, 286-351, 378-432). */
kp_status kp_mesh_finalize(kp_mesh *mesh);

kp_status kp_mesh_describe(const kp_mesh *mesh, kp_mesh_desc *desc);

/* Leaf block and cell containing x (cells are half-open [x_l,x_r)),
 * geo_t::idx semantics (geometry.h:120-140), x in mesh coordinates
 * ((r, theta, phi) on a spherical mesh).  ib = -1 if outside. */
kp_status kp_mesh_locate(const kp_mesh *mesh, const double x[3],
                         int32_t *ib, int32_t i[3]);

/* ================================================================== */
/* Transport engine                                                    */
/* ================================================================== */

typedef struct kp_engine kp_engine;

/* Copies the block table to the device (block_map_t::update,
 * block_map.h:53-103).  The engine is bound to `device`. */
kp_status kp_engine_create(const kp_mesh_desc *mesh, int32_t device,
                           kp_engine **out);
void kp_engine_destroy(kp_engine *eng);

/* Per-cell medium, host arrays of n_blocks*n_cell^3 floats in the block
 * table's cell order: kappa_rho (extinction per unit length, the rad_t
 * mfp_i field of rad_data.h:16) >= 0, albedo in [0,1], HG asymmetry g in
 * (-1,1) (may be NULL: then 0).  Blocking: waits for all work already
 * issued on the engine's device (any stream, graph replays included), then
 * copies the medium in place, so earlier launches read the old medium and
 * later ones the new. */
kp_status kp_set_media(kp_engine *eng, const float *kappa_rho,
                       const float *albedo, const float *g);

enum { KP_THOMSON = 0, KP_HG_RAYLEIGH = 1 };
enum { KP_F64 = 0, KP_F32 = 1 };
enum { KP_BAND_CELL_ALBEDO = 1, KP_BAND_CELL_G = 2 };

/* One wavelength (or source set).  Band b uses Philox stream b and tally
 * component b (rad_t::flx has n_hv components, radiation.h:24-35). */
typedef struct kp_band {
    double kappa_scale;  /* multiplies kappa_rho                           */
    double albedo;       /* used unless flags & KP_BAND_CELL_ALBEDO        */
    double g;            /* used unless flags & KP_BAND_CELL_G             */
    uint64_t n_packets;  /* packets of this band in this call              */
    uint32_t flags;
    uint32_t pad_;
} kp_band;

/* Isotropic, unpolarized point source. */
typedef struct kp_source {
    double x[3];
    double weight;       /* Stokes I of each emitted packet                */
} kp_source;

typedef struct kp_run {
    uint32_t seed;           /* Philox key word 0                          */
    int32_t scatter_model;   /* KP_THOMSON | KP_HG_RAYLEIGH                */
    double p_max;            /* peak linear polarization (HG-Rayleigh)     */
    int32_t n_bands;
    int32_t precision;       /* KP_F64: FP64 global coordinates;
                                KP_F32: FP32 block-local transport,
                                FP64 optical depth and tallies            */
    const kp_band *bands;    /* host array [n_bands]                       */
    uint64_t packet_offset;  /* id of the first packet of every band       */
    int32_t step_max;        /* march-step budget ([particle] step_max)    */
    int32_t n_incl;          /* observers (peel-off images)                */
    const double *incl_deg;  /* host [n_incl], azimuth 0                   */
    double rr_weight;        /* Russian roulette when I < rr_weight*weight */
    double rr_survive;       /* survival probability                       */
    int32_t image_n;         /* pixels per image side                      */
    int32_t pad_;
    double image_half;       /* half-width of the field of view            */
} kp_run;

/* Counters accumulated by kp_forward (uint64 each). */
enum {
    KP_C_CROSS_MAIN = 0,  /* march steps of packet paths                   */
    KP_C_CROSS_PEEL = 1,  /* march steps of peel-off rays                  */
    KP_C_EVENTS = 2,      /* interactions                                  */
    KP_C_ESCAPED = 3,
    KP_C_KILLED_RR = 4,
    KP_C_KILLED_STEPS = 5,
    KP_C_LOST = 6,        /* left through a face with no neighbour, or stuck */
    KP_C_PACKETS = 7,
    KP_NCOUNTERS = 8
};

/* Per-packet end record (doubles) written when end_state != NULL. */
enum {
    KP_ES_X = 0, KP_ES_D = 3, KP_ES_I = 6, KP_ES_Q = 7, KP_ES_U = 8,
    KP_ES_EVENTS = 9, KP_ES_CROSS = 10, KP_ES_STATUS = 11, KP_ES_IB = 12,
    KP_ES_CELL = 13, KP_ES_STRIDE = 16
};
enum { KP_END_ESCAPED = 1, KP_END_RR = 2, KP_END_STEPS = 3, KP_END_LOST = 4 };

/* Device outputs; each may be NULL.  kp_forward ADDS into J, image and
 * counters (chunked runs accumulate); end_state is overwritten. */
typedef struct kp_tallies {
    double *J;          /* [n_cells][n_bands]: sum of I*l/V (photon.h:73-83) */
    double *image;      /* [n_bands][n_incl][image_n][image_n][3] I,Q,U    */
    uint64_t *counters; /* [KP_NCOUNTERS]                                  */
    double *end_state;  /* [sum of n_packets][KP_ES_STRIDE], band-major    */
} kp_tallies;

/* Launch schedule of kp_forward.
 * KP_SCHED_HISTORY: one thread per packet, the integrator's launch shape
 *   (ceil(n/threads) blocks, integrator.h:40-46), the whole history in that
 *   thread (photon.h:104-125).
 * KP_SCHED_PERSISTENT: CUs*blocks_per_cu resident blocks; each lane runs a
 *   one-crossing-per-iteration state machine and refills from a wave slot
 *   range of `chunk` packets claimed with one atomic.  Same histories.
 * KP_SCHED_BLOCKQ: block-bucketed packet queues (the per-block pools of
 *   types/particle_pool.h:191-248 and the per-destination compaction of
 *   boundary/comm.h:118-127, with blocks as destinations).  Bands from
 *   queue_from_band on run as passes: the packets' records (144 B FP64,
 *   96 B FP32) are sorted by (band, block); a workgroup takes queue_chunk
 *   sorted packets, stages each of their blocks' kappa_rho and a J tile
 *   (FP32 in the mixed path, FP64 in the FP64 path) in LDS, and marches
 *   every packet until it leaves the block or ends, then adds the tile to J
 *   with one FP64 atomic per touched cell; packets that cross into another
 *   block are compacted into the next pass's records by wave ballot and
 *   prefix rank.  A packet that takes queue_yield steps in one pass without
 *   leaving its block goes back to the queue.  Each interaction queues what
 *   its peel-offs are made of (64 B FP32, 128 B FP64); a kernel of their own
 *   computes and marches them after each pass, one interaction per lane.
 *   After queue_passes passes the packets still alive finish one thread
 *   each.  Bands before queue_from_band run on the persistent kernel in the
 *   same launch.  Same histories; needs n_cell^3 <= 4096 and fewer than 2^31
 *   queued packets per launch.  Record memory: 2 x record size + 16 B per
 *   queued packet, kept by the engine. */
enum { KP_SCHED_HISTORY = 0, KP_SCHED_PERSISTENT = 1, KP_SCHED_BLOCKQ = 2 };
typedef struct kp_schedule {
    int32_t kind;
    int32_t block_threads;  /* 64, 128 or 256 (the reference's [particle] n_th) */
    int32_t blocks_per_cu;
    int32_t chunk;
    /* FP64 atomics on hot cells/pixels are spread over replicas of J and of
     * the images (workgroup b uses replica b % R), folded into the caller's
     * arrays after the launch.  0 = choose from a device-memory budget. */
    int32_t j_replicas;
    int32_t img_replicas;
    /* persistent kernel: a wave runs its pending interactions / peel-off
     * ends / packet ends together once this many lanes wait (1..64).
     * Forward launches also park lanes whose ray leaves its block and run
     * their block changes together (16 waiting, no lane marching, or with
     * such a phase); this does not change any history. */
    int32_t event_batch;
    /* register budget of the persistent kernel: minimum waves per SIMD the
     * compiler must fit (0 = default: FP32 4, FP64 unconstrained; 1 =
     * unconstrained; FP32: 4,5,6; FP64: 3,4) */
    int32_t waves_per_simd;
    /* 1 = march the packets of a launch in emission-direction order: each
     * packet's key is its band and its emission direction (up to 2^30
     * equal-solid-angle cells of the (mu, phi) square in Morton order, from
     * the emission draws), radix-sorted before the transport kernel, so the
     * lanes of a wave fly neighbouring rays: their steps stay in lockstep and
     * equal-cell tallies of neighbouring lanes are summed before one atomic.
     * With the mixed-precision persistent kind, a wave takes new packets
     * only once all its lanes are idle (lanes that start together stay in
     * step) and main-path J deposits are summed over runs of lanes without
     * the per-lane FP32 accumulator: faster on optically thin bands, slower
     * on thick ones (kp_autotune measures both).
     * Histories are unchanged (end_state stays indexed by slot); tallies
     * differ by summation order only.  Takes about 26 B of scratch per
     * packet; launches of 2^31 packets or more, or of more than 2^22 bands,
     * run in slot order.  0 = slot
     * order.  Engines on spherical meshes start with the one-thread-per-packet
     * schedule in this order. */
    int32_t coherent;
    /* kp_adjoint, persistent schedule, either precision: 1 (default) =
     * one-pass replay.  Each packet is marched once and its gradient deposits are
     * deferred: every main-path vertex (emission or scatter) is queued as a
     * record (88 B in FP32, 128 B in FP64: the interaction's cell, albedo
     * and scattering angle, the sums of the packet's peel-off contributions
     * before and after its peel-offs, and the start of the next segment),
     * every peel-off ray as a 32 B (FP32) or 48 B (FP64) record.  Once the packets' sums are known, one kernel deposits
     * the vertex terms and marches the segments, another marches the
     * peel-off rays in image-plane order, lanes in lockstep, summing
     * equal-cell deposits of neighbouring lanes before one atomic.  Packets
     * whose records do not fit are replayed by the two-pass kernel.
     * 0 = two-pass replay only: every packet is marched twice, the second
     * time with its deposits.  Same deposits, different summation order. */
    int32_t peel_queue;
    /* Queue capacity of the one-pass replay in main-path vertices per
     * packet (0 = 32), with one peel-off ray per vertex and observer: about
     * 2.8 KB + 1.3 KB per observer for each packet of a chunk in FP32
     * (4 KB + 1.8 KB in FP64).  Runs whose queues would take more than half
     * of the free device memory, or 48 GB, are processed in chunks of
     * packets.  The engine keeps the queues until it is destroyed. */
    int32_t vertex_slots;
    /* Largest chunk of the one-pass replay, in packets (0 = as large as the
     * queue memory allows; otherwise at least 1024). */
    int32_t adjoint_chunk;
    /* KP_SCHED_BLOCKQ: first band of the block queue (0 = every band; the
     * bands before it run on the persistent kernel), passes before the
     * packets still alive finish one thread each (0 = 32), sorted packets
     * per workgroup task (0 = 2048), and the march steps plus interactions a
     * packet may take in one pass before it goes back to the queue in its
     * block (0 = 64), which bounds the slowest lane of a pass. */
    int32_t queue_from_band;
    int32_t queue_passes;
    int32_t queue_chunk;
    int32_t queue_yield;
} kp_schedule;

kp_status kp_set_schedule(kp_engine *eng, const kp_schedule *sched);
kp_status kp_get_schedule(const kp_engine *eng, kp_schedule *sched);

/* Emit, march, scatter and peel off every packet of every band in one
 * launch on `stream` (asynchronous).
 * On an engine made from a spherical mesh this is the reference's spherical
 * photon (photon_sph.h:10-266 driven by photon.h:104-125): pure streaming
 * from the (Cartesian) source point with the path-length tally, in FP64;
 * kappa_scale must be 0, n_incl 0 and precision KP_F64, and the medium is not
 * needed.  end_state, counters, graphs and both schedules work as above. */
kp_status kp_forward(kp_engine *eng, const kp_source *src, const kp_run *run,
                     kp_tallies *out, void *stream);

/* Autotuner: times candidate schedules (kind and packet order, action-phase
 * batch, register budget, replica counts, blocks per CU, block size;
 * coordinate descent from the current schedule; then the block queue for
 * every band and for the longest tail of bands whose extinction scale is at
 * most 1, 0.5, 0.25 or 0.1, and, if one of them wins, its task size, pass
 * count, yield and blocks per CU) on a sample of about sample_packets
 * packets of `run` with HIP events on `stream`, installs the fastest and
 * returns it.  Results are cached per workload shape (precision, bands, observers, image size,
 * scattering model, mesh size, packet-count octave); force != 0 re-times.
 * Blocking call.  No reference counterpart (the reference takes n_th from
 * the input file, driver_base.h:30). */
kp_status kp_autotune(kp_engine *eng, const kp_source *src, const kp_run *run,
                      uint64_t sample_packets, int32_t force, kp_schedule *best,
                      void *stream);

/* hipGraph of a forward run: n_chunks chunk launches (chunk k draws packet
 * ids packet_offset + k*max_b(n_packets_b)) and the accumulator folds,
 * captured once with the engine's current schedule; kp_graph_launch replays
 * the whole sequence with one call and ADDS into the tallies given at
 * capture (same packet ids every replay).  The graph keeps its own band
 * table and launch scratch, so eager calls on the engine do not disturb it;
 * it reads the engine's medium, which kp_set_media rewrites in place.
 * end_state requires n_chunks == 1.  The engine must outlive its graphs.
 * Replaces the per-cycle host loop of the reference (driver.h:59-79). */
typedef struct kp_graph kp_graph;
kp_status kp_graph_create(kp_engine *eng, const kp_source *src, const kp_run *run,
                          kp_tallies *out, int32_t n_chunks, kp_graph **graph);
kp_status kp_graph_launch(kp_graph *graph, void *stream);
void kp_graph_destroy(kp_graph *graph);

/* ================================================================== */
/* Packet-pool compaction and per-destination binning                  */
/* ================================================================== */

/* Device workspace (bytes) that kp_pool_keep / kp_bin_by_key need for n
 * records.  No reference counterpart (the reference allocates inside its
 * drivers, particle_pool.h:77-114, comm.h:114). */
size_t kp_compact_workspace_bytes(int64_t n, int32_t n_bins);

/* pool_t::keep_pool + ker_pool_relocate (types/particle_pool.h:191-248):
 * records i with keep[i] == 0 are removed from the device pool of n records
 * of unit_bytes (multiple of 4); the survivors end up in [0, n_keep) and
 * only tail survivors move: the j-th hole of [0, n_keep) (ascending)
 * receives the j-th survivor of [n_keep, n) counted from the end (the
 * reference pairs them in atomic completion order).  *n_keep_dev (device
 * int64) receives n_keep.  Asynchronous on `stream`. */
kp_status kp_pool_keep(void *pool, int64_t n, int32_t unit_bytes, const uint8_t *keep,
                       int64_t *n_keep_dev, void *workspace, size_t workspace_bytes,
                       void *stream);

/* comm_t::reg (boundary/comm.h:118-127): record i with keys[i] in
 * [0, n_bins) gets slot offsets[keys[i]] + (its rank among the records of
 * that key in index order) -- a stable counting sort instead of the
 * reference's atomicAdd sequence numbers; other keys stay (slot -1).
 * offsets: device int64 [n_bins+1] (offsets[n_bins] = records binned);
 * order: device int32 [n] (order[s] = record in slot s, s < offsets[n_bins]),
 * slot: device int32 [n]; order and slot may be NULL.  n < 2^31. */
kp_status kp_bin_by_key(const int32_t *keys, int64_t n, int32_t n_bins, int64_t *offsets,
                        int32_t *order, int32_t *slot, void *workspace,
                        size_t workspace_bytes, void *stream);

/* Largest n that kp_bin_by_key_static takes. */
#define KP_BIN_STATIC_MAX (1 << 21)

/* kp_bin_by_key that allocates nothing and takes no workspace: for callers
 * that cannot allocate device memory (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Its workspace is a fixed device array inside the library,
 * about 34 MB on each GPU it runs on, held until the process ends; n above
 * KP_BIN_STATIC_MAX is rejected with KP_ERR_ARG.  Every call on one GPU
 * uses that GPU's array: calls that overlap on one GPU (other streams or
 * host threads) race and must be serialized by the caller.  Results equal
 * kp_bin_by_key's on the same input. */
kp_status kp_bin_by_key_static(const int32_t *keys, int64_t n, int32_t n_bins,
                               int64_t *offsets, int32_t *order, int32_t *slot, void *stream);

/* comm_t::cpy_d (boundary/comm.h:226-239): dst[s] = src[order[s]] for
 * s < min(*n_dev, n_max), records of unit_bytes (multiple of 4); n_dev is a
 * device int64 (e.g. offsets + n_bins of kp_bin_by_key). */
kp_status kp_gather(const void *src, int32_t unit_bytes, const int32_t *order,
                    const int64_t *n_dev, int64_t n_max, void *dst, void *stream);

/* Device gradient outputs of kp_adjoint, per cell (each may be NULL);
 * kp_adjoint ADDS into them.  jlin (device scalar) receives the primal
 * value sum(dL_dimage * image) of the replayed histories. */
typedef struct kp_grads {
    double *d_logrho;  /* d/d log(rho_c): kappa_rho of cell c scales with rho_c */
    double *d_albedo;  /* d/d albedo_c (bands with KP_BAND_CELL_ALBEDO)       */
    double *d_g;       /* d/d g_c      (KP_HG_RAYLEIGH, KP_BAND_CELL_G)        */
    double *jlin;
} kp_grads;

/* Gradient of Jlin = sum_pix dL_dimage[pix] . image[pix] (device array in the
 * kp_tallies image layout) by path replay: every history of the run is
 * replayed twice from its Philox key in the run's precision (KP_F64, or
 * KP_F32 with the persistent schedule: the histories of the mixed-precision
 * forward), with FP64 accumulation of A_k, B and the gradients.  Same packets
 * as kp_forward with the same run.  No reference counterpart (the reference
 * has no adjoint); it serves image-fitting callers. */
kp_status kp_adjoint(kp_engine *eng, const kp_source *src, const kp_run *run,
                     const double *dL_dimage, kp_grads *grads, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* KPOLRAD_H */
