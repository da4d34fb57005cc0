/*
 * gegs.h -- C ABI of the MI355X transport library (libgegs_hip.so).
 *
 * The reference has no plugin API: its hot path sits behind the host <->
 * device contract of launch_kernels()/runSimulation() (CUDA_EGS.cu:226,
 * CUDA_EGS.cu:367), with state published through cudaMemcpyToSymbol and
 * errors reported through exit().  This header replaces that contract with a
 * context object, plain pointers and status codes.
 *
 * Ownership: the caller owns every host buffer; the context owns the device
 * buffers it allocates.  One context per host thread; a context is not
 * re-entrant.  No function calls exit().
 */
#ifndef GEGS_H
#define GEGS_H
#include <stddef.h>
#include <stdint.h>
#include "gegs_media.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: reference error families (2xxx input, 8xxx phantom, 10xxx
 * media, output.c:51-89) plus codes the reference does not have. */
#define GEGS_OK 0
#define GEGS_ERR_ARG 2001            /* bad argument / missing input */
#define GEGS_ERR_PHANTOM 8001        /* inconsistent phantom definition */
#define GEGS_ERR_MEDIA 10029         /* media tables missing or invalid */
#define GEGS_ERR_DEVICE 58001        /* HIP runtime failure */
#define GEGS_ERR_QUEUE_OVERFLOW 90001 /* a particle bank overflowed (the reference drops, utility.cu:270) */
#define GEGS_ERR_STATE 90002         /* call out of order (e.g. run before load) */

/* Module ids (CUDA_EGS.h:396-404). */
#define GEGS_M_SECJAWY 1
#define GEGS_M_SECJAWX 2
#define GEGS_M_MLC 3
#define GEGS_M_WEDGE 4
#define GEGS_M_BLOCK 5
#define GEGS_M_PHANTOM 6

/* Secondary collimator jaws (SecJaws, read from the BEAMnrc JAWS input,
 * init.cu:612-688): two tungsten slabs, the Y pair at z in [zminy, zmaxy]
 * and the X pair at z in [zminx, zmaxx] (head frame), inside |r| < rmax.
 * The jaw faces are focused on the source; their openings come per field
 * from gegs_set_jaw_positions.  Regions in the module: 0 vacuum, 1 air,
 * 2 jaw tungsten.  A field starting in the jaws passes Y, then X, then the
 * MLC when one is loaded, then the phantom (CUDA_EGS.cu:1280-1298). */
typedef struct gegs_jaws {
    float rmax, zminy, zmaxy, zminx, zmaxx;
} gegs_jaws;

/* Block and wedge (two BEAMnrc PYRAMIDS modules, read_egsinp_block and
 * read_egsinp_wedge, init.cu:692-905).  Each is the intersection of two
 * prisms given by their cross-sections in the head frame: the block by two
 * quadrilaterals, block_xz (x, z) and block_yz (y, z); the wedge by a
 * polygon of wedge_nver (3..6) vertices across the wedge and the rectangle
 * wedge_bound (|r| <= rmax, the wedge's z range).  As built by the reference
 * (ROTATE_Block_XY and ROTATE_Wedge_XY, CUDA_EGS.h:28-29), the first polygon
 * is tested with the point (y, z) and the second with (x, z).  Regions: 1 air,
 * 4 block, 5 wedge.  Modules are passed in the order jaws, MLC, block, wedge,
 * phantom, skipping those not loaded (CUDA_EGS.cu:1279-1298). */
typedef struct gegs_block_wedge {
    gegs_float2 block_xz[4];
    gegs_float2 block_yz[4];
    gegs_float2 wedge[6];
    gegs_float2 wedge_bound[4];
    int32_t has_block, has_wedge, wedge_nver, pad;
} gegs_block_wedge;

/* Multileaf collimator (DYNVMLC-type, read_egsinp, init.cu:345-611).  Head
 * frame: z from the source; the MLC occupies z in [zmin, zmin + zthick].
 * cross_xz: leaf i's cross-section across the leaves, 14 vertices (x, z) at
 * [16 i, 16 i + 14), sides focused on the source.  cross_yz: the closed leaf
 * pair along the leaf travel (y, z): [0,1] the -rmax edge, [2,16] bank A's
 * rounded end, [17,31] bank B's, [32,33] the +rmax edge (leaf_end,
 * init.cu:256-343).  The scalars drive the leaf index lookup (indexOfLeaf,
 * geom.cu:346-410).  Regions in the module: 0 vacuum, 1 air, 3 MLC tungsten. */
#define GEGS_MLC_LEAVES 60
typedef struct gegs_mlc {
    float zmin, zthick, rmax;
    float x_start, x_end, x_mid_1, x_mid_2;
    float air_gap, wl_full, wl_target, wl_isocenter, wt;
    gegs_float2 cross_xz[GEGS_MLC_LEAVES * 16];
    gegs_float2 boundary_xz[4];
    gegs_float2 boundary_yz[4];
    gegs_float2 cross_yz[34];
} gegs_mlc;

/* voxels per phantom axis accepted by gegs_load_phantom (the reference caps at 128,
   EGSnrc_Parameters.h:140-142) */
#define GEGS_MAX_AXIS 4096
#define GEGS_MAX_CT_AXIS 4096       /* CT input voxels per axis (the reference: 512 x 512 x 270, EGSnrc_Parameters.h:137-139) */

/* One source particle in the head frame, as decoded from an egsphsp MODE0
 * record (CUDA_EGS.cu:40-158): total energy, position with z at the source
 * plane, direction cosines, weight, charge (-1, 0, +1). */
typedef struct gegs_source {
    float e, x, y, z, u, v, w, wt;
    int32_t charge;
} gegs_source;

/* A particle record as the event handlers see it (CUDA_EGS.h:294-322
 * particle_t, with the stream state of its Philox lineage). */
typedef struct gegs_particle {
    float e, x, y, z, u, v, w, wt, dnear;
    uint32_t region;
    int32_t charge;
    int32_t status;          /* event type, particle_status 0..9 */
    uint64_t lin;            /* random-stream lineage */
    uint32_t blk;            /* next Philox block of the lineage */
    uint32_t pad;
} gegs_particle;

/* Beam / field geometry (CUDA_EGS.h:103-104, :278-282; CUDA_EGS.cu:430-432). */
typedef struct gegs_field {
    int32_t change_xyz;          /* CT axis convention (Change_xyz) */
    int32_t module_first;        /* module of source particles (module_order[m_FirstMd]) */
    float source_to_isocenter;   /* SID, cm */
    float cos_angle, sin_angle;  /* gantry rotation */
    float iso_x, iso_y, iso_z;   /* isocentre in phantom coordinates */
} gegs_field;

/* One field of a plan for gegs_run_batched: its beam geometry, the plan whose
 * dose buffer receives its deposits, and the source records it transports.
 * The reference runs one field per runSimulation call (CUDA_EGS.cu:367-900)
 * with the gantry published per field (CUDA_EGS.cu:430-432); here all fields
 * of all plans share one event loop. */
typedef struct gegs_job {
    gegs_field field;            /* beam geometry of the field */
    uint32_t plan;               /* dose buffer: dose[plan * nvox + voxel] */
    uint32_t leaf_set;           /* row of gegs_set_leaf_positions (fields starting in the MLC) */
    uint64_t first;              /* first source record */
    uint64_t count;              /* records = histories of the field */
    uint64_t history_offset;     /* Philox history index of record `first` */
} gegs_job;

/* Options of gegs_run_plain. */
typedef struct gegs_run_options {
    uint64_t seed;               /* Philox-4x32-10 key */
    uint64_t history_offset;     /* global index of the first history */
    uint32_t n_batches;          /* time-sliced statistical batches (1 ... 4096; the batch variance needs >= 2) */
    uint32_t block_size;         /* plain method: threads per block, 0 = default */
    uint32_t grid_size;          /* plain method: blocks, 0 = default */
    uint32_t stack_depth;        /* plain method: per-history particle stack depth, 0 = default */
    uint64_t first;              /* first source record of the run */
    uint64_t count;              /* records to transport, 0 = all from first */
} gegs_run_options;

/* Launch shape and queue policy of the event-queued method; a zero field
 * takes the default.  The reference fixes these at compile time (block and
 * grid sizes and stack sizes in CUDA_EGS.h, init.cu:31-57); here an
 * autotuner (gegs_amd/autotune.py) picks them per workload. */
typedef struct gegs_tuning {
    uint32_t block;              /* threads per block of the single-pass event kernels: 64, 128 or 256 (256) */
    uint32_t grid;               /* blocks of the single-pass event kernels (2048) */
    uint32_t photon_grid;        /* blocks of the photon kernel (enough to fill every CU once) */
    uint32_t electron_grid;      /* blocks of the electron kernel (same) */
    uint32_t photon_chunk;       /* voxel crossings per photon and sweep (32; 0xFFFFFFFF: no limit) */
    uint32_t electron_chunk;     /* condensed-history steps per electron and sweep (16; 0xFFFFFFFF: no limit) */
    uint32_t inject;             /* primaries injected per refill (6 Mi) */
    uint32_t threshold;          /* refill when fewer particles are in flight (6 Mi) */
    uint32_t sweeps_per_graph;   /* sweeps captured in one graph replay (4) */
    uint32_t flight_block;       /* threads per block of the photon and electron kernels: 64 ... 512 (256) */
    uint32_t refill_min;         /* idle lanes of a wavefront that trigger a refill, 1 ... 64 (16) */
    uint32_t refill_slots;       /* bank slots a wavefront of the photon/electron kernels reserves per atomic: 64 ... 4096, multiple of 64 (64) */
    uint32_t sweep_streams;      /* streams a sweep runs on: 2 = the photon side (photon kernel and interactions) and
                                    the electron side concurrently, 1 = one after the other (2) */
    uint32_t tail_particles;     /* once every source record is injected and at most this many particles are in
                                    flight, one launch finishes them, each with its descendants in one thread
                                    (262144, the most; 0xFFFFFFFF: never, the sweeps run to the end) */
    uint64_t capacity;           /* particles per bank, rounded down to a multiple of 32: each bank is 32 equal
                                    append segments, and a full segment counts as overflow (0: twice the
                                    injection plus the threshold, each at most the run's histories, and at
                                    least 4 Mi) */
} gegs_tuning;

/* Event counters (diagnostics; replaces DO_STEP_COUNT / PRINT_PART_NUM). */
typedef struct gegs_counters {
    uint64_t photon_steps;       /* photon free-flight segments */
    uint64_t electron_steps;     /* condensed-history steps */
    uint64_t region_gathers;     /* region records read on voxel entry */
    uint64_t deposits;           /* escore calls with region >= 6 */
    uint64_t interactions[10];   /* pops per event type (particle_status 0..9) */
    uint64_t histories;
    uint64_t overflow;           /* particles lost to a full stack */
} gegs_counters;

typedef struct gegs_ctx gegs_ctx;

/* Lifecycle.  Replaces init()/free_all() (init.cu:1682, init.cu:1713). */
gegs_ctx *gegs_create(int device);
void gegs_destroy(gegs_ctx *ctx);
const char *gegs_last_error(gegs_ctx *ctx);
size_t gegs_sizeof_media_tables(void);

/* Physics tables (copyToDevice, copy.cu:3-277; init_user_photon upload,
 * media.cu:1030-1048).  GEGS_ERR_MEDIA for nmed outside 1..GEGS_MXMED and
 * for nbr_split < 1 (the reference then drops in-flight annihilations and
 * their energy, quirk Q5 at annih.cu:71-75). */
int gegs_load_media(gegs_ctx *ctx, const gegs_media_tables *tables);

/* Voxel phantom and region table (init_regions, init.cu:909-1051; phantom
 * bounds, init.cu:1199-1215).  regions has 6 + nx*ny*nz records: 0 vacuum,
 * 1..5 the head modules' media, then voxels x-fastest. */
int gegs_load_phantom(gegs_ctx *ctx, uint32_t nx, uint32_t ny, uint32_t nz,
                      const float *x_bounds, const float *y_bounds, const float *z_bounds,
                      const gegs_region *regions);

/* Host-only check of n region records, run by gegs_load_phantom: GEGS_OK, or
 * GEGS_ERR_PHANTOM for a medium index >= GEGS_MXMED (other than vacuum), or
 * GEGS_ERR_ARG for a flag bit outside GEGS_F_SUPPORTED.  The reference
 * (init_regions, init.cu:909-1051) sets the flags itself and checks nothing;
 * a caller may validate a region table without a device. */
int gegs_validate_regions(const gegs_region *regions, uint64_t n);

/* Field geometry (gantry symbols, CUDA_EGS.cu:430-432).  A field whose
 * module_first is GEGS_M_MLC starts its particles in the MLC, which needs
 * gegs_load_mlc and a leaf-position row (row 0 for gegs_run_plain/_event);
 * one with GEGS_M_SECJAWY starts them in the jaws (gegs_load_jaws and a
 * jaw-position row), then the MLC if one is loaded. */
int gegs_set_field(gegs_ctx *ctx, const gegs_field *field);

/* MLC geometry (init.cu:345-611). */
int gegs_load_mlc(gegs_ctx *ctx, const gegs_mlc *mlc);

/* Jaw geometry (SecJaws, init.cu:612-688) and the jaw openings of n_sets
 * fields, 4 values per set: x1, x2, y1, y2 in cm at the isocentre plane
 * (the field line of the .mlcinfo list, init.cu:1431).  Set k goes with
 * leaf set k: a job's leaf_set selects both.  The jaw polygons of a field
 * follow setup_secjaws (init.cu:1504-1592). */
int gegs_load_jaws(gegs_ctx *ctx, const gegs_jaws *jaws);
int gegs_set_jaw_positions(gegs_ctx *ctx, const float *x1x2y1y2, uint32_t n_sets);

/* Block and wedge geometry (init.cu:692-905, cudaMemcpyToSymbol of blockXZ,
 * blockYZ, goem_Wed, boundWed, nVer_Wed); has_block / has_wedge switch the
 * modules on (Block_on, Wedge_on, CUDA_EGS.cu:1231-1270).  Fields may then
 * start in GEGS_M_BLOCK or GEGS_M_WEDGE when no earlier module is loaded. */
int gegs_load_block_wedge(gegs_ctx *ctx, const gegs_block_wedge *bw);

/* Opt-in corrections of quirks of the reference (default 0: every quirk kept,
 * results as the reference's).  No counterpart in the reference, which has
 * them compiled in.
 *   GEGS_FIX_Q1: hownear_phantom measures from the particle's own voxel,
 *   region - 6, where the reference takes region - 5, the next voxel
 *   (geom.cu:1487-1490); electron steps are then limited by the true
 *   distance to the voxel boundary. */
#define GEGS_FIX_Q1 1u
/*   GEGS_FIX_Q8: a particle in the X jaw's tungsten heading into the opening
 *   steps by its distance to the opening, where the reference steps by the
 *   boundary distance of the other plane (geom.cu:992). */
#define GEGS_FIX_Q8 2u
/*   GEGS_FIX_Q7: range rejection (media with i_do_rr) compares the kinetic
 *   energy with e_max_rr, where the reference compares the total energy
 *   (electron.cu:1523). */
#define GEGS_FIX_Q7 4u
/*   GEGS_FIX_FP64_FRAME: the head -> phantom frame transform (change_xyz swap,
 *   gantry rotation, isocentre shift; geom.cu:59-109, binary32 in the
 *   reference) in binary64, the rotation from the field's cos/sin pair
 *   normalised to unit length, each coordinate rounded once to binary32.
 *   Not a quirk: an accuracy option (voxel indices can then differ from the
 *   reference's where a binary32 rounding crosses a voxel boundary). */
#define GEGS_FIX_FP64_FRAME 8u
/*   GEGS_FIX_Q2: Bethe-Heitler Coulomb-corrected pair sampling (E >= 50 MeV)
 *   reads bpar(1, medium) at bpar[2 * medium], not bpar[8 * medium]
 *   (pair.cu:137, a read past the medium's pair of values). */
#define GEGS_FIX_Q2 16u
/*   GEGS_FIX_Q3: the Klein-Nishina resampling loop of Compton scattering
 *   stops after 101 passes, as its br_loops cap intends (compton.cu:134-141,
 *   where the counter is never incremented). */
#define GEGS_FIX_Q3 32u
#define GEGS_FIX_ALL 63u
int gegs_set_fixes(gegs_ctx *ctx, uint32_t mask);

/* Leaf positions of n_sets fields, 120 values per set, in the order of the
 * .mlcinfo leaf files (init.cu:1480-1500), cm at the isocentre plane: leaf
 * pair j has its +x end at +pos[j] and its -x end at -pos[j+60] (the
 * reference's INVERSE_MLC_Bank_AB order, CUDA_EGS.h:27, init.cu:1607-1613);
 * a pair with pos[j] + pos[j+60] <= 0 stays closed.  The leaf ends follow
 * init_leaf_pos (init.cu:1594-1662). */
int gegs_set_leaf_positions(gegs_ctx *ctx, const float *positions, uint32_t n_sets);

/* Source particles (load_stack_from_phspfile, CUDA_EGS.cu:40-158): copied to
 * device memory; histories map 1:1 onto records. */
int gegs_set_source(gegs_ctx *ctx, const gegs_source *src, uint64_t n);

/* Plain per-history transport of the source records [first, first+count)
 * (history index = history_offset + record index): the baseline kernel.  dose and dose2 receive, per voxel, the energy deposited (MeV, FP64)
 * and the sum over batches of the squared batch energy, from which the
 * caller forms the batch variance. */
int gegs_run_plain(gegs_ctx *ctx, const gegs_run_options *opt, double *dose, double *dose2);

/* gegs_run_plain without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  dose and dose2 are DEVICE
 * buffers of nx*ny*nz doubles each (same contents as gegs_run_plain's); the
 * field, the head geometry, the batch dose row, the work counters and the
 * per-thread particle stacks live in a fixed device array of
 * GEGS_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until the
 * process ends.  Limit: the batch row (8 B per voxel) plus one block of
 * stacks (128 x stack_depth x 64 B) must fit; the grid shrinks to what fits
 * (histories do not depend on it), and an explicit grid_size that does not
 * fit is an error.  Per-history tallies are not available (GEGS_ERR_STATE if
 * enabled) and gegs_last_kernel_ms reports 0.  Calls on one GPU share its
 * array: calls that overlap on one GPU race.  Nothing is allocated on the
 * device by this call. */
#define GEGS_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int gegs_run_plain_noalloc(gegs_ctx *ctx, const gegs_run_options *opt, double *dose, double *dose2);

/* Same transport through the event-queued path: SoA particle banks in HBM,
 * one kernel per event type, wavefront-compacted queues. */
int gegs_run_event(gegs_ctx *ctx, const gegs_run_options *opt, double *dose, double *dose2);

/* Event-queued transport of several fields and plans in one loop (n_jobs <=
 * 512).  Every batch holds the b-th of opt->n_batches equal slices of every
 * job, so batch statistics hold per plan.  dose and dose2 receive n_plans *
 * nvox values (plan-major).  Of opt only seed and n_batches are used.
 * Per-history diagnostics, if enabled, are indexed job after job. */
int gegs_run_batched(gegs_ctx *ctx, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans,
                     const gegs_run_options *opt, double *dose, double *dose2);

/* Beamlet map for gegs_run_batched and gegs_adjoint: history slot k (job
 * after job, as in gegs_run_batched) belongs to beamlet beamlet[k] <
 * n_beamlets, and gegs_run_batched multiplies its source weight by
 * weight[beamlet[k]].  beamlet == NULL removes the map.  No counterpart in
 * the reference, which has no differentiable path. */
int gegs_set_beamlets(gegs_ctx *ctx, const uint32_t *beamlet, uint64_t n_slots, const float *weight,
                      uint32_t n_beamlets);

/* Gradient of an objective L(D) with respect to the beamlet weights.  The
 * histories of gegs_run_batched (same jobs, same opt->seed) are replayed from
 * their counter-based Philox streams at unit beamlet weight; every deposit E
 * in voxel v of plan p adds dLdD[p * nvox + v] * E to grad[beamlet].  The
 * transport is linear in the weights (they multiply deposits and do not
 * enter any sampling decision), so this is the exact gradient of the
 * estimator for those random numbers.  dLdD: n_plans * nvox values; grad:
 * n_beamlets values (FP64). */
int gegs_adjoint(gegs_ctx *ctx, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans,
                 const gegs_run_options *opt, const double *dLdD, double *grad);

/* Device-resident forms of gegs_run_batched, gegs_adjoint and the beamlet
 * weights, for callers that keep dose, dL/dD, weights and gradients in
 * device memory (the torch.autograd.Function binding): dose, dose2, dLdD,
 * grad and weight are DEVICE buffers of the same sizes; the calls return
 * once the results are in them.  The caller's writes to the inputs must be
 * complete before the call (the library works on its own stream).  No
 * counterpart in the reference, which has no differentiable path.
 * gegs_set_beamlet_weights_device replaces the weights of the map set by
 * gegs_set_beamlets (same n_beamlets). */
int gegs_run_batched_device(gegs_ctx *ctx, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans,
                            const gegs_run_options *opt, double *dose, double *dose2);
int gegs_adjoint_device(gegs_ctx *ctx, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans,
                        const gegs_run_options *opt, const double *dLdD, double *grad);
int gegs_set_beamlet_weights_device(gegs_ctx *ctx, const float *weight, uint32_t n_beamlets);

/* Tuning of gegs_run_event; NULL restores the defaults.  gegs_get_tuning
 * returns the values in force with defaults filled in (grid sizes that fill
 * the device are resolved at run time and reported as 0). */
int gegs_set_tuning(gegs_ctx *ctx, const gegs_tuning *tuning);
int gegs_get_tuning(gegs_ctx *ctx, gegs_tuning *tuning);

/* Per-history diagnostics of the last run (total energy deposited in the
 * phantom per history, FP64; number of transport steps per history), sized
 * by the number of source records (gegs_run_plain, gegs_run_event) or the
 * total histories of the jobs (gegs_run_batched); NULL pointers are skipped.  Available
 * only when the run was made with gegs_set_history_tally(ctx, 1). */
int gegs_set_history_tally(gegs_ctx *ctx, int enable);
int gegs_history_tally(gegs_ctx *ctx, double *edep, uint32_t *steps);

/* Sampler check: run the event handler in[i].status once on each of n
 * particles, with the random stream each particle's lineage defines, and
 * return the particles it pushes (at most max_out per input, in push order;
 * n_out[i] = count) and the energy it scored (edep[i]).  No counterpart in the
 * reference, whose samplers are reachable only through whole-stack kernels
 * (compton.cu:17, photo.cu:1, ...); used to check samplers uniform for
 * uniform against the oracle. */
int gegs_sample_events(gegs_ctx *ctx, const gegs_particle *in, uint32_t n, uint64_t seed,
                       gegs_particle *out, uint32_t max_out, uint32_t *n_out, double *edep);

/* Timing of the event path's photon kernel with HIP events on the context's
 * stream (event-record nodes around every photon kernel launch in the
 * captured sweep graphs).  Enabling resets the accumulators; gegs_kernel_timing
 * returns the summed launch durations (ms) and the number of launches since.
 * Replaces the reference's per-launch host timing (CUDA_EGS.cu:226-365). */
int gegs_set_kernel_timing(gegs_ctx *ctx, int enable);
int gegs_kernel_timing(gegs_ctx *ctx, double *photon_ms, uint64_t *photon_launches);

/* Measured FP64 atomic rate: `reps` passes of one hardware FP64 atomic add
 * per entry of idx (n voxel indices < nvox) into a zeroed nvox array, the
 * atomic the dose tally uses; *ms is the time of the timed passes (after one
 * untimed pass).  No counterpart in the reference; reports the tally's
 * deposits per second against the device's rate at the same conflict pattern
 * (escore's FP64 atomicAdd, scoring.cu:14-24, utility.cu:52-61). */
int gegs_atomic_probe(gegs_ctx *ctx, const uint32_t *idx, uint64_t n, uint32_t nvox, uint32_t reps, float *ms);

/* CT front end (host only, no device or context needed).  Volumes of CT
 * numbers are x-fastest int16: hu[i + nx * (j + ny * k)]; lengths in cm.
 *
 * gegs_ct_subsize replaces SubsizeCT (ctcreate.c:40-274): crops the volume in
 * place to the CT voxels that cover [box_min, box_max] (all six zero: the
 * whole volume; a box outside or empty along an axis keeps that axis whole),
 * updating size and offset (the lower corner).
 *
 * gegs_ct_resample_grid and gegs_ct_resample replace ResampleCT
 * (ctcreate.c:277-823): the output grid has int(extent / thickness) voxels per
 * axis starting at offset, thickness adjusted to fill the CT extent (returned
 * in place), or the reference's input errors 5002 (thickness <= 0), 5003
 * (thickness above the CT extent), 5004 (more than GEGS_MAX_AXIS voxels; the
 * reference's limit is 128, EGSnrc_Parameters.h:140-142).  The resampled
 * value of a voxel is the volume-weighted mean CT number, accumulated in
 * binary32 in the reference's order and truncated to an integer; xb, yb, zb
 * receive out_size + 1 bounds each. */
int gegs_ct_subsize(int32_t size[3], int16_t *hu, const float voxel[3], float offset[3],
                    const float box_min[3], const float box_max[3]);
int gegs_ct_resample_grid(const int32_t size[3], const float voxel[3], float thickness[3], int32_t out_size[3]);
int gegs_ct_resample(const int32_t size[3], const float voxel[3], const int16_t *hu, const float offset[3],
                     const float thickness[3], const int32_t out_size[3], float *xb, float *yb, float *zb,
                     int16_t *out);

/* Counters and timing of the last run. */
int gegs_get_counters(gegs_ctx *ctx, gegs_counters *out);
int gegs_last_kernel_ms(gegs_ctx *ctx, float *ms);

#ifdef __cplusplus
}
#endif
#endif
