/*
 * photonprop_hip.h - C ABI of libphotonprop_hip.so, the MI355X (gfx950) photon
 * propagation Monte Carlo.
 *
 * The library replaces the CUDA.jl propagation path of PhotonPropagation.jl v0.9.3:
 *   - run_photon_prop_no_local_cache!  (src/photon_prop_cuda.jl:889-959)
 *   - make_hit_buffers                  (src/photon_prop_cuda.jl:878-886)
 *   - cuda_propagate_photons!           (src/photon_prop_cuda.jl:429-549)
 * Everything is plain C: POD structs, raw pointers and sizes. Every function
 * returns an int status: PP_OK (0), PP_ERR_HIT_OVERFLOW (1, the reference's
 * PhotonPropOOMException, src/photon_prop_cuda.jl:40, :937-939) or a negative
 * error code; pp_last_error() gives a message for the calling thread.
 *
 * One context per host thread. A context owns its device buffers, its target
 * and spectrum tables and one HIP stream. Contexts are independent.
 */
#ifndef PHOTONPROP_HIP_H
#define PHOTONPROP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_ABI_VERSION 1

/* ---- status codes ---------------------------------------------------- */
#define PP_OK 0
#define PP_ERR_HIT_OVERFLOW 1 /* hit buffer full: PhotonPropOOMException */
#define PP_ERR_INVALID (-1)   /* bad argument                           */
#define PP_ERR_HIP (-2)       /* HIP runtime error                      */
#define PP_ERR_NOMEM (-3)     /* device or host allocation failed       */
#define PP_ERR_STATE (-4)     /* call sequence error (e.g. no targets)  */

/* ---- emitters: one tagged POD for every PhotonSource subtype ----------
 * Field layouts of src/lightyield.jl:503-804. Unused fields are ignored. */
enum pp_source_kind {
    PP_SRC_POINTLIKE_ISOTROPIC = 0,       /* lightyield.jl:549-553 */
    PP_SRC_POINTLIKE_TIME_RANGE = 1,      /* lightyield.jl:592-596 */
    PP_SRC_AXICONE = 2,                   /* lightyield.jl:503-510 */
    PP_SRC_COLLIMATED_ISOTROPIC = 3,      /* lightyield.jl:570-575 */
    PP_SRC_PENCIL = 4,                    /* lightyield.jl:527-534 */
    PP_SRC_EXTENDED_CHERENKOV = 5,        /* lightyield.jl:616-622 */
    PP_SRC_EXTENDED_CHERENKOV_NO_SMEAR = 6, /* lightyield.jl:670-676 */
    PP_SRC_EXTENDED_CHERENKOV_CUSTOM_SMEAR = 7, /* lightyield.jl:725-733 */
    PP_SRC_POINTLIKE_CHERENKOV = 8,       /* lightyield.jl:776-781 */
    PP_SRC_CHERENKOV_TRACK = 9,           /* lightyield.jl:798-804 */
    PP_SRC_KIND_COUNT = 10
};

typedef struct pp_source {
    int32_t kind;          /* enum pp_source_kind                           */
    int32_t reserved0;
    int64_t photons;       /* PhotonSource.photons                          */
    float position[3];     /* m                                             */
    float direction[3];    /* unit vector (axis of beam / cascade / track)  */
    float time;            /* ns; PointlikeTimeRange: time_range[1]         */
    float time_end;        /* PointlikeTimeRange: time_range[2]             */
    float angle;           /* Axicone opening angle (rad)                   */
    float beam_divergence; /* Axicone, Pencil (rad)                         */
    float cut_angle_cos;   /* CollimatedIsotropic                           */
    float length;          /* CherenkovTrack length (m)                     */
    float long_a;          /* LongitudinalParameterisation.a                */
    float long_b;          /* LongitudinalParameterisation.b                */
    float long_lrad;       /* LongitudinalParameterisation.lrad (m)         */
    float beta_a;          /* CustomSmear cherenkov_beta_a                  */
    float beta_b;          /* CustomSmear cherenkov_beta_b                  */
    float reserved1[3];
} pp_source; /* 96 bytes */

/* ---- media -------------------------------------------------------------
 * CASCADIA: CascadiaMediumProperties (src/water_properties.jl:43-100):
 *   Kopelevich scattering length, MixedHGES phase function, Quan-Fry
 *   dispersion. SIMPLE: SimpleMediumProperties (src/simple_medium.jl:3-15),
 *   constant lengths and an HG phase function. */
enum pp_medium_kind { PP_MEDIUM_SIMPLE = 0, PP_MEDIUM_CASCADIA = 1, PP_MEDIUM_ICE = 2 };
enum pp_phase_kind { PP_PHASE_HG = 0, PP_PHASE_MIXED_HG_ES = 1, PP_PHASE_MIXED_HG_SL = 2 };

typedef struct pp_medium {
    int32_t kind;          /* enum pp_medium_kind                     */
    int32_t phase_kind;    /* enum pp_phase_kind                      */
    float hg_g;            /* HG mean cosine                          */
    float es_b;            /* Einstein-Smoluchowski b                 */
    float hg_fraction;     /* probability of the HG branch            */
    float sca_scale;       /* multiplies the scattering length        */
    float abs_scale;       /* multiplies the absorption length (host) */
    /* SIMPLE */
    float sca_len;         /* m                                       */
    float group_velocity;  /* m/ns                                    */
    float phase_refractive_index;
    /* CASCADIA: Kopelevich volume concentrations and Quan-Fry inputs */
    float vol_conc_small;
    float vol_conc_large;
    float salinity;        /* permille                                */
    float temperature;     /* deg C                                   */
    float pressure;        /* atm                                     */
    /* ICE: HomogenousIceProperties (src/ice_properties.jl:26-177):
     *   scattering length = sca_scale (1 - hg_g) / (b_dust_400 (wl/400)^-alpha),
     *   phase/group index polynomials of the ice; phase function MIXED_HG_SL */
    float ice_b_dust_400;     /* 1/m                                  */
    float ice_alpha_sca_dust;
    float reserved[7];
} pp_medium; /* 96 bytes */

/* ---- targets ----------------------------------------------------------
 * Index order is kept: it defines the reference's first-hit rule
 * (src/photon_prop_cuda.jl:497-506).
 * SPHERE     check_intersection(::Spherical)   jl:309-351: centre, radius.
 * RECTANGLE  check_intersection(::Rectangular) jl:353-378: a horizontal
 *            rectangle in the plane z = position[2], centred on position,
 *            sides length[0] (x) and length[1] (y).
 * DISC       check_intersection(::Circular)    jl:380-404: a horizontal disc
 *            in the plane z = position[2], centre position, radius.
 * Planar targets are crossed at d = (z_t - z) / dir_z with 0 <= d <= step
 * and never when dir_z == 0. The reference declares dir_z > 0 through an
 * unchecked `assume`; this library evaluates the same formula for either
 * sign, so a plane is also crossed from above. A target list holding a
 * planar target is searched by testing every target in index order (the
 * reference's loop); the grid search takes spheres only. */
enum pp_shape_kind { PP_SHAPE_SPHERE = 0, PP_SHAPE_RECTANGLE = 1, PP_SHAPE_DISC = 2 };

typedef struct pp_target {
    float position[3];
    float radius;          /* SPHERE, DISC             */
    int32_t shape_kind;    /* enum pp_shape_kind       */
    int32_t module_id;     /* carried, not used on device */
    float length[2];       /* RECTANGLE: length_x, length_y */
} pp_target; /* 32 bytes */

/* ---- spectrum -----------------------------------------------------------
 * Monochromatic (src/spectrum.jl:66-72) or an inverse-CDF table of n knots
 * at u_k = k/(n-1) (src/spectrum.jl:75-83; 600 knots by default,
 * src/photon_propagation_setup.jl:18). */
enum pp_spectrum_mapping {
    PP_SPEC_KNOTS = 0,     /* x = u*(n-1): exact knots                     */
    PP_SPEC_TEXTURE = 1    /* x = u*n - 0.5, clamped: 1-D texture addressing */
};

/* ---- configuration ---------------------------------------------------- */
enum pp_search_kind { PP_SEARCH_AUTO = 0, PP_SEARCH_BRUTE = 1, PP_SEARCH_GRID = 2 };

/* How pp_propagate schedules a source list. The reference launches one
 * kernel per source (src/photon_prop_cuda.jl:925-944). All modes give the
 * same hits (same photon ids and random streams).
 *   PP_BATCH_PER_SOURCE  one launch per source
 *   PP_BATCH_PERSISTENT  one launch for the whole list: waves take chunks of
 *                        cfg.chunk_photons photons of one source from a
 *                        global counter (work stealing); with several
 *                        sources the last two chunks per resident wave go
 *                        out in pieces of 256 photons
 *   PP_BATCH_GRAPH       PP_BATCH_PERSISTENT as a hipGraph (counter clear +
 *                        propagation), instantiated once and re-parameterised
 *                        on later calls
 *   PP_BATCH_AUTO        PP_BATCH_PERSISTENT */
enum pp_batch_mode {
    PP_BATCH_AUTO = 0,
    PP_BATCH_PER_SOURCE = 1,
    PP_BATCH_PERSISTENT = 2,
    PP_BATCH_GRAPH = 3
};

typedef struct pp_config {
    int32_t block_size;         /* threads per block, multiple of 64; 0 = chosen per
                                   call: of 256, 384, 512, 768, 1024 the one that keeps
                                   the most waves resident (LDS of tables and walk
                                   queues), the smallest on a tie */
    int32_t waves_per_simd;     /* grid sizing: blocks = CUs*4*wps*64/block; 0 = auto */
    int32_t search;             /* enum pp_search_kind                         */
    float grid_cell;            /* grid cell edge (m); 0 = auto: the edge (and,
                                   with grid_cell_z = 0, the shape) of least
                                   modelled walk cost                          */
    int32_t deterministic_order;/* pp_copy_hits sorts by (source, photon id):
                                   0 = no, 1 = on the device, 2 = on the host */
    int32_t count_work;         /* instrumented kernel: counts steps, sphere
                                   tests and grid-cell visits (pp_work_counts) */
    int32_t batch_mode;         /* enum pp_batch_mode                          */
    int32_t chunk_photons;      /* photons per work chunk (multiple of 64);
                                   0 = auto: about 32 chunks per resident
                                   wave, at least 256 photons                  */
    float grid_cell_z;          /* grid cell edge along z (m): 0 = cubic cells
                                   (automatic shape if grid_cell is 0 too),
                                   < 0 = one layer over the targets' height
                                   (columns, suited to vertical strings)       */
    int32_t reserved[7];
} pp_config; /* 64 bytes */

/* ---- random stream layout (the reproducibility contract) ---------------
 * Every photon owns a Philox4x32-10 stream: key = (seed lo, seed hi),
 * counter = (word, photon id lo, photon id hi, source index). Words:
 *   PP_CTR_INIT                  draws at emission: [wavelength, e1, e2, e3]
 *                                (e1..e3: the emitter's draws in the order of
 *                                src/photon_prop_cuda.jl:158-276)
 *   PP_CTR_INIT_EXTRA | 1        Cherenkov direction: [phi, track u1, track u2, -]
 *   s = 1..n_steps               step s: [step eta, phase branch, phase cos, phi]
 *   PP_CTR_SAMPLER | sub<<24 | i rejection samplers, block i of sub-stream sub;
 *                                block PP_SAMPLER_BOOST_BLOCK holds the
 *                                shape<1 boost uniform.
 * Uniforms: [0,1) = (x>>8)*2^-24; (0,1] = ((x>>8)+1)*2^-24 where a log follows. */
#define PP_CTR_INIT 0x00000000u
#define PP_CTR_INIT_EXTRA 0x40000000u
#define PP_CTR_SAMPLER 0x80000000u
#define PP_SAMPLER_BOOST_BLOCK 0x00FFFFFFu
#define PP_SUB_LONG_GAMMA 0u
#define PP_SUB_BETA_X 1u
#define PP_SUB_BETA_Y 2u
#define PP_SUB_NORMAL 3u
#define PP_CTR_PMT (PP_CTR_INIT_EXTRA | 2u) /* PMT selection: [u_hit, u_pmt, -, -] */
#define PP_GAMMA_MAX_BLOCKS 64u
#define PP_MAX_STEPS 0x3FFFFFFF

typedef struct pp_ctx pp_ctx;

/* Version of this ABI (PP_ABI_VERSION). */
int pp_abi_version(void);

/* Message of the last error on the calling thread ("" if none). */
const char* pp_last_error(void);

/* Number of visible HIP devices (0 if none). */
int pp_device_count(void);

/* Total and free memory of a device in bytes (sizing of make_hit_buffers'
 * reserved_memory_fraction, src/photon_prop_cuda.jl:879-880). */
int pp_device_memory(int device, int64_t* total_bytes, int64_t* free_bytes);

/* Architecture name (e.g. "gfx950", NUL-terminated, at most arch_len bytes)
 * and compute-unit count of a device: the autotuner's cache key. */
int pp_device_info(int device, char* arch, int arch_len, int32_t* n_cu);

/* Vector-ALU issue rates of a device, for the roofline of the propagation
 * kernel (an addition with no counterpart in the reference): 8 independent
 * dependency chains per lane, 8 waves per SIMD, best of 3 timed launches.
 * *rate: FP32 FLOP/s of v_fma_f32 (no packed FP32), or operations/s of the
 * transcendental v_rcp_f32, or of v_mad_u64_u32 (the Philox product). */
enum pp_microbench_kind { PP_MB_FMA_F32 = 0, PP_MB_TRANS_F32 = 1, PP_MB_MAD_U64 = 2 };
int pp_microbench(int device, int which, double* rate);

/* make_hit_buffers (src/photon_prop_cuda.jl:878-886): create a context on
 * `device` with a device hit buffer of `hit_capacity` records. The host
 * buffer stays with the caller (see pp_copy_hits). cfg may be NULL. */
int pp_create(pp_ctx** out, int device, int64_t hit_capacity, const pp_config* cfg);
int pp_destroy(pp_ctx* ctx);

/* Device memory of one hit record in bytes (capacity sizing): 64, one
 * 64-byte aligned record per hit (photon ids below 2^40 per source, at most
 * 65535 targets). */
int64_t pp_hit_record_bytes(void);

/* Medium used by the next pp_propagate (by-value kernel argument in the
 * reference, src/photon_prop_cuda.jl:438). */
int pp_set_medium(pp_ctx* ctx, const pp_medium* medium);

/* Spectrum: if table == NULL, monochromatic `mono_wl` nm; else n >= 2 knots
 * of the inverse CDF, mapped with `mapping` (enum pp_spectrum_mapping). */
int pp_set_spectrum(pp_ctx* ctx, float mono_wl, const float* table, int n, int mapping);

/* Targets uploaded to the device (src/photon_prop_cuda.jl:911). Up to 65535
 * targets. Spheres are searched through a uniform grid held in LDS, or read
 * from device memory when the records and tables exceed it; lists with
 * planar targets are scanned from LDS (4096 targets at most, two 16-byte
 * records each beside a spectrum of up to 8192 knots), as is a list under
 * PP_SEARCH_BRUTE (about 8000 spheres, fewer with a long spectrum). */
int pp_set_targets(pp_ctx* ctx, const pp_target* targets, int n);

/* run_photon_prop_no_local_cache! (src/photon_prop_cuda.jl:889-959) without
 * the hit copy-back: propagates every source, each photon from its own
 * Philox4x32-10 stream (key = seed, counter = (draw, photon id, source)).
 * Photon count per source: floor(float(photons) * photon_scaling) as in
 * src/photon_prop_cuda.jl:456. Hits are reset at the start of each call.
 * Returns PP_ERR_HIT_OVERFLOW if the buffer overflowed; then *n_hits_out is
 * the capacity and *n_ph_sim_out counts only completed photons. */
int pp_propagate(pp_ctx* ctx, const pp_source* sources, int n_sources, uint64_t seed,
                 int32_t n_steps, float photon_scaling, int64_t* n_hits_out,
                 int64_t* n_ph_sim_out);

/* Copy hits [first, first+n) into caller-owned host SoA columns (the
 * PhotonHit fields, src/photon_prop_cuda.jl:407-416). Any pointer may be NULL
 * to skip that column. pos3/dir3/init_dir3 are n x 3 row-major. time is float
 * if time_f64 == 0, else double. module_ix is the 1-based target index.
 * photon_id / source_ix identify the emitting photon (deterministic key).
 * With cfg->deterministic_order the records are ordered by (source, photon):
 * with 1 the first such copy after a propagation sorts them on the device (a
 * radix sort into a buffer of about 88 bytes per hit that the context keeps
 * until pp_destroy; on the host beyond 2^24 sources or 2^31 hits or when that
 * buffer does not fit), with 2 on the host (every record is copied each
 * call); later copies of the same propagation reuse the order. */
int pp_copy_hits(pp_ctx* ctx, int64_t first, int64_t n, float* pos3, float* dir3,
                 float* init_dir3, float* wavelength, void* time, int time_f64,
                 float* dist_travelled, int64_t* module_ix, int32_t* n_steps,
                 uint64_t* photon_id, int32_t* source_ix);

/* Device time (ms, HIP events on the context stream) of the propagation
 * kernels of the last pp_propagate, and the number of kernel launches. */
int pp_last_timing(pp_ctx* ctx, float* kernel_ms, int32_t* n_launches);

/* Frees the buffers a context grows on demand and keeps between calls: the
 * device buffer of the ordered copy (about 88 bytes per hit), the pinned
 * staging of pp_copy_hits (at most 64 MB) and the device buffers of
 * pp_pom_pmt_hits_ctx. They are allocated again by the next call that needs
 * them. No counterpart in the reference (its buffers are the caller's). */
int pp_release_scratch(pp_ctx* ctx);

/* The context's hipStream_t, as an opaque pointer. */
void* pp_stream(pp_ctx* ctx);

/* Target search of the last pp_propagate (enum pp_search_kind) and, for the
 * grid, its cell edge (m), cell counts dims[3] and number of cell entries.
 * The grid replaces the reference's linear scan (src/photon_prop_cuda.jl:
 * 497-506) with the same first-in-index result. */
int pp_search_info(pp_ctx* ctx, int32_t* search, float* cell, int32_t* dims, int32_t* n_items);

/* The context's configuration (pp_create's cfg after defaults) and its
 * replacement; pp_set_config checks the same ranges as pp_create and takes
 * effect at the next pp_propagate. */
int pp_get_config(pp_ctx* ctx, pp_config* cfg);
int pp_set_config(pp_ctx* ctx, const pp_config* cfg);

/* ---- PMT-level detection of the POM (src/pom.jl:108-305) -----------------
 * check_pmt_hit(::POM{<:POMAcceptance}) for a batch of photon hits on the
 * host: every hit on a module's sphere is accepted by one of its 16 PMTs or
 * by none. With the hit position relative to the module centre in units of
 * its radius (rel_pos), for PMT j of group g(j) (1 near the equator, 2 at
 * +-57.5 deg), pt_j = acos(rel_pos . pmt_j):
 *   p_j = Rayleigh_pdf(pt_j; sigma_g) / (0.5 sin pt_j) * acc_g(wavelength) / 8
 * (0 where sin pt_j = 0; acc_g linear in the table, 0 outside). The hit is
 * detected if u_hit < sum_j p_j, by PMT j drawn with weights p_j from u_pmt.
 * The draws come from the photon's Philox stream (word PP_CTR_PMT) under
 * key `seed`, so a photon's fate does not depend on the batch. PMT
 * directions: orientation x R x (sin t cos f, sin t sin f, cos t) for the
 * reference's (t, f) of each PMT, R the rotation taking e_z to e_x. */
typedef struct pp_pom_model {
    float sigma_grp1, sigma_grp2;  /* Rayleigh widths (rad), groups 1 and 2 */
    float orientation[9];          /* module orientation, row-major         */
    int32_t n_wl;                  /* acceptance table length (>= 2)        */
    float reserved[4];
} pp_pom_model; /* 64 bytes */

/* PMT directions (16 x 3, row-major) of the model: the reference's
 * get_pmt_positions(target, orientation) for a POM. */
int pp_pom_pmt_directions(const pp_pom_model* model, float* dirs16x3);

/* pmt_id_out[i] = 1..16 for a detected hit, 0 otherwise. rel_pos3: n x 3;
 * wavelength (nm), photon_id and source_ix: n each (the photon's stream);
 * wl_grid, acc1, acc2: model->n_wl each (acceptances of groups 1 and 2).
 * *n_detected (may be NULL) counts the detected hits; *kernel_ms (may be
 * NULL) receives the device time of the detection kernel alone (HIP events;
 * the host copies around it excluded). Runs on `device`. */
int pp_pom_pmt_hits(int device, const pp_pom_model* model, const float* wl_grid,
                    const float* acc1, const float* acc2, int64_t n, const float* rel_pos3,
                    const float* wavelength, const uint64_t* photon_id,
                    const int32_t* source_ix, uint64_t seed, int32_t* pmt_id_out,
                    int64_t* n_detected, float* kernel_ms);

/* The same detection for the hits of the last pp_propagate of `ctx`, read
 * from the context's device hit buffer on its stream (no host round trip of
 * the hits): hits [first, first + n) in the order pp_copy_hits returns them;
 * each hit's module centre and radius come from the context's target list,
 * which must hold spheres. pmt_id_out (host, n int32) receives the PMT ids;
 * *n_detected (may be NULL) counts the detected hits of the range; *kernel_ms
 * (may be NULL) the detection kernel's device time. The context keeps the
 * device buffers (acceptance table, 4 bytes per hit), grown when a call needs
 * more; a call of a size already seen allocates nothing. Replaces the
 * per-module check_pmt_hit loop of make_hits_from_photons
 * (src/processing.jl:88-144, src/pom.jl:246-305). */
int pp_pom_pmt_hits_ctx(pp_ctx* ctx, const pp_pom_model* model, const float* wl_grid,
                        const float* acc1, const float* acc2, uint64_t seed, int64_t first,
                        int64_t n, int32_t* pmt_id_out, int64_t* n_detected, float* kernel_ms);

/* pp_pom_pmt_hits on device arrays, allocating nothing: no counterpart in
 * the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Inputs (rel_pos3: n x 3, wavelength, photon_id, source_ix) and the output
 * pmt_id_out (n int32) are device arrays of `device`; wl_grid, acc1, acc2 are
 * host arrays of model->n_wl <= 1024 knots. The acceptance table and the
 * detected-hit counter live in a static 12.25 KiB device array of the library
 * (one per GPU, held until the process ends); n_wl > 1024 is rejected.
 * Work runs on `stream` (a hipStream_t; NULL: the null stream); with
 * n_detected not NULL the call waits for the stream and writes the count,
 * otherwise it returns without waiting. Calls that overlap on one GPU share
 * that array and race: serialise them (one stream, or one call at a time).
 * Results equal pp_pom_pmt_hits's for the same hits and seed. */
int pp_pom_pmt_hits_noalloc(int device, const pp_pom_model* model, const float* wl_grid,
                            const float* acc1, const float* acc2, int64_t n,
                            const float* rel_pos3, const float* wavelength,
                            const uint64_t* photon_id, const int32_t* source_ix, uint64_t seed,
                            int32_t* pmt_id_out, int64_t* n_detected, void* stream);

/* Autotuner. The reference fixes its launch shape from the occupancy API
 * (src/photon_prop_cuda.jl:914-920); this library searches it instead.
 * Times pp_propagate on a sample of `sources` (photon counts scaled so that
 * the sample holds about sample_photons photons; photon_scaling as in
 * pp_propagate) by coordinate descent over
 *   block size        64, 128, 256, 512, 1024
 *   photons per lane  per work chunk: 1, 2, 4, 8, 16 (chunk_photons = 64 x),
 *                     or the automatic chunk
 *   grid cell edge    automatic, and 0.6x, 0.8x, 0.9x, 1.1x, 1.25x the automatic edge
 *                     (grid search only)
 *   grid cell shape   cubic, twice as tall as wide, or columns over the whole
 *                     height (grid only)
 * with the batch mode of the context (PP_BATCH_AUTO: persistent). Each
 * candidate runs once untimed and twice timed (HIP events); the faster run
 * counts. The fastest configuration is written to *best (if not NULL) and
 * becomes the context's configuration; *best_ms is its kernel time on the
 * sample. Hits of the tuning runs are discarded. Results never depend on
 * the configuration, only the speed does. */
int pp_autotune(pp_ctx* ctx, const pp_source* sources, int n_sources, uint64_t seed,
                int32_t n_steps, float photon_scaling, int64_t sample_photons,
                pp_config* best, float* best_ms);

/* Photon steps, sphere tests, grid-cell visits and grid walks (step
 * segments near the grid that were clipped and walked) executed by the last
 * pp_propagate of a context created with cfg->count_work (roofline
 * accounting; the hits are identical to the uninstrumented kernel's). Any
 * output may be NULL. */
int pp_work_counts(pp_ctx* ctx, int64_t* steps, int64_t* sphere_tests, int64_t* cell_visits,
                   int64_t* grid_walks);

/* ---- step-function evaluation (parity tests) ----------------------------
 * Evaluates one device step function over n explicit inputs on the GPU,
 * with the same inlined code the propagation kernel uses. `fn` selects the
 * function; `in` holds n*in_stride floats, `out` n*out_stride floats
 * (host pointers). Functions and their layouts:
 *   PP_FN_ISOTROPIC_DIR   in (u1,u2)             out (x,y,z)   jl:73-77
 *   PP_FN_ROT_FROM_EZ     in (ax,ay,az,bx,by,bz) out (x,y,z)   jl:92-95
 *   PP_FN_HG_COS          in (g,u)               out (cos)
 *   PP_FN_ES_COS          in (b,u)               out (cos)
 *   PP_FN_UPDATE_DIR      in (dx,dy,dz,cos,uphi) out (x,y,z)   jl:280-298
 *   PP_FN_SPHERE_ISEC     in (cx,cy,cz,r,px,py,pz,dx,dy,dz,step) out (hit,d) jl:309-351
 *   PP_FN_STEP_LEN        in (u,sca_len)         out (step)    jl:488-490
 *   PP_FN_CASCADIA        in (wl)                out (sca_len, c_grp, n_phase)
 *   PP_FN_TRACK_DIR       in (u1,u2)             out (x,y,z)
 *   PP_FN_PHILOX          in (c0..c3,k0,k1 as u32 bits) out (r0..r3 as u32 bits)
 *                         through the steps' per-photon form; all ones if it
 *                         differs from the emission's plain ten rounds
 *   PP_FN_GAMMA           in (k, theta, key_lo, key_hi, pid_lo, pid_hi, src as u32 bits)
 *                         out (draw, accepted normal variate)
 *   PP_FN_SPECTRUM        in (u)                 out (wl)  (uses the context's spectrum)
 *   PP_FN_SL_COS          in (g, u)              out (cos)  simplified Liu
 *   PP_FN_PLANE_ISEC      in (kind, cx,cy,cz, a, b, px,py,pz, dx,dy,dz, step) out (hit,d)
 *                         kind as u32 bits (RECTANGLE: a, b = length_x, length_y;
 *                         DISC: a = radius)                      jl:353-404
 *   PP_FN_MIXED_HGES      in (g, b, hg_fraction, u_choice, u) out (cos)  the
 *                         kernel's branch-free MixedHGES sampler (HG with
 *                         probability hg_fraction, else Einstein-Smoluchowski),
 *                         on the random words of u_choice and u, which must be
 *                         multiples of 2^-24 in [0, 1) (as the kernel's draws);
 *                         NaN if it differs from the float form's bits
 * PP_FN_CASCADIA evaluates the context's medium (any kind): scattering
 * length, group velocity and phase index at wl. */
enum pp_fn_id {
    PP_FN_ISOTROPIC_DIR = 0,
    PP_FN_ROT_FROM_EZ = 1,
    PP_FN_HG_COS = 2,
    PP_FN_ES_COS = 3,
    PP_FN_UPDATE_DIR = 4,
    PP_FN_SPHERE_ISEC = 5,
    PP_FN_STEP_LEN = 6,
    PP_FN_CASCADIA = 7,
    PP_FN_TRACK_DIR = 8,
    PP_FN_PHILOX = 9,
    PP_FN_GAMMA = 10,
    PP_FN_SPECTRUM = 11,
    PP_FN_SL_COS = 12,
    PP_FN_PLANE_ISEC = 13,
    PP_FN_MIXED_HGES = 14,
    PP_FN_COUNT = 15
};
int pp_eval_fn(pp_ctx* ctx, int fn, const float* in, int in_stride, float* out,
               int out_stride, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* PHOTONPROP_HIP_H */
