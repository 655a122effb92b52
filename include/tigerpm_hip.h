/*
 * tigerpm_hip.h -- C ABI of the MI355X TreePM kick library (libtigerpm_hip.so).
 *
 * The reference (dmarce1/tigerpm) drives the kick through module-global C++
 * functions over global particle arrays.  This ABI keeps the same steps and
 * argument meanings but hands state over explicitly: a context owns the
 * device-resident particles, chain mesh, PM mesh and scratch, and every call
 * returns a status instead of printing (the reference's CUDA_CHECK prints and
 * continues, cuda.hpp:14).  Plain pointers and sizes only.
 *
 * Position layout is the reference's fixed32 SoA (particles.hpp:17-25,
 * fixed.hpp:63-129): x = raw * 2^-32 in the unit periodic box.
 * Calls on one context are serialised on the context's HIP stream; different
 * contexts may be used from different host threads.
 */
#ifndef TIGERPM_HIP_H_
#define TIGERPM_HIP_H_

#include <stdint.h>

#if defined(__GNUC__)
#define TPM_API __attribute__((visibility("default")))
#else
#define TPM_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tpm_ctx tpm_ctx;
typedef int tpm_status;

#define TPM_OK 0
#define TPM_ERR_HIP (-1)      /* a HIP runtime call failed (message in tpm_last_error) */
#define TPM_ERR_INVALID (-2)  /* invalid argument or parameter combination */
#define TPM_ERR_STATE (-3)    /* call out of order (e.g. kick before chain mesh / PM) */
#define TPM_ERR_FFT (-4)      /* rocFFT failure */
#define TPM_ERR_BUDGET (-5)   /* device memory budget exceeded */

/* Long-range solver types (GRAVITY_LONG_PM / GRAVITY_LONG_PME). */
#define TPM_LONG_PM 0
#define TPM_LONG_PME 1

/* Mass-assignment orders. */
#define TPM_CIC 2
#define TPM_TSC 3
#define TPM_CLOUD4 4

/* Kick flags (bit set = stage skipped); 0 = the reference's full kick. */
#define TPM_SKIP_PP 1u   /* no short-range particle-particle term */
#define TPM_SKIP_PM 2u   /* no mesh interpolation term */
#define TPM_SKIP_KICK 4u /* store g / phi only: velocities and rungs untouched */
/* tpm_kick flag: kick_pme_kernel's pair law (kick_pme.cu:280-310) instead of
 * kick_treepm's -- every pair of the 27 cells with no cutoff test, plain
 * softened polynomials inside hsoft, phi starting at 0 (no phi0). */
#define TPM_KICK_PME 16u
/* tpm_kick flag: kick_treepm's tree walk (kick_treepm.cu:586-713) for the
 * short-range term -- per chain cell a tree of <= 32-particle leaves
 * (tree.cpp:7-245); per sink bucket of <= 128 particles, nodes beyond rcut
 * are dropped, nodes with R > (r_src + r_snk) / 0.4 and leaves beyond
 * r_src / 0.4 of every active sink act through their monopole + dipole
 * (compute_pc_interaction, kick_treepm.cu:226-240), the other leaves pair by
 * pair.  Builds the trees of the current chain mesh (tpm_tree_build). */
#define TPM_KICK_TREE 32u

/* Physical and numerical parameters.  Replaces the options struct
 * (options.hpp:5-24, derived in options.cpp:81-96) and the scalars that
 * kick_treepm copies into treepm_params (kick_treepm.cu:833-853). */
typedef struct {
    int mesh_dim;  /* M, PM mesh cells per side (reference four_dim) */
    int chain_dim; /* C, chain-mesh cells per side; 0 = floor(1/rcut) */
    int order;     /* TPM_CIC / TPM_TSC / TPM_CLOUD4 */
    int do_phi;    /* accumulate the potential (reference do_phi) */
    int long_type; /* TPM_LONG_PM: 1/k^2 only; TPM_LONG_PME: times exp(-k^2 rs^2)
                      (gravity_long_type, gravity_long.hpp:17-19) */
    int interlace; /* 1: add a deposit on the grid displaced by half a cell and
                      combine in k space (cancels odd aliases; not in the
                      reference, whose single deposit is interlace = 0) */
    double rs;     /* split scale in box units (reference rs = 1/four_dim) */
    double rcut;   /* short-range cutoff in box units (reference 1/chain_dim) */
    double hsoft;  /* softening length (reference 2/(25 parts_dim)) */
    double GM;     /* multiplies g and phi; 1/N for masses m = 1/N */
    double eta;    /* time-step parameter (reference 0.2/sqrt(2)) */
    double n_total;/* unit masses in the box (reference parts_dim^3) */
} tpm_params;

/* FP64 global sums over every resident particle of a box, with masses
 * m = GM (no reference counterpart; the reference prints no conserved
 * quantities): momentum m sum v, force m sum g, kinetic energy
 * m/2 sum |v|^2, potential energy m/2 sum phi.  Fixed reduction order: the
 * same particles in the same order give the same bits. */
typedef struct {
    double momentum[3];
    double force[3];
    double kinetic;
    double potential;
} tpm_moments;

/* Per-call statistics (the reference prints similar diagnostics,
 * kick_treepm.cu:1084-1087).  Times are HIP-event durations on the context
 * stream, milliseconds. */
typedef struct {
    double candidate_pairs;  /* W_sr = sum_c n_c * sum_{27 neighbours} n_c' */
    double executed_pairs;   /* pairs evaluated after bounding-box pruning */
    double ms_chainmesh;
    double ms_pm;
    double ms_pp;            /* the fused PP + interpolation + kick kernel */
    double ms_total;
    int64_t n_active;
    int64_t n_tiles;
    int max_rung;
    int pad_;
    double pc_interactions;  /* TPM_KICK_TREE: sink x multipole interactions */
    tpm_moments moments;     /* after the kick, over every resident particle */
} tpm_stats;

/* Launch-shape tuning (no reference counterpart: the reference fixes 32-thread
 * blocks and 128-sink buckets, tigerpm.hpp:29-32).  Results do not depend on
 * it beyond FP32 summation order.
 *   tile_sinks:    sinks per PP work item, 8, 16, 32 or 64 (one wave64 walks
 *                  64/tile_sinks source streams of one culled source list);
 *   deposit_block: chain cells per edge of an LDS-privatised deposit block,
 *                  0 = the largest edge <= 3 whose int64 footprint fits
 *                  80 KiB, k > 0 = the largest edge <= k that fits a
 *                  workgroup's 160 KiB, -1 = per-particle global atomics.
 *                  The integer grid is the same bits for every choice. */
typedef struct {
    int tile_sinks;
    int deposit_block;
    int pp_waves;  /* PP tiles per workgroup, one wave each: 1 or 4 (0 =
                      default); the waves of one workgroup share a CU and its
                      L1 cache, so neighbouring tiles' source loads meet there */
    int pp_lists;  /* 16-sink tiles: 1 = per-sink source lists (each sink
                      evaluates only the staged sources within the cull
                      distance of itself, not of the tile's box), 0 = one
                      list per tile, -1 = default */
    int pp_newton; /* whole-box kicks with every particle active and
                      chain_dim >= 8: 1 = each pair evaluated once, its
                      force and potential added to both ends (an extra
                      240 B per particle of device memory), 0 = from each
                      end as the reference does, -1 = default */
    int fft;       /* mesh FFTs: 0 = this library's LDS-tiled transform for
                      power-of-two meshes 16..1024 and rocFFT otherwise,
                      1 = rocFFT for every size.  Both agree to FP32
                      rounding; a batch uses the default. */
} tpm_tuning;

/* Context lifetime.  device = HIP device ordinal. */
TPM_API tpm_status tpm_ctx_create(int device, tpm_ctx** ctx);
TPM_API tpm_status tpm_ctx_destroy(tpm_ctx* ctx);
TPM_API const char* tpm_last_error(const tpm_ctx* ctx);
TPM_API tpm_status tpm_sync(tpm_ctx* ctx);

/* Parameters; validates and derives chain_dim.  Replaces get_options(). */
TPM_API tpm_status tpm_set_params(tpm_ctx* ctx, const tpm_params* params);
TPM_API tpm_status tpm_get_params(const tpm_ctx* ctx, tpm_params* params);
/* Changing tile_sinks invalidates the chain mesh (call tpm_chainmesh_create). */
TPM_API tpm_status tpm_set_tuning(tpm_ctx* ctx, const tpm_tuning* tuning);
TPM_API tpm_status tpm_get_tuning(const tpm_ctx* ctx, tpm_tuning* tuning);

/* Particle store (particles.hpp:17-25 SoA).  Host pointers; the copy is
 * synchronous.  vx/vy/vz/rung may be NULL (zeros). */
TPM_API tpm_status tpm_particles_upload(tpm_ctx* ctx, int64_t n, const uint32_t* x, const uint32_t* y,
                                const uint32_t* z, const float* vx, const float* vy, const float* vz,
                                const int8_t* rung);
/* Any output pointer may be NULL. */
TPM_API tpm_status tpm_particles_download(tpm_ctx* ctx, uint32_t* x, uint32_t* y, uint32_t* z, float* vx, float* vy,
                                  float* vz, int8_t* rung, float* gx, float* gy, float* gz, float* pot);
TPM_API int64_t tpm_particles_size(const tpm_ctx* ctx);

/* Particle order the caller sees.  The device always works in its own
 * order: after a chain mesh, cells z-major and inside a cell by 4x4x4
 * sub-cell Morton code (the PP kernel's tiles need compact groups of sinks).
 *   TPM_ORDER_DEVICE (default): the caller sees that order -- downloads,
 *     tpm_chainmesh_perm and the sample indices of tpm_ewald use it.
 *   TPM_ORDER_REFERENCE: every tpm_chainmesh_create also computes, on the
 *     device, the order in which chainmesh_create's recursive partition
 *     (chainmesh.cpp:168-213 with the two-pointer particles_sort,
 *     particles.cpp:414-445) leaves the caller's previous order, bit for
 *     bit; the caller sees that order from then on: downloads are gathered
 *     into it, tpm_ewald's sample indices refer to it and
 *     tpm_chainmesh_perm returns the reference's permutation.  Uploads and
 *     initial conditions define a new caller order (identical to the
 *     device's).  Costs 3 ceil(log2 C) scan passes over the particles per
 *     chain-mesh build.
 * The caller's order replaces particles.hpp's global arrays' order, which
 * chainmesh_create sorts in place in the reference. */
#define TPM_ORDER_DEVICE 0
#define TPM_ORDER_REFERENCE 1
TPM_API tpm_status tpm_set_host_order(tpm_ctx* ctx, int order);

/* chainmesh_create (chainmesh.cpp:139-166, sort :168-213): reorders the
 * resident particles into chain-cell order (cells z-major, the reference's
 * partition order; in-cell order by 4x4x4 sub-cell Morton code then previous
 * order) and builds the cell table.  Cell ranges are the reference's. */
TPM_API tpm_status tpm_chainmesh_create(tpm_ctx* ctx);
/* chainmesh_get (chainmesh.cpp:48-51): cell (ix, iy, iz), periodic. */
TPM_API tpm_status tpm_chainmesh_get(tpm_ctx* ctx, int ix, int iy, int iz, int64_t* pbegin, int64_t* pend);
/* All C^3 + 1 cell offsets, cell id = (iz*C + iy)*C + ix. */
TPM_API tpm_status tpm_chainmesh_cells(tpm_ctx* ctx, int64_t* cell_begin);
/* perm[new] = index before the last tpm_chainmesh_create, both in the
 * caller's order (TPM_ORDER_REFERENCE: the reference partition's
 * permutation). */
TPM_API tpm_status tpm_chainmesh_perm(tpm_ctx* ctx, int64_t* perm);

/* gravity_long_compute (gravity_long.cpp:36-48) with the PME filter
 * (type GRAVITY_LONG_PME): int64 fixed-point deposit, R2C FFT, Green's
 * function, C2R FFT; phi stays on the device. */
TPM_API tpm_status tpm_gravity_long_compute(tpm_ctx* ctx);
/* gravity_long_get_phi (gravity_long.cpp:19-34): the whole periodic mesh,
 * [x][y][z] z fastest, M^3 floats. */
TPM_API tpm_status tpm_gravity_long_get_phi(tpm_ctx* ctx, float* phi);
/* The integer density grid of the last deposit (M^3, Q32 unit masses). */
TPM_API tpm_status tpm_gravity_long_get_density(tpm_ctx* ctx, uint64_t* grid);
/* The FP32 Green's table in use, M*M*(M/2+1) floats. */
TPM_API tpm_status tpm_gravity_long_get_green(tpm_ctx* ctx, float* table);
/* fft3d_execute / fft3d_inv_execute (fft.cpp:51-84) on host arrays, with
 * the PM's transforms (the tuning's fft choice): r2c takes M^3 floats
 * [x][y][z] z fastest and gives M*M*(M/2+1) complex values (interleaved
 * float pairs), unnormalised; c2r the reverse, unnormalised (M^3 times the
 * inverse).  4 <= M <= 2048.  Uses the context's mesh buffers: the PM mesh
 * must be computed again afterwards. */
TPM_API tpm_status tpm_fft3d_r2c(tpm_ctx* ctx, int M, const float* in, float* out);
TPM_API tpm_status tpm_fft3d_c2r(tpm_ctx* ctx, int M, const float* in, float* out);

/* kick_treepm / kick_pme kernel step (kick_treepm.cu:460-759,
 * kick_pme.cu:128-357): per active sink (rung >= min_rung) the 6-point mesh
 * force, the short-range PP sum over the 27 neighbour chain cells and the
 * kick epilogue.  Needs tpm_chainmesh_create and tpm_gravity_long_compute.
 * stats may be NULL. */
TPM_API tpm_status tpm_kick(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call, unsigned flags,
                    tpm_stats* stats);

/* The FP64 momentum and energy sums of the resident particles' current v,
 * g and phi (the same values tpm_stats.moments holds after a kick). */
TPM_API tpm_status tpm_moments_get(tpm_ctx* ctx, tpm_moments* out);

/* kick_pme(range<int> box, ...) (kick_pme.hpp:15, kick_pme.cu:390-621) and
 * the box argument of kick_treepm (kick_treepm.hpp:14,
 * kick_treepm.cu:761-1089): tpm_kick restricted to the sinks of the chain
 * cells in [lo, hi) (cell coordinates x, y, z; 0 <= lo < hi <= C).  Sources
 * are every resident particle, as for the whole kick, so each sink gets the
 * same bits as in tpm_kick.  stats (may be NULL) count the box's sinks only;
 * stats->moments are over every resident particle. */
TPM_API tpm_status tpm_kick_box(tpm_ctx* ctx, const int32_t lo[3], const int32_t hi[3], int min_rung, double scale,
                                double t0, int first_call, unsigned flags, tpm_stats* stats);
/* Device budget of one kick (the reference's `mem_required > free_mem`
 * test, kick_treepm.cu:815-823, kick_pme.cu:422-429), in bytes; 0 (the
 * default) = no limit.  A kick's working set is what its launch touches:
 * 12 B per source particle of the box padded by one cell, 41 B per sink
 * (position, velocity, rung, force and potential), 16 B per sink tile,
 * 4 B per node of the potential mesh under the padded box, 8 B per cell of
 * the padded box, and on the tree path 48 B per tree node of the padded box
 * and 32 B per sink bucket.  While a box's working set exceeds the budget,
 * tpm_kick, tpm_kick_box and tpm_kick_treepm split it in halves on its
 * longest axis (range::split, range.hpp:118-133) and kick the halves one
 * after the other; every sink gets the same bits as unsplit.  A single cell
 * over the budget is TPM_ERR_BUDGET. */
TPM_API tpm_status tpm_set_budget(tpm_ctx* ctx, int64_t bytes);

/* Per-cell trees (tree_node / sink_bucket, tree.hpp:17-36).  Nodes of one
 * cell are stored in preorder (a node before its subtrees, child 0 first);
 * children are indices within the cell's own tree.  Particle ranges index
 * the tree order, a permutation of the chain-mesh order inside each cell. */
typedef struct {
    uint32_t x[3];      /* expansion centre, fixed32 */
    float radius;       /* bounds the node's particles, plus hsoft */
    int32_t child[2];   /* -1 for a leaf */
    int32_t begin, end; /* particle range, tree order */
    float m[4];         /* particle count and -(sum of offsets from the centre) */
} tpm_tree_node;

typedef struct {
    int32_t begin, end; /* sink range, tree order */
    uint32_t x[3];      /* centre of the sinks' bounding box */
    float radius;
    int32_t cell;       /* chain cell, z-major */
    int32_t pad_;
} tpm_tree_bucket;

/* tree_create (tree.cpp:216-245) for every chain cell, as the loop of
 * kick_treepm_begin (kick_treepm.cpp:26-44) does, on the device: each cell's
 * particles are split at the midpoint of the longest side of its box until
 * <= 32 remain, by the reference's two-pointer particles_sort
 * (particles.cpp:414-445: x < the fixed32 midpoint goes left) starting from
 * the caller's order; sink buckets are the first nodes of <= 128
 * particles.  Needs tpm_chainmesh_create.  A tree deeper than 50 levels is
 * TPM_ERR_INVALID (the reference aborts, tree.cpp:16-19). */
TPM_API tpm_status tpm_tree_build(tpm_ctx* ctx);
/* Sizes of the built trees (any pointer may be NULL). */
TPM_API tpm_status tpm_tree_info(tpm_ctx* ctx, int64_t* n_nodes, int64_t* n_buckets);
/* Copies the trees to the host: node_begin and bucket_begin get C^3 + 1
 * offsets, order[i] = index (in the caller's order) of the particle at tree
 * position i.  Any pointer may be NULL. */
TPM_API tpm_status tpm_tree_download(tpm_ctx* ctx, int64_t* node_begin, int64_t* bucket_begin, tpm_tree_node* nodes,
                                     tpm_tree_bucket* buckets, int32_t* order);

/* tree_create's in-place reordering (tree.cpp:101-107 calls particles_sort
 * on the caller's arrays): with TPM_ORDER_REFERENCE the caller's order
 * becomes the trees' order, cell by cell, as kick_treepm_begin leaves the
 * reference's global arrays (kick_treepm.cpp:26-44).  The trees are built
 * from the caller's order (tpm_tree_build), so their order is the
 * reference's two-pointer partitions of it.  Later downloads use it. */
TPM_API tpm_status tpm_tree_sort_particles(tpm_ctx* ctx);

/* One whole step: chain mesh + PM + kick (kick_treepm_test without IC and
 * comparison, test.cpp:237-290). */
TPM_API tpm_status tpm_kick_treepm(tpm_ctx* ctx, int min_rung, double scale, double t0, int first_call,
                           tpm_stats* stats);

/* Launch timing log (the reference prints wall-clock timers per phase,
 * kick_treepm.cu:826-1087).  enable != 0 clears the log and records HIP
 * events on the context stream around every later chain-mesh / PM / PP stage
 * of tpm_kick_treepm and around the PP kernel of tpm_kick; enable == 0 stops.
 * tpm_profile_read waits for the recorded work and returns sums in ms:
 * ms[0] chain mesh, ms[1] PM, ms[2] PP kernel, ms[3] whole steps, and the
 * number of PP launches and of whole steps recorded. */
TPM_API tpm_status tpm_profile(tpm_ctx* ctx, int enable);
TPM_API tpm_status tpm_profile_read(tpm_ctx* ctx, double* ms, int64_t* pp_launches, int64_t* steps);

/* Issue-rate ceiling of the pair law.  No counterpart in the reference: a
 * measurement aid for the roofline.  Times the direct kernel's pair body
 * alone (the same device function, 16-sink x 4-stream lane layout, flush
 * loop and 8-wave launch bound) over 256 sources held in LDS, with the
 * current parameters' constants, at waves_per_simd (1..8) waves per SIMD,
 * and returns pair evaluations per second.  Allocates a small buffer. */
TPM_API tpm_status tpm_probe_pair_rate(tpm_ctx* ctx, int waves_per_simd, double* pairs_per_second);

/* FP64 Ewald check (gravity_short_ewald_call_kernel, gravity_short.cu:10-91,
 * reformulated): for each sample (an index into the current particle
 * order) the periodic potential and force of all resident unit masses,
 * softened inside hsoft, excluding the sample itself, plus its own image /
 * background term.  alpha <= 0 picks the split automatically. */
TPM_API tpm_status tpm_ewald(tpm_ctx* ctx, int64_t nsamples, const int64_t* samples, double alpha, double* phi,
                     double* gx, double* gy, double* gz);
/* gravity_short_ewald_call_kernel(sinkx, sinky, sinkz) (gravity_short.hpp:
 * 25-26, gravity_short.cu:10-91) at arbitrary fixed32 positions: the FP64
 * periodic potential and force of all resident unit masses at each position,
 * every particle softened inside hsoft (a particle sitting exactly at the
 * position adds -35/16 h^-1 to phi and nothing to g), no particle left out.
 * Host arrays of nsamples raw positions. */
TPM_API tpm_status tpm_ewald_at(tpm_ctx* ctx, int64_t nsamples, const uint32_t* sx, const uint32_t* sy,
                                const uint32_t* sz, double alpha, double* phi, double* gx, double* gy, double* gz);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): tpm_ewald_at with the sample positions and the four outputs
 * in device memory, enqueued on the context's stream without waiting
 * (tpm_sync waits) and without allocating.  Its work space (wave vectors,
 * coefficients and the structure factor) is one fixed 640 KiB device array
 * per GPU, kept until the process ends: it holds up to 16384 wave vectors,
 * i.e. alpha below 12.08 (alpha <= 0 picks 9.6); a larger alpha, or more than
 * TPM_EWALD_NOALLOC_MAX_SAMPLES samples, is rejected with TPM_ERR_INVALID.
 * Every call on one GPU shares that array, so calls that overlap on one GPU
 * (from different contexts or threads) race. */
#define TPM_EWALD_NOALLOC_MAX_SAMPLES (1 << 20)
TPM_API tpm_status tpm_ewald_at_noalloc(tpm_ctx* ctx, int64_t nsamples, const uint32_t* d_sx, const uint32_t* d_sy,
                                        const uint32_t* d_sz, double alpha, double* d_phi, double* d_gx,
                                        double* d_gy, double* d_gz);

/* Zel'dovich initial conditions on the device: n_side^3 particles on the
 * lattice (i+1/2)/n_side displaced by a Gaussian field with power
 * P(k) ~ k^spectral_index exp(-(k/k_Nyq)^2), scaled so that the rms 3-D
 * displacement is rms_disp lattice cells.  Velocities and rungs zero.
 * Replaces initialize() (initialize.cpp:91-221) for the benchmark workloads. */
TPM_API tpm_status tpm_ic_zeldovich(tpm_ctx* ctx, int n_side, uint64_t seed, double rms_disp, double spectral_index);
/* n uniform random positions from a counter-based RNG (seeded). */
TPM_API tpm_status tpm_ic_uniform(tpm_ctx* ctx, int64_t n, uint64_t seed);

/* ---- Batches of independent boxes (ensemble and serving) ---------------
 * No reference counterpart: the reference kicks one box per call and
 * allocates per call (kick_treepm.cu:829, 1078-1082).  A batch keeps many
 * periodic boxes of different N resident in fixed device arenas sized once
 * (the HBM budget) and kicks all of them together: one chain-mesh sort, one
 * deposit, one Green's multiply and one fused PP + interpolation + kick
 * launch over the (box, tile) work items of every box, plus one pair of
 * FFTs per box.  Every box gets the same result, bit for bit, as the same
 * box kicked alone through tpm_kick_treepm with the same tile size.
 * Boxes join (tpm_batch_add*) and leave (tpm_batch_remove) between steps. */
typedef struct tpm_batch tpm_batch;

/* Arena capacities, fixed at creation. */
typedef struct {
    int64_t max_particles;  /* all resident boxes together */
    int64_t max_mesh_nodes; /* sum of mesh_dim^3 over resident boxes */
    int64_t max_cells;      /* sum of chain_dim^3; 0 = max_mesh_nodes / 64 + 4096 */
    int32_t max_boxes;
    int32_t order;          /* assignment order shared by the boxes */
    int32_t interlace;      /* shared interlace setting */
    int32_t pad_;
} tpm_batch_limits;

/* One box: its parameters and its time stepping.  A step kicks every
 * particle of the box with rung >= min_rung (kick epilogue with step t0,
 * first_call on the box's first step) and then, with TPM_BATCH_DRIFT, moves
 * every particle by v * t0 (a drift step, which the reference does not have:
 * repeated steps form a kick-drift-kick leapfrog). */
typedef struct {
    tpm_params params;
    double t0;
    double scale;
    int32_t min_rung;
    int32_t pad_;
} tpm_box_spec;

/* Launch shape of the batched step (results do not depend on it beyond FP32
 * summation order):
 *   tile_sinks:    8, 16, 32 or 64 sinks per PP work item;
 *   deposit_block: chain cells per deposit block edge, 0 = auto (as in
 *                  tpm_tuning);
 *   sched:         0 = one workgroup per tile (XCD-aware order),
 *                  1 = persistent workgroups taking tiles from per-XCD queues;
 *   graph:         1 = capture the step in a HIP graph and replay it while
 *                  the set of boxes and the tuning stay the same. */
typedef struct {
    int tile_sinks;
    int deposit_block;
    int sched;
    int graph;
    int pp_lists;  /* per-sink source lists in the PP kernel (as tpm_tuning) */
} tpm_batch_tuning;

typedef struct {
    double candidate_pairs; /* W_sr summed over the boxes */
    double executed_pairs;
    double ms_step;         /* HIP-event time of the step on the batch stream */
    double ms_chainmesh;    /* stage times of a step run without a graph; */
    double ms_pm;           /* -1 when the step was a graph replay        */
    double ms_pp;
    int64_t n_particles;
    int32_t n_boxes;
    int32_t max_rung;
} tpm_batch_stats;

#define TPM_BATCH_DRIFT 8u  /* tpm_batch_step flag: drift after the kick */

TPM_API tpm_status tpm_batch_create(int device, const tpm_batch_limits* limits, tpm_batch** batch);
TPM_API tpm_status tpm_batch_destroy(tpm_batch* batch);
TPM_API const char* tpm_batch_last_error(const tpm_batch* batch);
/* Device bytes the arenas of these limits take. */
TPM_API tpm_status tpm_batch_arena_bytes(const tpm_batch_limits* limits, int64_t* bytes);
/* Whether a box of n particles fits the free arena space now (1) or not (0). */
TPM_API tpm_status tpm_batch_fits(tpm_batch* batch, const tpm_box_spec* spec, int64_t n, int* fits);
/* Adds a box from host arrays (vx, vy, vz, rung may be NULL) or from
 * Zel'dovich initial conditions made on the device (as tpm_ic_zeldovich).
 * TPM_ERR_BUDGET when it does not fit.  *id identifies the box. */
TPM_API tpm_status tpm_batch_add(tpm_batch* batch, const tpm_box_spec* spec, int64_t n, const uint32_t* x,
                                 const uint32_t* y, const uint32_t* z, const float* vx, const float* vy,
                                 const float* vz, const int8_t* rung, int* id);
TPM_API tpm_status tpm_batch_add_zeldovich(tpm_batch* batch, const tpm_box_spec* spec, int n_side, uint64_t seed,
                                           double rms_disp, double spectral_index, int* id);
/* Makes what a box of this spec (and Zel'dovich lattice n_side, 0 = none)
 * needs before its first step -- the FFT plans of its mesh size and
 * scratch space for its initial conditions -- so that a server can prepare
 * the sizes it accepts before it takes requests. */
TPM_API tpm_status tpm_batch_prepare(tpm_batch* batch, const tpm_box_spec* spec, int n_side);
/* The box leaves; its arena space is reused after the next step. */
TPM_API tpm_status tpm_batch_remove(tpm_batch* batch, int id);
/* One step of every resident box (flags: TPM_SKIP_* as for tpm_kick, and
 * TPM_BATCH_DRIFT).  Asynchronous unless stats is non-NULL. */
TPM_API tpm_status tpm_batch_step(tpm_batch* batch, unsigned flags, tpm_batch_stats* stats);
TPM_API tpm_status tpm_batch_sync(tpm_batch* batch);
/* Particles and steps taken of a box; its arrays in the box's current order
 * (chain-mesh order after a step).  Output pointers may be NULL. */
TPM_API tpm_status tpm_batch_box_info(tpm_batch* batch, int id, int64_t* n, int* steps);
TPM_API tpm_status tpm_batch_download(tpm_batch* batch, int id, uint32_t* x, uint32_t* y, uint32_t* z, float* vx,
                                      float* vy, float* vz, int8_t* rung, float* gx, float* gy, float* gz,
                                      float* pot);
/* Copies a box's state into a context on the same device (parameters
 * included), e.g. for tpm_ewald or to kick it alone. */
TPM_API tpm_status tpm_batch_to_ctx(tpm_batch* batch, int id, tpm_ctx* ctx);
/* The FP64 momentum and energy sums of one box's particles after its last
 * step (tpm_moments, masses = the box's GM). */
TPM_API tpm_status tpm_batch_moments(tpm_batch* batch, int id, tpm_moments* out);
TPM_API tpm_status tpm_batch_set_tuning(tpm_batch* batch, const tpm_batch_tuning* tuning);
TPM_API tpm_status tpm_batch_get_tuning(const tpm_batch* batch, tpm_batch_tuning* tuning);
/* Autotuner: times the step over the resident boxes for every tile size and
 * schedule (then deposit block edge) with the kick epilogue and the drift
 * off, so the boxes' state does not advance, keeps the fastest tuning and
 * returns it with its step time.  reps = timed steps per candidate. */
TPM_API tpm_status tpm_batch_autotune(tpm_batch* batch, int reps, tpm_batch_tuning* best, double* ms_best);

#ifdef __cplusplus
}
#endif

#endif /* TIGERPM_HIP_H_ */
