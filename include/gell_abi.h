/*
 * gell_abi.h -- C ABI of the MI355X per-step hot path (mechanics, diffusion,
 * proliferation/death) of the Gell tumour-spheroid simulator.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no C++ types and
 * no exceptions cross this boundary. Every entry point returns a gell_status
 * (0 = ok, < 0 = error); the message of the last error on a context is read
 * with gell_last_error().
 *
 * Threading: one context is driven by one host thread; all work is enqueued
 * on the context's HIP stream, except that each substrate's LOD substeps may
 * run on a context-owned stream ordered against it
 * (GELL_KNOB_SUBSTRATE_STREAMS below). Entry points that return a host-side
 * count (gell_birth, gell_death) synchronise; the others are asynchronous.
 * Cell arrays are caller-owned device memory; all scratch is context-owned.
 *
 * Reference interfaces replaced (paths relative to the Gell sources):
 *   gell_fm_update   <- MechanicsMesh_struct::FM_update    include/FMkernel.cuh:181-202
 *   gell_prelod      <- Mesh_struct::PreLOD                include/LODkernel.cuh:270-278
 *   gell_lod_substep <- Mesh_struct::LODsolver             include/LODkernel.cuh:280-300
 *   gell_birth       <- CellBirth_kernel                   include/PDkernel.cuh:88-98
 *   gell_death       <- CellDeath_kernel                   include/PDkernel.cuh:100-117
 *   gell_create      <- MechanicsMesh_struct() + Mesh_struct(conc, D, decay)
 *                       include/FMkernel.cuh:167-172, include/LODkernel.cuh:205-225,239-259
 *   gell_field       <- public member Mesh_struct::p       include/LODkernel.cuh:199
 */
#ifndef GELL_ABI_H
#define GELL_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GELL_ABI_VERSION 1
#define GELL_MAX_SUBSTRATES 4

typedef struct gell_f3 {
    float x, y, z;
} gell_f3;

/*
 * The 72-byte agent record, field for field the reference's `struct Cell`
 * (include/Cells.cuh:9-47): bool sign @0 (3 pad bytes), int mesh_idx @4,
 * float3 pos @8, force @20, oldforce @32, float r @44, V @48, Vf @52,
 * Vcs @56, Vns @60, int phase @64, float cell_clock @68. sign is 0 or 1.
 * Phases: 0 quiescent, 1 premitotic, 2 postmitotic, 3 apoptotic,
 * 4 early necrotic, 5 late necrotic (include/Cells.cuh:25-30).
 */
typedef struct gell_cell {
    uint8_t sign;
    uint8_t pad_[3];
    int32_t mesh_idx;
    gell_f3 pos;
    gell_f3 force;
    gell_f3 oldforce;
    float r;
    float V;
    float Vf;
    float Vcs;
    float Vns;
    int32_t phase;
    float cell_clock;
} gell_cell;

/*
 * One diffusing substrate. The reference hard-codes oxygen
 * (include/Definitions.h:21-25); substrate 0 is the one the phase model
 * reads (PDkernel.cuh:37-39).
 */
typedef struct gell_substrate {
    float boundary;    /* Dirichlet value on the two outer layers (O2_Default_Concentration, 38 mmHg) */
    float D;           /* diffusion coefficient, um^2/min (O2_Diffusion_coef, 1e5) */
    float decay;       /* linear decay, 1/min (O2_Decay_rate, 0.1) */
    float uptake;      /* per-cell uptake per unit volume fraction, 1/min (Tumor_O2_consumption, -10) */
    float dead_factor; /* uptake factor for phases >= 3 (Dead_O2_consumption_factor, 0.1) */
    float initial;     /* initial field value (the reference starts at the boundary value) */
} gell_substrate;

/*
 * Runtime mirror of include/Definitions.h:7-47. The mechanics bins are the
 * diffusion voxels, as in the reference (FMkernel.cuh:25, Voxel_length).
 */
typedef struct gell_params {
    int32_t N;            /* voxels per axis (Voxel_num); 2 < N <= 1024 */
    float dx;             /* voxel edge, um (Voxel_length) */
    int32_t capacity;     /* cell-array capacity (Max_Cell_num) */
    int32_t n_substrates; /* 1..GELL_MAX_SUBSTRATES */
    float dt_bio;         /* Biology_dt, min */
    float dt_mech;        /* Mechanics_dt, min */
    float dt_diff;        /* Diffusion_dt, min */
    int32_t reserved_;
    uint64_t seed;        /* key of the Philox4x32-10 stream used by birth */
    gell_substrate sub[GELL_MAX_SUBSTRATES];
} gell_params;

typedef struct gell_ctx gell_ctx;

typedef enum gell_status {
    GELL_OK = 0,
    GELL_E_ARG = -1,         /* bad argument (null pointer, size out of range) */
    GELL_E_HIP = -2,         /* a HIP runtime call failed */
    GELL_E_NOMEM = -3,       /* device allocation failed */
    GELL_E_UNSUPPORTED = -4, /* parameter combination the library does not support */
    GELL_E_NODEVICE = -5     /* no usable gfx950 device */
} gell_status;

/* Fill *p with the reference's Definitions.h values (oxygen only). */
void gell_default_params(gell_params* p);

/* Build the context: cell-list arrays, fields (set to sub[s].initial),
 * uptake maps (zero), Thomas pre-factorisation, sort scratch. `stream` is a
 * hipStream_t (NULL = the null stream). */
int gell_create(const gell_params* params, void* stream, gell_ctx** out);
void gell_destroy(gell_ctx* ctx);
const char* gell_last_error(const gell_ctx* ctx);
const char* gell_status_string(int status);

/* Mechanics step on d_cells[0, n): Morton keys, stable sort of the records
 * by key, per-voxel [s, e] ranges of live cells, adhesion/repulsion forces,
 * Adams-Bashforth-2 movement. In place, asynchronous. */
int gell_fm_update(gell_ctx* ctx, gell_cell* d_cells, int32_t n);

/* Uptake map of one substrate from d_cells[0, n) (per-voxel sum of the
 * live cells' uptake rates, taken at mesh_idx). Asynchronous. */
int gell_prelod(gell_ctx* ctx, int32_t substrate, const gell_cell* d_cells, int32_t n);

/* The mechanics step followed by the uptake map of every substrate: the
 * sequence FM_update, then PreLOD per substrate, that Main.cu:115-120 runs
 * (FMkernel.cuh:181-202, LODkernel.cuh:270-278). The force/movement kernel
 * adds each moved cell's uptake rates to the maps' integer accumulators
 * while it holds the record, so the records are not read a second time.
 * Results equal gell_fm_update followed by gell_prelod for substrates
 * 0 .. n_substrates-1, bit for bit (the integer sums do not depend on
 * order). With a force variant that cannot accumulate (GELL_KNOB_FM_PREFETCH
 * below 7) it makes exactly those calls. Asynchronous. */
int gell_mech_step(gell_ctx* ctx, gell_cell* d_cells, int32_t n);

/* One diffusion substep of one substrate: double-layer Dirichlet boundary,
 * implicit uptake, LOD x/y/z Thomas sweeps. Asynchronous. */
int gell_lod_substep(gell_ctx* ctx, int32_t substrate);

/* Volume/phase update and stochastic division of d_cells[0, n) for biology
 * step `bio_step`; daughters are appended at [n, *n_out). Synchronises. */
int gell_birth(gell_ctx* ctx, gell_cell* d_cells, int32_t n, uint64_t bio_step, int32_t* n_out);

/* Stable removal of records with sign == 0 from d_cells[0, n). Synchronises. */
int gell_death(gell_ctx* ctx, gell_cell* d_cells, int32_t n, int32_t* n_out);

/* gell_death without a context and without allocating: no counterpart in
 * the reference. It is for callers that cannot allocate device memory (code
 * running while a stream is captured, tight serving loops, harnesses that
 * forbid hipMalloc): it calls no hipMalloc, hipFree or allocating library
 * routine. Its scratch (per-block counts, the live count and the compacted
 * records) is one static device array of the library, allocated with the
 * code object and held on every device the library runs on until the process
 * ends (about 72 B x GELL_DEATH_NOALLOC_MAX_CELLS, 9.4 MB). Same kernels and
 * same result as gell_death; `stream` is a hipStream_t (NULL = null stream);
 * synchronises it. n > GELL_DEATH_NOALLOC_MAX_CELLS returns GELL_E_ARG
 * (gell_last_error(NULL) says why). All calls on one device share the
 * array: calls that overlap on one device (other host threads, other
 * streams) race, so callers serialise them. */
#define GELL_DEATH_NOALLOC_MAX_CELLS 131072
int gell_death_noalloc(gell_cell* d_cells, int32_t n, void* stream, int32_t* n_out);

/* Device pointers owned by the context (valid until gell_destroy). Both
 * first order the substrate's own stream (GELL_KNOB_SUBSTRATE_STREAMS) back
 * into the context stream, so work enqueued afterwards on the context stream
 * (or on the null stream, for a context made on it) sees the finished field;
 * a caller on any other stream calls gell_sync first. The caller may write
 * the uptake map: the next gell_prelod of that substrate then clears the
 * whole map instead of only the voxels its previous call wrote. */
float* gell_field(gell_ctx* ctx, int32_t substrate);        /* N^3 floats, x fastest */
float* gell_uptake_field(gell_ctx* ctx, int32_t substrate); /* N^3 floats (p_rate) */

/* Cell-list state of the last gell_fm_update, device pointers:
 * sorted keys (Morton code, -1 outside; n entries), the permutation (sorted
 * slot -> input slot; n entries), and the per-voxel live ranges as
 * interleaved pairs (s[v], e[v]) for v < N^3 in the linear (x fastest)
 * layout -- (capacity+1, -1) when the voxel holds no live cell
 * (Mech_Mesh_s / Mech_Mesh_e, FMkernel.cuh:38-53). The library keeps the
 * ranges indexed by Morton code; gell_voxel_ranges converts them into a
 * context-owned buffer (allocated on first use), synchronises, returns it. */
const int32_t* gell_sorted_keys(gell_ctx* ctx);
const int32_t* gell_sorted_order(gell_ctx* ctx);
const int32_t* gell_voxel_ranges(gell_ctx* ctx);

/* Birth events of the last gell_birth: per input slot, the daughter slot
 * (>= n) or -1. Device pointer, capacity entries. */
const int32_t* gell_birth_slots(gell_ctx* ctx);

int gell_sync(gell_ctx* ctx);

/* Kernel-variant knobs (the autotuner's search space).
 *   GELL_KNOB_LOD_YZ_LDS   0 = y/z sweep lines fully in registers,
 *                          1 = first 64 forward values per line in LDS
 *                              (two waves per SIMD), for N > 64
 *   GELL_KNOB_FM_PREFETCH  0 = one stencil-range load per voxel in turn,
 *                              each candidate loaded one ahead of its
 *                              evaluation,
 *                          1 = all 27 stencil ranges loaded up front, the
 *                              load one candidate ahead running on across
 *                              voxel boundaries,
 *                          2 = each lane walks its own candidate list (a
 *                              27-bit mask of non-empty voxels), one pair per
 *                              lane per step of the wave's loop,
 *                          3 = per x-plane of the stencil, the first two
 *                              candidates of its nine voxels loaded together,
 *                          4 = as 3, with the block's records read and
 *                              written as coalesced 16-byte chunks staged in
 *                              LDS (needs 16-byte aligned cell arrays; 3
 *                              otherwise),
 *                          5 = one workgroup per aligned 8x8x8 Morton block:
 *                              the ranges of its 10^3 halo voxels and all
 *                              their candidates loaded into LDS once, each
 *                              cell's walk read from there,
 *                          6 = the stencil voxels that hold a live cell read
 *                              first from a one-bit-per-voxel occupancy map
 *                              (Morton-indexed: at most eight 8-byte reads),
 *                              only their ranges loaded, each while the
 *                              previous voxel is walked,
 *                          7 = as 6, with workgroups ordered so that each of
 *                              the eight XCDs (each with its own L2) walks
 *                              one contiguous run of the sorted cells, and
 *                              the records read and written as coalesced
 *                              16-byte chunks staged in LDS (needs 16-byte
 *                              aligned cell arrays; 8 otherwise),
 *                          8 = as 7 with per-lane record accesses,
 *                          9 = as 7 on a rank-compacted cell list: per
 *                              aligned 64-voxel Morton word its occupancy
 *                              bits and the rank of its first occupied voxel
 *                              (16 bytes), and per occupied voxel its first
 *                              live slot (4 bytes), instead of the 8-byte
 *                              Morton-indexed range of every voxel,
 *                          10 = 9 for at least 524288 cells (capacity, with
 *                              device-resident counts) and at most one cell
 *                              per 8 grid voxels, 7 otherwise: the list's
 *                              extra launches pay only on large sparse
 *                              tissue. Default 10.
 *                              Variants 7-10 also accumulate the uptake maps
 *                              in gell_mech_step.
 *   GELL_KNOB_LOD_PARTS    y/z sweeps: 1 = one lane solves a whole line (the
 *                              reference's serial recurrence, bit-identical
 *                              to it), 2, 4, 8 or 16 = each line split into
 *                              blocks solved by separate waves and joined
 *                              through the exact reduced system (same linear
 *                              solve, rounding differs at the 1e-7 relative
 *                              level); applies for N in {128, 192, 256, 512},
 *                              a block count the grid cannot split into
 *                              32-voxel chunks falls back to a smaller one;
 *                              at N = 64, 96 and 100 it is the z sweep's
 *                              block count (2 or 4).
 *                              Default 4.
 *   GELL_KNOB_LOD_X_PARTS  the same choice for the x sweep (at most 8); at
 *                              N = 64, 96 and 100 the block count (1, 2 or 4)
 *                              of the x and y sweeps, which run per z plane
 *                              in one workgroup.
 *                              Default 4.
 *   GELL_KNOB_LOD_FUSED    1 = for N in {128, 256, 512} with LOD_PARTS > 1,
 *                              the x sweep and the y blocks run in one pass
 *                              per (z plane, 64 rows), and the y coupling
 *                              is applied inside the z sweep: 20 bytes of
 *                              traffic per voxel instead of 28, x and y in
 *                              N / 64 blocks, z in LOD_PARTS blocks; results
 *                              equal the unfused kernels with those block
 *                              counts bit for bit. At N = 64, 96 and 100
 *                              with serial lines, 1 = the x and y sweeps of
 *                              each z plane in one workgroup, bit-identical
 *                              to the separate sweeps. Default 1.
 *   GELL_KNOB_SUBSTRATE_STREAMS 1 = each substrate's LOD substeps run on a
 *                              stream of its own, overlapping the other
 *                              substrates' substeps and the next mechanics
 *                              step (same results; off while profiling).
 *                              Field and uptake-map pointers are current
 *                              after gell_sync. Default 1.
 * The first two knobs never change results; gell_autotune sets all of them. */
enum {
    GELL_KNOB_LOD_YZ_LDS = 0,
    GELL_KNOB_FM_PREFETCH = 1,
    GELL_KNOB_LOD_PARTS = 2,
    GELL_KNOB_LOD_X_PARTS = 3,
    GELL_KNOB_LOD_FUSED = 4,
    GELL_KNOB_SUBSTRATE_STREAMS = 5,
    GELL_N_KNOBS = 6
};
int gell_tune(gell_ctx* ctx, int knob, int value);
/* Time every effective LOD variant on a scratch copy of substrate 0's field
 * (state is not changed) and keep the fastest; with GELL_TUNE_BIT_EXACT only
 * the variants bit-identical to the serial sweep are considered. The best
 * substep time (microseconds) is written to *lod_us_out when not NULL.
 * Synchronises; allocates one field-sized scratch buffer while it runs. */
#define GELL_TUNE_BIT_EXACT 1
int gell_autotune(gell_ctx* ctx, int flags, double* lod_us_out);
/* Coefficient table of the partitioned line solve for `parts` blocks
 * (GELL_KNOB_LOD_PARTS), host-side: El[N] | 1/Fl[N] | W[N] | V[N] |
 * Rinv[(2 parts)^2]. For inspection and CPU-side checks; needs no device. */
int gell_lod_partition_table(const gell_params* params, int32_t substrate, int32_t parts, float* out, int32_t out_len);
int gell_tune_get(const gell_ctx* ctx, int knob);

/* Per-phase device time, recorded with HIP events on the context stream
 * while profiling is on. Phases: 0 fm_keys, 1 fm_sort, 2 fm_gather,
 * 3 fm_force_move, 4 prelod, 5 lod, 6 birth, 7 death. ms_out and
 * calls_out hold GELL_N_PHASES entries; reading synchronises. */
#define GELL_N_PHASES 8
int gell_profile_enable(gell_ctx* ctx, int enable);
int gell_profile_read(gell_ctx* ctx, double* ms_out, int64_t* calls_out, int reset);

/* Thin device-memory helpers for hosts without their own HIP binding. */
void* gell_dev_alloc(size_t bytes);
void gell_dev_free(void* p);
int gell_copy_h2d(void* dst, const void* src, size_t bytes);
int gell_copy_d2h(void* dst, const void* src, size_t bytes);
int gell_copy_d2d(void* dst, const void* src, size_t bytes);
int gell_device_sync(void);
/* Device-to-device copy by the library's streaming kernel (16-byte accesses)
 * on `stream` (a hipStream_t, NULL = null stream); asynchronous. It is also
 * the known-size kernel the HBM byte counters are calibrated against. */
int gell_stream_copy(void* dst, const void* src, size_t bytes, void* stream);
int gell_device_name(char* buf, int len);

/* Trip count of the reference's float-accumulated time loops
 * `for (float t = 0; t < end; t += step)` (Main.cu:108,114,121): 61 for
 * (6, 0.1f), 11 for (0.1f, 0.01f). */
int32_t gell_float_loop_count(float end, float step);

/* ---------------------------------------------------------------------------
 * Output files and restart. The CSV writers take HOST arrays (copy the cells
 * or a field back first, as the reference does before saving) and print with
 * the reference's format strings. Each returns GELL_OK or GELL_E_ARG (bad
 * arguments, file not writable); gell_io_last_error() says which.
 */
const char* gell_io_last_error(void);
/* Cell_sphere_Initialization (Initialization.cuh:15-43), host-side: every
 * record of cells[0, capacity) set to Cell's defaults (Cells.cuh:10-33),
 * then n quiescent cells of radius `cell_radius` (V 2494, Vf 1871, Vns 135,
 * Vcs 489) placed uniformly in the ball of radius
 * cell_radius n^(1/3) / 1.612 * 1.87 around the domain centre, with
 * mesh_idx set. The draws come from SplitMix64 on `seed` (the reference
 * uses the C library's rand(), whose sequence is platform-specific). */
int gell_sphere_init(gell_cell* cells, int32_t capacity, int32_t n, const gell_params* params, float cell_radius,
                     uint64_t seed);
/* savecsv_cell (FileSave.cuh:24-33): "Index,Sign,Phase,x,y,z,r,V,Vns,Vcs",
 * one row per record of cells[0, n). */
int gell_write_cells_csv(const gell_cell* cells, int32_t n, const char* path);
/* savecsv_mesh (FileSave.cuh:35-49): "i,j,k,Concentration", N^3 rows, x
 * fastest. */
int gell_write_field_csv(const float* field, int32_t N, const char* path);
/* savecsv_meshslice (FileSave.cuh:50-64) of the plane k = N / 2 that
 * Slice_to_CPU copies (LODkernel.cuh:174-181, 301-306): "x,y,Concentration",
 * N^2 rows. `field` is the whole N^3 field. */
int gell_write_slice_csv(const float* field, int32_t N, const char* path);
/* savecsv_array (FileSave.cuh:66-75): "Index,Cell Num,Time", rows
 * i = 0 .. len-1 (the reference writes Max_Simulation_Iter + 1 of them). */
int gell_write_counts_csv(const int32_t* counts, const double* times, int32_t len, const char* path);
/* savecsv_gaparray (FileSave.cuh:77-86): rows i < n_iter / gap with
 * counts[(i + 1) gap - 1] and gap_times[i], the reference's indexing (it
 * passes Save_data_gap = 360, in minutes, as `gap`). */
int gell_write_gap_counts_csv(const int32_t* counts, const double* gap_times, int32_t n_iter, int32_t gap,
                              const char* path);

/* Binary checkpoint (no counterpart in the reference): the context's
 * parameters, the biology-step counter, d_cells[0, n) and every substrate's
 * field. That is the whole state a step depends on (the cell list and the
 * uptake maps are rebuilt by every mechanics step; the birth draws are keyed
 * on the step counter), so a run restarted from a checkpoint continues bit
 * for bit as the uninterrupted run. Load needs a context with the same N,
 * dx, substrate count, seed, time steps and substrate constants (D, decay,
 * uptake, dead factor, boundary) and room for the saved records, and refuses
 * the file otherwise; it writes the records to d_cells and the count and step
 * to *n and *bio_step.
 * Both synchronise; a failed load changes nothing on the device. */
int gell_checkpoint_save(gell_ctx* ctx, const gell_cell* d_cells, int32_t n, uint64_t bio_step, const char* path);
int gell_checkpoint_load(gell_ctx* ctx, gell_cell* d_cells, int32_t* n, uint64_t* bio_step, const char* path);

/* ---------------------------------------------------------------------------
 * Ensembles: many independent replicates (each its own N, dx, capacity,
 * seed and substrates; shared dt_bio / dt_mech / dt_diff) advanced together.
 * The reference runs one simulation per process (Main.cu:61-127); this is
 * its batched form. Replicate r is an ordinary context on a stream of its
 * own, so every per-context call (gell_field, gell_tune, ...) applies to
 * gell_ensemble_member(ens, r), and its results are bit-identical to
 * running that replicate alone through the per-context calls.
 * ------------------------------------------------------------------------- */
typedef struct gell_ensemble gell_ensemble;
enum {
    GELL_ENSEMBLE_STREAMS = 0, /* each mechanics step forked onto the replicate streams */
    GELL_ENSEMBLE_GRAPH = 1    /* the same fork/join step replayed as one hipGraph (default) */
};
/* params[0..count); `stream` (a hipStream_t) orders the ensemble's steps
 * against the caller's work; with NULL the ensemble makes a stream of its
 * own (the null stream cannot be captured into a graph). */
int gell_ensemble_create(const gell_params* params, int32_t count, void* stream, gell_ensemble** out);
void gell_ensemble_destroy(gell_ensemble* ens);
int32_t gell_ensemble_size(const gell_ensemble* ens);
gell_ctx* gell_ensemble_member(gell_ensemble* ens, int32_t r);
const char* gell_ensemble_last_error(const gell_ensemble* ens);
int gell_ensemble_set_mode(gell_ensemble* ens, int32_t mode);
/* gell_autotune once per distinct grid size; replicates of that size share
 * the result. */
int gell_ensemble_autotune(gell_ensemble* ens, int flags);
/* `steps` mechanics steps of every replicate with fixed counts: FM update,
 * uptake map of every substrate, then each substrate's LOD substeps
 * (Main.cu:114-124). cells[r] holds params[r].capacity records, n[r] the
 * live count. In graph mode the first step after a count change is issued
 * directly and the following ones replay a captured graph. Asynchronous. */
int gell_ensemble_mech_steps(gell_ensemble* ens, gell_cell* const* cells, const int32_t* n, int32_t steps);
/* One biology step of every replicate (Main.cu:108-127): birth with
 * Philox draws keyed by (params[r].seed, slot, bio_step), the mechanics
 * steps, death; n[r] is updated in place. Synchronises. */
int gell_ensemble_bio_step(gell_ensemble* ens, gell_cell* const* cells, int32_t* n, uint64_t bio_step);
int gell_ensemble_sync(gell_ensemble* ens);
/* Number of graph captures and graph launches so far. */
int gell_ensemble_stats(const gell_ensemble* ens, int64_t* captures, int64_t* launches);

/* ---------------------------------------------------------------------------
 * LOD surrogate (no counterpart in the reference, whose loop runs every
 * substep, Main.cu:121-123): a four-level 3-D U-Net (16..128 channels, BF16
 * matrix-core convolutions with FP32 accumulation) on the (N/2)^3 grid that
 * predicts the field after K substeps with the uptake map held fixed. With
 * the map fixed a substep is affine in the field, so the K-substep increment
 * is a linear map, depending only on the uptake map, of the first substep's
 * increment r = field1 - field: the network takes r in units of its interior
 * RMS s (plus the uptake map) and returns the K-substep increment in units of
 * s out_scale. Forward, gradients and Adam run on the device. In the loop
 * (gell_surrogate_advance) plain LOD stays the fallback and the judge: an
 * error budget decides whether a prediction stands in for the K substeps,
 * calibrated against the real K substeps every few calls.
 * ------------------------------------------------------------------------- */
typedef struct gell_surrogate gell_surrogate;
typedef struct {
    int32_t N;        /* fine grid (multiple of 128) */
    int32_t K;        /* substeps one prediction stands for */
    float rate_scale; /* input scaling of the uptake map (0.1) */
    float out_scale;  /* units of s in one head unit (8) */
    int32_t check_every; /* gell_surrogate_advance: one call in this many runs
                            the K real substeps to calibrate the gate (6) */
    uint64_t seed;    /* weight initialisation */
} gell_surrogate_config;
void gell_surrogate_default_config(gell_surrogate_config* cfg);
/* `stream` (a hipStream_t) is the stream every surrogate call runs on. */
int gell_surrogate_create(const gell_surrogate_config* cfg, void* stream, gell_surrogate** out);
void gell_surrogate_destroy(gell_surrogate* sur);
const char* gell_surrogate_last_error(const gell_surrogate* sur);
/* FP32 parameters, per layer W[cout][cin][27] (tap = (dz 3 + dy) 3 + dx) then b[cout]. */
int64_t gell_surrogate_param_count(const gell_surrogate* sur);
int gell_surrogate_get_params(gell_surrogate* sur, float* host, int64_t n);
int gell_surrogate_set_params(gell_surrogate* sur, const float* host, int64_t n);
/* out = predicted field after K substeps from `field` under the uptake map
 * `rate`; field1 = `field` after one real substep under it (device pointers
 * to N^3 floats; out may alias field1). The two outer layers of every face,
 * which each substep first resets to the boundary value, are field1's. */
int gell_surrogate_predict(gell_surrogate* sur, const float* field, const float* field1, const float* rate,
                           float* out);
/* Loss (mean squared interior error of the increment target - field, in
 * units of s out_scale, s the interior RMS of field1 - field) and parameter
 * gradients for one sample, replacing the stored gradients or, with
 * accumulate, added to them (mini-batches); the stored gradients are copied
 * to grad_host when not NULL (param_count floats). Synchronises. */
int gell_surrogate_gradients(gell_surrogate* sur, const float* field, const float* field1, const float* rate,
                             const float* target, int32_t accumulate, float* loss, float* grad_host);
/* One Adam step (0.9, 0.999, 1e-8) on the last gradients. */
int gell_surrogate_adam(gell_surrogate* sur, float lr);
/* Advance substrate `s` of ctx by K substeps. Each call runs the first
 * substep for real and measures the relative size of its increment,
 * rho = ||field1 - field|| (interior) / ||field|| (*residual). The first call
 * and every check_every-th call after also predict and then run the other
 * K - 1 real substeps, keeping their exact result; the prediction's true
 * relative L2 error e against them calibrates the ratio e / rho (the largest
 * seen). Between those checks a prediction is made and kept (*used = 1) only
 * while the estimated errors (1.5 ratio rho) of the predictions kept in the
 * current window add up to at most `tol`, relative L2; otherwise the other
 * K - 1 substeps run (*used = 0). tol = 0 makes every call exact (bit for
 * bit K substeps). On checks, and after a call that kept its prediction, the
 * network is queued before rho is read (a call that then falls back has run
 * it for nothing). ctx must have grid N and run on the surrogate's stream.
 * The host waits for rho; the call returns with the rest of its work queued
 * on the stream (a check waits for all of it). */
int gell_surrogate_advance(gell_surrogate* sur, gell_ctx* ctx, int32_t s, float tol, int32_t* used, float* residual);
/* Start a new error-budget window (e.g. at each biology step). */
int gell_surrogate_reset_budget(gell_surrogate* sur);
/* Calls of gell_surrogate_advance so far by outcome (exact check, prediction
 * kept, fallback to the K substeps), the calibrated error / rho ratio
 * (< 0 before the first check) and the error spent in the window. */
int gell_surrogate_gate_stats(const gell_surrogate* sur, int64_t* checked, int64_t* accepted, int64_t* fallback,
                              double* ratio, double* spent);
/* One of the network's 3x3x3 layers alone on an S^3 grid (S a multiple of
 * 8), for checking the kernels: x, y device BF16 tensors (channels-last);
 * w_master [cout][cin][27] and b[cout] on the host. With input_grad the call
 * computes the layer's input gradient instead: x = dy (cout channels), y =
 * dx (cin channels). Supported (cin, cout): the network's layers and their
 * reverses. Allocates; synchronises. */
int gell_nn_conv3(const uint16_t* x, const float* w_master, const float* b, uint16_t* y, int32_t cin, int32_t cout,
                  int32_t S, int32_t relu, int32_t input_grad);
/* gell_nn_conv3 (no input gradient) that also writes pooled ((S/2)^3 x cout)
 * = the 2x2x2 mean of y, rounded to BF16 as the network's pooling rounds it:
 * formed in the convolution's epilogue where its blocks allow (*in_kernel =
 * 1), else by the pooling kernel. For checking the fused form. Allocates;
 * synchronises. */
int gell_nn_conv3_pooled(const uint16_t* x, const float* w_master, const float* b, uint16_t* y, uint16_t* pooled,
                         int32_t cin, int32_t cout, int32_t S, int32_t relu, int32_t* in_kernel);
/* A layer that follows an upsampling (network layers 7, 9, 11: (cin, cout)
 * = (128, 64), (64, 32), (32, 16)) as inference runs it: y ((2S)^3 x cout) =
 * act(conv3x3x3(up2(x)) + b) (+ skip when not NULL) computed on x's S^3 grid
 * (S a multiple of 4) as 8 per-parity 2x2x2 convolutions. x, skip, y device
 * BF16 tensors; w_master, b on the host. Allocates; synchronises. */
int gell_nn_upconv3(const uint16_t* x, const float* w_master, const float* b, const uint16_t* skip, uint16_t* y,
                    int32_t cin, int32_t cout, int32_t S, int32_t relu);
/* The layer's weight gradient into host dw_master ([cout][cin][27]). */
int gell_nn_conv3_wgrad(const uint16_t* x, const uint16_t* dy, float* dw_master, int32_t cin, int32_t cout,
                        int32_t S);

#ifdef __cplusplus
}
#endif

#endif /* GELL_ABI_H */
