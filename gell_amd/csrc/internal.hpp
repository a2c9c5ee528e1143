// internal.hpp -- the context object behind the C ABI and the launchers each
// kernel file provides. Host-side, not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/gell_abi.h"

struct gell_ctx {
    gell_params P{};
    hipStream_t stream = nullptr;
    int dev = 0;
    int32_t N = 0;
    int64_t V = 0;
    std::string err;

    // Mechanics: sort buffers, cell list, compact neighbour array.
    uint32_t* keys[2] = {nullptr, nullptr};
    int32_t* vals[2] = {nullptr, nullptr};
    int32_t* hist = nullptr;       // radix histograms, digit-major
    int32_t* hist_tot = nullptr;   // per-digit totals
    int64_t hist_cap = 0;          // entries of hist
    int32_t* out_keys = nullptr;   // sorted signed keys of the last fm_update
    int32_t* out_order = nullptr;  // permutation of the last fm_update
    gell_cell* sorted = nullptr;   // records in key order
    float4* nb = nullptr;          // (x, y, z, r or dead tag) per sorted slot
    int2* se = nullptr;            // per voxel (s, e), indexed by Morton code
    int2* se_lin = nullptr;        // linear-layout copy for gell_voxel_ranges
    int64_t KS = 0;                // Morton key space, 2^(3 ceil(log2 N))
    int32_t* se_list = nullptr;    // voxels written by the last fm_update
    int2* mblk = nullptr;          // per aligned 8^3 Morton block: first / last sorted slot (force variant 5)
    int64_t nmblk = 0;             // blocks: max(1, KS / 512)
    int32_t* n_out = nullptr;      // device scalar: cells outside the domain (they sort first)
    // One bit per Morton code (bit v & 31 of word v / 32): set when voxel v
    // holds a live cell after the last fm_update (force variant 6).
    // max(KS / 32, 2) words.
    uint32_t* occ = nullptr;
    int64_t occ_words = 0;
    int32_t se_list_n = 0;
    // Rank-compacted cell list (force variant 9), instead of se and occ:
    // per aligned 64-voxel Morton word, (occupancy bits 0-31, bits 32-63, rank
    // of the word's first occupied voxel among all occupied voxels in Morton
    // order, unused); cs[r] = smallest live sorted slot of the r-th occupied
    // voxel, cs[R] = n. The r-th voxel's live cells lie in [cs[r], cs[r+1])
    // (with only dead cells between them), so the force walk reads dense
    // arrays (16 B per 64 voxels, 4 B per occupied voxel) instead of 8 B per
    // voxel of the Morton-indexed ranges.
    uint4* occr = nullptr;
    int64_t occr_words = 0;       // max(KS / 64, 1)
    int32_t* cs = nullptr;        // capacity + 2 entries
    int32_t* rank_tot = nullptr;  // per-block occupied counts, then their exclusive scan
    int64_t rank_blocks = 0;
    int cl_mode = 0;              // what the last gather maintained: 0 se + occ, 1 occr + cs

    // Device-resident counts (set while a graph-captured biology step is
    // built, gell_ensemble_bio_step): with devn the kernels read the cell
    // count and the list counts below from dn instead of launch arguments,
    // and the launch grids cover the capacity, so that a captured biology
    // step stays valid whatever birth and death do to the count.
    // dn[DN_CELLS] cells, dn[DN_LIST] entries of the last gather's voxel
    // list, dn[DN_MAP + s] entries of substrate s's uptake-map list,
    // dn[DN_STEP..+1] the biology step (uint64) birth draws with.
    int32_t* dn = nullptr;
    bool devn = false;

    // Diffusion, per substrate.
    float* field[GELL_MAX_SUBSTRATES] = {};
    float* rate[GELL_MAX_SUBSTRATES] = {};
    // Per x line (j + N k): first / last voxel index that may hold a nonzero
    // uptake rate ((N, -1): none); the partitioned x sweep reads no rate
    // outside it (a zero rate divides by exactly 1). Rebuilt from the map by
    // the first such substep after PreLOD or after the uptake-map pointer was
    // handed out (the caller may have written the map).
    int2* xspan[GELL_MAX_SUBSTRATES] = {};
    bool xspan_valid[GELL_MAX_SUBSTRATES] = {};
    float* E[GELL_MAX_SUBSTRATES] = {};
    float* F[GELL_MAX_SUBSTRATES] = {};
    float side[GELL_MAX_SUBSTRATES] = {};
    float* ptab[GELL_MAX_SUBSTRATES][4] = {};  // partitioned-solve tables, P = 2, 4, 8, 16
    // Uptake accumulators (int64, 2^-32 units), indexed by Morton code (KS
    // entries); zero except while a map is being built.
    long long* acc[GELL_MAX_SUBSTRATES] = {};
    int32_t* rate_list[GELL_MAX_SUBSTRATES] = {};
    int32_t rate_list_n[GELL_MAX_SUBSTRATES] = {};
    // The list of voxels (Morton codes, rate_list_n[s] entries) that may hold
    // a nonzero rate of substrate s: its own rate_list after gell_prelod, one
    // of the two mlist buffers after a fused mechanics step (gell_mech_step),
    // whose force kernel writes each moved cell's voxel there while
    // accumulating every substrate's uptake.
    const int32_t* rate_list_ptr[GELL_MAX_SUBSTRATES] = {};
    int32_t* mlist[2] = {nullptr, nullptr};
    // Set when gell_uptake_field handed out the map pointer since the last
    // PreLOD: the caller may have written any voxel, so neither the voxel list
    // nor the x-line spans bound the nonzero rates and accumulators any more,
    // and the next PreLOD clears the whole map (as the reference's
    // oxy_consumption_reset does every time, LODkernel.cuh:272).
    bool rate_handed_out[GELL_MAX_SUBSTRATES] = {};

    // Proliferation / death.
    int32_t* flags = nullptr;      // per slot
    int32_t* blk = nullptr;        // per-block counts, then offsets
    int32_t* slots = nullptr;      // daughter slot per mother
    int32_t* d_total = nullptr;    // device scalar
    int32_t* h_total = nullptr;    // pinned host scalar
    gell_cell* compact = nullptr;  // death scratch
    int64_t blk_cap = 0;

    // Kernel-variant knobs (GELL_KNOB_*).
    int tune[GELL_N_KNOBS] = {1, 10, 4, 4, 1, 1};

    // Every substrate runs its LOD substeps on a stream of its own when
    // GELL_KNOB_SUBSTRATE_STREAMS is set: forked from the context stream after
    // its uptake map (or at its first substep), joined back before anything
    // else touches that substrate (see gell::join_substrate). The mechanics
    // of the next step, which read cells only, overlap them.
    hipStream_t sub_stream[GELL_MAX_SUBSTRATES] = {};
    hipEvent_t sub_fork[GELL_MAX_SUBSTRATES] = {};
    hipEvent_t sub_join[GELL_MAX_SUBSTRATES] = {};
    bool sub_forked[GELL_MAX_SUBSTRATES] = {};   // fork event recorded, not yet waited on
    bool sub_pending[GELL_MAX_SUBSTRATES] = {};  // work on sub_stream not yet joined
    // y-coupling values of the fused LOD substep, one buffer per substrate
    // (substrates may run concurrently): 4 (N / 64) N^2 floats each
    float* coup[GELL_MAX_SUBSTRATES] = {};

    // Profiling.
    bool prof = false;
    struct Mark {
        int phase;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    std::vector<hipEvent_t> event_pool;  // recycled events
    double prof_ms[GELL_N_PHASES] = {};
    int64_t prof_calls[GELL_N_PHASES] = {};
};

namespace gell {

enum Phase { PH_FM_KEYS = 0, PH_FM_SORT, PH_FM_GATHER, PH_FM_FORCE, PH_PRELOD, PH_LOD, PH_BIRTH, PH_DEATH };

// Records the message of a failed context-free call (read back by
// gell_last_error(NULL)) and returns `code`.
int fail_global(int code, const std::string& msg);

enum DevCount { DN_CELLS = 0, DN_LIST = 1, DN_MAP = 2, DN_STEP = 2 + GELL_MAX_SUBSTRATES, DN_SIZE = DN_STEP + 2 };
// Count pointer for a kernel (null unless the counts are device-resident).
inline const int32_t* dev_count(const gell_ctx* c, int k = DN_CELLS) { return c->devn ? c->dn + k : nullptr; }
inline int32_t* dev_count_out(const gell_ctx* c, int k) { return c->devn ? c->dn + k : nullptr; }
// Cells a launch's grid must cover.
inline int32_t grid_count(const gell_ctx* c, int32_t n) { return c->devn ? c->P.capacity : n; }

// Each launcher enqueues on ctx->stream and returns a hipError_t.
hipError_t lod_substep_launch(gell_ctx* ctx, int s, hipStream_t st);
// Increment norms of an out-of-place substep, summed by its last kernel:
// per-block partials (sum over the interior of (dst - p0)^2, sum of p0^2)
// into part; nblk = how many blocks wrote them (0: this substep form does
// not sum them, and the caller must).
struct SubstepNorms {
    const float* p0 = nullptr;
    double* part = nullptr;
    int nblk = 0;
};
// One substep from the field in src into dst (both N^3; src == dst: in place;
// src may be the context's field and dst not, and the other way round).
hipError_t lod_substep_launch(gell_ctx* ctx, int s, const float* src, float* dst, hipStream_t st,
                              SubstepNorms* norms = nullptr);
// Largest nblk lod_substep_launch reports for grid side N.
inline int substep_norm_blocks(int N) { return N * N / 64; }
hipError_t xspan_rebuild(gell_ctx* ctx, int s, hipStream_t st);
// Make the context stream wait for substrate s's own stream (no-op when
// nothing is pending there); join_substrates does it for every substrate.
hipError_t join_substrate(gell_ctx* ctx, int s);
// Whether substrate s's substeps go to its own stream now.
bool substrate_stream_on(const gell_ctx* ctx, int s);
hipError_t join_substrates(gell_ctx* ctx);
// Whether the next substep takes the fused path (knobs and grid allow it).
bool lod_fused_active(const gell_ctx* ctx);
hipError_t lod_init_coeffs(gell_ctx* ctx, int s);
hipError_t lod_autotune(gell_ctx* ctx, bool bit_exact, double* best_us);

hipError_t fm_keys_launch(gell_ctx* ctx, const gell_cell* cells, int32_t n);
hipError_t fm_sort_launch(gell_ctx* ctx, int32_t n);
hipError_t fm_gather_launch(gell_ctx* ctx, const gell_cell* cells, int32_t n);
// With uptake_list the force kernel also accumulates every substrate's uptake
// rates of the moved cells into acc[] and writes each slot's voxel (Morton
// code, or -1) to uptake_list; only the variants fm_fuses_uptake() reports
// can.
hipError_t fm_force_move_launch(gell_ctx* ctx, gell_cell* cells, int32_t n, int32_t* uptake_list = nullptr);
bool fm_fuses_uptake(const gell_ctx* ctx);
hipError_t fm_init_ranges(gell_ctx* ctx);
hipError_t fm_ranges_linear(gell_ctx* ctx, int2* out);

hipError_t prelod_launch(gell_ctx* ctx, int s, const gell_cell* cells, int32_t n);
// Fused uptake maps (gell_mech_step): prelod_uptake_set describes every
// substrate's accumulators for the force kernel; prelod_fused_list picks the
// list buffer it writes; prelod_fused_finalize clears substrate s's previous
// map and writes the new one from the accumulators; prelod_fused_zero then
// zeroes every substrate's accumulators on the list.
struct UptakeSet;
UptakeSet prelod_uptake_set(const gell_ctx* ctx);
int32_t* prelod_fused_list(gell_ctx* ctx);
hipError_t prelod_fused_finalize(gell_ctx* ctx, int s, const int32_t* list, int32_t n);
hipError_t prelod_fused_zero(gell_ctx* ctx, const int32_t* list, int32_t n);

hipError_t birth_launch(gell_ctx* ctx, gell_cell* cells, int32_t n, uint64_t step, int32_t* n_out);
hipError_t death_launch(gell_ctx* ctx, gell_cell* cells, int32_t n, int32_t* n_out);
// Asynchronous halves of the two: *_begin enqueues the kernels and the count
// read-back on the context stream; *_finish waits for the stream and returns
// the new count.
hipError_t birth_begin(gell_ctx* ctx, gell_cell* cells, int32_t n, uint64_t step);
hipError_t birth_finish(gell_ctx* ctx, int32_t n, int32_t* n_out);
hipError_t death_begin(gell_ctx* ctx, gell_cell* cells, int32_t n);
hipError_t death_finish(gell_ctx* ctx, int32_t n, int32_t* n_out);

// Radix-sort planning: number of passes and bits per pass (at most
// max_bits, as even as possible) for a key width.
struct SortPlan {
    int passes;
    int bits[8];
    int shift[8];
};
SortPlan plan_sort(int key_bits, int max_bits);

inline int ceil_log2(int v) {
    int b = 0;
    while ((1 << b) < v) ++b;
    return b;
}

}  // namespace gell
