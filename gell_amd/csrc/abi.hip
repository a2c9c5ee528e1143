// abi.hip -- the extern "C" entry points of include/gell_abi.h: context
// construction (MechanicsMesh_struct() + Mesh_struct(...), FMkernel.cuh:167-172,
// LODkernel.cuh:205-259), the five per-step operations, accessors, profiling
// and device-memory helpers. No entry point throws or exits; errors come back
// as gell_status codes with a message kept on the context.
#include <cstdio>
#include <cstring>
#include <new>

#include "device_common.hpp"
#include "internal.hpp"

namespace {

thread_local std::string g_err;

int fail(gell_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    g_err = msg;
    return code;
}

int hip_fail(gell_ctx* ctx, hipError_t e, const char* where) {
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return fail(ctx, GELL_E_HIP, m);
}

template <class T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

int check_cells(gell_ctx* ctx, const void* cells, int32_t n) {
    if (n < 0 || n > ctx->P.capacity) return fail(ctx, GELL_E_ARG, "cell count out of range [0, capacity]");
    if (n > 0 && cells == nullptr) return fail(ctx, GELL_E_ARG, "null cell array");
    return GELL_OK;
}

hipEvent_t take_event(gell_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Brackets one phase with a pair of events on the context stream while
// profiling is on.
struct PhaseScope {
    gell_ctx* ctx;
    int phase;
    hipEvent_t a = nullptr;
    PhaseScope(gell_ctx* c, int ph) : ctx(c), phase(ph) {
        if (!ctx->prof) return;
        a = take_event(ctx);
        if (a && hipEventRecord(a, ctx->stream) != hipSuccess) {
            ctx->event_pool.push_back(a);
            a = nullptr;
        }
    }
    ~PhaseScope() {
        if (!a) return;
        hipEvent_t b = take_event(ctx);
        if (b && hipEventRecord(b, ctx->stream) == hipSuccess) {
            ctx->marks.push_back({phase, a, b});
        } else {
            ctx->event_pool.push_back(a);
            if (b) ctx->event_pool.push_back(b);
        }
    }
};

void release(gell_ctx* c) {
    if (!c) return;
    for (int s = 0; s < GELL_MAX_SUBSTRATES; ++s) {
        if (c->sub_stream[s]) {
            (void)hipStreamSynchronize(c->sub_stream[s]);
            (void)hipStreamDestroy(c->sub_stream[s]);
        }
        if (c->sub_fork[s]) (void)hipEventDestroy(c->sub_fork[s]);
        if (c->sub_join[s]) (void)hipEventDestroy(c->sub_join[s]);
    }
    for (auto& m : c->marks) {
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    c->marks.clear();
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    c->event_pool.clear();
    for (int q = 0; q < 2; ++q) {
        dfree(c->keys[q]);
        dfree(c->vals[q]);
    }
    dfree(c->hist);
    dfree(c->hist_tot);
    dfree(c->out_keys);
    dfree(c->out_order);
    dfree(c->sorted);
    dfree(c->nb);
    dfree(c->se);
    dfree(c->se_lin);
    dfree(c->se_list);
    dfree(c->mblk);
    dfree(c->occ);
    dfree(c->occr);
    dfree(c->cs);
    dfree(c->rank_tot);
    dfree(c->n_out);
    for (int s = 0; s < GELL_MAX_SUBSTRATES; ++s) {
        dfree(c->field[s]);
        dfree(c->rate[s]);
        dfree(c->E[s]);
        dfree(c->F[s]);
        dfree(c->ptab[s][0]);
        dfree(c->ptab[s][1]);
        dfree(c->ptab[s][2]);
        dfree(c->ptab[s][3]);
        dfree(c->acc[s]);
        dfree(c->rate_list[s]);
        dfree(c->xspan[s]);
    }
    for (int s = 0; s < GELL_MAX_SUBSTRATES; ++s) dfree(c->coup[s]);
    dfree(c->flags);
    dfree(c->blk);
    dfree(c->slots);
    dfree(c->d_total);
    dfree(c->dn);
    dfree(c->compact);
    dfree(c->mlist[0]);
    dfree(c->mlist[1]);
    if (c->h_total) (void)hipHostFree(c->h_total);
}

__global__ void fill_f32(float* p, int64_t n, float v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

extern "C" {

void gell_default_params(gell_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->N = 100;             // Voxel_num
    p->dx = 25.f;           // Voxel_length
    p->capacity = 2000000;  // Max_Cell_num
    p->n_substrates = 1;
    p->dt_bio = 6.f;        // Biology_dt
    p->dt_mech = 0.1f;      // Mechanics_dt
    p->dt_diff = 0.01f;     // Diffusion_dt
    p->seed = 123;          // the reference's curand seed (RandomFunctions.cuh:21)
    p->sub[0].boundary = 38.f;      // O2_Default_Concentration
    p->sub[0].D = 100000.f;         // O2_Diffusion_coef
    p->sub[0].decay = 0.1f;         // O2_Decay_rate
    p->sub[0].uptake = -10.f;       // Tumor_O2_consumption
    p->sub[0].dead_factor = 0.1f;   // Dead_O2_consumption_factor
    p->sub[0].initial = 38.f;
}

const char* gell_status_string(int status) {
    switch (status) {
        case GELL_OK: return "ok";
        case GELL_E_ARG: return "invalid argument";
        case GELL_E_HIP: return "HIP runtime error";
        case GELL_E_NOMEM: return "device allocation failed";
        case GELL_E_UNSUPPORTED: return "unsupported parameters";
        case GELL_E_NODEVICE: return "no usable device";
        default: return "unknown status";
    }
}

const char* gell_last_error(const gell_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int gell_create(const gell_params* params, void* stream, gell_ctx** out) {
    if (!params || !out) return fail(nullptr, GELL_E_ARG, "null argument");
    *out = nullptr;
    const gell_params& P = *params;
    if (P.N < 4 || P.N > 1024) return fail(nullptr, GELL_E_ARG, "N must be in [4, 1024] (10-bit Morton axes)");
    if (!(P.dx > 0.f)) return fail(nullptr, GELL_E_ARG, "dx must be positive");
    if (P.capacity < 1) return fail(nullptr, GELL_E_ARG, "capacity must be positive");
    if (P.n_substrates < 1 || P.n_substrates > GELL_MAX_SUBSTRATES)
        return fail(nullptr, GELL_E_ARG, "n_substrates out of range");
    if (!(P.dt_diff > 0.f) || !(P.dt_mech > 0.f) || !(P.dt_bio > 0.f))
        return fail(nullptr, GELL_E_ARG, "time steps must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, GELL_E_NODEVICE, "no HIP device");

    gell_ctx* c = new (std::nothrow) gell_ctx();
    if (!c) return fail(nullptr, GELL_E_NOMEM, "host allocation failed");
    c->P = P;
    c->stream = reinterpret_cast<hipStream_t>(stream);
    (void)hipGetDevice(&c->dev);
    c->N = P.N;
    c->V = (int64_t)P.N * P.N * P.N;
    c->KS = (int64_t)1 << (3 * gell::ceil_log2(P.N));
    const size_t cap = (size_t)P.capacity;
    const int64_t tiles = ((int64_t)cap + 4095) / 4096;
    c->hist_cap = 512 * tiles;
    c->blk_cap = ((int64_t)cap + 255) / 256;

    hipError_t e = hipSuccess;
    auto A = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    A(dalloc(&c->keys[0], cap));
    A(dalloc(&c->keys[1], cap));
    A(dalloc(&c->vals[0], cap));
    A(dalloc(&c->vals[1], cap));
    A(dalloc(&c->hist, (size_t)c->hist_cap));
    A(dalloc(&c->hist_tot, 512));
    A(dalloc(&c->out_keys, cap));
    A(dalloc(&c->out_order, cap));
    A(dalloc(&c->sorted, cap));
    A(dalloc(&c->nb, cap));
    A(dalloc(&c->se, (size_t)c->KS));
    A(dalloc(&c->se_list, cap));
    c->nmblk = c->KS >= 512 ? c->KS / 512 : 1;
    A(dalloc(&c->mblk, (size_t)c->nmblk));
    c->occ_words = c->KS >= 64 ? c->KS / 32 : 2;
    A(dalloc(&c->occ, (size_t)c->occ_words));
    c->occr_words = c->KS >= 64 ? c->KS / 64 : 1;
    A(dalloc(&c->occr, (size_t)c->occr_words));
    A(dalloc(&c->cs, cap + 2));
    c->rank_blocks = (c->occr_words + 4095) / 4096;
    A(dalloc(&c->rank_tot, (size_t)c->rank_blocks));
    A(dalloc(&c->n_out, 1));
    for (int s = 0; s < P.n_substrates; ++s) {
        A(dalloc(&c->field[s], (size_t)c->V));
        A(dalloc(&c->rate[s], (size_t)c->V));
        A(dalloc(&c->E[s], (size_t)P.N));
        A(dalloc(&c->F[s], (size_t)P.N));
        A(dalloc(&c->ptab[s][0], (size_t)(4 * P.N + 16)));
        A(dalloc(&c->ptab[s][1], (size_t)(4 * P.N + 64)));
        A(dalloc(&c->ptab[s][2], (size_t)(4 * P.N + 256)));
        A(dalloc(&c->ptab[s][3], (size_t)(4 * P.N + 1024)));
        A(dalloc(&c->acc[s], (size_t)c->KS));  // Morton-indexed
        A(dalloc(&c->rate_list[s], cap));
        c->rate_list_ptr[s] = c->rate_list[s];
        A(dalloc(&c->xspan[s], (size_t)P.N * P.N));
    }
    A(dalloc(&c->mlist[0], cap));
    A(dalloc(&c->mlist[1], cap));
    A(dalloc(&c->flags, cap));
    A(dalloc(&c->blk, (size_t)c->blk_cap));
    A(dalloc(&c->slots, cap));
    A(dalloc(&c->d_total, 1));
    A(dalloc(&c->dn, gell::DN_SIZE));
    A(dalloc(&c->compact, cap));
    if (P.N == 128 || P.N == 256 || P.N == 512)
        for (int s = 0; s < P.n_substrates; ++s) A(dalloc(&c->coup[s], (size_t)(4 * (P.N / 64)) * (size_t)P.N * P.N));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&c->h_total), sizeof(int32_t));
    for (int s = 0; s < P.n_substrates; ++s) {
        A(hipStreamCreateWithFlags(&c->sub_stream[s], hipStreamNonBlocking));
        A(hipEventCreateWithFlags(&c->sub_fork[s], hipEventDisableTiming));
        A(hipEventCreateWithFlags(&c->sub_join[s], hipEventDisableTiming));
    }
    if (e != hipSuccess) {
        release(c);
        delete c;
        (void)hipGetLastError();
        return fail(nullptr, GELL_E_NOMEM, std::string("device allocation: ") + hipGetErrorString(e));
    }
    const unsigned vb = (unsigned)((c->V + 255) / 256);
    for (int s = 0; s < P.n_substrates && e == hipSuccess; ++s) {
        hipLaunchKernelGGL(fill_f32, dim3(vb), dim3(256), 0, c->stream, c->field[s], c->V, P.sub[s].initial);
        A(hipMemsetAsync(c->rate[s], 0, sizeof(float) * (size_t)c->V, c->stream));
        A(hipMemsetAsync(c->acc[s], 0, sizeof(long long) * (size_t)c->KS, c->stream));
        A(gell::lod_init_coeffs(c, s));
    }
    A(gell::fm_init_ranges(c));
    A(hipStreamSynchronize(c->stream));
    if (e != hipSuccess) {
        int r = hip_fail(nullptr, e, "gell_create");
        release(c);
        delete c;
        return r;
    }
    *out = c;
    return GELL_OK;
}

void gell_destroy(gell_ctx* ctx) {
    if (!ctx) return;
    (void)gell::join_substrates(ctx);
    (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
    delete ctx;
}

}  // extern "C"

namespace {

// FM_update (FMkernel.cuh:181-202); with uptake_list the force kernel also
// accumulates the uptake maps (gell_mech_step).
int fm_step(gell_ctx* ctx, gell_cell* d_cells, int32_t n, int32_t* uptake_list) {
    hipError_t e;
    {
        PhaseScope ps(ctx, gell::PH_FM_KEYS);
        e = gell::fm_keys_launch(ctx, d_cells, n);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "fm keys");
    {
        PhaseScope ps(ctx, gell::PH_FM_SORT);
        e = gell::fm_sort_launch(ctx, n);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "fm sort");
    {
        PhaseScope ps(ctx, gell::PH_FM_GATHER);
        e = gell::fm_gather_launch(ctx, d_cells, n);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "fm gather");
    {
        PhaseScope ps(ctx, gell::PH_FM_FORCE);
        e = gell::fm_force_move_launch(ctx, d_cells, n, uptake_list);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "fm force/move");
    return GELL_OK;
}

}  // namespace

extern "C" {

int gell_fm_update(gell_ctx* ctx, gell_cell* d_cells, int32_t n) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    int r = check_cells(ctx, d_cells, n);
    if (r) return r;
    return fm_step(ctx, d_cells, n, nullptr);
}

int gell_mech_step(gell_ctx* ctx, gell_cell* d_cells, int32_t n) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    int r = check_cells(ctx, d_cells, n);
    if (r) return r;
    if (!gell::fm_fuses_uptake(ctx)) {
        // the chosen force variant cannot accumulate: the two calls it stands for
        r = fm_step(ctx, d_cells, n, nullptr);
        for (int s = 0; r == GELL_OK && s < ctx->P.n_substrates; ++s) r = gell_prelod(ctx, s, d_cells, n);
        return r;
    }
    int32_t* list = gell::prelod_fused_list(ctx);
    r = fm_step(ctx, d_cells, n, list);
    if (r) return r;
    hipError_t e = hipSuccess;
    for (int s = 0; s < ctx->P.n_substrates && e == hipSuccess; ++s) {
        // the substrate's previous substeps read the map about to be rewritten
        e = gell::join_substrate(ctx, s);
        if (e != hipSuccess) break;
        {
            PhaseScope ps(ctx, gell::PH_PRELOD);
            e = gell::prelod_fused_finalize(ctx, s, list, n);
        }
        if (e == hipSuccess && gell::substrate_stream_on(ctx, s)) {
            e = hipEventRecord(ctx->sub_fork[s], ctx->stream);
            ctx->sub_forked[s] = e == hipSuccess;
        }
    }
    if (e == hipSuccess) e = gell::prelod_fused_zero(ctx, list, n);
    return e == hipSuccess ? GELL_OK : hip_fail(ctx, e, "mech step uptake maps");
}

int gell_prelod(gell_ctx* ctx, int32_t substrate, const gell_cell* d_cells, int32_t n) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    if (substrate < 0 || substrate >= ctx->P.n_substrates) return fail(ctx, GELL_E_ARG, "substrate out of range");
    int r = check_cells(ctx, d_cells, n);
    if (r) return r;
    // the substrate's previous substeps read the uptake map about to be rewritten
    hipError_t e = gell::join_substrate(ctx, substrate);
    if (e != hipSuccess) return hip_fail(ctx, e, "prelod");
    {
        PhaseScope ps(ctx, gell::PH_PRELOD);
        e = gell::prelod_launch(ctx, substrate, d_cells, n);
    }
    // fork point of the substrate's own stream: right after its uptake map
    if (e == hipSuccess && gell::substrate_stream_on(ctx, substrate)) {
        e = hipEventRecord(ctx->sub_fork[substrate], ctx->stream);
        ctx->sub_forked[substrate] = e == hipSuccess;
    }
    return e == hipSuccess ? GELL_OK : hip_fail(ctx, e, "prelod");
}

int gell_lod_substep(gell_ctx* ctx, int32_t substrate) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    if (substrate < 0 || substrate >= ctx->P.n_substrates) return fail(ctx, GELL_E_ARG, "substrate out of range");
    hipError_t e = hipSuccess;
    hipStream_t st = ctx->stream;
    if (gell::substrate_stream_on(ctx, substrate)) {
        if (!ctx->sub_pending[substrate]) {
            if (!ctx->sub_forked[substrate]) e = hipEventRecord(ctx->sub_fork[substrate], ctx->stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(ctx->sub_stream[substrate], ctx->sub_fork[substrate], 0);
            ctx->sub_forked[substrate] = false;
            ctx->sub_pending[substrate] = e == hipSuccess;
        }
        st = ctx->sub_stream[substrate];
    } else {
        e = gell::join_substrate(ctx, substrate);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "lod substep");
    {
        PhaseScope ps(ctx, gell::PH_LOD);
        e = gell::lod_substep_launch(ctx, substrate, st);
    }
    return e == hipSuccess ? GELL_OK : hip_fail(ctx, e, "lod substep");
}

int gell_birth(gell_ctx* ctx, gell_cell* d_cells, int32_t n, uint64_t bio_step, int32_t* n_out) {
    if (!ctx || !n_out) return fail(ctx, GELL_E_ARG, "null argument");
    int r = check_cells(ctx, d_cells, n);
    if (r) return r;
    // birth reads substrate 0 (oxygen), whose substeps may still run on its stream
    hipError_t e = gell::join_substrate(ctx, 0);
    if (e != hipSuccess) return hip_fail(ctx, e, "birth");
    {
        PhaseScope ps(ctx, gell::PH_BIRTH);
        e = gell::birth_launch(ctx, d_cells, n, bio_step, n_out);
    }
    return e == hipSuccess ? GELL_OK : hip_fail(ctx, e, "birth");
}

int gell_death(gell_ctx* ctx, gell_cell* d_cells, int32_t n, int32_t* n_out) {
    if (!ctx || !n_out) return fail(ctx, GELL_E_ARG, "null argument");
    int r = check_cells(ctx, d_cells, n);
    if (r) return r;
    hipError_t e;
    {
        PhaseScope ps(ctx, gell::PH_DEATH);
        e = gell::death_launch(ctx, d_cells, n, n_out);
    }
    return e == hipSuccess ? GELL_OK : hip_fail(ctx, e, "death");
}

// Both accessors order the substrate's own stream back into the context
// stream first (so work the caller enqueues on the context stream, or on the
// null stream when the context was created on it, sees the finished field).
float* gell_field(gell_ctx* ctx, int32_t s) {
    if (!ctx || s < 0 || s >= ctx->P.n_substrates) return nullptr;
    if (gell::join_substrate(ctx, s) != hipSuccess) {
        hip_fail(ctx, hipGetLastError(), "gell_field");
        return nullptr;
    }
    return ctx->field[s];
}
float* gell_uptake_field(gell_ctx* ctx, int32_t s) {
    if (!ctx || s < 0 || s >= ctx->P.n_substrates) return nullptr;
    if (gell::join_substrate(ctx, s) != hipSuccess) {
        hip_fail(ctx, hipGetLastError(), "gell_uptake_field");
        return nullptr;
    }
    // The caller may write the map: the line spans are rebuilt before use and
    // the next PreLOD clears the whole map.
    ctx->xspan_valid[s] = false;
    ctx->rate_handed_out[s] = true;
    return ctx->rate[s];
}
const int32_t* gell_sorted_keys(gell_ctx* ctx) { return ctx ? ctx->out_keys : nullptr; }
const int32_t* gell_sorted_order(gell_ctx* ctx) { return ctx ? ctx->out_order : nullptr; }
const int32_t* gell_voxel_ranges(gell_ctx* ctx) {
    if (!ctx) return nullptr;
    if (!ctx->se_lin && dalloc(&ctx->se_lin, (size_t)ctx->V) != hipSuccess) {
        ctx->se_lin = nullptr;
        fail(ctx, GELL_E_NOMEM, "voxel range buffer");
        return nullptr;
    }
    if (gell::fm_ranges_linear(ctx, ctx->se_lin) != hipSuccess) return nullptr;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
    return reinterpret_cast<const int32_t*>(ctx->se_lin);
}
const int32_t* gell_birth_slots(gell_ctx* ctx) { return ctx ? ctx->slots : nullptr; }

int gell_sync(gell_ctx* ctx) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    hipError_t e = gell::join_substrates(ctx);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? GELL_OK : hip_fail(ctx, e, "sync");
}

int32_t gell_float_loop_count(float end, float step) {
    int32_t k = 0;
    for (volatile float t = 0.0f; t < end; t = t + step) ++k;
    return k;
}

int gell_tune(gell_ctx* ctx, int knob, int value) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    if (knob < 0 || knob >= GELL_N_KNOBS) return fail(ctx, GELL_E_ARG, "unknown knob");
    const bool parts = knob == GELL_KNOB_LOD_PARTS || knob == GELL_KNOB_LOD_X_PARTS;
    const bool ok = parts ? (value == 1 || value == 2 || value == 4 || value == 8 || value == 16)
                          : (value == 0 || value == 1 || (knob == GELL_KNOB_FM_PREFETCH && value >= 2 && value <= 10));
    if (!ok) return fail(ctx, GELL_E_ARG, "knob value out of range");
    const hipError_t e = gell::join_substrates(ctx);
    if (e != hipSuccess) return hip_fail(ctx, e, "tune");
    ctx->tune[knob] = value;
    return GELL_OK;
}

int gell_autotune(gell_ctx* ctx, int flags, double* lod_us_out) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    hipError_t e = gell::join_substrates(ctx);
    if (e == hipSuccess) e = gell::lod_autotune(ctx, (flags & GELL_TUNE_BIT_EXACT) != 0, lod_us_out);
    return e == hipSuccess ? GELL_OK : hip_fail(ctx, e, "autotune");
}

int gell_tune_get(const gell_ctx* ctx, int knob) {
    if (!ctx || knob < 0 || knob >= GELL_N_KNOBS) return -1;
    return ctx->tune[knob];
}

int gell_profile_enable(gell_ctx* ctx, int enable) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    const hipError_t e = gell::join_substrates(ctx);
    if (e != hipSuccess) return hip_fail(ctx, e, "profile");
    ctx->prof = enable != 0;
    return GELL_OK;
}

int gell_profile_read(gell_ctx* ctx, double* ms_out, int64_t* calls_out, int reset) {
    if (!ctx) return fail(nullptr, GELL_E_ARG, "null context");
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "profile sync");
    for (auto& m : ctx->marks) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
            ctx->prof_ms[m.phase] += ms;
            ctx->prof_calls[m.phase] += 1;
        }
        ctx->event_pool.push_back(m.a);
        ctx->event_pool.push_back(m.b);
    }
    ctx->marks.clear();
    for (int q = 0; q < GELL_N_PHASES; ++q) {
        if (ms_out) ms_out[q] = ctx->prof_ms[q];
        if (calls_out) calls_out[q] = ctx->prof_calls[q];
        if (reset) {
            ctx->prof_ms[q] = 0;
            ctx->prof_calls[q] = 0;
        }
    }
    return GELL_OK;
}

void* gell_dev_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}
void gell_dev_free(void* p) { dfree(p); }
int gell_copy_h2d(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? GELL_OK : GELL_E_HIP;
}
int gell_copy_d2h(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? GELL_OK : GELL_E_HIP;
}
int gell_copy_d2d(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? GELL_OK : GELL_E_HIP;
}
int gell_device_sync(void) { return hipDeviceSynchronize() == hipSuccess ? GELL_OK : GELL_E_HIP; }
int gell_device_name(char* buf, int len) {
    if (!buf || len <= 0) return GELL_E_ARG;
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return GELL_E_HIP;
    std::snprintf(buf, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    return GELL_OK;
}

}  // extern "C"

namespace gell {

int fail_global(int code, const std::string& msg) { return fail(nullptr, code, msg); }

bool substrate_stream_on(const gell_ctx* ctx, int s) {
    return s >= 0 && ctx->sub_stream[s] && ctx->tune[GELL_KNOB_SUBSTRATE_STREAMS] && !ctx->prof;
}

hipError_t join_substrate(gell_ctx* ctx, int s) {
    if (s < 0 || s >= GELL_MAX_SUBSTRATES || !ctx->sub_pending[s]) return hipSuccess;
    hipError_t e = hipEventRecord(ctx->sub_join[s], ctx->sub_stream[s]);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->sub_join[s], 0);
    if (e == hipSuccess) ctx->sub_pending[s] = false;
    return e;
}

hipError_t join_substrates(gell_ctx* ctx) {
    hipError_t e = hipSuccess;
    for (int s = 0; s < ctx->P.n_substrates && e == hipSuccess; ++s) e = join_substrate(ctx, s);
    return e;
}

}  // namespace gell
