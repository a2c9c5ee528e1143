// ensemble.hip -- many independent replicates (different N, cell counts and
// seeds) advanced together. The reference runs one simulation per process
// (Main.cu:61-127); here one context per replicate, each on its own stream,
// and one biology step of all replicates is
//
//   birth of every replicate (launched on every stream, then one wait each)
//   -> mechanics steps: each step forks every replicate's FM update, uptake
//      maps and LOD substeps onto its stream and joins them; from the second
//      step of a biology step on, that fork/join step is one hipGraph launch
//   -> death of every replicate (launched, then waited for).
//
// Cell counts are fixed between birth and death, so a graph captured for one
// set of counts serves every mechanics step of a biology step. Each replicate
// runs exactly the kernels of its own context, so results are bit-identical
// to running that replicate alone.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

struct gell_ensemble {
    hipStream_t main = nullptr;
    bool own_main = false;  // created here when the caller passed the null stream (which cannot be captured)
    std::vector<gell_ctx*> ctx;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    int mode = GELL_ENSEMBLE_GRAPH;
    int32_t diff_steps = 0;
    int32_t mech_steps = 0;
    // The graph and what it was captured for.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<gell_cell*> g_cells;
    std::vector<int32_t> g_n;
    std::vector<int> g_tune;
    int64_t graph_captures = 0;
    int64_t graph_launches = 0;
    // One graph per biology step (gell_ensemble_bio_step in graph mode), its
    // counts device-resident: what it was captured for, and the pinned
    // exchange area its first and last nodes copy the counts through
    // (per replicate DN_SIZE ints in, then DN_SIZE out).
    hipGraph_t bgraph = nullptr;
    hipGraphExec_t bexec = nullptr;
    std::vector<gell_cell*> b_cells;
    std::vector<int> b_tune;
    std::vector<uint8_t> b_state;
    int32_t* h_io = nullptr;
    std::string err;
};

namespace {

thread_local std::string g_ens_err;

int ens_fail(gell_ensemble* ens, int code, const std::string& msg) {
    if (ens) ens->err = msg;
    g_ens_err = msg;
    return code;
}

int ens_hip(gell_ensemble* ens, hipError_t e, const char* where) {
    return ens_fail(ens, GELL_E_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

void drop_bio_graph(gell_ensemble* ens) {
    if (ens->bexec) (void)hipGraphExecDestroy(ens->bexec);
    if (ens->bgraph) (void)hipGraphDestroy(ens->bgraph);
    ens->bexec = nullptr;
    ens->bgraph = nullptr;
    ens->b_cells.clear();
    ens->b_tune.clear();
    ens->b_state.clear();
}

// Host-side context state that decides which kernels a captured step
// launches (which clear an uptake map takes, which cell-list layout the
// keys kernel clears). A captured biology step is replayed only from the
// state it was captured in.
std::vector<uint8_t> state_snapshot(const gell_ensemble* ens) {
    std::vector<uint8_t> st;
    for (const gell_ctx* c : ens->ctx) {
        st.push_back((uint8_t)c->cl_mode);
        for (int s = 0; s < c->P.n_substrates; ++s) {
            st.push_back((uint8_t)c->xspan_valid[s]);
            st.push_back((uint8_t)c->rate_handed_out[s]);
            st.push_back((uint8_t)(c->rate_list_ptr[s] == c->rate_list[s]));
        }
    }
    return st;
}

void drop_graph(gell_ensemble* ens) {
    if (ens->exec) (void)hipGraphExecDestroy(ens->exec);
    if (ens->graph) (void)hipGraphDestroy(ens->graph);
    ens->exec = nullptr;
    ens->graph = nullptr;
    ens->g_cells.clear();
    ens->g_n.clear();
    ens->g_tune.clear();
}

std::vector<int> tune_snapshot(const gell_ensemble* ens) {
    std::vector<int> t;
    for (const gell_ctx* c : ens->ctx) t.insert(t.end(), c->tune, c->tune + GELL_N_KNOBS);
    return t;
}

// One mechanics step of replicate r on its own stream: FM update, uptake map
// of every substrate, then each substrate's LOD substeps (substrates do not
// interact inside a mechanics step, so substrate-major order gives the same
// fields as the reference's interleaved order, Main.cu:117-124).
int member_step(gell_ensemble* ens, int r, gell_cell* cells, int32_t n) {
    gell_ctx* c = ens->ctx[r];
    int rc = gell_fm_update(c, cells, n);
    for (int s = 0; rc == GELL_OK && s < c->P.n_substrates; ++s) rc = gell_prelod(c, s, cells, n);
    for (int s = 0; rc == GELL_OK && s < c->P.n_substrates; ++s)
        for (int d = 0; rc == GELL_OK && d < ens->diff_steps; ++d) rc = gell_lod_substep(c, s);
    if (rc != GELL_OK) return ens_fail(ens, rc, "replicate " + std::to_string(r) + ": " + c->err);
    // the replicate's own substrate streams join its stream before the join event
    const hipError_t e = gell::join_substrates(c);
    if (e != hipSuccess) return ens_hip(ens, e, "substrate join");
    return GELL_OK;
}

// Fork every replicate's step off `main`, then join them back onto it.
int fork_join_step(gell_ensemble* ens, gell_cell* const* cells, const int32_t* n) {
    hipError_t e = hipEventRecord(ens->fork, ens->main);
    if (e != hipSuccess) return ens_hip(ens, e, "fork event");
    for (size_t r = 0; r < ens->ctx.size(); ++r) {
        e = hipStreamWaitEvent(ens->streams[r], ens->fork, 0);
        if (e != hipSuccess) return ens_hip(ens, e, "fork wait");
        int rc = member_step(ens, (int)r, cells[r], n[r]);
        if (rc != GELL_OK) return rc;
        e = hipEventRecord(ens->join[r], ens->streams[r]);
        if (e != hipSuccess) return ens_hip(ens, e, "join event");
        e = hipStreamWaitEvent(ens->main, ens->join[r], 0);
        if (e != hipSuccess) return ens_hip(ens, e, "join wait");
    }
    return GELL_OK;
}

bool graph_matches(const gell_ensemble* ens, gell_cell* const* cells, const int32_t* n) {
    if (!ens->exec) return false;
    for (size_t r = 0; r < ens->ctx.size(); ++r) {
        if (ens->g_cells[r] != cells[r] || ens->g_n[r] != n[r]) return false;
        // the cell-list and uptake-map clears are sized by the previous call's
        // count; a replay is valid only once those equal n
        const gell_ctx* c = ens->ctx[r];
        if (n[r] > 0 && c->se_list_n != n[r]) return false;
        for (int s = 0; s < c->P.n_substrates; ++s)
            if (n[r] > 0 && c->rate_list_n[s] != n[r]) return false;
    }
    return ens->g_tune == tune_snapshot(ens);
}

int capture(gell_ensemble* ens, gell_cell* const* cells, const int32_t* n) {
    drop_graph(ens);
    hipError_t e = hipStreamBeginCapture(ens->main, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return ens_hip(ens, e, "begin capture");
    int rc = fork_join_step(ens, cells, n);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ens->main, &g);
    if (rc != GELL_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return ens_hip(ens, e, "end capture");
    ens->graph = g;
    e = hipGraphInstantiate(&ens->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph(ens);
        return ens_hip(ens, e, "graph instantiate");
    }
    ens->g_cells.assign(cells, cells + ens->ctx.size());
    ens->g_n.assign(n, n + ens->ctx.size());
    ens->g_tune = tune_snapshot(ens);
    ++ens->graph_captures;
    return GELL_OK;
}

// One biology step of every replicate (Main.cu:108-127) captured as one
// graph with device-resident counts: each replicate's chain on its own
// stream -- counts in from the pinned area, birth, the mechanics steps
// (FM update, uptake maps, substeps), death, counts out -- joined on the
// ensemble stream. No count crosses to the host inside the step, so the
// graph serves every biology step whatever births and deaths do.
int capture_bio_step(gell_ensemble* ens, gell_cell* const* cells, const int32_t* n) {
    drop_bio_graph(ens);
    const size_t R = ens->ctx.size();
    ens->b_state = state_snapshot(ens);
    for (gell_ctx* c : ens->ctx) c->devn = true;
    hipError_t e = hipStreamBeginCapture(ens->main, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        for (gell_ctx* c : ens->ctx) c->devn = false;
        return ens_hip(ens, e, "begin capture");
    }
    int rc = GELL_OK;
    e = hipEventRecord(ens->fork, ens->main);
    for (size_t r = 0; r < R && e == hipSuccess && rc == GELL_OK; ++r) {
        gell_ctx* c = ens->ctx[r];
        e = hipStreamWaitEvent(ens->streams[r], ens->fork, 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->dn, ens->h_io + (2 * r) * gell::DN_SIZE, sizeof(int32_t) * gell::DN_SIZE,
                               hipMemcpyHostToDevice, ens->streams[r]);
        if (e == hipSuccess) e = gell::join_substrates(c);
        if (e == hipSuccess) e = gell::birth_begin(c, cells[r], n[r], 0);
        for (int32_t k = 0; k < ens->mech_steps && e == hipSuccess && rc == GELL_OK; ++k)
            rc = member_step(ens, (int)r, cells[r], n[r]);
        if (e == hipSuccess && rc == GELL_OK) e = gell::death_begin(c, cells[r], n[r]);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ens->h_io + (2 * r + 1) * gell::DN_SIZE, c->dn, sizeof(int32_t) * gell::DN_SIZE,
                               hipMemcpyDeviceToHost, ens->streams[r]);
        if (e == hipSuccess) e = hipEventRecord(ens->join[r], ens->streams[r]);
        if (e == hipSuccess) e = hipStreamWaitEvent(ens->main, ens->join[r], 0);
    }
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(ens->main, &g);
    for (gell_ctx* c : ens->ctx) c->devn = false;
    if (rc != GELL_OK || e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        ens->b_state.clear();
        if (rc != GELL_OK) return rc;
        return ens_hip(ens, e != hipSuccess ? e : e2, "biology step capture");
    }
    ens->bgraph = g;
    e = hipGraphInstantiate(&ens->bexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_bio_graph(ens);
        return ens_hip(ens, e, "graph instantiate");
    }
    ens->b_cells.assign(cells, cells + R);
    ens->b_tune = tune_snapshot(ens);
    ++ens->graph_captures;
    return GELL_OK;
}

int bio_step_graph(gell_ensemble* ens, gell_cell* const* cells, int32_t* n, uint64_t bio_step) {
    const size_t R = ens->ctx.size();
    for (const gell_ctx* c : ens->ctx)
        if (c->prof) return ens_fail(ens, GELL_E_ARG, "graph mode needs profiling off on every replicate");
    // the step's counts in: cells, the two list counts, the biology step
    for (size_t r = 0; r < R; ++r) {
        const gell_ctx* c = ens->ctx[r];
        int32_t* in = ens->h_io + (2 * r) * gell::DN_SIZE;
        in[gell::DN_CELLS] = n[r];
        in[gell::DN_LIST] = c->se_list_n;
        for (int s = 0; s < GELL_MAX_SUBSTRATES; ++s) in[gell::DN_MAP + s] = s < c->P.n_substrates ? c->rate_list_n[s] : 0;
        in[gell::DN_STEP] = (int32_t)(uint32_t)bio_step;
        in[gell::DN_STEP + 1] = (int32_t)(uint32_t)(bio_step >> 32);
    }
    const bool match = ens->bexec && ens->b_cells == std::vector<gell_cell*>(cells, cells + R) &&
                       ens->b_tune == tune_snapshot(ens) && ens->b_state == state_snapshot(ens);
    if (!match) {
        int rc = capture_bio_step(ens, cells, n);
        if (rc != GELL_OK) return rc;
    }
    hipError_t e = hipGraphLaunch(ens->bexec, ens->main);
    if (e == hipSuccess) e = hipStreamSynchronize(ens->main);
    if (e != hipSuccess) return ens_hip(ens, e, "biology step graph");
    ++ens->graph_launches;
    // the counts out, back into the host-side state the per-context calls use
    for (size_t r = 0; r < R; ++r) {
        gell_ctx* c = ens->ctx[r];
        const int32_t* out = ens->h_io + (2 * r + 1) * gell::DN_SIZE;
        n[r] = out[gell::DN_CELLS];
        c->se_list_n = out[gell::DN_LIST];
        for (int s = 0; s < c->P.n_substrates; ++s) c->rate_list_n[s] = out[gell::DN_MAP + s];
    }
    return GELL_OK;
}

int check_args(gell_ensemble* ens, gell_cell* const* cells, const int32_t* n) {
    if (!ens) return ens_fail(nullptr, GELL_E_ARG, "null ensemble");
    if (!cells || !n) return ens_fail(ens, GELL_E_ARG, "null cell or count array");
    for (size_t r = 0; r < ens->ctx.size(); ++r) {
        if (n[r] < 0 || n[r] > ens->ctx[r]->P.capacity)
            return ens_fail(ens, GELL_E_ARG, "replicate " + std::to_string(r) + ": count out of range");
        if (n[r] > 0 && !cells[r]) return ens_fail(ens, GELL_E_ARG, "replicate " + std::to_string(r) + ": null cells");
    }
    return GELL_OK;
}

}  // namespace

extern "C" {

int gell_ensemble_create(const gell_params* params, int32_t count, void* stream, gell_ensemble** out) {
    if (!out) return ens_fail(nullptr, GELL_E_ARG, "null output pointer");
    *out = nullptr;
    if (!params || count <= 0) return ens_fail(nullptr, GELL_E_ARG, "need at least one replicate");
    for (int32_t r = 1; r < count; ++r)
        if (params[r].dt_bio != params[0].dt_bio || params[r].dt_mech != params[0].dt_mech ||
            params[r].dt_diff != params[0].dt_diff)
            return ens_fail(nullptr, GELL_E_ARG, "replicates must share dt_bio, dt_mech and dt_diff");
    gell_ensemble* ens = new (std::nothrow) gell_ensemble;
    if (!ens) return ens_fail(nullptr, GELL_E_NOMEM, "ensemble allocation");
    ens->main = reinterpret_cast<hipStream_t>(stream);
    int rc = GELL_OK;
    hipError_t e = hipSuccess;
    if (!ens->main) {
        e = hipStreamCreateWithFlags(&ens->main, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ens;
            return ens_hip(nullptr, e, "stream");
        }
        ens->own_main = true;
    }
    ens->mech_steps = gell_float_loop_count(params[0].dt_bio, params[0].dt_mech);
    e = hipHostMalloc(reinterpret_cast<void**>(&ens->h_io), sizeof(int32_t) * 2 * gell::DN_SIZE * (size_t)count);
    if (e != hipSuccess) {
        ens->h_io = nullptr;
        if (ens->own_main) (void)hipStreamDestroy(ens->main);
        delete ens;
        return ens_hip(nullptr, e, "pinned count area");
    }
    std::memset(ens->h_io, 0, sizeof(int32_t) * 2 * gell::DN_SIZE * (size_t)count);
    ens->diff_steps = gell_float_loop_count(params[0].dt_mech, params[0].dt_diff);
    e = hipEventCreateWithFlags(&ens->fork, hipEventDisableTiming);
    if (e != hipSuccess) rc = ens_hip(ens, e, "event");
    for (int32_t r = 0; rc == GELL_OK && r < count; ++r) {
        hipStream_t s = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) {
            rc = ens_hip(ens, e, "stream");
            break;
        }
        ens->streams.push_back(s);
        hipEvent_t j = nullptr;
        e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
        if (e != hipSuccess) {
            rc = ens_hip(ens, e, "event");
            break;
        }
        ens->join.push_back(j);
        gell_ctx* c = nullptr;
        rc = gell_create(&params[r], s, &c);
        if (rc != GELL_OK) {
            rc = ens_fail(ens, rc, "replicate " + std::to_string(r) + ": " + gell_last_error(nullptr));
            break;
        }
        ens->ctx.push_back(c);
    }
    if (rc != GELL_OK) {
        std::string msg = ens->err;
        gell_ensemble_destroy(ens);
        g_ens_err = msg;
        return rc;
    }
    *out = ens;
    return GELL_OK;
}

void gell_ensemble_destroy(gell_ensemble* ens) {
    if (!ens) return;
    for (hipStream_t s : ens->streams) (void)hipStreamSynchronize(s);
    drop_graph(ens);
    drop_bio_graph(ens);
    for (gell_ctx* c : ens->ctx) gell_destroy(c);
    if (ens->h_io) (void)hipHostFree(ens->h_io);
    for (hipEvent_t j : ens->join) (void)hipEventDestroy(j);
    for (hipStream_t s : ens->streams) (void)hipStreamDestroy(s);
    if (ens->fork) (void)hipEventDestroy(ens->fork);
    if (ens->own_main) {
        (void)hipStreamSynchronize(ens->main);
        (void)hipStreamDestroy(ens->main);
    }
    delete ens;
}

int32_t gell_ensemble_size(const gell_ensemble* ens) { return ens ? (int32_t)ens->ctx.size() : 0; }

gell_ctx* gell_ensemble_member(gell_ensemble* ens, int32_t r) {
    if (!ens || r < 0 || r >= (int32_t)ens->ctx.size()) return nullptr;
    return ens->ctx[r];
}

const char* gell_ensemble_last_error(const gell_ensemble* ens) {
    return ens ? ens->err.c_str() : g_ens_err.c_str();
}

int gell_ensemble_set_mode(gell_ensemble* ens, int32_t mode) {
    if (!ens) return ens_fail(nullptr, GELL_E_ARG, "null ensemble");
    if (mode != GELL_ENSEMBLE_STREAMS && mode != GELL_ENSEMBLE_GRAPH)
        return ens_fail(ens, GELL_E_ARG, "unknown ensemble mode");
    ens->mode = mode;
    if (mode != GELL_ENSEMBLE_GRAPH) {
        drop_graph(ens);
        drop_bio_graph(ens);
    }
    return GELL_OK;
}

int gell_ensemble_autotune(gell_ensemble* ens, int flags) {
    if (!ens) return ens_fail(nullptr, GELL_E_ARG, "null ensemble");
    // one search per grid size; the result is shared by the replicates of that size
    std::vector<int32_t> done_N;
    std::vector<const gell_ctx*> done_ctx;
    for (gell_ctx* c : ens->ctx) {
        auto it = std::find(done_N.begin(), done_N.end(), c->N);
        if (it == done_N.end()) {
            int rc = gell_autotune(c, flags, nullptr);
            if (rc != GELL_OK) return ens_fail(ens, rc, c->err);
            done_N.push_back(c->N);
            done_ctx.push_back(c);
        } else {
            const gell_ctx* src = done_ctx[it - done_N.begin()];
            for (int k = 0; k < GELL_N_KNOBS; ++k) {
                int rc = gell_tune(c, k, src->tune[k]);
                if (rc != GELL_OK) return ens_fail(ens, rc, c->err);
            }
        }
    }
    drop_graph(ens);
    drop_bio_graph(ens);
    return GELL_OK;
}

int gell_ensemble_mech_steps(gell_ensemble* ens, gell_cell* const* cells, const int32_t* n, int32_t steps) {
    int rc = check_args(ens, cells, n);
    if (rc != GELL_OK) return rc;
    if (steps < 0) return ens_fail(ens, GELL_E_ARG, "negative step count");
    for (int32_t k = 0; k < steps; ++k) {
        if (ens->mode == GELL_ENSEMBLE_GRAPH) {
            for (const gell_ctx* c : ens->ctx)
                if (c->prof) return ens_fail(ens, GELL_E_ARG, "graph mode needs profiling off on every replicate");
            if (graph_matches(ens, cells, n)) {
                hipError_t e = hipGraphLaunch(ens->exec, ens->main);
                if (e != hipSuccess) return ens_hip(ens, e, "graph launch");
                ++ens->graph_launches;
                continue;
            }
            // a step issued directly brings the clear counts to n; the graph
            // is captured from the next step on
            rc = fork_join_step(ens, cells, n);
            if (rc == GELL_OK && k + 1 < steps) rc = capture(ens, cells, n);
        } else {
            rc = fork_join_step(ens, cells, n);
        }
        if (rc != GELL_OK) return rc;
    }
    return GELL_OK;
}

int gell_ensemble_bio_step(gell_ensemble* ens, gell_cell* const* cells, int32_t* n, uint64_t bio_step) {
    int rc = check_args(ens, cells, n);
    if (rc != GELL_OK) return rc;
    if (ens->mode == GELL_ENSEMBLE_GRAPH) return bio_step_graph(ens, cells, n, bio_step);
    const size_t R = ens->ctx.size();
    // everything queued on the ensemble stream before this call comes first
    hipError_t e = hipEventRecord(ens->fork, ens->main);
    if (e != hipSuccess) return ens_hip(ens, e, "fork event");
    for (size_t r = 0; r < R; ++r) {
        e = hipStreamWaitEvent(ens->streams[r], ens->fork, 0);
        if (e == hipSuccess) e = gell::birth_begin(ens->ctx[r], cells[r], n[r], bio_step);
        if (e != hipSuccess) return ens_hip(ens, e, "birth");
    }
    for (size_t r = 0; r < R; ++r) {
        e = gell::birth_finish(ens->ctx[r], n[r], &n[r]);
        if (e != hipSuccess) return ens_hip(ens, e, "birth");
    }
    rc = gell_ensemble_mech_steps(ens, cells, n, ens->mech_steps);
    if (rc != GELL_OK) return rc;
    e = hipEventRecord(ens->fork, ens->main);
    if (e != hipSuccess) return ens_hip(ens, e, "fork event");
    for (size_t r = 0; r < R; ++r) {
        e = hipStreamWaitEvent(ens->streams[r], ens->fork, 0);
        if (e == hipSuccess) e = gell::death_begin(ens->ctx[r], cells[r], n[r]);
        if (e != hipSuccess) return ens_hip(ens, e, "death");
    }
    for (size_t r = 0; r < R; ++r) {
        e = gell::death_finish(ens->ctx[r], n[r], &n[r]);
        if (e != hipSuccess) return ens_hip(ens, e, "death");
    }
    return GELL_OK;
}

int gell_ensemble_sync(gell_ensemble* ens) {
    if (!ens) return ens_fail(nullptr, GELL_E_ARG, "null ensemble");
    hipError_t e = hipStreamSynchronize(ens->main);
    for (size_t r = 0; e == hipSuccess && r < ens->streams.size(); ++r) e = hipStreamSynchronize(ens->streams[r]);
    return e == hipSuccess ? GELL_OK : ens_hip(ens, e, "sync");
}

int gell_ensemble_stats(const gell_ensemble* ens, int64_t* captures, int64_t* launches) {
    if (!ens) return ens_fail(nullptr, GELL_E_ARG, "null ensemble");
    if (captures) *captures = ens->graph_captures;
    if (launches) *launches = ens->graph_launches;
    return GELL_OK;
}

}  // extern "C"
