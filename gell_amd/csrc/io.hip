// io.hip -- host-side code on either side of the per-step path: the seeded
// spheroid initializer (Initialization.cuh:15-43), the reference's CSV
// outputs (FileSave.cuh:24-86) and a binary checkpoint of a context's state
// (none exists in the reference).
//
// The CSV writers take host arrays, as the reference's take thrust host
// vectors after a device-to-host copy (Main.cu:141-150), and print with the
// same format strings, so the files match the reference's byte for byte for
// the same values.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

thread_local std::string g_io_err;

int io_fail(int code, const std::string& m) {
    g_io_err = m;
    return code;
}

struct File {
    FILE* f = nullptr;
    explicit File(const char* path, const char* mode) : f(path ? std::fopen(path, mode) : nullptr) {}
    ~File() {
        if (f) std::fclose(f);
    }
    bool close() {
        const bool ok = f && std::fclose(f) == 0;
        f = nullptr;
        return ok;
    }
};

constexpr char kMagic[8] = {'G', 'E', 'L', 'L', 'C', 'K', 'P', '1'};

// Checkpoint header, followed by n records and n_substrates N^3 fields.
struct CkptHeader {
    char magic[8];
    uint32_t version;
    uint32_t cell_bytes;
    gell_params params;
    int32_t n;
    int32_t reserved_;
    uint64_t bio_step;
};

}  // namespace

extern "C" {

const char* gell_io_last_error(void) { return g_io_err.c_str(); }

// Cell_sphere_Initialization, Initialization.cuh:15-43. The reference draws
// with the C library's rand() after srand(12345) (Main.cu:80), a sequence
// that differs between C libraries; this uses SplitMix64 (Steele, Lea and
// Flood, OOPSLA 2014) on `seed`, 24-bit uniforms, the same rejection of
// points outside the unit ball and the same record values.
int gell_sphere_init(gell_cell* cells, int32_t capacity, int32_t n, const gell_params* p, float cell_radius,
                     uint64_t seed) {
    if (!cells || !p || n < 0 || capacity < n || p->N < 1 || !(p->dx > 0.f) || !(cell_radius > 0.f))
        return io_fail(GELL_E_ARG, "bad arguments");
    for (int32_t i = 0; i < capacity; ++i) {  // Cell's member defaults, Cells.cuh:10-33
        gell_cell& c = cells[i];
        std::memset(&c, 0, sizeof(c));
        c.mesh_idx = -1;
        c.pos = {-1.f, -1.f, -1.f};
        c.r = c.V = c.Vf = c.Vcs = c.Vns = -1.f;
        c.phase = -1;
    }
    uint64_t state = seed;
    auto uniform = [&state]() {  // [0, 1), 24 bits
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)(z >> 40) * (1.f / 16777216.f);
    };
    const float middle = 0.5f * (float)p->N * p->dx;
    const float ballr = (float)((double)cell_radius * std::pow((double)n, 1.0 / 3) / 1.612 * 1.87);
    for (int32_t i = 0; i < n;) {
        const float x = 2.f * uniform() - 1.f, y = 2.f * uniform() - 1.f, z = 2.f * uniform() - 1.f;
        if (!(x * x + y * y + z * z < 1.f)) continue;
        gell_cell& c = cells[i++];
        c.sign = 1;
        c.pos = {middle + x * ballr, middle + y * ballr, middle + z * ballr};
        c.r = cell_radius;
        c.phase = 0;
        c.V = 2494.f;  // Default_V, Definitions.h:26
        c.Vf = 1871.f;
        c.Vns = 135.f;
        c.Vcs = 489.f;
        const int a = (int)(c.pos.x / p->dx), b = (int)(c.pos.y / p->dx), k = (int)(c.pos.z / p->dx);
        c.mesh_idx = a + p->N * (b + p->N * k);  // Find_Index(Find_Mesh(pos, dx), N)
    }
    return GELL_OK;
}

// savecsv_cell, FileSave.cuh:24-33
int gell_write_cells_csv(const gell_cell* cells, int32_t n, const char* path) {
    if ((!cells && n > 0) || n < 0) return io_fail(GELL_E_ARG, "bad cell array");
    File out(path, "w");
    if (!out.f) return io_fail(GELL_E_ARG, std::string("cannot open ") + (path ? path : "(null)"));
    std::fprintf(out.f, "%s,%s,%s,%s,%s,%s,%s,%s,%s,%s\n", "Index", "Sign", "Phase", "x", "y", "z", "r", "V", "Vns",
                 "Vcs");
    for (int32_t i = 0; i < n; ++i) {
        const gell_cell& c = cells[i];
        std::fprintf(out.f, "%d,%d,%d,%f,%f,%f,%f,%f,%f,%f\n", i, (int)(c.sign != 0), c.phase, c.pos.x, c.pos.y,
                     c.pos.z, c.r, c.V, c.Vns, c.Vcs);
    }
    return out.close() ? GELL_OK : io_fail(GELL_E_ARG, "write failed");
}

// savecsv_mesh, FileSave.cuh:35-49: every voxel, x fastest
int gell_write_field_csv(const float* field, int32_t N, const char* path) {
    if (!field || N < 1) return io_fail(GELL_E_ARG, "bad field");
    File out(path, "w");
    if (!out.f) return io_fail(GELL_E_ARG, std::string("cannot open ") + (path ? path : "(null)"));
    std::fprintf(out.f, "%s,%s,%s,%s\n", "i", "j", "k", "Concentration");
    int64_t idx = 0;
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) std::fprintf(out.f, "%d,%d,%d,%f\n", i, j, k, field[idx++]);
    return out.close() ? GELL_OK : io_fail(GELL_E_ARG, "write failed");
}

// savecsv_meshslice, FileSave.cuh:50-64, of the plane GPU_Get_Slice takes
// (k = N / 2, LODkernel.cuh:174-181)
int gell_write_slice_csv(const float* field, int32_t N, const char* path) {
    if (!field || N < 1) return io_fail(GELL_E_ARG, "bad field");
    File out(path, "w");
    if (!out.f) return io_fail(GELL_E_ARG, std::string("cannot open ") + (path ? path : "(null)"));
    std::fprintf(out.f, "%s,%s,%s\n", "x", "y", "Concentration");
    const int64_t k = N / 2;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i)
            std::fprintf(out.f, "%d,%d,%f\n", i, j, field[i + (int64_t)N * (j + (int64_t)N * k)]);
    return out.close() ? GELL_OK : io_fail(GELL_E_ARG, "write failed");
}

// savecsv_array, FileSave.cuh:66-75: entries 0 .. len-1
int gell_write_counts_csv(const int32_t* counts, const double* times, int32_t len, const char* path) {
    if ((!counts || !times) && len > 0) return io_fail(GELL_E_ARG, "bad arrays");
    File out(path, "w");
    if (!out.f) return io_fail(GELL_E_ARG, std::string("cannot open ") + (path ? path : "(null)"));
    std::fprintf(out.f, "%s,%s,%s\n", "Index", "Cell Num", "Time");
    for (int32_t i = 0; i < len; ++i) std::fprintf(out.f, "%d,%d,%f\n", i, counts[i], times[i]);
    return out.close() ? GELL_OK : io_fail(GELL_E_ARG, "write failed");
}

// savecsv_gaparray, FileSave.cuh:77-86
int gell_write_gap_counts_csv(const int32_t* counts, const double* gap_times, int32_t n_iter, int32_t gap,
                              const char* path) {
    if (gap < 1 || n_iter < 0 || ((!counts || !gap_times) && n_iter / gap > 0)) return io_fail(GELL_E_ARG, "bad arrays");
    File out(path, "w");
    if (!out.f) return io_fail(GELL_E_ARG, std::string("cannot open ") + (path ? path : "(null)"));
    std::fprintf(out.f, "%s,%s,%s\n", "Index", "Cell Num", "Time");
    for (int32_t i = 0; i < n_iter / gap; ++i)
        std::fprintf(out.f, "%d,%d,%f\n", i, counts[(int64_t)(i + 1) * gap - 1], gap_times[i]);
    return out.close() ? GELL_OK : io_fail(GELL_E_ARG, "write failed");
}

int gell_checkpoint_save(gell_ctx* ctx, const gell_cell* d_cells, int32_t n, uint64_t bio_step, const char* path) {
    if (!ctx || (!d_cells && n > 0) || n < 0 || n > ctx->P.capacity) return io_fail(GELL_E_ARG, "bad arguments");
    if (gell_sync(ctx) != GELL_OK) return io_fail(GELL_E_HIP, "sync failed");
    CkptHeader h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = 1;
    h.cell_bytes = sizeof(gell_cell);
    h.params = ctx->P;
    h.n = n;
    h.bio_step = bio_step;
    std::vector<gell_cell> cells((size_t)n);
    std::vector<float> field((size_t)ctx->V);
    if (n > 0 && hipMemcpy(cells.data(), d_cells, sizeof(gell_cell) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
        return io_fail(GELL_E_HIP, "copy of the cells failed");
    File out(path, "wb");
    if (!out.f) return io_fail(GELL_E_ARG, std::string("cannot open ") + (path ? path : "(null)"));
    bool ok = std::fwrite(&h, sizeof(h), 1, out.f) == 1;
    if (ok && n > 0) ok = std::fwrite(cells.data(), sizeof(gell_cell), (size_t)n, out.f) == (size_t)n;
    for (int s = 0; ok && s < ctx->P.n_substrates; ++s) {
        if (hipMemcpy(field.data(), ctx->field[s], sizeof(float) * field.size(), hipMemcpyDeviceToHost) != hipSuccess)
            return io_fail(GELL_E_HIP, "copy of a field failed");
        ok = std::fwrite(field.data(), sizeof(float), field.size(), out.f) == field.size();
    }
    return ok && out.close() ? GELL_OK : io_fail(GELL_E_ARG, "write failed");
}

int gell_checkpoint_load(gell_ctx* ctx, gell_cell* d_cells, int32_t* n, uint64_t* bio_step, const char* path) {
    if (!ctx || !d_cells || !n) return io_fail(GELL_E_ARG, "bad arguments");
    File in(path, "rb");
    if (!in.f) return io_fail(GELL_E_ARG, std::string("cannot open ") + (path ? path : "(null)"));
    CkptHeader h{};
    if (std::fread(&h, sizeof(h), 1, in.f) != 1 || std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0 ||
        h.version != 1 || h.cell_bytes != sizeof(gell_cell))
        return io_fail(GELL_E_ARG, "not a checkpoint of this format");
    const gell_params& P = ctx->P;
    if (h.params.N != P.N || h.params.n_substrates != P.n_substrates || h.params.dx != P.dx)
        return io_fail(GELL_E_ARG, "checkpoint grid or substrate count differ from this context's");
    // Everything else a step's results depend on must match too, or the
    // restarted run would silently leave the uninterrupted one's trajectory.
    // (capacity and each substrate's initial value do not enter a step.)
    if (h.params.seed != P.seed)
        return io_fail(GELL_E_ARG, "checkpoint seed differs from this context's");
    if (h.params.dt_bio != P.dt_bio || h.params.dt_mech != P.dt_mech || h.params.dt_diff != P.dt_diff)
        return io_fail(GELL_E_ARG, "checkpoint time steps differ from this context's");
    for (int s = 0; s < P.n_substrates; ++s) {
        const gell_substrate &a = h.params.sub[s], &b = P.sub[s];
        if (a.D != b.D || a.decay != b.decay || a.uptake != b.uptake || a.dead_factor != b.dead_factor ||
            a.boundary != b.boundary)
            return io_fail(GELL_E_ARG, "checkpoint substrate constants differ from this context's");
    }
    if (h.n < 0 || h.n > P.capacity) return io_fail(GELL_E_ARG, "checkpoint holds more cells than this context's capacity");
    // the whole file is read before anything on the device changes
    std::vector<gell_cell> cells((size_t)h.n);
    std::vector<float> fields((size_t)ctx->V * P.n_substrates);
    if (h.n > 0 && std::fread(cells.data(), sizeof(gell_cell), (size_t)h.n, in.f) != (size_t)h.n)
        return io_fail(GELL_E_ARG, "truncated checkpoint");
    if (std::fread(fields.data(), sizeof(float), fields.size(), in.f) != fields.size())
        return io_fail(GELL_E_ARG, "truncated checkpoint");
    if (gell_sync(ctx) != GELL_OK) return io_fail(GELL_E_HIP, "sync failed");
    for (int s = 0; s < P.n_substrates; ++s)
        if (hipMemcpy(ctx->field[s], fields.data() + (size_t)ctx->V * s, sizeof(float) * (size_t)ctx->V,
                      hipMemcpyHostToDevice) != hipSuccess)
            return io_fail(GELL_E_HIP, "copy of a field failed");
    if (h.n > 0 &&
        hipMemcpy(d_cells, cells.data(), sizeof(gell_cell) * (size_t)h.n, hipMemcpyHostToDevice) != hipSuccess)
        return io_fail(GELL_E_HIP, "copy of the cells failed");
    *n = h.n;
    if (bio_step) *bio_step = h.bio_step;
    return GELL_OK;
}

}  // extern "C"
