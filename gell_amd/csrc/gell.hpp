// gell.hpp -- header-only C++ host interface over the C ABI, shaped like the
// reference's host objects so a Main.cu-style time loop keeps its structure:
//
//   reference (Gell sources)                          here
//   MechanicsMesh_struct MecMesh;  (FMkernel.cuh:160)  gell::MechanicsMesh MecMesh(ctx);
//   MecMesh.FM_update(GpuCell, n)  (FMkernel.cuh:181)  MecMesh.FM_update(cells, n);
//   Mesh_struct O2Mesh(38, 1e5, .1) (LODkernel.cuh:205) gell::Mesh O2Mesh(ctx, 0);
//   O2Mesh.PreLOD(GC, n)            (LODkernel.cuh:270) O2Mesh.PreLOD(cells, n);
//   O2Mesh.LODsolver(GC, n)         (LODkernel.cuh:280) O2Mesh.LODsolver(cells, n);
//   O2Mesh.p                        (LODkernel.cuh:199) O2Mesh.p()
//   CellBirth_kernel(p, GpuCell, states, n, gn) (PDkernel.cuh:88)
//                                                       gell::CellBirth_kernel(ctx, cells, n, step)
//   CellDeath_kernel(GpuCell, states, n, gn)    (PDkernel.cuh:100)
//                                                       gell::CellDeath_kernel(ctx, cells, n)
//
// The substrate constants of Mesh_struct's constructor live in gell_params
// (one context holds every substrate). Errors are thrown as gell::Error on the
// host side of this header only; nothing throws across the C ABI.
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/gell_abi.h"

namespace gell {

struct Error : std::runtime_error {
    int status;
    Error(int st, const std::string& what) : std::runtime_error(what), status(st) {}
};

inline void check(gell_ctx* ctx, int st, const char* where) {
    if (st != GELL_OK) throw Error(st, std::string(where) + ": " + gell_last_error(ctx));
}

// Owns the context: the cell-list arrays, every substrate's field and uptake
// map, the Thomas factors and the scratch of the sort and PD steps.
class Context {
   public:
    explicit Context(const gell_params& p, void* stream = nullptr) {
        const int st = gell_create(&p, stream, &ctx_);
        if (st != GELL_OK) throw Error(st, std::string("gell_create: ") + gell_last_error(nullptr));
    }
    ~Context() { gell_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    gell_ctx* get() const { return ctx_; }
    void sync() { check(ctx_, gell_sync(ctx_), "gell_sync"); }
    void autotune(bool bit_exact = false) {
        check(ctx_, gell_autotune(ctx_, bit_exact ? GELL_TUNE_BIT_EXACT : 0, nullptr), "gell_autotune");
    }

   private:
    gell_ctx* ctx_ = nullptr;
};

// MechanicsMesh_struct (FMkernel.cuh:160-212).
class MechanicsMesh {
   public:
    explicit MechanicsMesh(Context& c) : ctx_(c.get()) {}
    void FM_update(gell_cell* d_cells, int32_t num) { check(ctx_, gell_fm_update(ctx_, d_cells, num), "FM_update"); }

   private:
    gell_ctx* ctx_;
};

// Mesh_struct (LODkernel.cuh:192-319) for substrate `s` of the context.
class Mesh {
   public:
    Mesh(Context& c, int32_t substrate) : ctx_(c.get()), s_(substrate) {}
    void PreLOD(const gell_cell* d_cells, int32_t num) {
        check(ctx_, gell_prelod(ctx_, s_, d_cells, num), "PreLOD");
    }
    // The reference's LODsolver takes the cells but does not use them.
    void LODsolver(const gell_cell* /*d_cells*/ = nullptr, int32_t /*num*/ = 0) {
        check(ctx_, gell_lod_substep(ctx_, s_), "LODsolver");
    }
    float* p() const { return gell_field(ctx_, s_); }
    float* p_rate() const { return gell_uptake_field(ctx_, s_); }

   private:
    gell_ctx* ctx_;
    int32_t s_;
};

// CellBirth_kernel (PDkernel.cuh:88-98): returns the new count. The draws
// come from the context's Philox stream at biology step `step`.
inline int32_t CellBirth_kernel(Context& c, gell_cell* d_cells, int32_t currentnum, uint64_t step) {
    int32_t n = currentnum;
    check(c.get(), gell_birth(c.get(), d_cells, currentnum, step, &n), "CellBirth_kernel");
    return n;
}

// CellDeath_kernel (PDkernel.cuh:100-117): returns the new count.
inline int32_t CellDeath_kernel(Context& c, gell_cell* d_cells, int32_t currentnum) {
    int32_t n = currentnum;
    check(c.get(), gell_death(c.get(), d_cells, currentnum, &n), "CellDeath_kernel");
    return n;
}

// The reference's output files (FileSave.cuh:24-86), from host arrays.
inline void io_check(int st, const char* where) {
    if (st != GELL_OK) throw Error(st, std::string(where) + ": " + gell_io_last_error());
}
inline void savecsv_cell(const gell_cell* host_cells, int32_t currentnum, const std::string& f) {
    io_check(gell_write_cells_csv(host_cells, currentnum, f.c_str()), "savecsv_cell");
}
inline void savecsv_mesh(const float* host_field, int32_t N, const std::string& f) {
    io_check(gell_write_field_csv(host_field, N, f.c_str()), "savecsv_mesh");
}
inline void savecsv_meshslice(const float* host_field, int32_t N, const std::string& f) {
    io_check(gell_write_slice_csv(host_field, N, f.c_str()), "savecsv_meshslice");
}
inline void savecsv_array(const int32_t* cnum, const double* time, int32_t len, const std::string& f) {
    io_check(gell_write_counts_csv(cnum, time, len, f.c_str()), "savecsv_array");
}
inline void savecsv_gaparray(const int32_t* cnum, const double* time, int32_t n_iter, int32_t gap,
                             const std::string& f) {
    io_check(gell_write_gap_counts_csv(cnum, time, n_iter, gap, f.c_str()), "savecsv_gaparray");
}

// Binary checkpoint of a context's state and the cells (no counterpart in the
// reference); a run restarted from it continues bit for bit.
inline void checkpoint_save(Context& c, const gell_cell* d_cells, int32_t n, uint64_t bio_step,
                            const std::string& f) {
    io_check(gell_checkpoint_save(c.get(), d_cells, n, bio_step, f.c_str()), "checkpoint_save");
}
inline void checkpoint_load(Context& c, gell_cell* d_cells, int32_t* n, uint64_t* bio_step, const std::string& f) {
    io_check(gell_checkpoint_load(c.get(), d_cells, n, bio_step, f.c_str()), "checkpoint_load");
}

// Trip count of `for (float t = 0; t < end; t += step)` (Main.cu:108,114,121).
inline int32_t float_loop_count(float end, float step) {
    int32_t c = 0;
    for (float t = 0.f; t < end; t += step) ++c;
    return c;
}

}  // namespace gell
