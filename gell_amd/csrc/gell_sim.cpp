// gell_sim.cpp -- a C++ driver with the structure of the reference's time
// loop (Main.cu:108-127), written against gell.hpp:
//
//   for each biology step:  birth
//       for each mechanics step (61):  FM_update; PreLOD per substrate;
//           for each diffusion step (11): LODsolver per substrate
//   death
//
// Usage: gell_sim [N] [cells] [bio_steps] [substrates]
// Seeds a spheroid of `cells` cells uniformly in a ball of 0.45 N dx around
// the domain centre (radii 4.5-5.5 um, the reference's volume split), runs the
// loop and prints one line per biology step: count, wall ms, agent-steps/s.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "gell.hpp"

namespace {

std::vector<gell_cell> seed_spheroid(int32_t n, int32_t capacity, int32_t N, float dx, uint64_t seed) {
    std::vector<gell_cell> cells(capacity);
    for (auto& c : cells) {  // Cell's default member values (Cells.cuh:10-33)
        std::memset(&c, 0, sizeof(c));
        c.mesh_idx = -1;
        c.pos = {-1.f, -1.f, -1.f};
        c.r = c.V = c.Vf = c.Vcs = c.Vns = -1.f;
        c.phase = -1;
    }
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> u(-1.f, 1.f), ur(4.5f, 5.5f);
    const float centre = 0.5f * N * dx, R = 0.45f * N * dx;
    for (int32_t i = 0; i < n;) {
        const float x = u(rng), y = u(rng), z = u(rng);
        if (x * x + y * y + z * z >= 1.f) continue;  // Initialization.cuh:22-40
        gell_cell& c = cells[i++];
        c.sign = 1;
        c.pos = {centre + x * R, centre + y * R, centre + z * R};
        c.r = ur(rng);
        c.V = 4.18879020f * c.r * c.r * c.r;
        c.Vf = c.V * (1871.f / 2495.f);
        c.Vns = c.V * (135.f / 2495.f);
        c.Vcs = c.V * (489.f / 2495.f);
        c.phase = 0;
        c.cell_clock = 0.f;
        const int a = (int)(c.pos.x / dx), b = (int)(c.pos.y / dx), k = (int)(c.pos.z / dx);
        c.mesh_idx = a + N * (b + N * k);
    }
    return cells;
}

}  // namespace

int main(int argc, char** argv) {
    const int32_t N = argc > 1 ? std::atoi(argv[1]) : 64;
    const int32_t n0 = argc > 2 ? std::atoi(argv[2]) : 10000;
    const int32_t bio_steps = argc > 3 ? std::atoi(argv[3]) : 2;
    const int32_t n_sub = argc > 4 ? std::atoi(argv[4]) : 1;

    gell_params P;
    gell_default_params(&P);
    P.N = N;
    P.dx = 10.f;
    P.capacity = 2 * n0;
    P.seed = 20240611;
    P.n_substrates = n_sub;
    if (n_sub > 1) P.sub[1] = gell_substrate{5.f, 4.0e4f, 0.f, -1.f, 0.1f, 5.f};

    try {
        gell::Context ctx(P);
        ctx.autotune();
        gell::MechanicsMesh mech(ctx);
        std::vector<gell::Mesh> meshes;
        for (int s = 0; s < n_sub; ++s) meshes.emplace_back(ctx, s);

        std::vector<gell_cell> host = seed_spheroid(n0, P.capacity, N, P.dx, P.seed);
        gell_cell* d_cells = nullptr;
        if (hipMalloc(&d_cells, sizeof(gell_cell) * P.capacity) != hipSuccess) return 2;
        if (hipMemcpy(d_cells, host.data(), sizeof(gell_cell) * P.capacity, hipMemcpyHostToDevice) != hipSuccess)
            return 2;

        const int32_t mech_steps = gell::float_loop_count(P.dt_bio, P.dt_mech);  // 61
        const int32_t diff_steps = gell::float_loop_count(P.dt_mech, P.dt_diff);  // 11
        int32_t n = n0;
        for (int32_t b = 0; b < bio_steps; ++b) {
            const auto t0 = std::chrono::steady_clock::now();
            n = gell::CellBirth_kernel(ctx, d_cells, n, (uint64_t)b);
            int64_t agents = 0;
            for (int32_t m = 0; m < mech_steps; ++m) {
                mech.FM_update(d_cells, n);
                for (auto& mesh : meshes) mesh.PreLOD(d_cells, n);
                for (int32_t d = 0; d < diff_steps; ++d)
                    for (auto& mesh : meshes) mesh.LODsolver(d_cells, n);
                agents += n;
            }
            n = gell::CellDeath_kernel(ctx, d_cells, n);
            ctx.sync();
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            std::printf("bio step %d: cells %d  %.2f ms  %.3e agent-steps/s\n", b, n, ms, agents / (ms * 1e-3));
        }
        (void)hipFree(d_cells);
    } catch (const gell::Error& e) {
        std::fprintf(stderr, "gell_sim: %s\n", e.what());
        return 1;
    }
    return 0;
}
