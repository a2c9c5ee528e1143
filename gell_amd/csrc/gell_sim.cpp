// gell_sim.cpp -- a C++ driver with the structure of the reference's time
// loop (Main.cu:108-127), written against gell.hpp:
//
//   for each biology step:  birth
//       for each mechanics step (61):  FM_update; PreLOD per substrate;
//           for each diffusion step (11): LODsolver per substrate
//   death
//
// Two modes.
//
//   gell_sim [N] [cells] [bio_steps] [substrates]
//     Seeds a spheroid of `cells` cells uniformly in a ball of 0.45 N dx
//     around the domain centre (radii 4.5-5.5 um), runs the loop and prints
//     one line per biology step: count, wall ms, agent-steps/s.
//
//   gell_sim hds [--out DIR] [--steps K] [--seed S] [--save-cells] [--save-mesh]
//                [--checkpoint-every M] [--restart FILE] [--exact] [--fm V]
//     The reference's own example (Definitions.h as shipped, Main.cu): a
//     100^3 grid of 25 um voxels, 2347 cells of radius 8.4 um seeded by
//     Cell_sphere_Initialization, 19 days of biology steps (4560), capacity
//     2M. Writes the reference's files into DIR: Gell_0.csv,
//     Mesh_slice_0.csv, Mesh_slice_<i>.csv every 360 min of biology
//     (Main.cu:132-170), Gell_Time_record.csv and Gell_GapTime_record.csv
//     (Main.cu:206-207); with --save-cells also Gell_<i>.csv at each gap and
//     Gell_Simuend.csv, with --save-mesh Mesh_Simuend.csv. Every M steps the
//     state goes to DIR/checkpoint.bin; --restart continues from such a file.
//     --exact restricts the autotuner to kernel variants that are bit-identical
//     to the serial sweeps; --fm V sets the force-kernel variant
//     (GELL_KNOB_FM_PREFETCH). Prints one line per 360 min of biology.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
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

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int run_bench(int argc, char** argv) {
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
        const double ms = 1e3 * seconds_since(t0);
        std::printf("bio step %d: cells %d  %.2f ms  %.3e agent-steps/s\n", b, n, ms, agents / (ms * 1e-3));
    }
    (void)hipFree(d_cells);
    return 0;
}

struct HdsOptions {
    std::string out = ".";
    int32_t steps = -1;  // default: the whole Max_Simulation_Time
    uint64_t seed = 12345;
    bool save_cells = false, save_mesh = false, exact = false;
    int32_t checkpoint_every = 0;
    int32_t fm_variant = -1;
    std::string restart;
};

bool parse_hds(int argc, char** argv, HdsOptions& o) {
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "gell_sim hds: %s needs a value\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--out") o.out = next("--out");
        else if (a == "--steps") o.steps = std::atoi(next("--steps"));
        else if (a == "--seed") o.seed = std::strtoull(next("--seed"), nullptr, 10);
        else if (a == "--save-cells") o.save_cells = true;
        else if (a == "--save-mesh") o.save_mesh = true;
        else if (a == "--exact") o.exact = true;
        else if (a == "--fm") o.fm_variant = std::atoi(next("--fm"));
        else if (a == "--checkpoint-every") o.checkpoint_every = std::atoi(next("--checkpoint-every"));
        else if (a == "--restart") o.restart = next("--restart");
        else {
            std::fprintf(stderr, "gell_sim hds: unknown option %s\n", a.c_str());
            return false;
        }
    }
    return true;
}

int run_hds(int argc, char** argv) {
    HdsOptions o;
    if (!parse_hds(argc, argv, o)) return 2;

    // Definitions.h as shipped: Voxel_num 100, Voxel_length 25, Max_Cell_num
    // 2M, Cell_Init_num 2347, Tumor_Cell_Radius 8.4, Max_Simulation_Time 19 d,
    // Save_data_gap 360 min, and the oxygen constants of gell_default_params.
    gell_params P;
    gell_default_params(&P);
    const int32_t init_num = 2347;
    const float cell_radius = 8.4f, max_time = 19.f * 24 * 60;
    const int32_t save_gap_min = 360;
    const int32_t max_iter = (int32_t)(max_time / P.dt_bio);       // Max_Simulation_Iter, 4560
    const int32_t iter_savegap = (int32_t)(save_gap_min / P.dt_bio);  // 60
    const int32_t steps = o.steps < 0 ? gell::float_loop_count(max_time, P.dt_bio) : o.steps;
    const std::string dir = o.out + "/";

    gell::Context ctx(P);
    ctx.autotune(o.exact);
    if (o.fm_variant >= 0)
        gell::check(ctx.get(), gell_tune(ctx.get(), GELL_KNOB_FM_PREFETCH, o.fm_variant), "gell_tune");
    gell::MechanicsMesh MecMesh(ctx);
    gell::Mesh O2Mesh(ctx, 0);

    gell_cell* d_cells = nullptr;
    if (hipMalloc(&d_cells, sizeof(gell_cell) * P.capacity) != hipSuccess) return 2;
    std::vector<gell_cell> host(P.capacity);
    std::vector<float> field((size_t)P.N * P.N * P.N);
    auto copy_field = [&]() {
        ctx.sync();
        if (hipMemcpy(field.data(), O2Mesh.p(), sizeof(float) * field.size(), hipMemcpyDeviceToHost) != hipSuccess)
            throw gell::Error(GELL_E_HIP, "field copy failed");
    };
    auto copy_cells = [&](int32_t n) {
        ctx.sync();
        if (hipMemcpy(host.data(), d_cells, sizeof(gell_cell) * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
            throw gell::Error(GELL_E_HIP, "cell copy failed");
    };

    int32_t currentnum = init_num;
    uint64_t first = 0;
    if (o.restart.empty()) {
        gell::io_check(gell_sphere_init(host.data(), P.capacity, init_num, &P, cell_radius, o.seed), "sphere_init");
        if (hipMemcpy(d_cells, host.data(), sizeof(gell_cell) * P.capacity, hipMemcpyHostToDevice) != hipSuccess)
            return 2;
        gell::savecsv_cell(host.data(), currentnum, dir + "Gell_0.csv");
        copy_field();
        gell::savecsv_meshslice(field.data(), P.N, dir + "Mesh_slice_0.csv");
    } else {
        std::vector<gell_cell> blank(P.capacity);
        gell::io_check(gell_sphere_init(blank.data(), P.capacity, 0, &P, cell_radius, 0), "sphere_init");
        if (hipMemcpy(d_cells, blank.data(), sizeof(gell_cell) * P.capacity, hipMemcpyHostToDevice) != hipSuccess)
            return 2;
        gell::checkpoint_load(ctx, d_cells, &currentnum, &first, o.restart);
        std::printf("restart from %s: biology step %llu, %d cells\n", o.restart.c_str(), (unsigned long long)first,
                    currentnum);
    }

    std::vector<int32_t> cellnum((size_t)std::max(max_iter, steps + (int32_t)first) + 1, 0);
    std::vector<double> itertime(cellnum.size(), 0.0);
    std::vector<double> gaptime((size_t)(max_time / save_gap_min) + 1, 0.0);
    cellnum[first] = currentnum;

    const int32_t mech_steps = gell::float_loop_count(P.dt_bio, P.dt_mech);  // 61
    const int32_t diff_steps = gell::float_loop_count(P.dt_mech, P.dt_diff);  // 11
    const auto start = std::chrono::steady_clock::now();
    double gapduration = 0.0;
    int64_t agent_steps = 0;
    std::printf("HDS: N %d, dx %.0f um, %d cells, %d biology steps from step %llu\n", P.N, P.dx, currentnum, steps,
                (unsigned long long)first);
    std::fflush(stdout);
    for (int32_t b = (int32_t)first; b < (int32_t)first + steps; ++b) {
        const int32_t SimuIter = b + 1;
        const auto t0 = std::chrono::steady_clock::now();
        currentnum = gell::CellBirth_kernel(ctx, d_cells, currentnum, (uint64_t)b);
        for (int32_t m = 0; m < mech_steps; ++m) {
            MecMesh.FM_update(d_cells, currentnum);
            O2Mesh.PreLOD(d_cells, currentnum);
            for (int32_t d = 0; d < diff_steps; ++d) O2Mesh.LODsolver(d_cells, currentnum);
        }
        agent_steps += (int64_t)mech_steps * currentnum;
        currentnum = gell::CellDeath_kernel(ctx, d_cells, currentnum);
        ctx.sync();
        const double iterduration = seconds_since(t0);
        gapduration += iterduration;
        cellnum[SimuIter] = currentnum;
        itertime[SimuIter] = iterduration;

        if ((SimuIter + 1) % iter_savegap == 0) {  // Main.cu:134-181
            const int32_t filesaveidx = (SimuIter + 1) / iter_savegap;
            if (filesaveidx - 1 < (int32_t)gaptime.size()) gaptime[filesaveidx - 1] = gapduration;
            if (o.save_cells) {
                copy_cells(currentnum);
                gell::savecsv_cell(host.data(), currentnum, dir + "Gell_" + std::to_string(filesaveidx) + ".csv");
            }
            copy_field();
            gell::savecsv_meshslice(field.data(), P.N, dir + "Mesh_slice_" + std::to_string(filesaveidx) + ".csv");
            std::printf("%3d h: %8d cells  gap %.2f s  total %.1f s  %.3e agent-steps/s\n",
                        (int)((filesaveidx - 1) * save_gap_min / 60), currentnum, gapduration, seconds_since(start),
                        agent_steps / seconds_since(start));
            std::fflush(stdout);
            gapduration = 0.0;
        }
        if (o.checkpoint_every > 0 && SimuIter % o.checkpoint_every == 0)
            gell::checkpoint_save(ctx, d_cells, currentnum, (uint64_t)SimuIter, dir + "checkpoint.bin");
    }
    const double total = seconds_since(start);
    std::printf("HDS finished: %d cells after %d biology steps (%.1f h), %.1f s, %.3e agent-steps/s\n", currentnum,
                (int)first + steps, ((int)first + steps) * P.dt_bio / 60.0, total, agent_steps / total);

    const auto save0 = std::chrono::steady_clock::now();
    if (o.save_cells) {
        copy_cells(currentnum);
        gell::savecsv_cell(host.data(), currentnum, dir + "Gell_Simuend.csv");
    }
    copy_field();
    if (o.save_mesh) gell::savecsv_mesh(field.data(), P.N, dir + "Mesh_Simuend.csv");
    gell::savecsv_meshslice(field.data(), P.N, dir + "Mesh_slice_Simuend.csv");
    gell::savecsv_array(cellnum.data(), itertime.data(), std::min((int32_t)cellnum.size(), (int32_t)first + steps + 1),
                        dir + "Gell_Time_record.csv");
    gell::savecsv_gaparray(cellnum.data(), gaptime.data(), std::min(max_iter, (int32_t)cellnum.size() - 1),
                           save_gap_min, dir + "Gell_GapTime_record.csv");
    std::printf("files written in %.2f s\n", seconds_since(save0));
    (void)hipFree(d_cells);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc > 1 && std::strcmp(argv[1], "hds") == 0) return run_hds(argc, argv);
        return run_bench(argc, argv);
    } catch (const gell::Error& e) {
        std::fprintf(stderr, "gell_sim: %s\n", e.what());
        return 1;
    }
}
