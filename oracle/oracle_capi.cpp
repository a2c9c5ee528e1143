// oracle_capi.cpp -- extern "C" face of the CPU oracle, loaded by the tests
// (ctypes) and timed by bench.py's CPU-baseline leg. Test infrastructure only;
// the product never loads this library.
#include <chrono>
#include <cstdio>
#include <vector>

#include "gell_oracle.hpp"

using namespace gell_oracle;

namespace {

template <class R>
std::vector<CellT<R>> load_cells(const gell_cell* c, int32_t count) {
    std::vector<CellT<R>> v(count);
    for (int32_t i = 0; i < count; ++i) v[i] = cell_cast<R>(c[i]);
    return v;
}

template <class R>
void store_cells(const std::vector<CellT<R>>& v, gell_cell* c, int32_t count) {
    for (int32_t i = 0; i < count; ++i) c[i] = cell_to_abi(v[i]);
}

double now_ms() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double, std::milli>(clk::now().time_since_epoch()).count();
}

// One mechanics step of Main.cu:114-124 plus its diffusion substeps, for every
// substrate. `times` accumulates {fm, prelod, lod} milliseconds.
template <class R>
void mech_step(const gell_params* P, std::vector<CellT<R>>& cells, int32_t n, std::vector<std::vector<R>>& fields,
               std::vector<std::vector<R>>& rates, const std::vector<std::vector<R>>& E,
               const std::vector<std::vector<R>>& F, int32_t diff_steps, FMState<R>& st, double* times) {
    double t0 = now_ms();
    fm_update<R>(cells, n, P->N, (R)P->dx, (R)P->dt_mech, P->capacity, st);
    double t1 = now_ms();
    for (int s = 0; s < P->n_substrates; ++s)
        prelod<R>(cells, n, P->N, (R)P->dx, (R)P->sub[s].uptake, (R)P->sub[s].dead_factor, rates[s]);
    double t2 = now_ms();
    for (int d = 0; d < diff_steps; ++d)
        for (int s = 0; s < P->n_substrates; ++s)
            lod_substep<R>(fields[s], rates[s], P->N, (R)P->dt_diff, E[s].data(), F[s].data(), (R)P->sub[s].boundary);
    double t3 = now_ms();
    if (times) {
        times[0] += t1 - t0;
        times[1] += t2 - t1;
        times[2] += t3 - t2;
    }
}

// The time loop of Main.cu:108-127 for `bio_steps` biology steps, with the
// mechanics loop cut to `mech_limit` steps when mech_limit >= 0 and PD switched
// off when with_pd == 0 (used by drift runs, whose event sets would otherwise
// diverge between precisions).
template <class R>
int32_t run(const gell_params* P, gell_cell* io_cells, int32_t n, void* io_fields, int32_t bio_steps,
            int32_t mech_limit, int32_t with_pd, uint64_t first_step, double* times) {
    const int64_t V = (int64_t)P->N * P->N * P->N;
    std::vector<CellT<R>> cells = load_cells<R>(io_cells, P->capacity);
    std::vector<std::vector<R>> fields(P->n_substrates), rates(P->n_substrates), E(P->n_substrates),
        F(P->n_substrates);
    R* fio = (R*)io_fields;
    for (int s = 0; s < P->n_substrates; ++s) {
        fields[s].assign(fio + s * V, fio + (s + 1) * V);
        rates[s].assign(V, R(0));
        lod_coeffs<R>(P->N, (R)P->dx, (R)P->dt_diff, (R)P->sub[s].D, (R)P->sub[s].decay, E[s], F[s]);
    }
    const int32_t mech_steps = float_loop_count(P->dt_bio, P->dt_mech);
    const int32_t diff_steps = float_loop_count(P->dt_mech, P->dt_diff);
    FMState<R> st;
    std::vector<int32_t> slots;
    for (int32_t b = 0; b < bio_steps; ++b) {
        if (with_pd)
            n = birth<R>(cells, n, fields[0].data(), P->N, (R)P->dx, (R)P->dt_bio, P->capacity, P->seed,
                         first_step + b, slots);
        const int32_t m_end = mech_limit >= 0 ? std::min(mech_limit, mech_steps) : mech_steps;
        for (int32_t m = 0; m < m_end; ++m) mech_step<R>(P, cells, n, fields, rates, E, F, diff_steps, st, times);
        if (with_pd) n = death<R>(cells, n);
    }
    store_cells<R>(cells, io_cells, P->capacity);
    for (int s = 0; s < P->n_substrates; ++s) std::copy(fields[s].begin(), fields[s].end(), fio + s * V);
    return n;
}

// Exactly `steps` mechanics steps (FM, uptake maps, diffusion substeps), no PD.
template <class R>
int32_t run_mech(const gell_params* P, gell_cell* io_cells, int32_t n, void* io_fields, int32_t steps) {
    const int64_t V = (int64_t)P->N * P->N * P->N;
    std::vector<CellT<R>> cells = load_cells<R>(io_cells, n);
    std::vector<std::vector<R>> fields(P->n_substrates), rates(P->n_substrates), E(P->n_substrates),
        F(P->n_substrates);
    R* fio = (R*)io_fields;
    for (int s = 0; s < P->n_substrates; ++s) {
        fields[s].assign(fio + s * V, fio + (s + 1) * V);
        rates[s].assign(V, R(0));
        lod_coeffs<R>(P->N, (R)P->dx, (R)P->dt_diff, (R)P->sub[s].D, (R)P->sub[s].decay, E[s], F[s]);
    }
    const int32_t diff_steps = float_loop_count(P->dt_mech, P->dt_diff);
    FMState<R> st;
    for (int32_t m = 0; m < steps; ++m) mech_step<R>(P, cells, n, fields, rates, E, F, diff_steps, st, nullptr);
    store_cells<R>(cells, io_cells, n);
    for (int s = 0; s < P->n_substrates; ++s) std::copy(fields[s].begin(), fields[s].end(), fio + s * V);
    return n;
}

}  // namespace

extern "C" {

int32_t orc_run_mech(const gell_params* P, gell_cell* cells, int32_t n, float* fields, int32_t steps) {
    return run_mech<float>(P, cells, n, fields, steps);
}

int32_t orc_run_mech_f64(const gell_params* P, gell_cell* cells, int32_t n, double* fields, int32_t steps) {
    return run_mech<double>(P, cells, n, fields, steps);
}

int orc_omp_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    U4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
    U4 r = philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void orc_birth_uniforms(uint64_t seed, uint32_t slot, uint64_t step, float* out) {
    birth_uniforms(seed, slot, step, out);
}

int32_t orc_morton(int32_t i, int32_t j, int32_t k) { return morton_encode(i, j, k); }

void orc_morton_decode(int32_t key, int32_t* ijk) { morton_decode(key, ijk[0], ijk[1], ijk[2]); }

int32_t orc_float_loop_count(float end, float step) { return float_loop_count(end, step); }

// Key pass only (updateCellMeshIdx): keys[i] and the records' sign/mesh_idx.
void orc_fm_keys(const gell_params* P, gell_cell* cells, int32_t n, int32_t* keys) {
    std::vector<CellT<float>> c = load_cells<float>(cells, n);
    for (int32_t i = 0; i < n; ++i) keys[i] = cell_key(c[i], P->N, P->dx);
    store_cells(c, cells, n);
}

// FM_update on cells[0, n); outputs sorted keys, order (sorted slot -> input
// slot) and the per-voxel ranges s, e (N^3 each).
void orc_fm_update(const gell_params* P, gell_cell* cells, int32_t n, int32_t* keys, int32_t* order, int32_t* s,
                   int32_t* e) {
    std::vector<CellT<float>> c = load_cells<float>(cells, n);
    FMState<float> st;
    fm_update<float>(c, n, P->N, P->dx, P->dt_mech, P->capacity, st);
    store_cells(c, cells, n);
    std::copy(st.keys.begin(), st.keys.end(), keys);
    std::copy(st.order.begin(), st.order.end(), order);
    std::copy(st.s.begin(), st.s.end(), s);
    std::copy(st.e.begin(), st.e.end(), e);
}

// Force of every live cell of an already sorted array with given ranges
// (forcekernel alone); dead cells get 0.
void orc_forces(const gell_params* P, const gell_cell* cells, int32_t n, const int32_t* s, const int32_t* e,
                float* f3) {
    std::vector<CellT<float>> c = load_cells<float>(cells, n);
    for (int32_t i = 0; i < n; ++i) {
        V3<float> f{0, 0, 0};
        if (c[i].sign) f = cell_force(c.data(), i, s, e, P->N, P->dx);
        f3[3 * i] = f.x;
        f3[3 * i + 1] = f.y;
        f3[3 * i + 2] = f.z;
    }
}

// movementkernel alone on cells[0, n).
void orc_move(const gell_params* P, gell_cell* cells, int32_t n) {
    std::vector<CellT<float>> c = load_cells<float>(cells, n);
    for (int32_t i = 0; i < n; ++i)
        if (c[i].sign) cell_move(c[i], P->dt_mech, P->N, P->dx);
    store_cells(c, cells, n);
}

void orc_prelod(const gell_params* P, int32_t sub, const gell_cell* cells, int32_t n, float* p_rate) {
    std::vector<CellT<float>> c = load_cells<float>(cells, n);
    std::vector<float> r;
    prelod<float>(c, n, P->N, P->dx, P->sub[sub].uptake, P->sub[sub].dead_factor, r);
    std::copy(r.begin(), r.end(), p_rate);
}

float orc_cell_uptake(const gell_params* P, int32_t sub, const gell_cell* cell) {
    CellT<float> c = cell_cast<float>(*cell);
    return cell_uptake<float>(c, P->dx, P->sub[sub].uptake, P->sub[sub].dead_factor);
}

void orc_lod_coeffs(const gell_params* P, int32_t sub, float* E, float* F) {
    std::vector<float> e, f;
    lod_coeffs<float>(P->N, P->dx, P->dt_diff, P->sub[sub].D, P->sub[sub].decay, e, f);
    std::copy(e.begin(), e.end(), E);
    std::copy(f.begin(), f.end(), F);
}

void orc_lod_coeffs_f64(const gell_params* P, int32_t sub, double* E, double* F) {
    std::vector<double> e, f;
    lod_coeffs<double>(P->N, P->dx, P->dt_diff, P->sub[sub].D, P->sub[sub].decay, e, f);
    std::copy(e.begin(), e.end(), E);
    std::copy(f.begin(), f.end(), F);
}

// One Thomas line (stride 1) with the given coefficients.
void orc_thomas_line(float* p, int32_t N, const float* E, const float* F, float bnd) {
    std::vector<float> Y(N);
    thomas_line<float>(p, 0, 1, N, E, F, bnd, Y.data());
}

void orc_lod_substep(const gell_params* P, int32_t sub, float* p, const float* p_rate) {
    const int64_t V = (int64_t)P->N * P->N * P->N;
    std::vector<float> e, f;
    lod_coeffs<float>(P->N, P->dx, P->dt_diff, P->sub[sub].D, P->sub[sub].decay, e, f);
    std::vector<float> pv(p, p + V), rv(p_rate, p_rate + V);
    lod_substep<float>(pv, rv, P->N, P->dt_diff, e.data(), f.data(), P->sub[sub].boundary);
    std::copy(pv.begin(), pv.end(), p);
}

void orc_lod_substep_f64(const gell_params* P, int32_t sub, double* p, const double* p_rate) {
    const int64_t V = (int64_t)P->N * P->N * P->N;
    std::vector<double> e, f;
    lod_coeffs<double>(P->N, P->dx, P->dt_diff, P->sub[sub].D, P->sub[sub].decay, e, f);
    std::vector<double> pv(p, p + V), rv(p_rate, p_rate + V);
    lod_substep<double>(pv, rv, P->N, P->dt_diff, e.data(), f.data(), P->sub[sub].boundary);
    std::copy(pv.begin(), pv.end(), p);
}

void orc_volume_update(const gell_params* P, gell_cell* cell) {
    CellT<float> c = cell_cast<float>(*cell);
    volume_update<float>(c, P->dt_bio);
    *cell = cell_to_abi(c);
}

int32_t orc_phase_update(const gell_params* P, gell_cell* cell, float rnd, float O2) {
    CellT<float> c = cell_cast<float>(*cell);
    bool pr = phase_update<float>(c, rnd, O2, P->dt_bio);
    *cell = cell_to_abi(c);
    return pr ? 1 : 0;
}

void orc_random_dir(float u, float v, float* out) {
    V3<float> d = random_dir<float>(u, v);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

// CellBirth_kernel on cells (capacity records); slots gets n entries.
int32_t orc_birth(const gell_params* P, const float* p0, gell_cell* cells, int32_t n, uint64_t step,
                  int32_t* slots) {
    std::vector<CellT<float>> c = load_cells<float>(cells, P->capacity);
    std::vector<int32_t> sl;
    int32_t nn = birth<float>(c, n, p0, P->N, P->dx, P->dt_bio, P->capacity, P->seed, step, sl);
    store_cells(c, cells, P->capacity);
    std::copy(sl.begin(), sl.end(), slots);
    return nn;
}

int32_t orc_death(gell_cell* cells, int32_t n) {
    std::vector<CellT<float>> c = load_cells<float>(cells, n);
    int32_t nn = death<float>(c, n);
    store_cells(c, cells, n);
    return nn;
}

// Full time loop (float). `fields` holds n_substrates * N^3 floats in and out;
// `cells` holds capacity records. times (may be NULL) gets {fm, prelod, lod} ms.
int32_t orc_run(const gell_params* P, gell_cell* cells, int32_t n, float* fields, int32_t bio_steps,
                int32_t mech_limit, int32_t with_pd, uint64_t first_step, double* times) {
    return run<float>(P, cells, n, fields, bio_steps, mech_limit, with_pd, first_step, times);
}

// Same loop in double precision; fields are doubles, cells are rounded to
// float on the way out.
int32_t orc_run_f64(const gell_params* P, gell_cell* cells, int32_t n, double* fields, int32_t bio_steps,
                    int32_t mech_limit, int32_t with_pd, uint64_t first_step) {
    return run<double>(P, cells, n, fields, bio_steps, mech_limit, with_pd, first_step, nullptr);
}

}  // extern "C"
