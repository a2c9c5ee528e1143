// gell_oracle.hpp -- CPU restatement of the Gell per-step hot path.
//
// TEST INFRASTRUCTURE ONLY. This is the correctness oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as
// the checker or the timed CPU baseline. The product (gell_amd/) never links,
// loads or calls it.
//
// It follows the reference kernels statement by statement (paths relative to
// the Gell sources), templated on the scalar type: R = float is the parity
// oracle (compile with -ffp-contract=off so no fused multiply-adds appear),
// R = double is the high-precision run that drift is measured against.
//
// Deterministic replacements of the reference's nondeterministic parts
// (each is the semantics the HIP path implements too):
//   * sort_by_key (FMkernel.cuh:190): std::stable_sort, ties keep prior order.
//   * PreLOD float atomicAdd (LODkernel.cuh:21): exact fixed-point sum
//     (int64, 2^-32 units) per voxel, rounded once to float -- independent of
//     order, so equal to any serial order up to that one rounding.
//   * cuRAND XORWOW per slot (RandomFunctions.cuh:13-32, PDkernel.cuh:32,56):
//     counter-based Philox4x32-10 keyed by (seed), counter (slot, step).
//   * daughter slots from atomicAdd (PDkernel.cuh:51-54): exclusive prefix
//     count of dividing mothers in slot order; capacity clamp in that order.
//
// Pinning (what checks each function; see the tests named):
//   morton_encode      brute-force bijection + bit-loop closed form  (test_oracle_basic)
//   philox4x32_10      Random123 published KAT + rocRAND host engine vectors (test_oracle_basic)
//   lod_coeffs/sweep   dense numpy solve of the same tridiagonal system (test_oracle_lod)
//   lod_substep        fixed point p == boundary, implicit-uptake closed form,
//                      separable sine-mode decay factor (test_oracle_lod)
//   fm keys/sort/s,e   library stable sort, Check_se invariant, brute force (test_oracle_fm)
//   force              equal-radius brute-force O(n^2) sum, antisymmetry (test_oracle_fm)
//   move               closed-form AB2 on hand-set forces, out-of-domain flag (test_oracle_fm)
//   prelod             closed form -uptake*sum(V)/dx^3 (test_oracle_pd)
//   volume/phase       closed-form relaxation, hand-set rand thresholds (test_oracle_pd)
//   birth/death        n' = n + births - deaths, daughter geometry,
//                      std::stable_partition (test_oracle_pd)
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "../include/gell_abi.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace gell_oracle {

// ---------------------------------------------------------------------------
// Records
// ---------------------------------------------------------------------------
template <class R>
struct V3 {
    R x, y, z;
};

// Cells.cuh:9-47; for R = float this is byte-for-byte gell_cell.
template <class R>
struct CellT {
    uint8_t sign;
    uint8_t pad_[3];
    int32_t mesh_idx;
    V3<R> pos, force, oldforce;
    R r, V, Vf, Vcs, Vns;
    int32_t phase;
    R cell_clock;
};
static_assert(sizeof(CellT<float>) == 72, "Cell record must be 72 bytes");
static_assert(sizeof(CellT<float>) == sizeof(gell_cell), "oracle and ABI records differ");

template <class R>
inline CellT<R> cell_cast(const gell_cell& c) {
    CellT<R> o;
    std::memset(&o, 0, sizeof(o));
    o.sign = c.sign;
    o.mesh_idx = c.mesh_idx;
    o.pos = {R(c.pos.x), R(c.pos.y), R(c.pos.z)};
    o.force = {R(c.force.x), R(c.force.y), R(c.force.z)};
    o.oldforce = {R(c.oldforce.x), R(c.oldforce.y), R(c.oldforce.z)};
    o.r = c.r; o.V = c.V; o.Vf = c.Vf; o.Vcs = c.Vcs; o.Vns = c.Vns;
    o.phase = c.phase;
    o.cell_clock = c.cell_clock;
    return o;
}

template <class R>
inline gell_cell cell_to_abi(const CellT<R>& c) {
    gell_cell o;
    std::memset(&o, 0, sizeof(o));
    o.sign = c.sign;
    o.mesh_idx = c.mesh_idx;
    o.pos = {float(c.pos.x), float(c.pos.y), float(c.pos.z)};
    o.force = {float(c.force.x), float(c.force.y), float(c.force.z)};
    o.oldforce = {float(c.oldforce.x), float(c.oldforce.y), float(c.oldforce.z)};
    o.r = float(c.r); o.V = float(c.V); o.Vf = float(c.Vf); o.Vcs = float(c.Vcs); o.Vns = float(c.Vns);
    o.phase = c.phase;
    o.cell_clock = float(c.cell_clock);
    return o;
}

// ---------------------------------------------------------------------------
// Index helpers (BasicFunctions.cuh:18-66)
// ---------------------------------------------------------------------------

// (int)(x) of Find_Mesh (BasicFunctions.cuh:60-66): truncation toward zero.
// Values outside the int range (and NaN) map to INT32_MIN, which every
// inside-test rejects; the reference leaves that case undefined.
template <class R>
inline int32_t trunc_index(R q) {
    if (!(q > R(-2147483648.0) && q < R(2147483647.0))) return INT32_MIN;
    return (int32_t)q;
}

template <class R>
inline void find_mesh(const V3<R>& p, R dx, int32_t& i, int32_t& j, int32_t& k) {
    i = trunc_index<R>(p.x / dx);
    j = trunc_index<R>(p.y / dx);
    k = trunc_index<R>(p.z / dx);
}

// Check_Inside, BasicFunctions.cuh:18-28
inline bool check_inside(int32_t i, int32_t j, int32_t k, int32_t N) {
    return !(i < 0 || i >= N || j < 0 || j >= N || k < 0 || k >= N);
}

// Find_Index, BasicFunctions.cuh:31-36 (x fastest)
inline int64_t lin_index(int32_t i, int32_t j, int32_t k, int32_t N) {
    return (int64_t)i + (int64_t)N * ((int64_t)j + (int64_t)N * (int64_t)k);
}

// Morton_encode, BasicFunctions.cuh:111-137: 10 bits per axis, x in bit 0,
// y in bit 1, z in bit 2 of every triple. Written here as the plain per-bit
// interleave so that it pins the device's magic-mask version.
inline int32_t morton_encode(int32_t i, int32_t j, int32_t k) {
    uint32_t key = 0;
    for (int b = 0; b < 10; ++b) {
        key |= (uint32_t)((i >> b) & 1) << (3 * b);
        key |= (uint32_t)((j >> b) & 1) << (3 * b + 1);
        key |= (uint32_t)((k >> b) & 1) << (3 * b + 2);
    }
    return (int32_t)key;
}

inline void morton_decode(int32_t key, int32_t& i, int32_t& j, int32_t& k) {
    i = j = k = 0;
    for (int b = 0; b < 10; ++b) {
        i |= ((key >> (3 * b)) & 1) << b;
        j |= ((key >> (3 * b + 1)) & 1) << b;
        k |= ((key >> (3 * b + 2)) & 1) << b;
    }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11), the counter-based generator that replaces the
// reference's per-slot XORWOW states.
// ---------------------------------------------------------------------------
struct U4 {
    uint32_t v[4];
};

inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += W0; k1 += W1; }
        uint64_t p0 = (uint64_t)M0 * ctr.v[0];
        uint64_t p1 = (uint64_t)M1 * ctr.v[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
    }
    return ctr;
}

// Uniform in (0, 1] from the top 24 bits -- the range of curand_uniform that
// the reference's `rand <= p` tests assume (PDkernel.cuh:32, Cells.cuh:204-213).
inline float u01(uint32_t w) { return (float)((w >> 8) + 1u) * (1.0f / 16777216.0f); }

// Stream layout: key = seed, counter = (slot, step lo, step hi, tag).
// Tag 0 = birth draws: word 0 -> phase draw (PDkernel.cuh:32), words 1, 2 ->
// the two direction draws u, v of randomvect (RandomFunctions.cuh:24-25).
inline void birth_uniforms(uint64_t seed, uint32_t slot, uint64_t step, float out[3]) {
    U4 c{{slot, (uint32_t)step, (uint32_t)(step >> 32), 0u}};
    U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    out[0] = u01(r.v[0]);
    out[1] = u01(r.v[1]);
    out[2] = u01(r.v[2]);
}

// ---------------------------------------------------------------------------
// Mechanics (FMkernel.cuh)
// ---------------------------------------------------------------------------
template <class R>
struct FMState {
    std::vector<int32_t> keys;   // sorted keys
    std::vector<int32_t> order;  // sorted slot -> input slot
    std::vector<int32_t> s, e;   // per-voxel live ranges
};

// updateCellMeshIdx, FMkernel.cuh:22-36 (all slots, dead or alive).
template <class R>
inline int32_t cell_key(CellT<R>& c, int32_t N, R dx) {
    int32_t i, j, k;
    find_mesh(c.pos, dx, i, j, k);
    if (!check_inside(i, j, k, N)) {
        c.sign = 0;
        return -1;  // Sort_Key_Last, Definitions.h:16
    }
    c.mesh_idx = (int32_t)lin_index(i, j, k, N);
    return morton_encode(i, j, k);
}

// forcekernel, FMkernel.cuh:56-118. Returns the force on live cell i.
template <class R>
inline V3<R> cell_force(const CellT<R>* cell, int32_t i, const int32_t* s, const int32_t* e,
                        int32_t N, R dx) {
    const V3<R> pos1 = cell[i].pos;
    const R r1 = cell[i].r;
    int32_t px, py, pz;
    find_mesh(pos1, dx, px, py, pz);
    R fx = 0, fy = 0, fz = 0;
    R thres = R(25.0f);  // Tumor_CC_Max_Dist, carried across pairs (FMkernel.cuh:72,91,99)
    const R Cccr = R(10.f), Ccca = R(0.4f);
    for (int32_t xi = px - 1; xi <= px + 1; ++xi)
        for (int32_t yi = py - 1; yi <= py + 1; ++yi)
            for (int32_t zi = pz - 1; zi <= pz + 1; ++zi) {
                if (!check_inside(xi, yi, zi, N)) continue;
                const int64_t m = lin_index(xi, yi, zi, N);
                const int32_t cs = s[m], ce = e[m];
                if (ce < 0) continue;
                for (int32_t j = cs; j <= ce; ++j) {
                    if (!cell[j].sign) continue;
                    const V3<R> pos2 = cell[j].pos;
                    const R ddx = pos1.x - pos2.x, ddy = pos1.y - pos2.y, ddz = pos1.z - pos2.z;
                    const R ss = ddx * ddx + ddy * ddy + ddz * ddz;
                    // FMkernel.cuh:89-90: float sum, + 0.0001 in double, sqrtf of the float.
                    R distance;
                    if (std::is_same<R, float>::value)
                        distance = (R)std::sqrt((float)((double)ss + 0.0001));
                    else
                        distance = (R)std::sqrt((double)ss + 0.0001);
                    if (distance >= thres) continue;
                    const R r2 = cell[j].r;
                    const R rsum = r2 + r1;
                    const R Rs = rsum > R(0.01f) ? rsum : R(0.01f);  // bigger(), BasicFunctions.cuh:11-13
                    thres = (R)((double)Rs * 1.25);              // Ra_ratio is a double literal
                    if (distance >= thres) continue;
                    const R qr = R(1) - distance / Rs;
                    const R qa = R(1) - distance / thres;
                    const R coefr = (distance <= Rs) ? Cccr * (qr * qr / distance) : R(0);
                    const R coefa = (distance <= thres) ? Ccca * (qa * qa / distance) : R(0);
                    const R c = coefr - coefa;
                    fx += ddx * c;
                    fy += ddy * c;
                    fz += ddz * c;
                }
            }
    return {fx, fy, fz};
}

// movementkernel, FMkernel.cuh:130-158 (live cells only).
template <class R>
inline void cell_move(CellT<R>& c, R dt, int32_t N, R dx) {
    const R h = dt / R(2.f);
    const V3<R> f = c.force, of = c.oldforce;
    V3<R> p = c.pos;
    p.x += h * (R(3.f) * f.x - of.x);
    p.y += h * (R(3.f) * f.y - of.y);
    p.z += h * (R(3.f) * f.z - of.z);
    int32_t i, j, k;
    find_mesh(p, dx, i, j, k);
    if (!check_inside(i, j, k, N)) {
        c.sign = 0;  // pos, force and oldforce are left as they are (FMkernel.cuh:148-151)
        return;
    }
    c.pos = p;
    c.oldforce = f;
    c.force = {R(0), R(0), R(0)};
    c.mesh_idx = (int32_t)lin_index(i, j, k, N);
}

// MechanicsMesh_struct::FM_update, FMkernel.cuh:181-202.
template <class R>
void fm_update(std::vector<CellT<R>>& cells, int32_t n, int32_t N, R dx, R dt, int32_t capacity,
               FMState<R>& st) {
    const int64_t V = (int64_t)N * N * N;
    // K1 Mech_Mesh_se_Reset (FMkernel.cuh:38-43)
    st.s.assign(V, capacity + 1);
    st.e.assign(V, -1);
    // K2 updateCellMeshIdx (FMkernel.cuh:22-36)
    std::vector<int32_t> key(n);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) key[i] = cell_key(cells[i], N, dx);
    // L1 sort_by_key (FMkernel.cuh:190): ascending signed keys, stable.
    st.order.resize(n);
    for (int32_t i = 0; i < n; ++i) st.order[i] = i;
    std::stable_sort(st.order.begin(), st.order.end(),
                     [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    std::vector<CellT<R>> sorted(n);
    st.keys.resize(n);
    for (int32_t i = 0; i < n; ++i) {
        sorted[i] = cells[st.order[i]];
        st.keys[i] = key[st.order[i]];
    }
    std::copy(sorted.begin(), sorted.end(), cells.begin());
    // K3 Mech_Mesh_se_Get (FMkernel.cuh:45-53): min/max sorted slot of live cells.
    for (int32_t i = 0; i < n; ++i) {
        if (!cells[i].sign) continue;
        const int32_t m = cells[i].mesh_idx;
        if (m < 0 || m >= V) continue;
        st.s[m] = std::min(st.s[m], i);
        st.e[m] = std::max(st.e[m], i);
    }
    // K4 forcekernel (FMkernel.cuh:56-118): writes force of live cells.
    std::vector<V3<R>> force(n);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < n; ++i)
        if (cells[i].sign) force[i] = cell_force(cells.data(), i, st.s.data(), st.e.data(), N, dx);
    for (int32_t i = 0; i < n; ++i)
        if (cells[i].sign) cells[i].force = force[i];
    // K5 movementkernel (FMkernel.cuh:130-158)
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i)
        if (cells[i].sign) cell_move(cells[i], dt, N, dx);
}

// ---------------------------------------------------------------------------
// Uptake map (LODkernel.cuh:12-22,183-190,270-278; Cells.cuh:74-89)
// ---------------------------------------------------------------------------

// Cell::O2_consume(1.f), Cells.cuh:74-89, with the substrate's uptake and
// dead factor in place of the oxygen constants.
template <class R>
inline R cell_uptake(const CellT<R>& c, R dx, R uptake, R dead_factor) {
    const R vf = c.V / dx / dx / dx;
    R rate = 0;
    if (c.phase >= 0 && c.phase <= 2)
        rate = uptake * vf * R(1.f);
    else if (c.phase >= 3)
        rate = uptake * vf * R(1.f) * dead_factor;
    return rate;
}

// Fixed-point unit of the deterministic per-voxel sum.
// Scaled values are clamped to +-2^62 so that the conversion is defined.
inline int64_t to_fixed(float c) {
    double x = (double)(c * 4294967296.0f);
    if (!(x > -4.611686018427388e18)) x = (x != x) ? 0.0 : -4.611686018427388e18;
    if (x > 4.611686018427388e18) x = 4.611686018427388e18;
    return (int64_t)std::llrint(x);
}
inline float from_fixed(int64_t q) { return (float)q * (1.0f / 4294967296.0f); }

template <class R>
void prelod(const std::vector<CellT<R>>& cells, int32_t n, int32_t N, R dx, R uptake, R dead_factor,
            std::vector<R>& p_rate) {
    const int64_t V = (int64_t)N * N * N;
    if (std::is_same<R, float>::value) {
        std::vector<int64_t> acc(V, 0);
        for (int32_t i = 0; i < n; ++i) {
            if (!cells[i].sign) continue;
            const int32_t m = cells[i].mesh_idx;
            if (m < 0 || m >= V) continue;  // the reference would write out of bounds
            acc[m] += to_fixed((float)cell_uptake(cells[i], dx, uptake, dead_factor));
        }
        p_rate.resize(V);
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < V; ++v) p_rate[v] = (R)from_fixed(acc[v]);
    } else {
        p_rate.assign(V, R(0));
        for (int32_t i = 0; i < n; ++i) {
            if (!cells[i].sign) continue;
            const int32_t m = cells[i].mesh_idx;
            if (m < 0 || m >= V) continue;
            p_rate[m] += cell_uptake(cells[i], dx, uptake, dead_factor);
        }
    }
}

// ---------------------------------------------------------------------------
// Diffusion (LODkernel.cuh)
// ---------------------------------------------------------------------------

// Mesh_struct::Mesh_initialize, LODkernel.cuh:239-259: Thomas pre-factorisation
// of the corner-modified tridiagonal system.
template <class R>
void lod_coeffs(int32_t N, R dx, R dt, R D, R decay, std::vector<R>& E, std::vector<R>& F) {
    const R side = -dt * D / (dx * dx);
    const R corner = R(1.f) + dt * decay / R(3.f) - side;
    const R diag = corner - side;
    E.assign(N, R(0));
    F.assign(N, R(0));
    E[0] = 0;
    F[0] = corner;
    for (int32_t i = 1; i < N - 1; ++i) {
        E[i] = side / F[i - 1];
        F[i] = diag - E[i] * side;
    }
    E[N - 1] = side / F[N - 2];
    F[N - 1] = corner - E[N - 1] * side;
}

// One line of LODkernel_x/_y/_z (LODkernel.cuh:34-50,73-110): elements
// p[base + t*stride], t = 0..N-1.
template <class R>
inline void thomas_line(R* p, int64_t base, int64_t stride, int32_t N, const R* E, const R* F, R bnd,
                        R* Y) {
    const R side = E[1] * F[0];
    p[base + (int64_t)(N - 1) * stride] = bnd;
    Y[0] = bnd;
    for (int32_t t = 1; t < N; ++t) Y[t] = p[base + t * stride] - E[t] * Y[t - 1];
    for (int32_t t = N - 2; t > 0; --t)
        p[base + t * stride] = (Y[t] - side * p[base + (t + 1) * stride]) / F[t];
    p[base] = bnd;
}

// Mesh_struct::LODsolver, LODkernel.cuh:280-300.
template <class R>
void lod_substep(std::vector<R>& p, const std::vector<R>& p_rate, int32_t N, R dt, const R* E, const R* F,
                 R bnd) {
    const int64_t N1 = N, N2 = (int64_t)N * N;
    // Setboundary, LODkernel.cuh:112-136: the two outer layers of every face.
#pragma omp parallel for schedule(static)
    for (int32_t b = 0; b < N; ++b)
        for (int32_t a = 0; a < N; ++a) {
            const int32_t layers[4] = {0, 1, N - 2, N - 1};
            for (int32_t q = 0; q < 4; ++q) {
                const int32_t l = layers[q];
                p[lin_index(l, a, b, N)] = bnd;
                p[lin_index(a, l, b, N)] = bnd;
                p[lin_index(a, b, l, N)] = bnd;
            }
        }
    // meshupdatekernel, LODkernel.cuh:24-31: implicit uptake.
    const int64_t V = N2 * N;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) p[v] = p[v] / (R(1.f) - p_rate[v] * dt);
    // LODkernel_x: lines (j, k) along i.
#pragma omp parallel
    {
        std::vector<R> Y(N);
#pragma omp for schedule(static)
        for (int64_t l = 0; l < N2; ++l) thomas_line(p.data(), l * N1, 1, N, E, F, bnd, Y.data());
        // LODkernel_y: lines (i, k) along j.
#pragma omp for schedule(static)
        for (int64_t l = 0; l < N2; ++l) {
            const int64_t i = l % N1, k = l / N1;
            thomas_line(p.data(), i + k * N2, N1, N, E, F, bnd, Y.data());
        }
        // LODkernel_z: lines (i, j) along k. The reference's back loop also
        // visits k = 0, whose value is then overwritten with the boundary.
#pragma omp for schedule(static)
        for (int64_t l = 0; l < N2; ++l) thomas_line(p.data(), l, N2, N, E, F, bnd, Y.data());
    }
}

// ---------------------------------------------------------------------------
// Proliferation and death (Cells.cuh:91-224, PDkernel.cuh)
// ---------------------------------------------------------------------------

// Cell::Volume_update, Cells.cuh:91-142.
template <class R>
inline void volume_update(CellT<R>& c, R dt_bio) {
    R ff = R(0.7502);
    R rf = R(3.f) / R(60.f);
    R rcs = R(0.33f) / R(60.f);
    R rns = R(0.33f) / R(60.f);
    R Vns_0 = R(135.f);
    R Vf_0;
    R Vcs_0;
    if (c.phase == 1) {
        Vns_0 = Vns_0 * R(2);
        rcs = R(0.27 / 60);
    } else if (c.phase == 3) {
        ff = 0;
        Vns_0 = 0;
        rns = R(0.35 / 60);
        rcs = R(1.f) / R(60);
    } else if (c.phase == 4) {
        ff = 1;
        rf = R(0.67 / 60);
        Vns_0 = 0;
        rns = R(0.013 / 60);
        rcs = R(0.0032 / 60);
    } else if (c.phase == 5) {
        ff = 0;
        rf = R(0.05 / 60);
        Vns_0 = 0;
        rns = R(0.013 / 60);
        rcs = R(0.0032 / 60);
    }
    Vf_0 = c.V * ff;
    Vcs_0 = Vns_0 * R(3.615f);
    c.Vf = c.Vf * (R(1.f) - dt_bio * rf) + rf * dt_bio * Vf_0;
    c.Vns = c.Vns * (R(1.f) - dt_bio * rns) + rns * dt_bio * Vns_0;
    c.Vcs = c.Vcs * (R(1.f) - dt_bio * rcs) + rcs * dt_bio * Vcs_0;
    c.V = c.Vf + c.Vns + c.Vcs;
    c.r = (R)std::pow(0.23873241 * (double)c.V, 1.0 / 3);
}

// Cell::Phase_update_ki67adv, Cells.cuh:145-224. Probabilities follow the
// reference's float/double promotion chain (Definitions.h:38-40 are double
// expressions). Returns `proliferate`.
template <class R>
inline bool phase_update(CellT<R>& c, float rnd, R O2, R dt_bio) {
    using D = typename std::conditional<std::is_same<R, float>::value, double, R>::type;
    const D apo_rate = (D)1.f / 7.0 / 24.0 / 60.0;
    const D pro_rate = (D)1.f / 8.5 / 60.0;
    const D nec_rate = (D)1.f / 6.0 / 60.0;
    c.cell_clock = c.cell_clock + dt_bio;
    bool proliferate = false;
    const R apoptosis_prob = (R)(apo_rate * (D)dt_bio);
    R proliferation_prob =
        (R)((D)(R(1.f) * (O2 - R(5.f)) / (R(10.f) - R(5.f))) * pro_rate * (D)dt_bio);
    if (O2 > R(10.f))
        proliferation_prob = (R)(pro_rate * (D)dt_bio);
    else if (O2 < R(5.f))
        proliferation_prob = 0;
    R necrosis_prob = (R)((D)(R(1.f) * (R(5.f) - O2) / (R(5.f) - R(2.5f))) * nec_rate * (D)dt_bio);
    if (O2 > R(5.f))
        necrosis_prob = 0;
    else if (O2 < R(2.5f))
        necrosis_prob = (R)(nec_rate * (D)dt_bio);

    const R r = (R)rnd;
    if (c.phase >= 3) {
        if (c.phase == 4 && c.V >= R(2494.f) * R(2)) {
            c.phase = 5;
            c.cell_clock = 0;
        }
        if (c.Vcs <= R(0.05f) * R(488)) {
            c.sign = 0;
            c.cell_clock = 0;
        }
    } else {
        if (c.phase == 1 && c.cell_clock >= R(13.f) * R(60)) {
            c.phase = 2;
            proliferate = true;
            c.cell_clock = 0;
        }
        if (c.phase == 2 && c.cell_clock >= R(2.5f) * R(60)) {
            c.phase = 0;
            c.cell_clock = 0;
        }
        if (c.phase == 0) {
            if (r <= proliferation_prob) {
                c.phase = 1;
                c.cell_clock = 0;
            } else if (r <= proliferation_prob + apoptosis_prob) {
                c.phase = 3;
                c.cell_clock = 0;
            } else if (r <= proliferation_prob + apoptosis_prob + necrosis_prob) {
                c.phase = 4;
                c.cell_clock = 0;
            }
        }
    }
    return proliferate;
}

// randomvect, RandomFunctions.cuh:23-32 (u for theta, v for phi).
template <class R>
inline V3<R> random_dir(float u, float v) {
    if (std::is_same<R, float>::value) {
        const float theta = 2.f * u * 3.1415926f;
        const float phi = std::acos(2.f * v - 1.f);
        return {(R)(std::sin(theta) * std::sin(phi)), (R)(std::cos(theta) * std::sin(phi)), (R)std::cos(phi)};
    } else {
        const double theta = 2.0 * (double)u * 3.1415926;
        const double phi = std::acos(2.0 * (double)v - 1.0);
        return {(R)(std::sin(theta) * std::sin(phi)), (R)(std::cos(theta) * std::sin(phi)), (R)std::cos(phi)};
    }
}

// Birthkernel + CellBirth_kernel, PDkernel.cuh:27-98. `p0` is substrate 0
// (oxygen). slots[i] = daughter slot of mother i, or -1. Returns the new count.
template <class R>
int32_t birth(std::vector<CellT<R>>& cells, int32_t n, const R* p0, int32_t N, R dx, R dt_bio,
              int32_t capacity, uint64_t seed, uint64_t step, std::vector<int32_t>& slots) {
    slots.assign(std::max<int32_t>(n, 0), -1);
    if (n >= capacity) return n;  // PDkernel.cuh:31, evaluated once at launch
    std::vector<uint8_t> divide(n, 0);
    std::vector<float> uu(3 * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        CellT<R>& c = cells[i];
        if (!c.sign) continue;
        float u[3];
        birth_uniforms(seed, (uint32_t)i, step, u);
        uu[3 * i + 1] = u[1];
        uu[3 * i + 2] = u[2];
        int32_t a, b, k;
        find_mesh(c.pos, dx, a, b, k);
        // PDkernel.cuh:37-39 reads p at Find_Mesh(pos) without a bounds check;
        // live cells are inside, and the index is clamped for the rest.
        a = std::min(std::max(a, 0), N - 1);
        b = std::min(std::max(b, 0), N - 1);
        k = std::min(std::max(k, 0), N - 1);
        const R O2 = p0[lin_index(a, b, k, N)];
        volume_update(c, dt_bio);
        divide[i] = phase_update(c, u[0], O2, dt_bio) ? 1 : 0;
    }
    // Deterministic slot assignment in mother order; PDkernel.cuh:51-54 admits
    // a division while the running count is below capacity - 1.
    int32_t count = n;
    for (int32_t i = 0; i < n; ++i) {
        if (!divide[i]) continue;
        if (count < capacity - 1) slots[i] = count++;
    }
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        const int32_t idx = slots[i];
        if (idx < 0) continue;
        CellT<R>& m = cells[i];
        const V3<R> dir = random_dir<R>(uu[3 * i + 1], uu[3 * i + 2]);
        const V3<R> posm0 = m.pos;
        R rm = m.r;
        const R distance = R(0.2063f) * rm;
        rm *= R(0.7937f);
        const V3<R> posd = {posm0.x + dir.x * distance, posm0.y + dir.y * distance, posm0.z + dir.z * distance};
        const V3<R> posm = {posm0.x - dir.x * distance, posm0.y - dir.y * distance, posm0.z - dir.z * distance};
        m.r = rm;
        m.pos = posm;
        m.phase = 2;
        m.cell_clock = 0;
        m.V = m.V / R(2);
        m.Vf = m.Vf / R(2);
        m.Vcs = m.Vcs / R(2);
        m.Vns = m.Vns / R(2);
        CellT<R> d;  // Cell's default member values, Cells.cuh:10-33
        std::memset(&d, 0, sizeof(d));
        d.sign = 1;
        d.mesh_idx = -1;
        d.pos = posd;
        d.force = {0, 0, 0};
        d.oldforce = {0, 0, 0};
        d.phase = 2;
        d.cell_clock = 0;
        d.r = rm;
        d.V = m.V;
        d.Vf = m.Vf;
        d.Vcs = m.Vcs;
        d.Vns = m.Vns;
        cells[idx] = d;
    }
    return count;
}

// CellDeath_kernel, PDkernel.cuh:100-117: stable remove_if(sign == false).
template <class R>
int32_t death(std::vector<CellT<R>>& cells, int32_t n) {
    int32_t w = 0;
    for (int32_t i = 0; i < n; ++i)
        if (cells[i].sign) cells[w++] = cells[i];
    return w;
}

// Trip count of `for (float t = 0; t < end; t += step)` (Main.cu:108,114,121).
inline int32_t float_loop_count(float end, float step) {
    int32_t c = 0;
    for (float t = 0.f; t < end; t += step) ++c;
    return c;
}

}  // namespace gell_oracle
