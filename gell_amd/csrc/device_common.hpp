// device_common.hpp -- device-side building blocks shared by the gfx950
// kernels: the 72-byte cell record, voxel index math, Morton keys, the
// Philox4x32-10 generator and the per-cell biology of the Gell model.
//
// All arithmetic keeps the reference's operation order and its float/double
// promotions so that, compiled with -ffp-contract=off (no fused multiply-add)
// and IEEE division/sqrt (hipcc's default for fp32), each value is the one
// the reference's own expression gives.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gell_abi.h"

namespace gell {

static_assert(sizeof(gell_cell) == 72, "gell_cell must be the reference's 72-byte Cell");

constexpr int kWave = 64;

// The dead marker in the compact neighbour array (x, y, z, r): a radius slot
// holding this bit pattern means sign == false.
constexpr uint32_t kDeadTag = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Record access: the record is 18 dwords, 8-byte aligned.
// ---------------------------------------------------------------------------
struct CellRegs {
    uint32_t w[18];
};

__device__ __forceinline__ CellRegs load_cell(const gell_cell* base, int64_t i) {
    const uint2* src = reinterpret_cast<const uint2*>(base + i);
    CellRegs c;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        uint2 v = src[q];
        c.w[2 * q] = v.x;
        c.w[2 * q + 1] = v.y;
    }
    return c;
}

__device__ __forceinline__ void store_cell(gell_cell* base, int64_t i, const CellRegs& c) {
    uint2* dst = reinterpret_cast<uint2*>(base + i);
#pragma unroll
    for (int q = 0; q < 9; ++q) dst[q] = make_uint2(c.w[2 * q], c.w[2 * q + 1]);
}

// Dword offsets of the record fields (include/Cells.cuh:9-47).
enum CellWord : int {
    W_SIGN = 0, W_MESH = 1, W_PX = 2, W_PY = 3, W_PZ = 4, W_FX = 5, W_FY = 6, W_FZ = 7,
    W_OX = 8, W_OY = 9, W_OZ = 10, W_R = 11, W_V = 12, W_VF = 13, W_VCS = 14, W_VNS = 15,
    W_PHASE = 16, W_CLOCK = 17
};

__device__ __forceinline__ float cf(const CellRegs& c, int w) { return __uint_as_float(c.w[w]); }
__device__ __forceinline__ void cset(CellRegs& c, int w, float v) { c.w[w] = __float_as_uint(v); }
__device__ __forceinline__ bool csign(const CellRegs& c) { return (c.w[W_SIGN] & 0xFFu) != 0; }
__device__ __forceinline__ void cset_sign(CellRegs& c, bool s) { c.w[W_SIGN] = (c.w[W_SIGN] & ~0xFFu) | (s ? 1u : 0u); }

// ---------------------------------------------------------------------------
// Voxel index math (BasicFunctions.cuh:18-66)
// ---------------------------------------------------------------------------

// (int)(q) with out-of-range values and NaN sent to INT32_MIN (outside).
__device__ __forceinline__ int32_t trunc_index(float q) {
    if (!(q > -2147483648.0f && q < 2147483647.0f)) return INT32_MIN;
    return (int32_t)q;
}

__device__ __forceinline__ bool inside(int32_t i, int32_t j, int32_t k, int32_t N) {
    return (uint32_t)i < (uint32_t)N && (uint32_t)j < (uint32_t)N && (uint32_t)k < (uint32_t)N;
}

__device__ __forceinline__ int64_t lin(int32_t i, int32_t j, int32_t k, int32_t N) {
    return (int64_t)i + (int64_t)N * ((int64_t)j + (int64_t)N * (int64_t)k);
}

// Spread the low 10 bits of v to every third bit (bit b -> bit 3b).
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Inverse of spread3: every third bit of v (bit 3b -> bit b), 10 bits.
__device__ __forceinline__ uint32_t compact3(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0x030000FFu;
    v = (v | (v >> 16)) & 0x3FFu;
    return v;
}

// Morton key, x in bit 0 (BasicFunctions.cuh:129-137).
__device__ __forceinline__ uint32_t morton3(int32_t i, int32_t j, int32_t k) {
    return spread3((uint32_t)i) | (spread3((uint32_t)j) << 1) | (spread3((uint32_t)k) << 2);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): key = seed, counter = (slot, step, tag).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Uniform in (0, 1] from the top 24 bits (curand_uniform's range).
__device__ __forceinline__ float u01(uint32_t w) { return (float)((w >> 8) + 1u) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------------------
// Cell biology (Cells.cuh:74-224, RandomFunctions.cuh:23-32)
// ---------------------------------------------------------------------------

// Cell::O2_consume(1.f) with a substrate's uptake and dead factor (Cells.cuh:74-89).
__device__ __forceinline__ float cell_uptake(float V, int32_t phase, float dx, float uptake, float dead_factor) {
    const float vf = V / dx / dx / dx;
    float rate = 0.f;
    if (phase >= 0 && phase <= 2)
        rate = uptake * vf * 1.f;
    else if (phase >= 3)
        rate = uptake * vf * 1.f * dead_factor;
    return rate;
}

// Exact fixed-point (2^-32 units) form of an uptake rate for the
// order-independent per-voxel sum; clamped to +-2^62.
__device__ __forceinline__ long long to_fixed(float c) {
    double x = (double)(c * 4294967296.0f);
    if (!(x > -4.611686018427388e18)) x = (x != x) ? 0.0 : -4.611686018427388e18;
    if (x > 4.611686018427388e18) x = 4.611686018427388e18;
    return __double2ll_rn(x);
}
__device__ __forceinline__ float from_fixed(long long q) { return __ll2float_rn(q) * (1.0f / 4294967296.0f); }

// Cell counts can live on the device (graph-captured biology steps, where no
// count comes back to the host between birth and death): a kernel given a
// non-null count pointer takes its count from there instead of from its
// launch argument, and its grid then covers the capacity.
__device__ __forceinline__ int32_t count_of(int32_t n, const int32_t* np) { return np ? *np : n; }

// Linear voxel index (x fastest) of a Morton code.
__device__ __forceinline__ int64_t lin_of_code(uint32_t code, int32_t N) {
    return lin((int32_t)compact3(code), (int32_t)compact3(code >> 1), (int32_t)compact3(code >> 2), N);
}

// The uptake accumulators of up to GELL_MAX_SUBSTRATES substrates: int64
// sums in 2^-32 units, indexed by Morton code (so that cells in sorted order
// add to neighbouring words).
struct UptakeSet {
    int32_t S;
    long long* acc[GELL_MAX_SUBSTRATES];
    float uptake[GELL_MAX_SUBSTRATES];
    float dead_factor[GELL_MAX_SUBSTRATES];
};

// oxy_consumption_update (LODkernel.cuh:12-22) for one cell per lane; every
// lane of the wave calls it. A live cell whose mesh_idx v lies in [0, N^3)
// adds Cell::O2_consume() of every substrate at v, and `out_code` is set to
// v's Morton code (-1 for a cell that adds nothing). Lanes with the same
// code next to each other (cells in sorted order) are summed in the wave
// first, so a run issues one atomic per substrate. Integer sums: the result
// does not depend on the order or grouping of the additions.
__device__ __forceinline__ void uptake_accumulate_wave(bool take, int32_t v, float V, int32_t phase, int32_t N,
                                                       float dx, const UptakeSet& u, int32_t& out_code) {
    int32_t code = -1;
    if (take) {
        const int32_t i = v % N, j = (v / N) % N, k = v / (N * N);
        code = (int32_t)morton3(i, j, k);
    }
    out_code = code;
    const int lane = (int)(threadIdx.x & 63);
    const int32_t prev = __shfl_up(code, 1, 64), next = __shfl_down(code, 1, 64);
    const bool tail = lane == 63 || next != code;
    const uint64_t heads = __ballot(lane == 0 || prev != code);
    // first lane of this lane's run
    const int first = 63 - (int)__builtin_clzll(heads & ((2ull << lane) - 1ull));
    for (int s = 0; s < u.S; ++s) {
        long long q = take ? to_fixed(cell_uptake(V, phase, dx, u.uptake[s], u.dead_factor[s])) : 0;
        // inclusive scan within runs (after the step with offset d a lane holds
        // the sum over the last 2d lanes of its run), only as many steps as the
        // wave's longest run needs: at ~1 cell per voxel, none or one
        for (int d = 1; __any(lane - d >= first); d <<= 1) {
            const long long y = __shfl_up(q, d, 64);
            if (lane - d >= first) q += y;
        }
        if (tail && code >= 0)
            atomicAdd(reinterpret_cast<unsigned long long*>(u.acc[s] + code), (unsigned long long)q);
    }
}

struct Bio {
    float V, Vf, Vcs, Vns, r, clock;
    int32_t phase;
    bool sign;
};

// Cell::Volume_update, Cells.cuh:91-142.
__device__ __forceinline__ void volume_update(Bio& c, float dt_bio) {
    float ff = (float)0.7502;
    float rf = 3.f / 60.f;
    float rcs = 0.33f / 60.f;
    float rns = 0.33f / 60.f;
    float Vns_0 = 135.f;
    if (c.phase == 1) {
        Vns_0 = Vns_0 * 2.f;
        rcs = (float)(0.27 / 60);
    } else if (c.phase == 3) {
        ff = 0.f;
        Vns_0 = 0.f;
        rns = (float)(0.35 / 60);
        rcs = 1.f / 60.f;
    } else if (c.phase == 4) {
        ff = 1.f;
        rf = (float)(0.67 / 60);
        Vns_0 = 0.f;
        rns = (float)(0.013 / 60);
        rcs = (float)(0.0032 / 60);
    } else if (c.phase == 5) {
        ff = 0.f;
        rf = (float)(0.05 / 60);
        Vns_0 = 0.f;
        rns = (float)(0.013 / 60);
        rcs = (float)(0.0032 / 60);
    }
    const float Vf_0 = c.V * ff;
    const float Vcs_0 = Vns_0 * 3.615f;
    c.Vf = c.Vf * (1.f - dt_bio * rf) + rf * dt_bio * Vf_0;
    c.Vns = c.Vns * (1.f - dt_bio * rns) + rns * dt_bio * Vns_0;
    c.Vcs = c.Vcs * (1.f - dt_bio * rcs) + rcs * dt_bio * Vcs_0;
    c.V = c.Vf + c.Vns + c.Vcs;
    c.r = (float)pow(0.23873241 * (double)c.V, 1.0 / 3);
}

// Cell::Phase_update_ki67adv, Cells.cuh:145-224; returns `proliferate`.
__device__ __forceinline__ bool phase_update(Bio& c, float rnd, float O2, float dt_bio) {
    const double apo_rate = 1.0 / 7.0 / 24.0 / 60.0;
    const double pro_rate = 1.0 / 8.5 / 60.0;
    const double nec_rate = 1.0 / 6.0 / 60.0;
    c.clock = c.clock + dt_bio;
    bool proliferate = false;
    const float apoptosis_prob = (float)(apo_rate * (double)dt_bio);
    float proliferation_prob = (float)((double)(1.f * (O2 - 5.f) / (10.f - 5.f)) * pro_rate * (double)dt_bio);
    if (O2 > 10.f)
        proliferation_prob = (float)(pro_rate * (double)dt_bio);
    else if (O2 < 5.f)
        proliferation_prob = 0.f;
    float necrosis_prob = (float)((double)(1.f * (5.f - O2) / (5.f - 2.5f)) * nec_rate * (double)dt_bio);
    if (O2 > 5.f)
        necrosis_prob = 0.f;
    else if (O2 < 2.5f)
        necrosis_prob = (float)(nec_rate * (double)dt_bio);

    if (c.phase >= 3) {
        if (c.phase == 4 && c.V >= 2494.f * 2.f) {
            c.phase = 5;
            c.clock = 0.f;
        }
        if (c.Vcs <= 0.05f * 488.f) {
            c.sign = false;
            c.clock = 0.f;
        }
    } else {
        if (c.phase == 1 && c.clock >= 13.f * 60.f) {
            c.phase = 2;
            proliferate = true;
            c.clock = 0.f;
        }
        if (c.phase == 2 && c.clock >= 2.5f * 60.f) {
            c.phase = 0;
            c.clock = 0.f;
        }
        if (c.phase == 0) {
            if (rnd <= proliferation_prob) {
                c.phase = 1;
                c.clock = 0.f;
            } else if (rnd <= proliferation_prob + apoptosis_prob) {
                c.phase = 3;
                c.clock = 0.f;
            } else if (rnd <= proliferation_prob + apoptosis_prob + necrosis_prob) {
                c.phase = 4;
                c.clock = 0.f;
            }
        }
    }
    return proliferate;
}

// ---------------------------------------------------------------------------
// Wave helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

}  // namespace gell
