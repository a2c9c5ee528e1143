// lod.hip -- one locally-one-dimensional diffusion substep on gfx950
// (Mesh_struct::LODsolver, LODkernel.cuh:280-300).
//
// Layout: one FP32 field of N^3 voxels per substrate, x fastest
// (Find_Index, BasicFunctions.cuh:31-36). Every line is solved by one lane
// with the serial Thomas recurrence of LODkernel_x/_y/_z, so each value is the
// reference's own:
//   Y[0] = b; Y[t] = p[t] - E[t] Y[t-1]  (p[N-1] = b first)
//   p[t] = (Y[t] - s p[t+1]) / F[t], t = N-2..1;  p[0] = p[N-1] = b
//
// Kernels (one 64-lane wave per workgroup, 64 lines per wave):
//   lod_x_reg<N>   Setboundary + implicit uptake (LODkernel.cuh:24-31,112-136)
//                  fused into the x sweep. x lines are contiguous in memory,
//                  so the wave loads 64 lines x 32 voxels as coalesced 128-byte
//                  rows into LDS and each lane reads its own line out of the
//                  transposed tile into registers; the store goes back the same
//                  way.
//   lod_yz_reg<N>  y and z sweeps: neighbouring lanes own neighbouring lines,
//                  so each step of the recurrence is one coalesced 256-byte
//                  wave access; the whole line stays in registers.
//   lod_sweep_any  any N: one lane per line, Y kept in place in the field.
// Bytes per voxel-update (algorithmic): x 12 (p, p_rate in; p out), y 8, z 8.
#include "device_common.hpp"
#include "internal.hpp"

namespace gell {
namespace {

__device__ __forceinline__ bool on_boundary(int i, int j, int k, int N) {
    return i < 2 || i >= N - 2 || j < 2 || j >= N - 2 || k < 2 || k >= N - 2;
}

// Serial Thomas recurrence on a register-resident line.
template <int N>
__device__ __forceinline__ void thomas_regs(float (&v)[N], const float* __restrict__ E, const float* __restrict__ F,
                                            float side, float bnd) {
    v[N - 1] = bnd;
    v[0] = bnd;
#pragma unroll
    for (int t = 1; t < N; ++t) v[t] = v[t] - E[t] * v[t - 1];
    v[N - 1] = bnd;
#pragma unroll
    for (int t = N - 2; t > 0; --t) v[t] = (v[t] - side * v[t + 1]) / F[t];
    v[0] = bnd;
}

constexpr int kChunk = 32;           // voxels per line per transpose step
constexpr int kTileLd = kChunk + 1;  // padded LDS row

// Buffer descriptor over a whole field (gfx9 layout: raw 32-bit data).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const float* p, int64_t count) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, (int)(count * 4), 0x00020000);
}

template <int N>
__global__ __launch_bounds__(64) void lod_x_reg(float* __restrict__ p, const float* __restrict__ rate,
                                                const float* __restrict__ E, const float* __restrict__ F, float side,
                                                float bnd, float dt) {
    static_assert(N % kChunk == 0, "x sweep needs N divisible by the chunk");
    constexpr int NC = N / kChunk;
    constexpr int Q = kChunk / 4;  // float4 per line per chunk
    constexpr int LPI = 64 / Q;    // lines covered by one wave-wide float4 access
    constexpr int64_t NN = (int64_t)N * N;
    __shared__ float tile[64 * kTileLd];
    const int lane = threadIdx.x;
    const int64_t L0 = (int64_t)blockIdx.x * 64;
    const __amdgpu_buffer_rsrc_t rp = field_rsrc(p, NN * N);
    const __amdgpu_buffer_rsrc_t rr = field_rsrc(rate, NN * N);
    // Lane's share of each access: line (lane / Q), float4 (lane % Q); the
    // rest of the address (wave base, access index, chunk) is uniform.
    const int lq = lane / Q, part = lane % Q;
    const int voff = (lq * N + part * 4) * 4;
    const int sbase = (int)(L0 * N * 4);
    float v[N];

#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int it = 0; it < Q; ++it) {
            const int l = it * LPI + lq;
            const int64_t L = L0 + l;
            const int j = (int)(L % N), k = (int)(L / N);
            const int i0 = c * kChunk + part * 4;
            const int so = sbase + (it * LPI * N + c * kChunk) * 4;
            const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rp, voff, so, 0));
            const float4 r = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rr, voff, so, 0));
            const float av[4] = {a.x, a.y, a.z, a.w};
            const float rv[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float x = av[e];
                if (on_boundary(i0 + e, j, k, N)) x = bnd;
                x = x / (1.f - rv[e] * dt);
                tile[l * kTileLd + part * 4 + e] = x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kChunk; ++t) v[c * kChunk + t] = tile[lane * kTileLd + t];
        __syncthreads();
    }

    thomas_regs<N>(v, E, F, side, bnd);

#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int t = 0; t < kChunk; ++t) tile[lane * kTileLd + t] = v[c * kChunk + t];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < Q; ++it) {
            const int l = it * LPI + lq;
            const int so = sbase + (it * LPI * N + c * kChunk) * 4;
            const float* src = &tile[l * kTileLd + part * 4];
            const float4 o = make_float4(src[0], src[1], src[2], src[3]);
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned int, o),
                                                   rp, voff, so, 0);
        }
        __syncthreads();
    }
}

// DIR 1: y lines (i, k), DIR 2: z lines (i, j); lane = flat line index.
// Raw buffer accesses keep one per-lane byte offset for the whole line and
// put the step t * stride in the scalar offset, so no per-element addresses
// are held in registers. With SPILL > 0 the forward-sweep values Y[1..SPILL)
// live in LDS ([t][lane], conflict-free) instead of registers, which lets two
// waves share a SIMD; the arithmetic is the same either way.
template <int N, int DIR, int SPILL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(SPILL > 0 ? 2 : 1))) void lod_yz_reg(float* __restrict__ p, const float* __restrict__ E,
                                                 const float* __restrict__ F, float side, float bnd) {
    constexpr int64_t NN = (int64_t)N * N;
    constexpr int SP = SPILL > 0 ? SPILL : 1;
    __shared__ float ys[SP][64];
    const int lane = threadIdx.x;
    const int64_t L = (int64_t)blockIdx.x * 64 + lane;
    const bool active = L < NN;
    const int64_t Lc = active ? L : 0;
    const int64_t base = (DIR == 1) ? (Lc % N) + NN * (Lc / N) : Lc;
    constexpr int stride_b = (int)((DIR == 1) ? N * 4 : NN * 4);
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(p, NN * N);
    const int voff = (int)(base * 4);
    float v[N];
#pragma unroll
    for (int t = 1; t < N - 1; ++t)
        v[t] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, t * stride_b, 0));
    // Forward sweep (p[N-1] = bnd is not needed: Y[N-1] is never used).
    float y = bnd;
#pragma unroll
    for (int t = 1; t < N - 1; ++t) {
        y = v[t] - E[t] * y;
        if (t < SPILL)
            ys[t][lane] = y;
        else
            v[t] = y;
    }
    if (!active) return;
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, bnd), rs, voff, (N - 1) * stride_b, 0);
    float nxt = bnd;
#pragma unroll
    for (int t = N - 2; t > 0; --t) {
        const float yt = (t < SPILL) ? ys[t][lane] : v[t];
        nxt = (yt - side * nxt) / F[t];
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, nxt), rs, voff, t * stride_b, 0);
    }
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, bnd), rs, voff, 0, 0);
}

// Any N, any direction (dir 0 = x with fused boundary + uptake).
__global__ __launch_bounds__(256) void lod_sweep_any(float* __restrict__ p, const float* __restrict__ rate,
                                                     const float* __restrict__ E, const float* __restrict__ F,
                                                     float side, float bnd, float dt, int N, int dir) {
    const int64_t NN = (int64_t)N * N;
    const int64_t L = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= NN) return;
    int64_t base, stride;
    const int a = (int)(L % N), b = (int)(L / N);
    if (dir == 0) {
        base = L * N;
        stride = 1;
    } else if (dir == 1) {
        base = a + NN * b;
        stride = N;
    } else {
        base = L;
        stride = NN;
    }
    // Forward sweep, Y[t] stored in place of p[t] (t = 1..N-2).
    float y = bnd;
    for (int t = 1; t < N; ++t) {
        float x;
        if (t == N - 1) {
            x = bnd;
        } else {
            const int64_t o = base + t * stride;
            x = p[o];
            if (dir == 0) {
                if (on_boundary(t, a, b, N)) x = bnd;
                x = x / (1.f - rate[o] * dt);
            }
        }
        y = x - E[t] * y;
        if (t < N - 1) p[base + t * stride] = y;
    }
    float nxt = bnd;
    p[base + (int64_t)(N - 1) * stride] = bnd;
    for (int t = N - 2; t > 0; --t) {
        const int64_t o = base + t * stride;
        nxt = (p[o] - side * nxt) / F[t];
        p[o] = nxt;
    }
    p[base] = bnd;
}

template <int N>
hipError_t launch_reg(float* p, const float* rate, const float* E, const float* F, float side, float bnd, float dt,
                      int spill, hipStream_t st) {
    const int64_t lines = (int64_t)N * N;
    const unsigned blocks = (unsigned)((lines + 63) / 64);
    hipLaunchKernelGGL((lod_x_reg<N>), dim3(blocks), dim3(64), 0, st, p, rate, E, F, side, bnd, dt);
    if (spill > 0 && N > 64) {
        constexpr int S = N > 64 ? 64 : 1;
        hipLaunchKernelGGL((lod_yz_reg<N, 1, S>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
        hipLaunchKernelGGL((lod_yz_reg<N, 2, S>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
    } else {
        hipLaunchKernelGGL((lod_yz_reg<N, 1, 0>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
        hipLaunchKernelGGL((lod_yz_reg<N, 2, 0>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t lod_substep_launch(gell_ctx* ctx, int s) {
    const int N = ctx->N;
    float* p = ctx->field[s];
    const float* rate = ctx->rate[s];
    const float* E = ctx->E[s];
    const float* F = ctx->F[s];
    const float side = ctx->side[s], bnd = ctx->P.sub[s].boundary, dt = ctx->P.dt_diff;
    hipStream_t st = ctx->stream;
    const int spill = ctx->tune[GELL_KNOB_LOD_YZ_LDS];
    switch (N) {
        case 32: return launch_reg<32>(p, rate, E, F, side, bnd, dt, spill, st);
        case 64: return launch_reg<64>(p, rate, E, F, side, bnd, dt, spill, st);
        case 96: return launch_reg<96>(p, rate, E, F, side, bnd, dt, spill, st);
        case 128: return launch_reg<128>(p, rate, E, F, side, bnd, dt, spill, st);
        case 192: return launch_reg<192>(p, rate, E, F, side, bnd, dt, spill, st);
        case 256: return launch_reg<256>(p, rate, E, F, side, bnd, dt, spill, st);
        default: break;
    }
    const int64_t lines = (int64_t)N * N;
    const unsigned blocks = (unsigned)((lines + 255) / 256);
    for (int dir = 0; dir < 3; ++dir)
        hipLaunchKernelGGL(lod_sweep_any, dim3(blocks), dim3(256), 0, st, p, rate, E, F, side, bnd, dt, N, dir);
    return hipGetLastError();
}

// Mesh_struct::Mesh_initialize, LODkernel.cuh:239-259: the Thomas
// pre-factorisation, computed on the host in float and copied once.
hipError_t lod_init_coeffs(gell_ctx* ctx, int s) {
    const int N = ctx->N;
    const gell_substrate& sb = ctx->P.sub[s];
    const float dt = ctx->P.dt_diff, dx = ctx->P.dx;
    std::vector<float> E(N, 0.f), F(N, 0.f);
    volatile float side = -dt * sb.D / (dx * dx);
    volatile float corner = 1.f + dt * sb.decay / 3.f - side;
    volatile float diag = corner - side;
    E[0] = 0.f;
    F[0] = corner;
    for (int i = 1; i < N - 1; ++i) {
        E[i] = side / F[i - 1];
        volatile float prod = E[i] * side;
        F[i] = diag - prod;
    }
    E[N - 1] = side / F[N - 2];
    volatile float prod = E[N - 1] * side;
    F[N - 1] = corner - prod;
    volatile float ts = E[1] * F[0];  // thomas_side, LODkernel.cuh:39
    ctx->side[s] = ts;
    hipError_t e = hipMemcpyAsync(ctx->E[s], E.data(), N * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(ctx->F[s], F.data(), N * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ctx->stream);
}

}  // namespace gell
