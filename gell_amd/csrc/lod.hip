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
//   lod_x_flat<N>  the x sweep for N = 100 or 200 (not multiples of the
//                  32-voxel chunk): 64 lines staged whole through LDS.
//   lod_xy_plane<N> x and y sweeps of one z plane in one workgroup (N = 64,
//                  96, 100); lod_xy_plane_part<N, P> the same with
//                  partitioned line solves.
//   lod_sweep_any  any other N: one lane per line, Y kept in place in the
//                  field.
// Bytes per voxel-update (algorithmic): x 12 (p, p_rate in; p out), y 8, z 8.
#include <cmath>

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
            const float* src = &tile[l * kTileLd + part * 4];
            // 16-byte stores go through global (FLAT) stores: the compiler's
            // hazard model covers a VALU overwrite of their data registers.
            *reinterpret_cast<float4*>(p + (L0 + l) * N + c * kChunk + part * 4) =
                make_float4(src[0], src[1], src[2], src[3]);
        }
        __syncthreads();
    }
}

// x sweep for an N that is not a multiple of the chunk (the reference's own
// 100, say; N % 4 == 0). The 64 lines of a wave are 64 N contiguous floats,
// read and written as coalesced 16-byte accesses (lane-consecutive float4s)
// through an LDS tile of odd row pitch, so that each lane's reads of its own
// row are bank-conflict free. The loads are issued in unrolled batches of kB
// float4 per array, so a wave has them all in flight at once: with one wave
// per CU at these grid sizes, latency is the cost. Same arithmetic as
// lod_x_reg.
template <int N>
__global__ __launch_bounds__(64) void lod_x_flat(float* __restrict__ p, const float* __restrict__ rate,
                                                 const float* __restrict__ E, const float* __restrict__ F,
                                                 float side, float bnd, float dt) {
    static_assert(N % 4 == 0 && N * N < (1 << 30), "line span must be whole float4s");
    constexpr int NN = N * N;
    constexpr int Q = N / 4;  // float4 per line = float4 per lane per array
    constexpr int kB = 13;
    constexpr int LD = N | 1;
    __shared__ float tile[64 * LD];
    const int lane = threadIdx.x;
    const int L0 = blockIdx.x * 64;
    const int nl = NN - L0 < 64 ? NN - L0 : 64;
    const int n4 = nl * Q;
    const float4* __restrict__ pb = reinterpret_cast<const float4*>(p + (int64_t)L0 * N);
    const float4* __restrict__ rb = reinterpret_cast<const float4*>(rate + (int64_t)L0 * N);
#pragma unroll
    for (int b0 = 0; b0 < Q; b0 += kB) {
        float4 a[kB], r[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int q = (b0 + u) * 64 + lane;
            if (b0 + u < Q && q < n4) {
                a[u] = pb[q];
                r[u] = rb[q];
            }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int q = (b0 + u) * 64 + lane;
            if (b0 + u < Q && q < n4) {
                const int l = q / Q, t0 = 4 * (q - l * Q);
                const int L = L0 + l, j = L % N, k = L / N;
                const float av[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
                const float rv[4] = {r[u].x, r[u].y, r[u].z, r[u].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float x = av[e];
                    if (on_boundary(t0 + e, j, k, N)) x = bnd;
                    tile[l * LD + t0 + e] = x / (1.f - rv[e] * dt);
                }
            }
        }
    }
    __syncthreads();
    float v[N];
#pragma unroll
    for (int t = 0; t < N; ++t) v[t] = tile[lane * LD + t];
    thomas_regs<N>(v, E, F, side, bnd);  // rows of lanes >= nl hold no line and are never stored
#pragma unroll
    for (int t = 0; t < N; ++t) tile[lane * LD + t] = v[t];
    __syncthreads();
    float4* __restrict__ po = reinterpret_cast<float4*>(p + (int64_t)L0 * N);
#pragma unroll
    for (int u = 0; u < Q; ++u) {
        const int q = u * 64 + lane;
        if (q < n4) {
            const int l = q / Q, t0 = 4 * (q - l * Q);
            const float* src = &tile[l * LD + t0];
            po[q] = make_float4(src[0], src[1], src[2], src[3]);
        }
    }
}

// x and y sweeps of one z plane in one workgroup, for N whose plane fits in
// LDS (N <= 128; used at N = 100): the plane is read once with coalesced
// float4 accesses (boundary and implicit uptake applied on the way in, as in
// lod_x_reg), lane j solves x line j out of the odd-pitch tile, then lane i
// solves y line i (lane-consecutive addresses, conflict free), and the plane
// is written back once. The y sweep keeps lod_yz_reg's operation order, so
// the result equals the separate x and y kernels bit for bit, with half the
// launches and one pass over the field instead of two.
template <int N>
__global__ __launch_bounds__(128) void lod_xy_plane(float* __restrict__ p, const float* __restrict__ rate,
                                                    const float* __restrict__ E, const float* __restrict__ F,
                                                    float side, float bnd, float dt) {
    static_assert(N % 4 == 0 && N <= 128, "plane must fit in LDS as whole float4 rows");
    constexpr int Q = N / 4;
    constexpr int NQ = N * Q;                // float4 per plane
    constexpr int PER = (NQ + 127) / 128;    // float4 per thread
    constexpr int kB = 10;
    constexpr int LD = N | 1;
    __shared__ float tile[N * LD];
    const int tid = threadIdx.x;
    const int k = blockIdx.x;
    float* __restrict__ plane = p + (int64_t)k * N * N;
    const float4* __restrict__ pb = reinterpret_cast<const float4*>(plane);
    const float4* __restrict__ rb = reinterpret_cast<const float4*>(rate + (int64_t)k * N * N);
#pragma unroll
    for (int b0 = 0; b0 < PER; b0 += kB) {
        float4 a[kB], r[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int q = (b0 + u) * 128 + tid;
            if (b0 + u < PER && q < NQ) {
                a[u] = pb[q];
                r[u] = rb[q];
            }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int q = (b0 + u) * 128 + tid;
            if (b0 + u < PER && q < NQ) {
                const int j = q / Q, i0 = 4 * (q - j * Q);
                const float av[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
                const float rv[4] = {r[u].x, r[u].y, r[u].z, r[u].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float x = av[e];
                    if (on_boundary(i0 + e, j, k, N)) x = bnd;
                    tile[j * LD + i0 + e] = x / (1.f - rv[e] * dt);
                }
            }
        }
    }
    __syncthreads();
    if (tid < N) {  // x line j = tid
        float v[N];
#pragma unroll
        for (int t = 0; t < N; ++t) v[t] = tile[tid * LD + t];
        thomas_regs<N>(v, E, F, side, bnd);
#pragma unroll
        for (int t = 0; t < N; ++t) tile[tid * LD + t] = v[t];
    }
    __syncthreads();
    if (tid < N) {  // y line i = tid, as lod_yz_reg<N, 1, 0>
        float v[N];
#pragma unroll
        for (int t = 1; t < N - 1; ++t) v[t] = tile[t * LD + tid];
        float y = bnd;
#pragma unroll
        for (int t = 1; t < N - 1; ++t) {
            y = v[t] - E[t] * y;
            v[t] = y;
        }
        tile[(N - 1) * LD + tid] = bnd;
        float nxt = bnd;
#pragma unroll
        for (int t = N - 2; t > 0; --t) {
            nxt = (v[t] - side * nxt) / F[t];
            tile[t * LD + tid] = nxt;
        }
        tile[tid] = bnd;
    }
    __syncthreads();
    float4* __restrict__ po = reinterpret_cast<float4*>(plane);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int q = u * 128 + tid;
        if (q < NQ) {
            const int j = q / Q, i0 = 4 * (q - j * Q);
            const float* src = &tile[j * LD + i0];
            po[q] = make_float4(src[0], src[1], src[2], src[3]);
        }
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

// ---------------------------------------------------------------------------
// Partitioned line solve (P segments per line, one wave per segment).
//
// The line system of the reference's sweep is A x = d with
//   row 0      [corner, side]            d = b  (the injected Y[0] = b)
//   rows 1..   [side, diag, side]        d = p
//   row N-1    [0 ... 0, 1]              d = b  (p[N-1] = b)
// and x[0] overwritten with b afterwards. Splitting the rows into P blocks of
// m = N / P, each block is solved alone with its own Thomas factors (El, Fl;
// the table holds iFl = 1 / Fl, so the back sweep multiplies),
// giving y_k; the exact solution is x_k = y_k - u_{k-1} W_k - t_{k+1} V_k,
// where W_k, V_k are the block's spikes (its solution for the coupling
// coefficient placed on its first / last row) and t_k, u_k are the first and
// last unknowns of block k, obtained from the 2P boundary values of y through
// the inverse of the reduced 2P x 2P system. Spikes and the reduced inverse
// depend only on the matrix, which every line shares, so they are computed
// once on the host (in double) and read as wave-uniform scalars.
// Table layout (floats): El[N] | iFl[N] | W[N] | V[N] | Rinv[(2P)^2].
// ---------------------------------------------------------------------------
// Block-local forward / backward sweep with the block's own Thomas factors
// (iFl = reciprocal pivots).
template <int m>
__device__ __forceinline__ void part_local(float (&v)[m], const float* __restrict__ El, const float* __restrict__ iFl,
                                           float side) {
#pragma unroll
    for (int i = 1; i < m; ++i) v[i] = v[i] - El[i] * v[i - 1];
    v[m - 1] = v[m - 1] * iFl[m - 1];
#pragma unroll
    for (int i = m - 2; i >= 0; --i) v[i] = (v[i] - side * v[i + 1]) * iFl[i];
}

// bt, bb: P x LS floats of LDS (LS lines per block of the launch).
template <int N, int P, int LS = 64>
__device__ __forceinline__ void part_solve(float (&v)[N / P], const float* __restrict__ tab, int k, float side,
                                           float bnd, float* __restrict__ bt, float* __restrict__ bb, int lane) {
    constexpr int m = N / P;
    const float* W = tab + 2 * N + k * m;
    const float* Vs = tab + 3 * N + k * m;
    const float* Rinv = tab + 4 * N;
    part_local<m>(v, tab + k * m, tab + N + k * m, side);
    // exchange block boundary values through LDS
    bt[k * LS + lane] = v[0];
    bb[k * LS + lane] = v[m - 1];
    __syncthreads();
    float um1 = 0.f, tp1 = 0.f;
    if (k > 0) {
        const float* row = Rinv + (2 * (k - 1) + 1) * (2 * P);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            um1 += row[2 * q] * bt[q * LS + lane];
            um1 += row[2 * q + 1] * bb[q * LS + lane];
        }
    }
    if (k < P - 1) {
        const float* row = Rinv + (2 * (k + 1)) * (2 * P);
#pragma unroll
        for (int q = 0; q < P; ++q) {
            tp1 += row[2 * q] * bt[q * LS + lane];
            tp1 += row[2 * q + 1] * bb[q * LS + lane];
        }
    }
#pragma unroll
    for (int i = 0; i < m; ++i) v[i] = v[i] - um1 * W[i] - tp1 * Vs[i];
    if (k == 0) v[0] = bnd;
}

template <int N, int DIR, int P>
__global__ __launch_bounds__(64 * P) void lod_yz_part(float* __restrict__ p, const float* __restrict__ tab,
                                                      float side, float bnd) {
    constexpr int64_t NN = (int64_t)N * N;
    constexpr int m = N / P;
    __shared__ float bt[P * 64], bb[P * 64];
    const int lane = threadIdx.x & 63;
    const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t L = (int64_t)blockIdx.x * 64 + lane;
    const bool active = L < NN;
    const int64_t Lc = active ? L : 0;
    const int64_t base = (DIR == 1) ? (Lc % N) + NN * (Lc / N) : Lc;
    constexpr int stride_b = (int)((DIR == 1) ? N * 4 : NN * 4);
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(p, NN * N);
    const int voff = (int)(base * 4);
    const int so0 = k * m * stride_b;
    float v[m];
#pragma unroll
    for (int i = 0; i < m; ++i)
        v[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, so0 + i * stride_b, 0));
    if (k == 0) v[0] = bnd;
    if (k == P - 1) v[m - 1] = bnd;
    part_solve<N, P>(v, tab, k, side, bnd, bt, bb, lane);
    if (!active) return;
#pragma unroll
    for (int i = 0; i < m; ++i)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v[i]), rs, voff, so0 + i * stride_b, 0);
}

// x and y sweeps of one z plane with partitioned line solves (P blocks per
// line of m = N / P voxels), for N whose plane fits in LDS and is not a
// multiple of the 32-voxel chunk (N = 100). The plane is read once as in
// lod_xy_plane; thread (k, j) solves block k of x line j out of the odd-pitch
// tile and the blocks are joined through the reduced system (part_solve),
// then the same for the y lines (thread (k, i): block k of y line i), and the
// plane is written once. Rounding differs from the serial sweep at the 1e-7
// relative level, as for lod_x_part / lod_yz_part; each chain is N / P long
// instead of N, and P times as many lanes work on a plane.
template <int N, int P>
__global__ __launch_bounds__(128 * P) void lod_xy_plane_part(float* __restrict__ p, const float* __restrict__ rate,
                                                             const float* __restrict__ tab, float side, float bnd,
                                                             float dt) {
    static_assert(N % 4 == 0 && N <= 128 && N % P == 0, "plane must fit in LDS, lines split evenly");
    constexpr int m = N / P;
    constexpr int Q = N / 4;
    constexpr int NQ = N * Q;
    constexpr int T = 128 * P;
    constexpr int PER = (NQ + T - 1) / T;
    constexpr int kB = 10;
    constexpr int LD = N | 1;
    __shared__ float tile[N * LD];
    __shared__ float bt[P * 128], bb[P * 128];
    const int tid = threadIdx.x;
    const int k = __builtin_amdgcn_readfirstlane(tid >> 7);
    const int j = tid & 127;
    const int kp = blockIdx.x;
    float* __restrict__ plane = p + (int64_t)kp * N * N;
    const float4* __restrict__ pb = reinterpret_cast<const float4*>(plane);
    const float4* __restrict__ rb = reinterpret_cast<const float4*>(rate + (int64_t)kp * N * N);
#pragma unroll
    for (int b0 = 0; b0 < PER; b0 += kB) {
        float4 a[kB], r[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int q = (b0 + u) * T + tid;
            if (b0 + u < PER && q < NQ) {
                a[u] = pb[q];
                r[u] = rb[q];
            }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int q = (b0 + u) * T + tid;
            if (b0 + u < PER && q < NQ) {
                const int jj = q / Q, i0 = 4 * (q - jj * Q);
                const float av[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
                const float rv[4] = {r[u].x, r[u].y, r[u].z, r[u].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float x = av[e];
                    if (on_boundary(i0 + e, jj, kp, N)) x = bnd;
                    tile[jj * LD + i0 + e] = x / (1.f - rv[e] * dt);
                }
            }
        }
    }
    __syncthreads();
    const bool act = j < N;
    float v[m];
#pragma unroll
    for (int i = 0; i < m; ++i) v[i] = act ? tile[j * LD + k * m + i] : 0.f;  // x line j, block k
    if (k == 0) v[0] = bnd;
    if (k == P - 1) v[m - 1] = bnd;
    part_solve<N, P, 128>(v, tab, k, side, bnd, bt, bb, j);
    if (act) {
#pragma unroll
        for (int i = 0; i < m; ++i) tile[j * LD + k * m + i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < m; ++t) v[t] = act ? tile[(k * m + t) * LD + j] : 0.f;  // y line j, block k
    if (k == 0) v[0] = bnd;
    if (k == P - 1) v[m - 1] = bnd;
    part_solve<N, P, 128>(v, tab, k, side, bnd, bt, bb, j);
    if (act) {
#pragma unroll
        for (int t = 0; t < m; ++t) tile[(k * m + t) * LD + j] = v[t];
    }
    __syncthreads();
    float4* __restrict__ po = reinterpret_cast<float4*>(plane);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int q = u * T + tid;
        if (q < NQ) {
            const int jj = q / Q, i0 = 4 * (q - jj * Q);
            const float* src = &tile[jj * LD + i0];
            po[q] = make_float4(src[0], src[1], src[2], src[3]);
        }
    }
}

// x sweep with fused boundary + uptake for the 64 x lines L0.. (P waves per
// line, wave k owning columns [k m, (k + 1) m)); each wave loads its segment
// through its LDS transpose tile and leaves its solved segment in v.
// xspan (per x line, see gell_ctx) limits the rate loads to the voxels that
// may hold a nonzero rate; the others use 0, which is what they hold.
template <int N, int P>
__device__ __forceinline__ void x_lines_solve(float (&v)[N / P], const float* __restrict__ p,
                                              const float* __restrict__ rate, const int2* __restrict__ xspan,
                                              const float* __restrict__ tab, float side, float bnd, float dt,
                                              int64_t L0, int k, int lane, float* __restrict__ tile,
                                              float* __restrict__ bt, float* __restrict__ bb) {
    constexpr int m = N / P;
    static_assert(m % kChunk == 0, "segment must be a whole number of chunks");
    constexpr int NC = m / kChunk;
    constexpr int Q = kChunk / 4;
    constexpr int LPI = 64 / Q;
    constexpr int64_t NN = (int64_t)N * N;
    const __amdgpu_buffer_rsrc_t rp = field_rsrc(p, NN * N);
    const __amdgpu_buffer_rsrc_t rr = field_rsrc(rate, NN * N);
    const int lq = lane / Q, part = lane % Q;
    const int voff = (lq * N + part * 4) * 4;
    const int sbase = (int)(L0 * N * 4) + k * m * 4;
    // With at most two chunks every load of the segment is issued before the
    // first transpose, so the second chunk does not wait for a round trip of
    // its own behind the first chunk's barriers. Not with eight waves per
    // workgroup (N = 512): the preload's 128 registers would leave room for
    // one such workgroup per CU instead of two.
    constexpr bool kPre = NC <= 2 && P < 8;
    int2 sp[Q];
#pragma unroll
    for (int it = 0; it < Q; ++it) sp[it] = xspan[L0 + it * LPI + lq];
    // the lane's four voxels [i0, i0 + 3] of line `it` may hold a nonzero rate
    auto rate_needed = [&](int it, int i0) { return i0 <= sp[it].y && i0 + 3 >= sp[it].x; };
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 pa[kPre ? NC : 1][Q], pr[kPre ? NC : 1][Q];
    if constexpr (kPre) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int it = 0; it < Q; ++it) {
                const int so = sbase + (it * LPI * N + c * kChunk) * 4;
                pa[c][it] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rp, voff, so, 0));
                pr[c][it] = rate_needed(it, k * m + c * kChunk + part * 4)
                                ? __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rr, voff, so, 0))
                                : zero4;
            }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int it = 0; it < Q; ++it) {
            const int l = it * LPI + lq;
            const int64_t L = L0 + l;
            const int j = (int)(L % N), kk = (int)(L / N);
            const int i0 = k * m + c * kChunk + part * 4;
            const int so = sbase + (it * LPI * N + c * kChunk) * 4;
            float4 a, r;
            if constexpr (kPre) {
                a = pa[c][it];
                r = pr[c][it];
            } else {
                a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rp, voff, so, 0));
                r = rate_needed(it, i0) ? __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rr, voff, so, 0))
                                        : zero4;
            }
            const float av[4] = {a.x, a.y, a.z, a.w};
            const float rv[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float x = av[e];
                if (on_boundary(i0 + e, j, kk, N)) x = bnd;
                x = x / (1.f - rv[e] * dt);
                tile[l * kTileLd + part * 4 + e] = x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kChunk; ++t) v[c * kChunk + t] = tile[lane * kTileLd + t];
        __syncthreads();
    }
    if (k == 0) v[0] = bnd;
    if (k == P - 1) v[m - 1] = bnd;
    part_solve<N, P>(v, tab, k, side, bnd, bt, bb, lane);
}

template <int N, int P>
__global__ __launch_bounds__(64 * P) void lod_x_part(float* __restrict__ p, const float* __restrict__ rate,
                                                     const int2* __restrict__ xspan, const float* __restrict__ tab,
                                                     float side, float bnd, float dt) {
    constexpr int m = N / P;
    constexpr int NC = m / kChunk;
    constexpr int Q = kChunk / 4;
    constexpr int LPI = 64 / Q;
    __shared__ float tiles[P][64 * kTileLd];
    __shared__ float bt[P * 64], bb[P * 64];
    const int lane = threadIdx.x & 63;
    const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float* tile = tiles[k];
    const int64_t L0 = (int64_t)blockIdx.x * 64;
    const int lq = lane / Q, part = lane % Q;
    float v[m];
    x_lines_solve<N, P>(v, p, rate, xspan, tab, side, bnd, dt, L0, k, lane, tile, bt, bb);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int t = 0; t < kChunk; ++t) tile[lane * kTileLd + t] = v[c * kChunk + t];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < Q; ++it) {
            const int l = it * LPI + lq;
            const float* src = &tile[l * kTileLd + part * 4];
            // 16-byte stores go through global (FLAT) stores: the compiler's
            // hazard model covers a VALU overwrite of their data registers.
            *reinterpret_cast<float4*>(p + (L0 + l) * N + k * m + c * kChunk + part * 4) =
                make_float4(src[0], src[1], src[2], src[3]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Fused substep for N = 64 P (P = 2, 4, 8): 20 bytes of HBM traffic per voxel
// instead of 28, with the same arithmetic as lod_x_part<N, P> followed by
// lod_yz_part<N, 1, P> and lod_yz_part<N, 2, PZ> (bit-identical results).
//
//   lod_xy_part    one workgroup per (z plane, block of 64 y rows): the x
//                  sweep of those 64 lines (P waves, as lod_x_part), the block
//                  into LDS, then one thread per column runs the y block's
//                  local sweep; writes the local y solution to the field and
//                  its first / last values to the coupling buffer.
//   lod_y_reduce   per y line: the P blocks' first / last values through the
//                  reduced inverse -> the two coupling coefficients per block.
//   lod_z_part_ycorr  z sweep (PZ waves per 64 lines) whose loads apply the
//                  y correction x = y - um1 W[j] - tp1 V[j] of their row.
//
// Coupling buffer (floats, 4 P N^2): first | last | um1 | tp1, each indexed
// [(k P + block) N + i] for the y line (i, k).
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(N) void lod_xy_part(const float* src, float* dst, const float* __restrict__ rate,
                                                 const int2* __restrict__ xspan, const float* __restrict__ tab,
                                                 float side, float bnd, float dt, float* __restrict__ coup) {
    constexpr int P = N / 64;
    constexpr int m = 64;  // x segment per wave = y rows per block
    constexpr int64_t NN = (int64_t)N * N;
    constexpr int XT = P * 64 * kTileLd;
    __shared__ float smem[XT];
    __shared__ float bt[P * 64], bb[P * 64];
    const int lane = threadIdx.x & 63;
    const int kx = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t L0 = (int64_t)blockIdx.x * 64;  // x lines (j, k), j = L % N
    const int kz = (int)(L0 / N), yb = (int)((L0 % N) / 64);
    float v[m];
    x_lines_solve<N, P>(v, src, rate, xspan, tab, side, bnd, dt, L0, kx, lane, smem + kx * 64 * kTileLd, bt, bb);
    // Column i = kx m + lane of the block comes from this wave's own 64 x 64
    // segment (row r = lane r), so each wave transposes it through its own
    // 64 x 33 tile in two halves of 32 columns. (Every wave is past its
    // transpose-tile reads: part_solve synchronises.)
    float* tw = smem + kx * 64 * kTileLd;
    float col[m];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int t = 0; t < kChunk; ++t) tw[lane * kTileLd + t] = v[h * kChunk + t];
        __syncthreads();
        if ((lane >> 5) == h) {
#pragma unroll
            for (int r = 0; r < m; ++r) col[r] = tw[r * kTileLd + (lane & 31)];
        }
        __syncthreads();
    }
    const int i = threadIdx.x;
#pragma unroll
    for (int t = 0; t < m; ++t) v[t] = col[t];
    if (yb == 0) v[0] = bnd;
    if (yb == P - 1) v[m - 1] = bnd;
    part_local<m>(v, tab + yb * m, tab + N + yb * m, side);
    coup[(kz * P + yb) * N + i] = v[0];
    coup[P * NN + (kz * P + yb) * N + i] = v[m - 1];
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(dst, NN * N);
    const int so0 = (int)(L0 * N * 4);
#pragma unroll
    for (int t = 0; t < m; ++t)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v[t]), rs, i * 4, so0 + t * N * 4, 0);
}

template <int N>
__global__ __launch_bounds__(256) void lod_y_reduce(const float* __restrict__ tab, float* __restrict__ coup) {
    constexpr int P = N / 64;
    constexpr int64_t NN = (int64_t)N * N;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= NN) return;
    const int kz = (int)(t / N), i = (int)(t % N);
    const float* Rinv = tab + 4 * N;
    float f[P], l[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        f[q] = coup[(kz * P + q) * N + i];
        l[q] = coup[P * NN + (kz * P + q) * N + i];
    }
#pragma unroll
    for (int b = 0; b < P; ++b) {
        float um1 = 0.f, tp1 = 0.f;
        if (b > 0) {
            const float* row = Rinv + (2 * (b - 1) + 1) * (2 * P);
#pragma unroll
            for (int q = 0; q < P; ++q) {
                um1 += row[2 * q] * f[q];
                um1 += row[2 * q + 1] * l[q];
            }
        }
        if (b < P - 1) {
            const float* row = Rinv + (2 * (b + 1)) * (2 * P);
#pragma unroll
            for (int q = 0; q < P; ++q) {
                tp1 += row[2 * q] * f[q];
                tp1 += row[2 * q + 1] * l[q];
            }
        }
        coup[2 * P * NN + (kz * P + b) * N + i] = um1;
        coup[3 * P * NN + (kz * P + b) * N + i] = tp1;
    }
}

// With NORMS the kernel also sums, per block in a fixed order, (p - p0)^2
// over the interior (voxels two or more from every face) and p0^2 over all
// voxels, in double, into part[2 block], part[2 block + 1]: the increment
// norms of a first substep run out of place (p0 its input), as incr_sums
// would compute them from the two fields, without reading p again.
template <int N, int PZ, bool NORMS = false>
__global__ __launch_bounds__(64 * PZ) void lod_z_part_ycorr(float* __restrict__ p, const float* __restrict__ tabz,
                                                            const float* __restrict__ taby,
                                                            const float* __restrict__ coup, float side, float bnd,
                                                            const float* __restrict__ p0 = nullptr,
                                                            double* __restrict__ part = nullptr) {
    constexpr int PY = N / 64;
    constexpr int64_t NN = (int64_t)N * N;
    constexpr int m = N / PZ;
    __shared__ float bt[PZ * 64], bb[PZ * 64];
    const int lane = threadIdx.x & 63;
    const int kb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t L = (int64_t)blockIdx.x * 64 + lane;  // z line (i, j)
    const int j = __builtin_amdgcn_readfirstlane((int)(((int64_t)blockIdx.x * 64) / N));
    const int i = (int)(L % N);
    const int yb = j / 64;
    const float Wj = taby[2 * N + j], Vj = taby[3 * N + j];
    // coupling pair of row block yb for the y line (i, k): um1 at
    // [2 PY N^2 + (k PY + yb) N + i], tp1 PY N^2 floats further
    const __amdgpu_buffer_rsrc_t rc = field_rsrc(coup, 4 * PY * NN);
    const int coff = (int)((2 * PY * NN + yb * N + i) * 4);
    constexpr int ct_b = (int)(PY * NN * 4);
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(p, NN * N);
    const int voff = (int)(L * 4);
    constexpr int stride_b = (int)(NN * 4);
    const int so0 = kb * m * stride_b;
    float v[m];
#pragma unroll
    for (int t = 0; t < m; ++t) {
        const int kz = kb * m + t;
        const float y = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, so0 + t * stride_b, 0));
        const int cso = kz * PY * N * 4;
        const float um1 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rc, coff, cso, 0));
        const float tp1 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rc, coff, cso + ct_b, 0));
        float x = y - um1 * Wj;
        x = x - tp1 * Vj;
        v[t] = j == 0 ? bnd : x;
    }
    if (kb == 0) v[0] = bnd;
    if (kb == PZ - 1) v[m - 1] = bnd;
    part_solve<N, PZ>(v, tabz, kb, side, bnd, bt, bb, lane);
#pragma unroll
    for (int t = 0; t < m; ++t)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v[t]), rs, voff, so0 + t * stride_b, 0);
    if constexpr (NORMS) {
        const __amdgpu_buffer_rsrc_t r0 = field_rsrc(p0, NN * N);
        const bool line_in = i >= 2 && i < N - 2 && j >= 2 && j < N - 2;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int t = 0; t < m; ++t) {
            const int kz = kb * m + t;
            const double a =
                (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r0, voff, so0 + t * stride_b, 0));
            s1 += a * a;
            if (line_in && kz >= 2 && kz < N - 2) {
                const double d = (double)v[t] - a;
                s0 += d * d;
            }
        }
        // block sums in a fixed order: the waves' lanes by tree in LDS
        __shared__ double r0s[64 * PZ], r1s[64 * PZ];
        r0s[threadIdx.x] = s0;
        r1s[threadIdx.x] = s1;
        __syncthreads();
        for (int w = 32 * PZ; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                r0s[threadIdx.x] += r0s[threadIdx.x + w];
                r1s[threadIdx.x] += r1s[threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[2 * blockIdx.x] = r0s[0];
            part[2 * blockIdx.x + 1] = r1s[0];
        }
    }
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

// Serial line kernels for an N that is not a multiple of the chunk (x sweep
// staged whole through LDS, y and z in registers).
template <int N>
hipError_t launch_reg_flat(float* p, const float* rate, const float* E, const float* F, float side, float bnd,
                           float dt, int spill, hipStream_t st) {
    const unsigned blocks = (unsigned)(((int64_t)N * N + 63) / 64);
    hipLaunchKernelGGL((lod_x_flat<N>), dim3(blocks), dim3(64), 0, st, p, rate, E, F, side, bnd, dt);
    if (spill > 0) {
        hipLaunchKernelGGL((lod_yz_reg<N, 1, 64>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
        hipLaunchKernelGGL((lod_yz_reg<N, 2, 64>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
    } else {
        hipLaunchKernelGGL((lod_yz_reg<N, 1, 0>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
        hipLaunchKernelGGL((lod_yz_reg<N, 2, 0>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
    }
    return hipGetLastError();
}

template <int N, int P>
void launch_x_part(float* p, const float* rate, const int2* xspan, const float* tab, float side, float bnd, float dt,
                   hipStream_t st) {
    const unsigned blocks = (unsigned)(((int64_t)N * N + 63) / 64);
    hipLaunchKernelGGL((lod_x_part<N, P>), dim3(blocks), dim3(64 * P), 0, st, p, rate, xspan, tab, side, bnd, dt);
}

template <int N, int DIR, int P>
void launch_yz_part(float* p, const float* tab, float side, float bnd, hipStream_t st) {
    const unsigned blocks = (unsigned)(((int64_t)N * N + 63) / 64);
    hipLaunchKernelGGL((lod_yz_part<N, DIR, P>), dim3(blocks), dim3(64 * P), 0, st, p, tab, side, bnd);
}

// N in {64, 96, 100}: the x and y sweeps run per z plane (xp = 2 or 4:
// partitioned blocks; 1: serial lines, lod_xy_plane), the z sweep
// partitioned (yzp = 2 or 4, lod_yz_part) or serial (lod_yz_reg).
template <int N>
hipError_t launch_plane_part(float* p, const float* rate, const float* E, const float* F, float* const* tabs,
                             float side, float bnd, float dt, int xp, int yzp, int spill, hipStream_t st) {
    const unsigned blocks = (unsigned)(((int64_t)N * N + 63) / 64);
    if (xp >= 4)
        hipLaunchKernelGGL((lod_xy_plane_part<N, 4>), dim3(N), dim3(512), 0, st, p, rate, tabs[1], side, bnd, dt);
    else if (xp >= 2)
        hipLaunchKernelGGL((lod_xy_plane_part<N, 2>), dim3(N), dim3(256), 0, st, p, rate, tabs[0], side, bnd, dt);
    else
        hipLaunchKernelGGL((lod_xy_plane<N>), dim3(N), dim3(128), 0, st, p, rate, E, F, side, bnd, dt);
    if (yzp >= 4)
        launch_yz_part<N, 2, 4>(p, tabs[1], side, bnd, st);
    else if (yzp >= 2)
        launch_yz_part<N, 2, 2>(p, tabs[0], side, bnd, st);
    else if (spill > 0)
        hipLaunchKernelGGL((lod_yz_reg<N, 2, 64>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
    else
        hipLaunchKernelGGL((lod_yz_reg<N, 2, 0>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
    return hipGetLastError();
}

// One substep with per-sweep choices: xp / yzp = 1 (serial line per lane),
// 2, 4, 8 or 16 (partitioned; a choice the grid does not support -- N / P
// not a whole number of 32-voxel chunks -- falls back to the next smaller
// one; the x sweep stops at 8 blocks, whose transpose tiles already take
// 66 KB of LDS); spill selects the LDS-spilled serial y/z kernel. Above
// N = 256 a whole line does not fit in registers and the serial choice runs
// the generic kernel.
template <int N>
hipError_t launch_mixed(float* p, const float* rate, const int2* xspan, const float* E, const float* F, float* const* tabs,
                        float side, float bnd, float dt, int xp, int yzp, int spill, hipStream_t st) {
    constexpr bool ok2 = N % (2 * kChunk) == 0, ok4 = N % (4 * kChunk) == 0, ok8 = N % (8 * kChunk) == 0,
                   ok16 = N % (16 * kChunk) == 0;
    const unsigned blocks = (unsigned)(((int64_t)N * N + 63) / 64);
    const unsigned gblocks = (unsigned)(((int64_t)N * N + 255) / 256);
    if (xp >= 8 && ok8)
        launch_x_part<N, ok8 ? 8 : 2>(p, rate, xspan, tabs[2], side, bnd, dt, st);
    else if (xp >= 4 && ok4)
        launch_x_part<N, ok4 ? 4 : 2>(p, rate, xspan, tabs[1], side, bnd, dt, st);
    else if (xp >= 2 && ok2)
        launch_x_part<N, 2>(p, rate, xspan, tabs[0], side, bnd, dt, st);
    else if constexpr (N <= 256)
        hipLaunchKernelGGL((lod_x_reg<N>), dim3(blocks), dim3(64), 0, st, p, rate, E, F, side, bnd, dt);
    else
        hipLaunchKernelGGL(lod_sweep_any, dim3(gblocks), dim3(256), 0, st, p, rate, E, F, side, bnd, dt, N, 0);
    if (yzp >= 16 && ok16) {
        launch_yz_part<N, 1, ok16 ? 16 : 2>(p, tabs[3], side, bnd, st);
        launch_yz_part<N, 2, ok16 ? 16 : 2>(p, tabs[3], side, bnd, st);
    } else if (yzp >= 8 && ok8) {
        launch_yz_part<N, 1, ok8 ? 8 : 2>(p, tabs[2], side, bnd, st);
        launch_yz_part<N, 2, ok8 ? 8 : 2>(p, tabs[2], side, bnd, st);
    } else if (yzp >= 4 && ok4) {
        launch_yz_part<N, 1, ok4 ? 4 : 2>(p, tabs[1], side, bnd, st);
        launch_yz_part<N, 2, ok4 ? 4 : 2>(p, tabs[1], side, bnd, st);
    } else if (yzp >= 2 && ok2) {
        launch_yz_part<N, 1, 2>(p, tabs[0], side, bnd, st);
        launch_yz_part<N, 2, 2>(p, tabs[0], side, bnd, st);
    } else if constexpr (N <= 256) {
        if (spill) {
            hipLaunchKernelGGL((lod_yz_reg<N, 1, 64>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
            hipLaunchKernelGGL((lod_yz_reg<N, 2, 64>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
        } else {
            hipLaunchKernelGGL((lod_yz_reg<N, 1, 0>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
            hipLaunchKernelGGL((lod_yz_reg<N, 2, 0>), dim3(blocks), dim3(64), 0, st, p, E, F, side, bnd);
        }
    } else {
        for (int dir = 1; dir < 3; ++dir)
            hipLaunchKernelGGL(lod_sweep_any, dim3(gblocks), dim3(256), 0, st, p, rate, E, F, side, bnd, dt, N, dir);
    }
    return hipGetLastError();
}

template <int N, int PZ>
void launch_z_ycorr(float* p, const float* tabz, const float* taby, const float* coup, float side, float bnd,
                    hipStream_t st, SubstepNorms* norms) {
    const dim3 grid((unsigned)((int64_t)N * N / 64));
    if (norms) {
        hipLaunchKernelGGL((lod_z_part_ycorr<N, PZ, true>), grid, dim3(64 * PZ), 0, st, p, tabz, taby, coup, side, bnd,
                           norms->p0, norms->part);
        norms->nblk = (int)grid.x;
    } else {
        hipLaunchKernelGGL((lod_z_part_ycorr<N, PZ>), grid, dim3(64 * PZ), 0, st, p, tabz, taby, coup, side, bnd,
                           nullptr, nullptr);
    }
}

// Fused substep (see lod_xy_part); zp = z blocks, falling back to the next
// smaller count the grid supports.
template <int N>
hipError_t launch_fused(const float* src, float* p, const float* rate, const int2* xspan, float* const* tabs,
                        float side, float bnd, float dt, int zp, float* coup, hipStream_t st,
                        SubstepNorms* norms = nullptr) {
    constexpr int P = N / 64;
    static_assert(P == 2 || P == 4 || P == 8, "fused substep needs N in {128, 256, 512}");
    const float* taby = tabs[P == 2 ? 0 : (P == 4 ? 1 : 2)];
    // the x + y kernel reads src and writes p (src == p: in place); the z
    // sweep then works on p
    hipLaunchKernelGGL((lod_xy_part<N>), dim3((unsigned)((int64_t)N * N / 64)), dim3(N), 0, st, src, p, rate, xspan,
                       taby, side, bnd, dt, coup);
    hipLaunchKernelGGL((lod_y_reduce<N>), dim3((unsigned)(((int64_t)N * N + 255) / 256)), dim3(256), 0, st, taby,
                       coup);
    constexpr bool ok4 = N % (4 * kChunk) == 0, ok8 = N % (8 * kChunk) == 0, ok16 = N % (16 * kChunk) == 0;
    if (zp >= 16 && ok16)
        launch_z_ycorr<N, ok16 ? 16 : 2>(p, tabs[3], taby, coup, side, bnd, st, norms);
    else if (zp >= 8 && ok8)
        launch_z_ycorr<N, ok8 ? 8 : 2>(p, tabs[2], taby, coup, side, bnd, st, norms);
    else if (zp >= 4 && ok4)
        launch_z_ycorr<N, ok4 ? 4 : 2>(p, tabs[1], taby, coup, side, bnd, st, norms);
    else
        launch_z_ycorr<N, 2>(p, tabs[0], taby, coup, side, bnd, st, norms);
    return hipGetLastError();
}

}  // namespace

bool lod_fused_active(const gell_ctx* ctx) {
    const int N = ctx->N;
    return ctx->tune[GELL_KNOB_LOD_FUSED] && ctx->tune[GELL_KNOB_LOD_PARTS] > 1 && ctx->coup[0] &&
           (N == 128 || N == 256 || N == 512);
}

hipError_t lod_substep_launch(gell_ctx* ctx, int s, hipStream_t st) {
    return lod_substep_launch(ctx, s, ctx->field[s], ctx->field[s], st);
}

// One substep from the field in src to the field in dst (src == dst: in
// place). The fused substep reads src in its first kernel; the other forms
// copy src to dst first and work in place on dst.
hipError_t lod_substep_launch(gell_ctx* ctx, int s, const float* src, float* dst, hipStream_t st,
                              SubstepNorms* norms) {
    const int N = ctx->N;
    float* p = dst;
    if (norms) norms->nblk = 0;
    const float* rate = ctx->rate[s];
    const float* E = ctx->E[s];
    const float* F = ctx->F[s];
    const float side = ctx->side[s], bnd = ctx->P.sub[s].boundary, dt = ctx->P.dt_diff;
    const int spill = ctx->tune[GELL_KNOB_LOD_YZ_LDS];
    const int yzp = ctx->tune[GELL_KNOB_LOD_PARTS];
    const int xp = ctx->tune[GELL_KNOB_LOD_X_PARTS];
    const bool fused = ctx->tune[GELL_KNOB_LOD_FUSED] != 0;
    const int2* xspan = ctx->xspan[s];
    // the partitioned x sweep (fused, or xp > 1) reads the line spans of the uptake map
    const bool spans = lod_fused_active(ctx) || (xp > 1 && (N == 128 || N == 192 || N == 256 || N == 512));
    if (spans && !ctx->xspan_valid[s]) {
        const hipError_t e = xspan_rebuild(ctx, s, st);
        if (e != hipSuccess) return e;
    }
    if (lod_fused_active(ctx)) {
        float* const* tabs = ctx->ptab[s];
        if (N == 128) return launch_fused<128>(src, p, rate, xspan, tabs, side, bnd, dt, yzp, ctx->coup[s], st, norms);
        if (N == 256) return launch_fused<256>(src, p, rate, xspan, tabs, side, bnd, dt, yzp, ctx->coup[s], st, norms);
        return launch_fused<512>(src, p, rate, xspan, tabs, side, bnd, dt, yzp, ctx->coup[s], st, norms);
    }
    if (src != p) {
        const hipError_t e = hipMemcpyAsync(p, src, sizeof(float) * (size_t)ctx->V, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    if ((xp > 1 || yzp > 1) && (N == 128 || N == 192 || N == 256 || N == 512)) {
        float* const* tabs = ctx->ptab[s];
        if (N == 128) return launch_mixed<128>(p, rate, xspan, E, F, tabs, side, bnd, dt, xp, yzp, spill, st);
        if (N == 192) return launch_mixed<192>(p, rate, xspan, E, F, tabs, side, bnd, dt, xp, yzp, spill, st);
        if (N == 256) return launch_mixed<256>(p, rate, xspan, E, F, tabs, side, bnd, dt, xp, yzp, spill, st);
        return launch_mixed<512>(p, rate, xspan, E, F, tabs, side, bnd, dt, xp, yzp, spill, st);
    }
    switch (N) {
        case 32: return launch_reg<32>(p, rate, E, F, side, bnd, dt, spill, st);
        case 64:
            if (xp > 1 || yzp > 1 || fused)
                return launch_plane_part<64>(p, rate, E, F, ctx->ptab[s], side, bnd, dt, xp, yzp, spill, st);
            return launch_reg<64>(p, rate, E, F, side, bnd, dt, spill, st);
        case 96:
            if (xp > 1 || yzp > 1 || fused)
                return launch_plane_part<96>(p, rate, E, F, ctx->ptab[s], side, bnd, dt, xp, yzp, spill, st);
            return launch_reg<96>(p, rate, E, F, side, bnd, dt, spill, st);
        case 128: return launch_reg<128>(p, rate, E, F, side, bnd, dt, spill, st);
        case 192: return launch_reg<192>(p, rate, E, F, side, bnd, dt, spill, st);
        case 256: return launch_reg<256>(p, rate, E, F, side, bnd, dt, spill, st);
        case 100:
            if (xp > 1 || yzp > 1 || fused)
                return launch_plane_part<100>(p, rate, E, F, ctx->ptab[s], side, bnd, dt, xp, yzp, spill, st);
            return launch_reg_flat<100>(p, rate, E, F, side, bnd, dt, spill, st);
        case 200: return launch_reg_flat<200>(p, rate, E, F, side, bnd, dt, spill, st);
        default: break;
    }
    const int64_t lines = (int64_t)N * N;
    const unsigned blocks = (unsigned)((lines + 255) / 256);
    for (int dir = 0; dir < 3; ++dir)
        hipLaunchKernelGGL(lod_sweep_any, dim3(blocks), dim3(256), 0, st, p, rate, E, F, side, bnd, dt, N, dir);
    return hipGetLastError();
}

namespace {

// Tables of the partitioned solve (see part_solve) for P blocks, in double.
std::vector<float> part_table(int N, int P, double side, double corner, double diag) {
    const int m = N / P;
    std::vector<double> sub(N), dia(N), sup(N);
    for (int i = 0; i < N; ++i) {
        sub[i] = (i == 0 || i == N - 1) ? 0.0 : side;
        dia[i] = i == 0 ? corner : (i == N - 1 ? 1.0 : diag);
        sup[i] = i == N - 1 ? 0.0 : side;
    }
    std::vector<double> El(N), Fl(N), W(N, 0.0), Vs(N, 0.0);
    auto local_solve = [&](int k, std::vector<double>& rhs) {  // in place, block k
        const int a = k * m;
        std::vector<double> Y(m);
        Y[0] = rhs[0];
        for (int i = 1; i < m; ++i) Y[i] = rhs[i] - El[a + i] * Y[i - 1];
        rhs[m - 1] = Y[m - 1] / Fl[a + m - 1];
        for (int i = m - 2; i >= 0; --i) rhs[i] = (Y[i] - sup[a + i] * rhs[i + 1]) / Fl[a + i];
    };
    for (int k = 0; k < P; ++k) {
        const int a = k * m;
        El[a] = 0.0;
        Fl[a] = dia[a];
        for (int i = 1; i < m; ++i) {
            El[a + i] = sub[a + i] / Fl[a + i - 1];
            Fl[a + i] = dia[a + i] - El[a + i] * sup[a + i - 1];
        }
        std::vector<double> w(m, 0.0), v(m, 0.0);
        if (k > 0) {
            w[0] = sub[a];
            local_solve(k, w);
        }
        if (k < P - 1) {
            v[m - 1] = sup[a + m - 1];
            local_solve(k, v);
        }
        for (int i = 0; i < m; ++i) {
            W[a + i] = w[i];
            Vs[a + i] = v[i];
        }
    }
    // reduced system R z = g, z = (t0, u0, t1, u1, ...)
    const int R2 = 2 * P;
    std::vector<double> R(R2 * R2, 0.0), Ri(R2 * R2, 0.0);
    for (int k = 0; k < P; ++k) {
        const int a = k * m;
        R[(2 * k) * R2 + 2 * k] = 1.0;
        R[(2 * k + 1) * R2 + 2 * k + 1] += 1.0;
        if (k > 0) {
            R[(2 * k) * R2 + 2 * (k - 1) + 1] += W[a];
            R[(2 * k + 1) * R2 + 2 * (k - 1) + 1] += W[a + m - 1];
        }
        if (k < P - 1) {
            R[(2 * k) * R2 + 2 * (k + 1)] += Vs[a];
            R[(2 * k + 1) * R2 + 2 * (k + 1)] += Vs[a + m - 1];
        }
    }
    for (int i = 0; i < R2; ++i) Ri[i * R2 + i] = 1.0;
    for (int c = 0; c < R2; ++c) {  // Gauss-Jordan with partial pivoting
        int piv = c;
        for (int r = c + 1; r < R2; ++r)
            if (std::fabs(R[r * R2 + c]) > std::fabs(R[piv * R2 + c])) piv = r;
        for (int q = 0; q < R2; ++q) {
            std::swap(R[c * R2 + q], R[piv * R2 + q]);
            std::swap(Ri[c * R2 + q], Ri[piv * R2 + q]);
        }
        const double d = R[c * R2 + c];
        for (int q = 0; q < R2; ++q) {
            R[c * R2 + q] /= d;
            Ri[c * R2 + q] /= d;
        }
        for (int r = 0; r < R2; ++r) {
            if (r == c) continue;
            const double f = R[r * R2 + c];
            if (f == 0.0) continue;
            for (int q = 0; q < R2; ++q) {
                R[r * R2 + q] -= f * R[c * R2 + q];
                Ri[r * R2 + q] -= f * Ri[c * R2 + q];
            }
        }
    }
    std::vector<float> t(4 * N + R2 * R2);
    for (int i = 0; i < N; ++i) {
        t[i] = (float)El[i];
        t[N + i] = (float)(1.0 / Fl[i]);
        t[2 * N + i] = (float)W[i];
        t[3 * N + i] = (float)Vs[i];
    }
    for (int i = 0; i < R2 * R2; ++i) t[4 * N + i] = (float)Ri[i];
    return t;
}

}  // namespace

// Times every effective LOD variant on a scratch copy of substrate 0's field
// (the context's own field and uptake map are left as they were) and sets
// the fastest. bit_exact restricts the search to the serial-line kernels.
hipError_t lod_autotune(gell_ctx* ctx, bool bit_exact, double* best_us) {
    const size_t bytes = sizeof(float) * (size_t)ctx->V;
    float* scratch = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&scratch), bytes);
    if (e != hipSuccess) return e;
    float* const real = ctx->field[0];
    e = hipMemcpyAsync(scratch, real, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    hipEvent_t a = nullptr, b = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    // candidate (x parts, y/z parts, LDS spill, fused)
    struct Cand {
        int xp, yzp, sp, fused;
    };
    std::vector<Cand> cands;
    const int N = ctx->N;
    const bool part_ok = (N == 128 || N == 192 || N == 256 || N == 512) && !bit_exact;
    const int cand[5] = {1, 2, 4, 8, 16};
    for (int xi = 0; xi < (part_ok ? 4 : 1); ++xi)
        for (int yi = 0; yi < (part_ok ? 5 : 1); ++yi)
            for (int sp = 0; sp < 2; ++sp) {
                if ((cand[yi] > 1 || N > 256) && sp == 1) continue;  // the spill exists for serial y/z only
                const int xp = cand[xi], yzp = cand[yi];
                if ((xp > 1 && N % (xp * 32) != 0) || (yzp > 1 && N % (yzp * 32) != 0)) continue;
                cands.push_back({xp, yzp, sp, 0});
            }
    if (part_ok && ctx->coup[0])
        for (int yi = 1; yi < 5; ++yi)
            if (N % (cand[yi] * 32) == 0) cands.push_back({N / 64, cand[yi], 0, 1});
    if (N == 64 || N == 96 || N == 100) {  // x + y per z plane: serial lines (bit-exact), then partitioned blocks
        for (int sp = 0; sp < 2; ++sp) cands.push_back({1, 1, sp, 1});
        if (!bit_exact)
            for (int xp : {1, 2, 4})
                for (int yzp : {1, 2, 4})
                    if (xp > 1 || yzp > 1) cands.push_back({xp, yzp, 0, 1});
    }
    const int knobs[4] = {GELL_KNOB_LOD_X_PARTS, GELL_KNOB_LOD_PARTS, GELL_KNOB_LOD_YZ_LDS, GELL_KNOB_LOD_FUSED};
    int saved[4], best[4];
    for (int q = 0; q < 4; ++q) saved[q] = best[q] = ctx->tune[knobs[q]];
    double best_t = 1e30;
    ctx->field[0] = scratch;
    for (const Cand& c : cands) {
        if (e != hipSuccess) break;
        const int val[4] = {c.xp, c.yzp, c.sp, c.fused};
        for (int q = 0; q < 4; ++q) ctx->tune[knobs[q]] = val[q];
        for (int w = 0; w < 2 && e == hipSuccess; ++w) e = lod_substep_launch(ctx, 0, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
        const int reps = 5;
        for (int r = 0; r < reps && e == hipSuccess; ++r) e = lod_substep_launch(ctx, 0, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        const double us = 1e3 * ms / reps;
        if (e == hipSuccess && us < best_t) {
            best_t = us;
            for (int q = 0; q < 4; ++q) best[q] = val[q];
        }
    }
    ctx->field[0] = real;
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = e2;
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(scratch);
    for (int q = 0; q < 4; ++q) ctx->tune[knobs[q]] = e == hipSuccess ? best[q] : saved[q];
    if (e == hipSuccess && best_us) *best_us = best_t;
    return e;
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
    const double side64 = -(double)dt * sb.D / ((double)dx * dx);
    const double corner64 = 1.0 + (double)dt * sb.decay / 3.0 - side64;
    const double diag64 = corner64 - side64;
    // The host tables must outlive the asynchronous copies: they are kept
    // until the stream synchronisation below.
    const int parts[4] = {2, 4, 8, 16};
    std::vector<float> tabs[4];
    for (int q = 0; q < 4; ++q) {
        if (N % parts[q] != 0 || N / parts[q] < 2) continue;
        tabs[q] = part_table(N, parts[q], side64, corner64, diag64);
        hipError_t e = hipMemcpyAsync(ctx->ptab[s][q], tabs[q].data(), tabs[q].size() * sizeof(float),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipMemcpyAsync(ctx->E[s], E.data(), N * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(ctx->F[s], F.data(), N * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ctx->stream);
}

}  // namespace gell

extern "C" int gell_lod_partition_table(const gell_params* P, int32_t substrate, int32_t parts, float* out,
                                        int32_t out_len) {
    if (!P || !out || substrate < 0 || substrate >= P->n_substrates) return GELL_E_ARG;
    if (parts < 2 || P->N % parts != 0 || P->N / parts < 2) return GELL_E_ARG;
    const int N = P->N;
    if (out_len < 4 * N + 4 * parts * parts) return GELL_E_ARG;
    const gell_substrate& sb = P->sub[substrate];
    const double side = -(double)P->dt_diff * sb.D / ((double)P->dx * P->dx);
    const double corner = 1.0 + (double)P->dt_diff * sb.decay / 3.0 - side;
    const std::vector<float> t = gell::part_table(N, parts, side, corner, corner - side);
    for (size_t i = 0; i < t.size(); ++i) out[i] = t[i];
    return GELL_OK;
}
