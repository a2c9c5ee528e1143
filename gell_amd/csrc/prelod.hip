// prelod.hip -- uptake map of one substrate (Mesh_struct::PreLOD,
// LODkernel.cuh:270-278; oxy_consumption_update, LODkernel.cuh:12-22).
//
// The reference zeroes all N^3 rates and float-atomicAdds each live cell's
// Cell::O2_consume() into its voxel, so its sums depend on the order the
// atomics land in. Here every rate is turned into an exact 2^-32 fixed-point
// integer and summed with 64-bit integer atomics (order-independent), then
// rounded to float once. Only voxels that hold cells are touched: the voxels
// the previous call wrote are cleared first (O(n) instead of O(N^3)) -- along
// each x line's span of nonzero rates, with coalesced stores, when the spans
// of the previous map are at hand, else voxel by voxel from the previous
// call's list. (A voxel's accumulator is nonzero only where its rate is.)
// After the map pointer was handed out to the caller, who may have written
// any voxel, the whole map and its accumulators are cleared instead.
#include "device_common.hpp"
#include "internal.hpp"

namespace gell {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void prelod_clear(const int32_t* __restrict__ list, int32_t m,
                                                       long long* __restrict__ acc, float* __restrict__ rate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t v = list[i];
    if (v >= 0) {
        acc[v] = 0;
        rate[v] = 0.f;
    }
}

// The same from the x-line spans of the map (every nonzero rate lies inside
// its line's span): one wave per line, consecutive lanes on consecutive voxels.
__global__ __launch_bounds__(kBlock) void prelod_clear_spans(const int2* __restrict__ xspan,
                                                             long long* __restrict__ acc, float* __restrict__ rate,
                                                             int32_t N) {
    const int64_t line = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
    if (line >= (int64_t)N * N) return;  // uniform per wave
    const int2 sp = xspan[line];
    float* r = rate + line * N;
    long long* a = acc + line * N;
    for (int i = sp.x + lane_id(); i <= sp.y; i += 64) {
        a[i] = 0;
        r[i] = 0.f;
    }
}

// Whole-map clear (no hipMemsetAsync: the ensemble captures this launch into
// a graph).
__global__ __launch_bounds__(kBlock) void prelod_clear_all(long long* __restrict__ acc, float* __restrict__ rate,
                                                           int64_t V) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) {
        acc[i] = 0;
        rate[i] = 0.f;
    }
}

__global__ __launch_bounds__(kBlock) void prelod_accumulate(const gell_cell* __restrict__ cells, int32_t n,
                                                            int64_t V, float dx, float uptake, float dead_factor,
                                                            long long* __restrict__ acc, int32_t* __restrict__ list) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(cells + i);
    const bool live = (w[W_SIGN] & 0xFFu) != 0;
    const int32_t v = (int32_t)w[W_MESH];
    int32_t out = -1;
    if (live && v >= 0 && v < V) {
        const float rate = cell_uptake(__uint_as_float(w[W_V]), (int32_t)w[W_PHASE], dx, uptake, dead_factor);
        atomicAdd(reinterpret_cast<unsigned long long*>(acc + v), (unsigned long long)to_fixed(rate));
        out = v;
    }
    list[i] = out;
}

__global__ __launch_bounds__(kBlock) void prelod_finalize(const int32_t* __restrict__ list, int32_t n,
                                                          const long long* __restrict__ acc,
                                                          float* __restrict__ rate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = list[i];
    if (v >= 0) rate[v] = from_fixed(acc[v]);
}

// x-line spans of the uptake map: one wave per x line, first / last nonzero
// voxel (one coalesced read of the map; per-cell atomics on the span of a
// line contended too much).
__global__ __launch_bounds__(kBlock) void xspan_scan(const float* __restrict__ rate, int2* __restrict__ xspan,
                                                     int32_t N) {
    const int64_t line = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
    if (line >= (int64_t)N * N) return;  // uniform per wave
    const int lane = lane_id();
    const float* r = rate + line * N;
    int lo = N, hi = -1;
    if ((N & 3) == 0) {  // lines are 16-byte aligned: one float4 per lane per step
        for (int i4 = lane; i4 < N / 4; i4 += 64) {
            const float4 q = reinterpret_cast<const float4*>(r)[i4];
            const float qv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (qv[e] != 0.f) {
                    const int i = 4 * i4 + e;
                    lo = i < lo ? i : lo;
                    hi = i;
                }
        }
    } else {
        for (int i = lane; i < N; i += 64)
            if (r[i] != 0.f) {
                lo = i < lo ? i : lo;
                hi = i;
            }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const int olo = __shfl_xor(lo, d, 64), ohi = __shfl_xor(hi, d, 64);
        lo = olo < lo ? olo : lo;
        hi = ohi > hi ? ohi : hi;
    }
    if (lane == 0) xspan[line] = make_int2(lo, hi);
}

}  // namespace

hipError_t prelod_launch(gell_ctx* ctx, int s, const gell_cell* cells, int32_t n) {
    hipStream_t st = ctx->stream;
    const int32_t m = ctx->rate_list_n[s];
    const int64_t lines = (int64_t)ctx->N * ctx->N;
    if (ctx->rate_handed_out[s])
        hipLaunchKernelGGL(prelod_clear_all, dim3((unsigned)((ctx->V + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                           ctx->acc[s], ctx->rate[s], ctx->V);
    else if (m > 0 && ctx->xspan_valid[s])
        hipLaunchKernelGGL(prelod_clear_spans, dim3((unsigned)((lines + kBlock / 64 - 1) / (kBlock / 64))),
                           dim3(kBlock), 0, st, ctx->xspan[s], ctx->acc[s], ctx->rate[s], ctx->N);
    else if (m > 0)
        hipLaunchKernelGGL(prelod_clear, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, st, ctx->rate_list[s], m,
                           ctx->acc[s], ctx->rate[s]);
    ctx->rate_list_n[s] = 0;
    ctx->rate_handed_out[s] = false;
    if (n > 0) {
        const gell_substrate& sb = ctx->P.sub[s];
        const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(prelod_accumulate, dim3(blocks), dim3(kBlock), 0, st, cells, n, ctx->V, ctx->P.dx,
                           sb.uptake, sb.dead_factor, ctx->acc[s], ctx->rate_list[s]);
        hipLaunchKernelGGL(prelod_finalize, dim3(blocks), dim3(kBlock), 0, st, ctx->rate_list[s], n, ctx->acc[s],
                           ctx->rate[s]);
        ctx->rate_list_n[s] = n;
    }
    ctx->xspan_valid[s] = false;  // rebuilt by the first substep that reads the spans
    return hipGetLastError();
}

hipError_t xspan_rebuild(gell_ctx* ctx, int s, hipStream_t st) {
    const int64_t lines = (int64_t)ctx->N * ctx->N;
    hipLaunchKernelGGL(xspan_scan, dim3((unsigned)((lines + kBlock / 64 - 1) / (kBlock / 64))), dim3(kBlock), 0, st,
                       ctx->rate[s], ctx->xspan[s], ctx->N);
    ctx->xspan_valid[s] = true;
    return hipGetLastError();
}

}  // namespace gell
