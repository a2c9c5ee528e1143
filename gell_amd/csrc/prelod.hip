// prelod.hip -- uptake map of one substrate (Mesh_struct::PreLOD,
// LODkernel.cuh:270-278; oxy_consumption_update, LODkernel.cuh:12-22).
//
// The reference zeroes all N^3 rates and float-atomicAdds each live cell's
// Cell::O2_consume() into its voxel, so its sums depend on the order the
// atomics land in. Here every rate is turned into an exact 2^-32 fixed-point
// integer and summed with 64-bit integer atomics (order-independent), then
// rounded to float once.
//
// The accumulators are indexed by Morton code, so that cells in sorted order
// add to neighbouring words (lanes of one voxel are summed in the wave
// first), and they are zero between maps: each map's voxels are zeroed right
// after the map is written from them. Only voxels that hold cells are
// touched: the previous map is cleared first (O(n) instead of O(N^3)) --
// along each x line's span of nonzero rates, with coalesced stores, when the
// spans of the previous map are at hand, else voxel by voxel from the list
// that built it. After the map pointer was handed out to the caller, who may
// have written any voxel, the whole map is cleared instead.
//
// gell_mech_step uses the same maps with the accumulation done by the
// mechanics step's force/movement kernel (fm.hip): only the clear, the
// finalize and the accumulator reset run here.
#include "device_common.hpp"
#include "internal.hpp"

namespace gell {
namespace {

constexpr int kBlock = 256;

// Rates of the voxels on a list of Morton codes back to zero.
__global__ __launch_bounds__(kBlock) void prelod_clear(const int32_t* __restrict__ list, int32_t m,
                                                       float* __restrict__ rate, int32_t N,
                                                       const int32_t* __restrict__ mp) {
    m = count_of(m, mp);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t c = list[i];
    if (c >= 0) rate[lin_of_code((uint32_t)c, N)] = 0.f;
}

// The same from the x-line spans of the map (every nonzero rate lies inside
// its line's span): one wave per line, consecutive lanes on consecutive voxels.
__global__ __launch_bounds__(kBlock) void prelod_clear_spans(const int2* __restrict__ xspan, float* __restrict__ rate,
                                                             int32_t N) {
    const int64_t line = (int64_t)blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
    if (line >= (int64_t)N * N) return;  // uniform per wave
    const int2 sp = xspan[line];
    float* r = rate + line * N;
    for (int i = sp.x + lane_id(); i <= sp.y; i += 64) r[i] = 0.f;
}

// Whole-map clear (no hipMemsetAsync: the ensemble captures this launch into
// a graph).
__global__ __launch_bounds__(kBlock) void prelod_clear_all(float* __restrict__ rate, int64_t V) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < V) rate[i] = 0.f;
}

__global__ __launch_bounds__(kBlock) void prelod_accumulate(const gell_cell* __restrict__ cells, int32_t n,
                                                            int32_t N, float dx, UptakeSet u,
                                                            int32_t* __restrict__ list,
                                                            const int32_t* __restrict__ np) {
    n = count_of(n, np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t V = (int64_t)N * N * N;
    bool take = false;
    int32_t v = -1, phase = 0;
    float Vc = 0.f;
    if (i < n) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(cells + i);
        v = (int32_t)w[W_MESH];
        take = (w[W_SIGN] & 0xFFu) != 0 && v >= 0 && v < V;
        Vc = __uint_as_float(w[W_V]);
        phase = (int32_t)w[W_PHASE];
    }
    int32_t code;
    uptake_accumulate_wave(take, v, Vc, phase, N, dx, u, code);  // every lane of the wave
    if (i < n) list[i] = code;
}

// rate[v] = the voxel's sum, rounded to float once, for the voxels on the
// list (a voxel listed several times is written with the same value).
__global__ __launch_bounds__(kBlock) void prelod_finalize(const int32_t* __restrict__ list, int32_t n,
                                                          const long long* __restrict__ acc, float* __restrict__ rate,
                                                          int32_t N, const int32_t* __restrict__ np,
                                                          int32_t* __restrict__ map_np) {
    n = count_of(n, np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (map_np && i == 0) *map_np = n;  // the next clear of this map reads that many list entries
    if (i >= n) return;
    const int32_t c = list[i];
    if (c >= 0) rate[lin_of_code((uint32_t)c, N)] = from_fixed(acc[c]);
}

// Accumulators of the voxels on the list back to zero, every substrate of
// the set (after the finalize has read them).
__global__ __launch_bounds__(kBlock) void prelod_zero_acc(const int32_t* __restrict__ list, int32_t m, UptakeSet u,
                                                          const int32_t* __restrict__ mp) {
    m = count_of(m, mp);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t c = list[i];
    if (c >= 0)
        for (int s = 0; s < u.S; ++s) u.acc[s][c] = 0;
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

// Zero substrate s's previous map: the whole map after it was handed out,
// else along the x-line spans when they are current, else the voxels of the
// list that built it.
void clear_map(gell_ctx* ctx, int s) {
    hipStream_t st = ctx->stream;
    // with device-resident counts the list clear runs over the device count
    const int32_t m = ctx->devn ? ctx->P.capacity : ctx->rate_list_n[s];
    const int64_t lines = (int64_t)ctx->N * ctx->N;
    if (ctx->rate_handed_out[s])
        hipLaunchKernelGGL(prelod_clear_all, dim3((unsigned)((ctx->V + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                           ctx->rate[s], ctx->V);
    else if (m > 0 && ctx->xspan_valid[s])
        hipLaunchKernelGGL(prelod_clear_spans, dim3((unsigned)((lines + kBlock / 64 - 1) / (kBlock / 64))),
                           dim3(kBlock), 0, st, ctx->xspan[s], ctx->rate[s], ctx->N);
    else if (m > 0)
        hipLaunchKernelGGL(prelod_clear, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, st, ctx->rate_list_ptr[s],
                           ctx->rate_list_n[s], ctx->rate[s], ctx->N, dev_count(ctx, DN_MAP + s));
    ctx->rate_list_n[s] = 0;
    ctx->rate_handed_out[s] = false;
}

UptakeSet uptake_set(const gell_ctx* ctx, int first, int count) {
    UptakeSet u{};
    u.S = count;
    for (int q = 0; q < count; ++q) {
        u.acc[q] = ctx->acc[first + q];
        u.uptake[q] = ctx->P.sub[first + q].uptake;
        u.dead_factor[q] = ctx->P.sub[first + q].dead_factor;
    }
    return u;
}

}  // namespace

hipError_t prelod_launch(gell_ctx* ctx, int s, const gell_cell* cells, int32_t n) {
    hipStream_t st = ctx->stream;
    clear_map(ctx, s);
    ctx->rate_list_ptr[s] = ctx->rate_list[s];
    const int32_t gn = grid_count(ctx, n);
    if (gn > 0) {
        const unsigned blocks = (unsigned)((gn + kBlock - 1) / kBlock);
        const UptakeSet u = uptake_set(ctx, s, 1);
        const int32_t* np = dev_count(ctx);
        hipLaunchKernelGGL(prelod_accumulate, dim3(blocks), dim3(kBlock), 0, st, cells, n, ctx->N, ctx->P.dx, u,
                           ctx->rate_list[s], np);
        hipLaunchKernelGGL(prelod_finalize, dim3(blocks), dim3(kBlock), 0, st, ctx->rate_list[s], n, ctx->acc[s],
                           ctx->rate[s], ctx->N, np, dev_count_out(ctx, DN_MAP + s));
        hipLaunchKernelGGL(prelod_zero_acc, dim3(blocks), dim3(kBlock), 0, st, ctx->rate_list[s], n, u, np);
        ctx->rate_list_n[s] = n;
    }
    ctx->xspan_valid[s] = false;  // rebuilt by the first substep that reads the spans
    return hipGetLastError();
}

UptakeSet prelod_uptake_set(const gell_ctx* ctx) { return uptake_set(ctx, 0, ctx->P.n_substrates); }

int32_t* prelod_fused_list(gell_ctx* ctx) {
    // the buffer no substrate's current map refers to (at most one does: a
    // fused step moves every substrate onto the same list)
    bool used0 = false;
    for (int s = 0; s < ctx->P.n_substrates; ++s) used0 |= ctx->rate_list_ptr[s] == ctx->mlist[0];
    return used0 ? ctx->mlist[1] : ctx->mlist[0];
}

hipError_t prelod_fused_finalize(gell_ctx* ctx, int s, const int32_t* list, int32_t n) {
    clear_map(ctx, s);
    const int32_t gn = grid_count(ctx, n);
    if (gn > 0)
        hipLaunchKernelGGL(prelod_finalize, dim3((gn + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, list, n,
                           ctx->acc[s], ctx->rate[s], ctx->N, dev_count(ctx), dev_count_out(ctx, DN_MAP + s));
    ctx->rate_list_ptr[s] = list;
    ctx->rate_list_n[s] = n;
    ctx->xspan_valid[s] = false;
    return hipGetLastError();
}

hipError_t prelod_fused_zero(gell_ctx* ctx, const int32_t* list, int32_t n) {
    const int32_t gn = grid_count(ctx, n);
    if (gn <= 0) return hipSuccess;
    hipLaunchKernelGGL(prelod_zero_acc, dim3((gn + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, list, n,
                       prelod_uptake_set(ctx), dev_count(ctx));
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
