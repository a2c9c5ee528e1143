// fm.hip -- the mechanics step on gfx950 (MechanicsMesh_struct::FM_update,
// FMkernel.cuh:181-202).
//
// Reference: key kernel -> thrust::sort_by_key of the 72-byte records ->
// atomic min/max cell ranges (after a full reset of all N^3 ranges) -> force
// kernel -> movement kernel. Here:
//   fm_keys      Morton key per slot (biased by +1 so that the outside
//                sentinel -1 becomes 0 and still sorts first); clears the
//                ranges the previous call wrote (O(n), not O(N^3)).
//   radix_*      stable LSD radix sort of (key, slot) pairs: per pass a tile
//                histogram, a per-digit column scan, and a scatter whose
//                stable in-tile rank comes from wave64 ballots.
//   fm_gather    one 72-byte gather of the records into key order; applies the
//                key kernel's record updates (sign, mesh_idx), builds the
//                per-voxel live ranges with atomic min/max (order-independent
//                result) and the compact (x, y, z, r) neighbour array.
//                The ranges are indexed by Morton code, not by the linear
//                voxel index: cells arrive in Morton order, so both the
//                atomics and the force kernel's 27-voxel stencil lookups fall
//                into a few compact bricks of the array. gell_voxel_ranges
//                converts them to the reference's linear layout on request.
//   fm_force_move  forcekernel + movementkernel fused: reads the sorted
//                record once, visits the 27-voxel stencil in the reference's
//                order (x outer, z inner, ascending slot) through the compact
//                array, moves the cell and writes the record back in place.
#include "device_common.hpp"
#include "internal.hpp"

namespace gell {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Keys
// ---------------------------------------------------------------------------
// 32-bit occupancy word holding Morton code v: word v / 32 of the plain
// bitmap, or the matching half of the 16-byte (bits, bits, rank, -) record of
// the rank-compacted list. Non-decreasing in v either way.
__device__ __forceinline__ uint32_t occ_word32(uint32_t v, bool ranked) {
    return ranked ? (((v >> 6) << 2) | ((v >> 5) & 1u)) : (v >> 5);
}

__global__ __launch_bounds__(kBlock) void fm_keys(const gell_cell* __restrict__ cells, int32_t n, int32_t N, float dx,
                                                  uint32_t* __restrict__ keys, const int32_t* __restrict__ clear_list,
                                                  int32_t clear_n, int2* __restrict__ se, int32_t s_empty,
                                                  uint32_t* __restrict__ occ, int ranked,
                                                  const int32_t* __restrict__ np, const int32_t* __restrict__ clear_np) {
    n = count_of(n, np);
    clear_n = count_of(clear_n, clear_np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < clear_n) {
        // what the previous gather wrote: its ranges (plain list only) and its
        // occupancy words (the whole word: each voxel set in it is on the list)
        const int32_t v = clear_list[i];
        if (v >= 0) {
            if (!ranked) se[v] = make_int2(s_empty, -1);
            occ[occ_word32((uint32_t)v, ranked)] = 0u;
        }
    }
    if (i < n) {
        const float* pos = reinterpret_cast<const float*>(cells + i) + W_PX;
        const int32_t a = trunc_index(pos[0] / dx), b = trunc_index(pos[1] / dx), c = trunc_index(pos[2] / dx);
        keys[i] = inside(a, b, c, N) ? morton3(a, b, c) + 1u : 0u;
    }
}

// ---------------------------------------------------------------------------
// Block-wide exclusive scan of one int per thread (256 threads = 4 waves).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_inclusive_scan(int x) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ int block_exclusive_scan(int x, int* lds4, int& total) {
    const int lane = lane_id(), w = threadIdx.x / 64;
    const int inc = wave_inclusive_scan(x);
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) {
        const int s = lds4[q];
        if (q < w) off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + inc - x;
}

// Lanes of the wave whose digit equals this lane's (BITS ballots).
template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
    uint64_t peers = valid;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

// ---------------------------------------------------------------------------
// Radix sort passes
// ---------------------------------------------------------------------------
template <int BITS>
__global__ __launch_bounds__(kSortThreads) void radix_hist(const uint32_t* __restrict__ keys, int32_t n, int shift,
                                                           int32_t* __restrict__ hist, int nblocks,
                                                           const int32_t* __restrict__ np) {
    n = count_of(n, np);
    constexpr int BINS = 1 << BITS;
    __shared__ int h[BINS];
    for (int b = threadIdx.x; b < BINS; b += kSortThreads) h[b] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    const int lane = lane_id();
#pragma unroll 4
    for (int it = 0; it < kSortItems; ++it) {
        const int64_t idx = base + (int64_t)it * kSortThreads + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t d = valid ? (keys[idx] >> shift) & (BINS - 1) : 0u;
        const uint64_t vmask = __ballot(valid);
        const uint64_t peers = match_digit<BITS>(d, vmask);
        if (valid && lane == 63 - __clzll(peers)) atomicAdd(&h[d], __popcll(peers));
    }
    __syncthreads();
    for (int b = threadIdx.x; b < BINS; b += kSortThreads) hist[(int64_t)b * nblocks + blockIdx.x] = h[b];
}

// One workgroup per digit: exclusive scan of that digit's per-tile counts.
__global__ __launch_bounds__(kSortThreads) void radix_colscan(int32_t* __restrict__ hist, int nblocks,
                                                              int32_t* __restrict__ tot) {
    __shared__ int lds4[kSortThreads / 64];
    int32_t* col = hist + (int64_t)blockIdx.x * nblocks;
    int carry = 0;
    for (int base = 0; base < nblocks; base += kSortThreads * 4) {
        int v[4], s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = base + threadIdx.x * 4 + q;
            v[q] = idx < nblocks ? col[idx] : 0;
            s += v[q];
        }
        int total;
        int ex = block_exclusive_scan(s, lds4, total) + carry;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int idx = base + threadIdx.x * 4 + q;
            if (idx < nblocks) col[idx] = ex;
            ex += v[q];
        }
        carry += total;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

template <int BITS>
__global__ __launch_bounds__(kSortThreads) void radix_scatter(const uint32_t* __restrict__ kin,
                                                              const int32_t* __restrict__ vin,
                                                              uint32_t* __restrict__ kout, int32_t* __restrict__ vout,
                                                              int32_t* __restrict__ signed_out, int32_t n, int shift,
                                                              const int32_t* __restrict__ hist,
                                                              const int32_t* __restrict__ tot, int nblocks,
                                                              const int32_t* __restrict__ np) {
    n = count_of(n, np);
    constexpr int BINS = 1 << BITS;
    constexpr int PER = (BINS + kSortThreads - 1) / kSortThreads;
    constexpr int NW = kSortThreads / 64;
    __shared__ int cnt[NW][BINS];
    __shared__ int lbase[BINS];
    __shared__ int gbase[BINS];
    __shared__ uint32_t sk[kSortTile];
    __shared__ int32_t sv[kSortTile];
    __shared__ int lds4[NW];

    const int tid = threadIdx.x, lane = lane_id(), w = tid / 64;
    const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
    const int tile_n = (int)min<int64_t>(kSortTile, (int64_t)n - tile0);

    // Global base of each digit for this tile: digits below + earlier tiles.
    {
        int v[PER], s = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int b = tid * PER + q;
            v[q] = b < BINS ? tot[b] : 0;
            s += v[q];
        }
        int total;
        int ex = block_exclusive_scan(s, lds4, total);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int b = tid * PER + q;
            if (b < BINS) gbase[b] = ex + hist[(int64_t)b * nblocks + blockIdx.x];
            ex += v[q];
        }
    }
    for (int b = tid; b < BINS * NW; b += kSortThreads) (&cnt[0][0])[b] = 0;
    __syncthreads();

    // Stable rank inside each wave's contiguous sub-tile.
    uint32_t k[kSortItems];
    int32_t val[kSortItems];
    int rk[kSortItems];
    const int64_t wbase = tile0 + (int64_t)w * 64 * kSortItems;
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int64_t idx = wbase + it * 64 + lane;
        const bool valid = idx < n;
        k[it] = valid ? kin[idx] : 0u;
        val[it] = valid ? (vin ? vin[idx] : (int32_t)idx) : 0;
        const uint32_t d = (k[it] >> shift) & (BINS - 1);
        const uint64_t vmask = __ballot(valid);
        const uint64_t peers = match_digit<BITS>(d, vmask);
        const int before = valid ? cnt[w][d] : 0;
        rk[it] = before + __popcll(peers & lanemask_lt());
        if (valid && lane == 63 - __clzll(peers)) cnt[w][d] = before + __popcll(peers);
    }
    __syncthreads();

    // Per digit: offsets of the waves, and the tile-local digit base.
    {
        int v[PER], s = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int b = tid * PER + q;
            int acc = 0;
            if (b < BINS) {
#pragma unroll
                for (int ww = 0; ww < NW; ++ww) {
                    const int c = cnt[ww][b];
                    cnt[ww][b] = acc;
                    acc += c;
                }
            }
            v[q] = acc;
            s += acc;
        }
        int total;
        int ex = block_exclusive_scan(s, lds4, total);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int b = tid * PER + q;
            if (b < BINS) lbase[b] = ex;
            ex += v[q];
        }
    }
    __syncthreads();

#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int64_t idx = wbase + it * 64 + lane;
        if (idx < n) {
            const uint32_t d = (k[it] >> shift) & (BINS - 1);
            const int pos = lbase[d] + cnt[w][d] + rk[it];
            sk[pos] = k[it];
            sv[pos] = val[it];
        }
    }
    __syncthreads();

    for (int q = tid; q < tile_n; q += kSortThreads) {
        const uint32_t kk = sk[q];
        const uint32_t d = (kk >> shift) & (BINS - 1);
        const int64_t g = gbase[d] + (q - lbase[d]);
        kout[g] = kk;
        vout[g] = sv[q];
        if (signed_out) signed_out[g] = (int32_t)kk - 1;
    }
}

// Small counts: the whole sort in one workgroup and one launch. The tiled
// passes above cost three launches each, which at a few thousand keys is
// most of the sort's time. Same stable LSD order: within a wave, ranks come
// from the digit ballots in index order; wave w holds keys [w kSmallPer,
// (w + 1) kSmallPer); per digit the waves are offset in order. Between passes
// the keys and values go through the two global ping-pong buffers (the
// workgroup barrier makes one pass's writes visible to the next).
constexpr int kSmallThreads = 1024, kSmallItems = 16, kSmallWaves = kSmallThreads / 64;
constexpr int kSmallPer = 64 * kSmallItems;
constexpr int kSmallMax = kSmallThreads * kSmallItems;
// Digits of up to 11 bits: per-wave counts of 2048 bins take 128 KB of LDS,
// and the keys of grids up to 128^3 (22 bits) sort in two passes.
constexpr int kSmallBits = 11, kSmallBins = 1 << kSmallBits;

__device__ __forceinline__ uint64_t match_digit_rt(uint32_t d, uint64_t valid, int bits) {
    uint64_t peers = valid;
    for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
    }
    return peers;
}

__global__ __launch_bounds__(kSmallThreads) void radix_sort_small(uint32_t* __restrict__ k0, int32_t* __restrict__ v0,
                                                                  uint32_t* __restrict__ k1, int32_t* __restrict__ v1,
                                                                  int32_t* __restrict__ order_out,
                                                                  int32_t* __restrict__ signed_out, int32_t n,
                                                                  SortPlan plan, const int32_t* __restrict__ np) {
    n = count_of(n, np);
    __shared__ int cnt[kSmallWaves][kSmallBins];
    __shared__ int dbase[kSmallBins];
    __shared__ int lds16[kSmallWaves];
    const int tid = threadIdx.x, lane = lane_id(), w = tid / 64;
    uint32_t* kin = k0;
    uint32_t* kout = k1;
    int32_t* vin = nullptr;
    int32_t* vbuf[2] = {v0, v1};
    int cur = 0;
    for (int q = 0; q < plan.passes; ++q) {
        const int bits = plan.bits[q], shift = plan.shift[q], bins = 1 << bits;
        const bool last = q == plan.passes - 1;
        int32_t* vout = last ? order_out : vbuf[1 - cur];
        for (int b = tid; b < kSmallWaves * kSmallBins; b += kSmallThreads) (&cnt[0][0])[b] = 0;
        __syncthreads();
        uint32_t k[kSmallItems];
        int32_t val[kSmallItems];
        int rk[kSmallItems];
#pragma unroll
        for (int it = 0; it < kSmallItems; ++it) {
            const int idx = w * kSmallPer + it * 64 + lane;
            const bool valid = idx < n;
            // agent-scope loads read L2, where the previous pass's stores went
            // (the vector L1 may still hold lines of this buffer from two passes ago)
            k[it] = valid ? __hip_atomic_load(&kin[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
            val[it] = valid ? (vin ? __hip_atomic_load(&vin[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : idx)
                            : 0;
            const uint32_t d = (k[it] >> shift) & (uint32_t)(bins - 1);
            const uint64_t peers = match_digit_rt(d, __ballot(valid), bits);
            const int before = valid ? cnt[w][d] : 0;
            rk[it] = before + __popcll(peers & lanemask_lt());
            if (valid && lane == 63 - __clzll(peers)) cnt[w][d] = before + __popcll(peers);
        }
        __syncthreads();
        // per digit: offsets of the waves (in place) and the digit's total;
        // thread t owns digits 2t and 2t + 1
        int acc[2] = {0, 0};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int dg = 2 * tid + h;
            if (dg < bins) {
#pragma unroll 4
                for (int ww = 0; ww < kSmallWaves; ++ww) {
                    const int c = cnt[ww][dg];
                    cnt[ww][dg] = acc[h];
                    acc[h] += c;
                }
            }
        }
        // exclusive scan of the totals over digits
        const int pair = acc[0] + acc[1];
        const int inc = wave_inclusive_scan(pair);
        if (lane == 63) lds16[w] = inc;
        __syncthreads();
        int off = 0;
        for (int ww = 0; ww < w; ++ww) off += lds16[ww];
        if (2 * tid < bins) dbase[2 * tid] = off + inc - pair;
        if (2 * tid + 1 < bins) dbase[2 * tid + 1] = off + inc - pair + acc[0];
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kSmallItems; ++it) {
            const int idx = w * kSmallPer + it * 64 + lane;
            if (idx < n) {
                const uint32_t d = (k[it] >> shift) & (uint32_t)(bins - 1);
                const int pos = dbase[d] + cnt[w][d] + rk[it];
                kout[pos] = k[it];
                vout[pos] = val[it];
                if (last) signed_out[pos] = (int32_t)k[it] - 1;
            }
        }
        __syncthreads();
        uint32_t* t = kin;
        kin = kout;
        kout = t;
        vin = vout;
        cur = 1 - cur;
    }
}

template <int BITS>
hipError_t sort_pass(gell_ctx* ctx, const uint32_t* kin, const int32_t* vin, uint32_t* kout, int32_t* vout,
                     int32_t* signed_out, int32_t n, int shift) {
    const int nblocks = (int)((grid_count(ctx, n) + kSortTile - 1) / kSortTile);
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL((radix_hist<BITS>), dim3(nblocks), dim3(kSortThreads), 0, st, kin, n, shift, ctx->hist,
                       nblocks, dev_count(ctx));
    hipLaunchKernelGGL(radix_colscan, dim3(1 << BITS), dim3(kSortThreads), 0, st, ctx->hist, nblocks, ctx->hist_tot);
    hipLaunchKernelGGL((radix_scatter<BITS>), dim3(nblocks), dim3(kSortThreads), 0, st, kin, vin, kout, vout,
                       signed_out, n, shift, ctx->hist, ctx->hist_tot, nblocks, dev_count(ctx));
    return hipGetLastError();
}

typedef hipError_t (*PassFn)(gell_ctx*, const uint32_t*, const int32_t*, uint32_t*, int32_t*, int32_t*, int32_t, int);
const PassFn kPass[10] = {nullptr,         sort_pass<1>, sort_pass<2>, sort_pass<3>, sort_pass<4>,
                          sort_pass<5>,    sort_pass<6>, sort_pass<7>, sort_pass<8>, sort_pass<9>};

// ---------------------------------------------------------------------------
// Gather + cell list
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void fm_gather(const gell_cell* __restrict__ cells,
                                                    const int32_t* __restrict__ order, int32_t n, int32_t N, float dx,
                                                    gell_cell* __restrict__ sorted, float4* __restrict__ nb,
                                                    int2* __restrict__ se, int32_t* __restrict__ se_list,
                                                    const int32_t* __restrict__ skeys, int2* __restrict__ blk,
                                                    int32_t* __restrict__ n_out, uint32_t* __restrict__ occ,
                                                    int ranked, const int32_t* __restrict__ np,
                                                    int32_t* __restrict__ list_np) {
    n = count_of(n, np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (list_np && i == 0) *list_np = n;  // the next call's keys kernel clears this many list entries
    // Voxel (Morton code; -1 outside, INT32_MAX past n: non-decreasing along
    // the wave), 32-bit occupancy word and bit, and the slot if live. With
    // `ranked` (rank-compacted list) se is null: no ranges are kept here.
    int32_t kv = INT32_MAX, ow = INT32_MAX;
    uint32_t obit = 0u;
    int32_t lmin = INT32_MAX, lmax = -1;
    if (i < n) {
        if (blk) {
            // Boundaries of the aligned 8^3 Morton blocks in sorted order (keys
            // are sorted, so each boundary has exactly one writer); the outside
            // cells (key -1) sort first and their count goes to *n_out.
            const auto block_of = [&](int64_t q) { return skeys[q] >= 0 ? skeys[q] >> 9 : -1; };
            const int32_t b = block_of(i);
            const int32_t bp = i > 0 ? block_of(i - 1) : -2, bn = i + 1 < n ? block_of(i + 1) : -2;
            if (b >= 0) {
                if (bp != b) blk[b].x = (int32_t)i;
                if (bn != b) blk[b].y = (int32_t)i;
            } else if (bn != b) {
                *n_out = (int32_t)i + 1;
            }
        }
        CellRegs c = load_cell(cells, order[i]);
        const float px = cf(c, W_PX), py = cf(c, W_PY), pz = cf(c, W_PZ);
        const int32_t a = trunc_index(px / dx), b = trunc_index(py / dx), k = trunc_index(pz / dx);
        // updateCellMeshIdx's record writes (FMkernel.cuh:26-35).
        const bool in = inside(a, b, k, N);
        if (!in)
            cset_sign(c, false);
        else
            c.w[W_MESH] = (uint32_t)lin(a, b, k, N);
        store_cell(sorted, i, c);
        const bool live = csign(c);
        const uint32_t code = in ? morton3(a, b, k) : 0u;
        const int32_t v = live ? (int32_t)code : -1;
        se_list[i] = v;
        nb[i] = make_float4(px, py, pz, live ? cf(c, W_R) : __uint_as_float(kDeadTag));
        kv = in ? (int32_t)code : -1;
        ow = in ? (int32_t)occ_word32(code, ranked) : -1;
        obit = live ? 1u << (code & 31u) : 0u;
        lmin = live ? (int32_t)i : INT32_MAX;
        lmax = live ? (int32_t)i : -1;
    }
    // Slots are in key order, so the lanes of one voxel, and those of one
    // occupancy word, are adjacent runs. Segmented scans give the first lane
    // of each run the run's smallest live slot and OR of bits, and the last
    // lane its largest live slot; only those lanes issue atomics (runs cut by
    // a wave boundary combine through the atomics).
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int d = 1; d < 64; d *= 2) {
        const uint32_t b2 = __shfl_down(obit, d);
        const int32_t w2 = __shfl_down(ow, d), k2 = __shfl_down(kv, d), m2 = __shfl_down(lmin, d);
        const int32_t ku = __shfl_up(kv, d), mu = __shfl_up(lmax, d);
        if (lane + d < 64) {
            if (w2 == ow) obit |= b2;
            if (k2 == kv) lmin = m2 < lmin ? m2 : lmin;
        }
        if (lane >= d && ku == kv) lmax = mu > lmax ? mu : lmax;
    }
    const int32_t wprev = __shfl_up(ow, 1), kprev = __shfl_up(kv, 1), knext = __shfl_down(kv, 1);
    // Mech_Mesh_se_Get, FMkernel.cuh:45-53: min/max of the live slots.
    if (se && lmin != INT32_MAX && (lane == 0 || kprev != kv)) atomicMin(&se[kv].x, lmin);
    if (se && lmax >= 0 && (lane == 63 || knext != kv)) atomicMax(&se[kv].y, lmax);
    if (obit != 0u && (lane == 0 || wprev != ow)) atomicOr(&occ[ow], obit);
}

// ---------------------------------------------------------------------------
// Rank-compacted cell list (force variant 9): ranks of the occupied voxels
// in Morton order, then the first live slot of each.
// ---------------------------------------------------------------------------
constexpr int kRankPer = 16;                      // words per thread
constexpr int kRankWords = kBlock * kRankPer;     // words per workgroup (4096)

__device__ __forceinline__ int occr_count(const uint4* __restrict__ occr, int64_t w, int64_t words) {
    if (w >= words) return 0;
    const uint4 r = occr[w];
    return __popc(r.x) + __popc(r.y);
}

// Occupied voxels per workgroup's 4096 words (coalesced: thread t reads words
// t, t + 256, ...).
__global__ __launch_bounds__(kBlock) void rank_count(const uint4* __restrict__ occr, int64_t words,
                                                     int32_t* __restrict__ tot) {
    __shared__ int lds4[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kRankWords;
    int c = 0;
#pragma unroll 4
    for (int k = 0; k < kRankPer; ++k) c += occr_count(occr, base + (int64_t)k * kBlock + threadIdx.x, words);
    int total;
    block_exclusive_scan(c, lds4, total);
    if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

// Exclusive scan of the per-workgroup counts (one workgroup); the sentinel
// cs[R] = n after the last occupied voxel.
__global__ __launch_bounds__(kBlock) void rank_scan(int32_t* __restrict__ tot, int32_t m, int32_t* __restrict__ cs,
                                                    int32_t n, const int32_t* __restrict__ np) {
    n = count_of(n, np);
    __shared__ int lds4[kBlock / 64];
    int carry = 0;
    for (int base = 0; base < m; base += kBlock) {
        const int idx = base + (int)threadIdx.x;
        const int x = idx < m ? tot[idx] : 0;
        int total;
        const int ex = block_exclusive_scan(x, lds4, total);
        if (idx < m) tot[idx] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) cs[carry] = n;
}

// Each word's rank base into its record, in the same word order as
// rank_count; the cs entries of its voxels start at INT32_MAX (fm_cs takes
// the minimum of the live slots).
__global__ __launch_bounds__(kBlock) void rank_write(uint4* __restrict__ occr, int64_t words,
                                                     const int32_t* __restrict__ tot, int32_t* __restrict__ cs) {
    __shared__ int lds4[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kRankWords;
    int carry = tot[blockIdx.x];
    for (int k = 0; k < kRankPer; ++k) {
        const int64_t w = base + (int64_t)k * kBlock + threadIdx.x;
        const int c = occr_count(occr, w, words);
        int total;
        const int r = carry + block_exclusive_scan(c, lds4, total);
        if (w < words) {
            reinterpret_cast<uint32_t*>(occr + w)[2] = (uint32_t)r;
            for (int j = 0; j < c; ++j) cs[r + j] = INT32_MAX;
        }
        carry += total;
    }
}

// Rank of occupied voxel `code` from its word's record.
__device__ __forceinline__ int32_t voxel_rank(uint4 rec, uint32_t code) {
    const uint64_t bits = ((uint64_t)rec.y << 32) | rec.x;
    return (int32_t)rec.z + __popcll(bits & ((1ull << (code & 63u)) - 1ull));
}

// cs[rank(v)] = smallest live slot of voxel v: a segmented minimum over the
// wave's lanes (sorted slots, so a voxel's lanes are adjacent), one atomicMin
// per run (runs cut by a wave boundary combine through it).
__global__ __launch_bounds__(kBlock) void fm_cs(const int32_t* __restrict__ skeys, const float4* __restrict__ nb,
                                                int32_t n, const uint4* __restrict__ occr,
                                                int32_t* __restrict__ cs, const int32_t* __restrict__ np) {
    n = count_of(n, np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t kv = INT32_MAX, lmin = INT32_MAX;
    if (i < n) {
        kv = skeys[i];
        const bool live = __float_as_uint(nb[i].w) != kDeadTag;
        lmin = live ? (int32_t)i : INT32_MAX;
    }
    const int lane = (int)(threadIdx.x & 63);
#pragma unroll
    for (int d = 1; d < 64; d *= 2) {
        const int32_t k2 = __shfl_down(kv, d), m2 = __shfl_down(lmin, d);
        if (lane + d < 64 && k2 == kv) lmin = m2 < lmin ? m2 : lmin;
    }
    const int32_t kprev = __shfl_up(kv, 1);
    if (lmin != INT32_MAX && kv >= 0 && (lane == 0 || kprev != kv))
        atomicMin(&cs[voxel_rank(occr[(uint32_t)kv >> 6], (uint32_t)kv)], lmin);
}

// ---------------------------------------------------------------------------
// Force + movement
// ---------------------------------------------------------------------------
// One 16-byte neighbour entry. The raw buffer load cannot be split: a plain
// float4 load followed by the dead-tag test on .w was compiled into a 4-byte
// load, a wait, and a conditional 12-byte load -- two memory round trips per
// candidate.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t nb_rsrc(const float4* nb) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(nb), (short)0, 0x7FFFFFFF, 0x00020000);
}
__device__ __forceinline__ float4 load_nb(__amdgpu_buffer_rsrc_t rs, int32_t j) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, j * 16, 0, 0));
}

// thres of forcekernel (Tumor_CC_Max_Dist at first, then 1.25 (r1 + r2) of
// the last accepted pair; FMkernel.cuh:72, 96) with a squared-distance bound
// ss_max = thres^2 (1 + 2^-16), rounded: for ss >= ss_max the pair test's
// distance, sqrtf of ss + 0.0001 rounded to float, is at least
// thres (1 + 7.5e-6) before its final rounding and so >= thres after it. Far
// candidates are rejected on ss alone, without the square root; every
// candidate the bound lets through takes the exact test.
struct Cutoff {
    float thres, ss_max;
    __device__ Cutoff() { set(25.f); }
    __device__ void set(float t) {
        thres = t;
        ss_max = t * t * 1.0000153f;
    }
};

// One candidate pair of forcekernel (FMkernel.cuh:86-109); the cutoff carries
// over between calls.
__device__ __forceinline__ void force_pair(float4 o, float p1x, float p1y, float p1z, float r1, Cutoff& cut,
                                           float& fx, float& fy, float& fz) {
    if (__float_as_uint(o.w) == kDeadTag) return;
    const float ddx = p1x - o.x, ddy = p1y - o.y, ddz = p1z - o.z;
    const float ss = ddx * ddx + ddy * ddy + ddz * ddz;
    if (ss >= cut.ss_max) return;
    const float distance = sqrtf((float)((double)ss + 0.0001));
    if (distance >= cut.thres) return;
    const float rsum = o.w + r1;
    const float R = rsum > 0.01f ? rsum : 0.01f;
    cut.set((float)((double)R * 1.25));
    const float thres = cut.thres;
    if (distance >= thres) return;
    const float qr = 1.f - distance / R;
    const float qa = 1.f - distance / thres;
    const float coefr = (distance <= R) ? 10.f * (qr * qr / distance) : 0.f;
    const float coefa = (distance <= thres) ? 0.4f * (qa * qa / distance) : 0.f;
    const float cc = coefr - coefa;
    fx += ddx * cc;
    fy += ddy * cc;
    fz += ddz * cc;
}

// Grouped variant: all 27 ranges are loaded at once; then, for each of the
// three x-planes of the stencil, the first two candidates of its nine voxels
// are loaded together before the plane's pairs are evaluated in order. A
// voxel's third and later candidates are loaded as they are reached. Each
// wave thus waits on ~5 memory round trips instead of one per candidate.
__device__ __forceinline__ void cell_force_grouped(const float4* __restrict__ nb, const int2* __restrict__ se,
                                                   int32_t N, float dx, float p1x, float p1y, float p1z, float r1,
                                                   float& fx, float& fy, float& fz) {
    const int32_t px = trunc_index(p1x / dx), py = trunc_index(p1y / dx), pz = trunc_index(p1z / dx);
    Cutoff cut;  // thres = Tumor_CC_Max_Dist, carried across pairs (FMkernel.cuh:72)
    const __amdgpu_buffer_rsrc_t rs = nb_rsrc(nb);
    int2 rr[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int xi = px - 1 + q / 9, yi = py - 1 + (q / 3) % 3, zi = pz - 1 + q % 3;
        rr[q] = inside(xi, yi, zi, N) ? se[morton3(xi, yi, zi)] : make_int2(0, -1);
    }
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        float4 c0[9], c1[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int2 r = rr[g * 9 + t];
            c0[t] = r.y >= 0 ? load_nb(rs, r.x) : make_float4(0.f, 0.f, 0.f, 0.f);
            c1[t] = r.y > r.x ? load_nb(rs, r.x + 1) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int2 r = rr[g * 9 + t];
            if (r.y < 0) continue;
            force_pair(c0[t], p1x, p1y, p1z, r1, cut, fx, fy, fz);
            if (r.y > r.x) force_pair(c1[t], p1x, p1y, p1z, r1, cut, fx, fy, fz);
            if (r.y > r.x + 1) {
                float4 o = load_nb(rs, r.x + 2);
                for (int j = r.x + 2; j <= r.y; ++j) {
                    const float4 nxt = load_nb(rs, j < r.y ? j + 1 : j);  // one candidate ahead
                    force_pair(o, p1x, p1y, p1z, r1, cut, fx, fy, fz);
                    o = nxt;
                }
            }
        }
    }
}

// Force on live cell i (forcekernel, FMkernel.cuh:56-118). Candidates are
// loaded one ahead of their evaluation. With PREFETCH the 27 stencil ranges
// are loaded before the pair loop, so their latencies overlap, and the load
// ahead runs on across voxel boundaries; the visiting order (x outer, z
// inner, ascending slot) is the same either way.
template <bool PREFETCH>
__device__ __forceinline__ void cell_force(const float4* __restrict__ nb, const int2* __restrict__ se, int32_t N,
                                           float dx, float p1x, float p1y, float p1z, float r1, float& fx, float& fy,
                                           float& fz) {
    const int32_t px = trunc_index(p1x / dx), py = trunc_index(p1y / dx), pz = trunc_index(p1z / dx);
    Cutoff cut;  // thres = Tumor_CC_Max_Dist, carried across pairs (FMkernel.cuh:72)
    const __amdgpu_buffer_rsrc_t rs = nb_rsrc(nb);
    if (PREFETCH) {
        // last[q]: last candidate of stencil voxel q (-1: empty); nf[q]: first
        // candidate of the next non-empty voxel after q (-1: none). A voxel's
        // first candidate is where the previous one's load ahead pointed, so
        // the walk needs no other range data and the load one candidate ahead
        // runs on across voxel boundaries.
        int last[27], nf[27], next = -1;
        {
            int2 rr[27];
#pragma unroll
            for (int q = 0; q < 27; ++q) {
                const int xi = px - 1 + q / 9, yi = py - 1 + (q / 3) % 3, zi = pz - 1 + q % 3;
                rr[q] = inside(xi, yi, zi, N) ? se[morton3(xi, yi, zi)] : make_int2(0, -1);
            }
#pragma unroll
            for (int q = 26; q >= 0; --q) {
                nf[q] = next;
                last[q] = rr[q].y;
                if (rr[q].y >= 0) next = rr[q].x;
            }
        }
        if (next < 0) return;
        int j = next;
        float4 o = load_nb(rs, j);
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int e = last[q];
            if (e < 0) continue;
            for (; j <= e; ++j) {
                const int ni = j < e ? j + 1 : (nf[q] >= 0 ? nf[q] : j);
                const float4 nxt = load_nb(rs, ni);
                force_pair(o, p1x, p1y, p1z, r1, cut, fx, fy, fz);
                o = nxt;
            }
            j = nf[q];
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int xi = px - 1 + q / 9, yi = py - 1 + (q / 3) % 3, zi = pz - 1 + q % 3;
        if (!inside(xi, yi, zi, N)) continue;
        const int2 r = se[morton3(xi, yi, zi)];
        if (r.y < 0) continue;
        // FMkernel.cuh:89-109, one candidate's load ahead of its evaluation
        float4 o = load_nb(rs, r.x);
        for (int j = r.x; j <= r.y; ++j) {
            const float4 nxt = load_nb(rs, j < r.y ? j + 1 : j);
            force_pair(o, p1x, p1y, p1z, r1, cut, fx, fy, fz);
            o = nxt;
        }
    }
}

// Flattened variant: each lane keeps a 27-bit mask of its non-empty stencil
// voxels and walks its own candidate list (voxel by voxel in the reference's
// order, slots ascending); every iteration of the wave's loop evaluates one
// pair per active lane. The loop runs for the largest candidate count in the
// wave instead of for every (voxel, candidate) position any lane occupies.
// All lanes of the wave must call it (the loop exit is a wave vote); lanes
// with live == false have no candidates and load nothing.
__device__ __forceinline__ void cell_force_flat(const float4* __restrict__ nb, const int2* __restrict__ se,
                                                int32_t N, float dx, bool live, float p1x, float p1y, float p1z,
                                                float r1, float& fx, float& fy, float& fz) {
    const int32_t px = trunc_index(p1x / dx), py = trunc_index(p1y / dx), pz = trunc_index(p1z / dx);
    uint32_t mask = 0;
    if (live) {
        int2 rr[27];
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            const int xi = px - 1 + q / 9, yi = py - 1 + (q / 3) % 3, zi = pz - 1 + q % 3;
            rr[q] = inside(xi, yi, zi, N) ? se[morton3(xi, yi, zi)] : make_int2(0, -1);
        }
#pragma unroll
        for (int q = 0; q < 27; ++q) mask |= (rr[q].y >= 0 ? 1u : 0u) << q;
    }
    Cutoff cut;  // thres = Tumor_CC_Max_Dist, carried across pairs (FMkernel.cuh:72)
    const __amdgpu_buffer_rsrc_t rs = nb_rsrc(nb);
    int j = 0, je = -1;
    while (true) {
        if (j > je && mask != 0) {
            const int q = __builtin_ctz(mask);
            mask &= mask - 1;
            const int xi = px - 1 + q / 9, yi = py - 1 + (q / 3) % 3, zi = pz - 1 + q % 3;
            const int2 r = se[morton3(xi, yi, zi)];
            j = r.x;
            je = r.y;
        }
        const bool active = j <= je;
        if (!__any(active)) break;
        if (!active) continue;
        const float4 o = load_nb(rs, j);
        ++j;
        force_pair(o, p1x, p1y, p1z, r1, cut, fx, fy, fz);  // FMkernel.cuh:89-109
    }
}

// Morton codes of the stencil's coordinates, per axis in dilated form (the
// axis's bits spread to every third bit): code(x, y, z) = cx | cy | cz. The
// neighbours' codes come from dilated increments and decrements of the
// cell's own (FMkernel.cuh:22-36 for the code), not from three spreads each.
struct StencilCodes {
    uint32_t cx[3], cy[3], cz[3];
};

template <uint32_t M>
__device__ __forceinline__ uint32_t dilated_inc(uint32_t s) {
    return ((s | ~M) + (M & (0u - M))) & M;
}
template <uint32_t M>
__device__ __forceinline__ uint32_t dilated_dec(uint32_t s) {
    return (s - (M & (0u - M))) & M;
}

__device__ __forceinline__ StencilCodes stencil_codes(int32_t px, int32_t py, int32_t pz) {
    constexpr uint32_t MX = 0x09249249u, MY = MX << 1, MZ = MX << 2;
    const uint32_t sx = spread3((uint32_t)px), sy = spread3((uint32_t)py) << 1, sz = spread3((uint32_t)pz) << 2;
    StencilCodes c;
    c.cx[0] = dilated_dec<MX>(sx);
    c.cx[1] = sx;
    c.cx[2] = dilated_inc<MX>(sx);
    c.cy[0] = dilated_dec<MY>(sy);
    c.cy[1] = sy;
    c.cy[2] = dilated_inc<MY>(sy);
    c.cz[0] = dilated_dec<MZ>(sz);
    c.cz[1] = sz;
    c.cz[2] = dilated_inc<MZ>(sz);
    return c;
}

// Stencil voxels that hold a live cell, as a 27-bit mask (bit q = 9 xo +
// 3 yo + zo for voxel (px - 1 + xo, py - 1 + yo, pz - 1 + zo)), read from the
// occupancy bits: one bit per Morton code, so a 64-bit word (code / 64) is an
// aligned 4 x 4 x 4 block and the stencil touches at most 8 such words.
__device__ __forceinline__ uint32_t occ_mask(const uint32_t* __restrict__ occ, int32_t N, int32_t px, int32_t py,
                                             int32_t pz, const StencilCodes& c, uint32_t word_bytes = 8) {
    const __amdgpu_buffer_rsrc_t ro =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(occ), (short)0, 0x7FFFFFFF, 0x00020000);
    const bool vx[3] = {px > 0, true, px + 1 < N}, vy[3] = {py > 0, true, py + 1 < N}, vz[3] = {pz > 0, true, pz + 1 < N};
    // first and last stencil coordinate inside the grid, per axis
    const uint32_t xs = vx[0] ? c.cx[0] : c.cx[1], xe = vx[2] ? c.cx[2] : c.cx[1];
    const uint32_t ys = vy[0] ? c.cy[0] : c.cy[1], ye = vy[2] ? c.cy[2] : c.cy[1];
    const uint32_t zs = vz[0] ? c.cz[0] : c.cz[1], ze = vz[2] ? c.cz[2] : c.cz[1];
    uint64_t W[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const uint32_t wi = ((a ? xe : xs) | (b ? ye : ys) | (d ? ze : zs)) >> 6;
                const uint2 w =
                    __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(ro, (int)(wi * word_bytes), 0, 0));
                W[a][b][d] = ((uint64_t)w.y << 32) | w.x;
            }
    uint32_t mask = 0;
#pragma unroll
    for (int xo = 0; xo < 3; ++xo) {
        if (!vx[xo]) continue;
        const int a = (c.cx[xo] >> 6) != (xs >> 6);
#pragma unroll
        for (int yo = 0; yo < 3; ++yo) {
            if (!vy[yo]) continue;
            const int b = (c.cy[yo] >> 6) != (ys >> 6);
            const uint64_t w0 = a ? (b ? W[1][1][0] : W[1][0][0]) : (b ? W[0][1][0] : W[0][0][0]);
            const uint64_t w1 = a ? (b ? W[1][1][1] : W[1][0][1]) : (b ? W[0][1][1] : W[0][0][1]);
            const uint32_t cxy = c.cx[xo] | c.cy[yo];
#pragma unroll
            for (int zo = 0; zo < 3; ++zo) {
                if (!vz[zo]) continue;
                const uint64_t w = ((c.cz[zo] >> 6) != (zs >> 6)) ? w1 : w0;
                mask |= (uint32_t)((w >> ((cxy | c.cz[zo]) & 63u)) & 1u) << (9 * xo + 3 * yo + zo);
            }
        }
    }
    return mask;
}

// The same 27-bit mask as occ_mask, built a row at a time. In a word (an
// aligned 4^3 block) the bit of local voxel (x, y, z) is x0 + 2 y0 + 4 z0 +
// 8 x1 + 16 y1 + 32 z1, so shifting the word right by a row's x and y part
// (the low six bits of its dilated x and y codes, masked to 0x1B) leaves the
// row's four z voxels at bits 0, 4, 32 and 36. The two z-words of a row (the
// blocks of the first and the last stencil z) give a line of eight z voxels,
// from which the stencil's three are one bit-field extract. Stencil voxels
// outside the grid read a word of the stencil (the loads stay at in-grid
// coordinates) and are cleared at the end.
__device__ __forceinline__ uint32_t occ_mask_rows(const uint32_t* __restrict__ occ, int32_t N, int32_t px, int32_t py,
                                                  int32_t pz, const StencilCodes& c, uint32_t word_bytes) {
    const __amdgpu_buffer_rsrc_t ro =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(occ), (short)0, 0x7FFFFFFF, 0x00020000);
    const bool vx0 = px > 0, vx2 = px + 1 < N, vy0 = py > 0, vy2 = py + 1 < N, vz0 = pz > 0, vz2 = pz + 1 < N;
    const uint32_t xs = vx0 ? c.cx[0] : c.cx[1], xe = vx2 ? c.cx[2] : c.cx[1];
    const uint32_t ys = vy0 ? c.cy[0] : c.cy[1], ye = vy2 ? c.cy[2] : c.cy[1];
    const uint32_t zs = vz0 ? c.cz[0] : c.cz[1], ze = vz2 ? c.cz[2] : c.cz[1];
    uint64_t W[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const uint32_t wi = ((a ? xe : xs) | (b ? ye : ys) | (d ? ze : zs)) >> 6;
                const uint2 w =
                    __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(ro, (int)(wi * word_bytes), 0, 0));
                W[a][b][d] = ((uint64_t)w.y << 32) | w.x;
            }
    // z voxels 0..3 of a shifted word (bits 0, 4, 32, 36) as bits 0..3
    const auto nibble = [](uint64_t w) {
        const uint32_t t = ((uint32_t)w & 0x11u) | (((uint32_t)(w >> 32) & 0x11u) << 2);  // z0 z2 z1 z3 at 0 2 4 6
        return (t | (t >> 3)) & 0xFu;
    };
    // which of the two x (y) blocks stencil column 1 and 2 lie in (column 0
    // lies in the first whenever it is inside the grid)
    const bool a1 = (c.cx[1] >> 6) != (xs >> 6), a2 = (c.cx[2] >> 6) != (xs >> 6);
    const bool b1 = (c.cy[1] >> 6) != (ys >> 6), b2 = (c.cy[2] >> 6) != (ys >> 6);
    // bit of stencil z 0 in the z line below (the line starts one bit up, so
    // that a first stencil z below the grid starts at bit 0)
    const uint32_t start = vz0 ? (uint32_t)((pz - 1) & 3) + 1u : 0u;
    uint32_t rows = 0;
#pragma unroll
    for (int xo = 0; xo < 3; ++xo) {
        const bool a = xo == 0 ? false : (xo == 1 ? a1 : a2);
        const uint64_t X00 = a ? W[1][0][0] : W[0][0][0], X01 = a ? W[1][0][1] : W[0][0][1];
        const uint64_t X10 = a ? W[1][1][0] : W[0][1][0], X11 = a ? W[1][1][1] : W[0][1][1];
#pragma unroll
        for (int yo = 0; yo < 3; ++yo) {
            const bool b = yo == 0 ? false : (yo == 1 ? b1 : b2);
            const uint32_t base = (c.cx[xo] | c.cy[yo]) & 0x1Bu;
            const uint32_t line = (nibble((b ? X10 : X00) >> base) | (nibble((b ? X11 : X01) >> base) << 4)) << 1;
            rows |= __builtin_amdgcn_ubfe(line, start, 3u) << (9 * xo + 3 * yo);
        }
    }
    uint32_t valid = 0x7FFFFFFu;
    if (!vx0) valid &= ~0x1FFu;
    if (!vx2) valid &= ~0x7FC0000u;
    if (!vy0) valid &= ~0x1C0E07u;
    if (!vy2) valid &= ~0x70381C0u;
    if (!vz0) valid &= ~0x1249249u;
    if (!vz2) valid &= ~0x4924924u;
    return rows & valid;
}

// Occupancy-mask variant: only the ranges of the stencil voxels that hold a
// live cell are loaded (at ~0.1 cells per voxel, about 4 of 27). Each range is
// loaded while the previous voxel's candidates are walked, and the load one
// candidate ahead runs on into the next voxel. Visiting order as in the other
// variants: x outer, z inner, slots ascending.
__device__ __forceinline__ void cell_force_occ(const float4* __restrict__ nb, const int2* __restrict__ se,
                                               const uint32_t* __restrict__ occ, int32_t N, float dx, float p1x,
                                               float p1y, float p1z, float r1, float& fx, float& fy, float& fz) {
    const int32_t px = trunc_index(p1x / dx), py = trunc_index(p1y / dx), pz = trunc_index(p1z / dx);
    Cutoff cut;  // thres = Tumor_CC_Max_Dist, carried across pairs (FMkernel.cuh:72)
    const __amdgpu_buffer_rsrc_t rs = nb_rsrc(nb);
    const StencilCodes sc = stencil_codes(px, py, pz);
    uint32_t mask = occ_mask(occ, N, px, py, pz, sc);
    if (mask == 0) return;
    const auto pick = [](const uint32_t (&v)[3], uint32_t o) { return o == 0 ? v[0] : (o == 1 ? v[1] : v[2]); };
    const auto range_of = [&](uint32_t q) {
        return se[pick(sc.cx, q / 9) | pick(sc.cy, (q / 3) % 3) | pick(sc.cz, q % 3)];
    };
    int2 cur = range_of((uint32_t)__builtin_ctz(mask));
    mask &= mask - 1;
    int j = cur.x;
    float4 o = load_nb(rs, j);
    while (true) {
        const bool more = mask != 0;
        const int2 nxt = more ? range_of((uint32_t)__builtin_ctz(mask)) : make_int2(0, -1);
        mask &= mask - 1;
        const int e = cur.y;
        for (; j <= e; ++j) {
            // FMkernel.cuh:89-109, one candidate's load ahead of its evaluation
            const float4 ahead = load_nb(rs, j < e ? j + 1 : (more ? nxt.x : j));
            force_pair(o, p1x, p1y, p1z, r1, cut, fx, fy, fz);
            o = ahead;
        }
        if (!more) break;
        cur = nxt;
        j = cur.x;
    }
}

// The pair test of forcekernel (FMkernel.cuh:89-99) decided without the
// square root. distance = sqrtf(x), x = (float)(ss + 0.0001), is correctly
// rounded, so for a positive float t: distance < t  <=>  sqrt(x) < m, with m
// the midpoint of t and its float predecessor. m has at most 26 significant
// bits, so m^2 is exact in double and never a float: x < m^2 decides it with
// no tie. A NaN x passes, as it passes the reference's `distance >= thres`.
struct CutoffSq {
    float thres, ss_max;
    double m2;
    __device__ CutoffSq() { set(25.f); }  // Tumor_CC_Max_Dist (FMkernel.cuh:72)
    __device__ void set(float t) {
        thres = t;
        ss_max = t * t * 1.0000153f;
        const double m = ((double)__uint_as_float(__float_as_uint(t) - 1u) + (double)t) * 0.5;
        m2 = m * m;
    }
    __device__ bool below(float x) const { return !((double)x >= m2); }
};

// Acceptance half of force_pair: true when the pair contributes a force; the
// cutoff carries over exactly as in force_pair.
__device__ __forceinline__ bool pair_accepts(float4 o, float p1x, float p1y, float p1z, float r1, CutoffSq& cut) {
    if (__float_as_uint(o.w) == kDeadTag) return false;
    const float ddx = p1x - o.x, ddy = p1y - o.y, ddz = p1z - o.z;
    const float ss = ddx * ddx + ddy * ddy + ddz * ddz;
    if (ss >= cut.ss_max) return false;
    const float x = (float)((double)ss + 0.0001);
    if (!cut.below(x)) return false;  // distance >= thres
    const float rsum = o.w + r1;
    const float R = rsum > 0.01f ? rsum : 0.01f;
    cut.set((float)((double)R * 1.25));
    return cut.below(x);  // distance < 1.25 R
}

// Force of an accepted pair (FMkernel.cuh:94-109), i.e. what force_pair adds
// after its tests passed: the same operations in the same order, with
// thres = 1.25 R of this pair.
__device__ __forceinline__ float3 pair_force(float4 o, float p1x, float p1y, float p1z, float r1) {
    const float ddx = p1x - o.x, ddy = p1y - o.y, ddz = p1z - o.z;
    const float ss = ddx * ddx + ddy * ddy + ddz * ddz;
    const float distance = sqrtf((float)((double)ss + 0.0001));
    const float rsum = o.w + r1;
    const float R = rsum > 0.01f ? rsum : 0.01f;
    const float thres = (float)((double)R * 1.25);
    const float qr = 1.f - distance / R;
    const float qa = 1.f - distance / thres;
    const float coefr = (distance <= R) ? 10.f * (qr * qr / distance) : 0.f;
    const float coefa = (distance <= thres) ? 0.4f * (qa * qa / distance) : 0.f;
    const float cc = coefr - coefa;
    return make_float3(ddx * cc, ddy * cc, ddz * cc);
}

__device__ __forceinline__ float lane_f(float v, int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}
__device__ __forceinline__ uint32_t lane_u(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}

// The occupancy walk of variants 7-9 (forcekernel's stencil loop,
// FMkernel.cuh:74-112): the stencil's occupied voxels come from the
// occupancy words, visited in the reference's order (x outer, z inner, slots
// ascending). RANK (variant 9): a voxel's candidates are the slots
// [cs[rank], cs[rank + 1]) of the rank-compacted list -- its live cells in
// ascending order with, at most, dead cells after them (skipped by the dead
// tag, as the reference skips dead cells inside a range, FMkernel.cuh:86);
// otherwise (7, 8) its Morton-indexed range se[code] = [first, last].
//
// The whole wave calls it (live == false: no candidates). Each lane tests its
// candidates in order with pair_accepts, one per iteration of a loop that
// runs until every lane is done. A pair with zero displacement (the cell
// itself) adds +-0 * cc with cc finite, which leaves a force that started at
// +0 unchanged (a float sum that starts at +0 never becomes -0), so only its
// cutoff update counts. Every other accepted pair goes to a queue of the
// wave's accepted pairs (64 entries in LDS: slot and owner lane, plus a bit
// per entry in the owner's mask). When an iteration's pairs would overflow
// it (before they join) and at the end, the queue is drained: lane k
// computes entry k's force (pair_force), then every owner adds its entries'
// forces in queue order, which is its own acceptance order. The sums are
// therefore force_pair's, while the pair arithmetic runs once per 64
// accepted pairs of the wave instead of at every candidate position any of
// its lanes reaches.
constexpr int kBalCap = 448;  // stencil voxels of a wave whose ranges are looked up before the walk
template <bool RANK>
__device__ __forceinline__ void cell_force_queued(const float4* __restrict__ nb, const int2* __restrict__ se,
                                                  const uint32_t* __restrict__ occ, const uint4* __restrict__ occr,
                                                  const int32_t* __restrict__ cs, int32_t N, float dx, bool live,
                                                  float p1x, float p1y, float p1z, float r1, int2* wq, int2* rq,
                                                  float& fx, float& fy, float& fz) {
    const __amdgpu_buffer_rsrc_t rs = nb_rsrc(nb);
    const __amdgpu_buffer_rsrc_t ro =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint4*>(occr), (short)0, 0x7FFFFFFF, 0x00020000);
    const __amdgpu_buffer_rsrc_t rc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(cs), (short)0, 0x7FFFFFFF, 0x00020000);
    const int lane = (int)__lane_id();
    // the stencil's dilated coordinates as scalars: selected by value below
    // (an array picked by a runtime index would live in scratch)
    uint32_t cx0 = 0, cx1 = 0, cx2 = 0, cy0 = 0, cy1 = 0, cy2 = 0, cz0 = 0, cz1 = 0, cz2 = 0, mask = 0;
    if (live) {
        const int32_t px = trunc_index(p1x / dx), py = trunc_index(p1y / dx), pz = trunc_index(p1z / dx);
        const StencilCodes sc = stencil_codes(px, py, pz);
        mask = RANK ? occ_mask_rows(reinterpret_cast<const uint32_t*>(occr), N, px, py, pz, sc, 16)
                    : occ_mask_rows(occ, N, px, py, pz, sc, 8);
        cx0 = sc.cx[0], cx1 = sc.cx[1], cx2 = sc.cx[2];
        cy0 = sc.cy[0], cy1 = sc.cy[1], cy2 = sc.cy[2];
        cz0 = sc.cz[0], cz1 = sc.cz[1], cz2 = sc.cz[2];
    }
    // [first, end) candidate slots of stencil voxel q. Its word's record is
    // read again (a cache hit: the eight were just loaded) rather than picked
    // from the registers above, which a runtime index would move to scratch.
    const auto code_of = [](uint32_t q, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t y0, uint32_t y1,
                            uint32_t y2, uint32_t z0, uint32_t z1, uint32_t z2) {
        const uint32_t xo = (q * 57u) >> 9, r = q - 9u * xo, yo = (r * 11u) >> 5, zo = r - 3u * yo;
        return (xo == 0 ? x0 : 0u) | (xo == 1 ? x1 : 0u) | (xo == 2 ? x2 : 0u) | (yo == 0 ? y0 : 0u) |
               (yo == 1 ? y1 : 0u) | (yo == 2 ? y2 : 0u) | (zo == 0 ? z0 : 0u) | (zo == 1 ? z1 : 0u) |
               (zo == 2 ? z2 : 0u);
    };
    const auto range_at = [&](uint32_t code) {
        if (!RANK) {
            const int2 r = se[code];  // [first, last]
            return make_int2(r.x, r.y + 1);
        }
        const uint4 rec =
            __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ro, (int)((code >> 6) * 16u), 0, 0));
        const int32_t rk = voxel_rank(rec, code);
        const int32_t s = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rc, rk * 4, 0, 0);
        const int32_t e = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(rc, rk * 4 + 4, 0, 0);
        return make_int2(s, e);
    };
    // Ranges looked up before the walk (RANK with an LDS region rq, when the
    // wave's stencil voxels fit): every lane lists its occupied stencil
    // voxels (owner lane, q) in order, the wave's lanes then resolve the list
    // 64 entries at a time, and a lane's walk reads its ranges from rq. A lane
    // reaches its next voxel at a different iteration of the walk than its
    // neighbours, so looked up inside the walk the chain (word record, rank,
    // cell starts) runs in almost every iteration for some lane.
    bool listed = false;
    int t_next = 0, t_end = 0;
    if (RANK && rq) {
        const int v = __popc(mask);
        int incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        const int V = __shfl(incl, 63, 64);
        listed = V <= kBalCap;
        if (listed) {
            uint16_t* keys = reinterpret_cast<uint16_t*>(rq + kBalCap);
            t_next = incl - v;
            t_end = incl;
            int t = t_next;
            for (uint32_t mm = mask; mm != 0; mm &= mm - 1) keys[t++] = (uint16_t)(((uint32_t)lane << 5) | __builtin_ctz(mm));
            for (int k0 = 0; k0 < V; k0 += 64) {
                const int k = k0 + lane;
                const uint32_t key = k < V ? (uint32_t)keys[k] : ((uint32_t)lane << 5);
                const int ow = (int)(key >> 5);
                const uint32_t code =
                    code_of(key & 31u, lane_u(cx0, ow), lane_u(cx1, ow), lane_u(cx2, ow), lane_u(cy0, ow),
                            lane_u(cy1, ow), lane_u(cy2, ow), lane_u(cz0, ow), lane_u(cz1, ow), lane_u(cz2, ow));
                if (k < V) rq[k] = range_at(code);
            }
        }
    }
    // the range of this lane's next occupied stencil voxel ({0, 0}: none)
    const auto next_range = [&]() {
        int2 r = make_int2(0, 0);
        if (listed) {
            if (t_next < t_end) r = rq[t_next++];
        } else if (mask != 0) {
            r = range_at(code_of((uint32_t)__builtin_ctz(mask), cx0, cx1, cx2, cy0, cy1, cy2, cz0, cz1, cz2));
            mask &= mask - 1;
        }
        return r;
    };
    // current voxel's candidates [j, e); the next voxel's range nxt (nxt.y ==
    // 0: none); o = the candidate at j, loaded one iteration ahead
    int j = 0, e = 0;
    int2 nxt = make_int2(0, 0);
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if (mask != 0) {
        const int2 cur = next_range();
        j = cur.x;
        e = cur.y;
        nxt = next_range();
        o = load_nb(rs, j);
    }
    CutoffSq cut;
    uint64_t mine = 0;  // queue entries this lane owns
    int queued = 0;     // entries in the queue (the same in every lane)
    const auto drain = [&]() {
        // lane k: entry k's force
        const int2 ent = wq[lane < queued ? lane : 0];
        const int src = lane < queued ? ent.y : lane;
        const float qx = lane_f(p1x, src), qy = lane_f(p1y, src), qz = lane_f(p1z, src), qr = lane_f(r1, src);
        float3 f = make_float3(0.f, 0.f, 0.f);
        if (lane < queued) f = pair_force(load_nb(rs, ent.x), qx, qy, qz, qr);
        // owners add their entries' forces in queue order
        while (__any(mine != 0)) {
            const int k = mine != 0 ? (int)__builtin_ctzll(mine) : lane;
            const float gx = lane_f(f.x, k), gy = lane_f(f.y, k), gz = lane_f(f.z, k);
            if (mine != 0) {
                fx += gx;
                fy += gy;
                fz += gz;
                mine &= mine - 1;
            }
        }
        queued = 0;
    };
    // One candidate per lane: oc is the candidate at j, on receives the one
    // after it. The loop below alternates the two registers, so a load is
    // consumed one iteration after it was issued (a copy of it at the end of
    // each iteration would wait for it at once). False when every lane is done.
    const auto step = [&](float4& oc, float4& on) {
        const bool active = j < e;
        if (!__any(active)) return false;
        bool enq = false;
        const int slot = j;
        if (active) {
            // on the voxel's last candidate, the range of the voxel after the
            // next one first (its loads are then older than the candidate's)
            const bool last = j + 1 >= e;
            int2 nn = nxt;
            if (last) nn = next_range();
            // the next candidate: this voxel's next slot, else the next voxel's first
            on = load_nb(rs, !last ? j + 1 : (nxt.y > 0 ? nxt.x : j));
            // FMkernel.cuh:86-99
            if (pair_accepts(oc, p1x, p1y, p1z, r1, cut))
                enq = !(p1x - oc.x == 0.f && p1y - oc.y == 0.f && p1z - oc.z == 0.f);
            ++j;
            if (last) {
                j = nxt.x;
                e = nxt.y;
                nxt = nn;
            }
        }
        const uint64_t bal = __ballot(enq);
        if (queued + __popcll(bal) > 64) drain();  // older entries first: their forces come first
        if (enq) {
            const int pos = queued + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            wq[pos] = make_int2(slot, lane);
            mine |= 1ull << pos;
        }
        queued += __popcll(bal);
        return true;
    };
    float4 o2 = o;
    while (step(o, o2) && step(o2, o)) {
    }
    if (queued > 0) drain();
}

// movementkernel, FMkernel.cuh:130-158: AB2 step of a live cell; outside the
// domain the cell is flagged dead and keeps its position.
__device__ __forceinline__ void move_cell(CellRegs& c, float fx, float fy, float fz, int32_t N, float dx, float dt) {
    const float p1x = cf(c, W_PX), p1y = cf(c, W_PY), p1z = cf(c, W_PZ);
    const float h = dt / 2.f;
    const float ox = cf(c, W_OX), oy = cf(c, W_OY), oz = cf(c, W_OZ);
    const float nx = p1x + h * (3.f * fx - ox);
    const float ny = p1y + h * (3.f * fy - oy);
    const float nz = p1z + h * (3.f * fz - oz);
    const int32_t a = trunc_index(nx / dx), b = trunc_index(ny / dx), k = trunc_index(nz / dx);
    if (!inside(a, b, k, N)) {
        cset_sign(c, false);
        cset(c, W_FX, fx);
        cset(c, W_FY, fy);
        cset(c, W_FZ, fz);
    } else {
        cset(c, W_PX, nx);
        cset(c, W_PY, ny);
        cset(c, W_PZ, nz);
        cset(c, W_OX, fx);
        cset(c, W_OY, fy);
        cset(c, W_OZ, fz);
        cset(c, W_FX, 0.f);
        cset(c, W_FY, 0.f);
        cset(c, W_FZ, 0.f);
        c.w[W_MESH] = (uint32_t)lin(a, b, k, N);
    }
}

// VARIANT 0: stencil voxels in turn; 1: their ranges loaded up front;
// 2: flattened per-lane candidate walk; 3: grouped candidate loads;
// 6: non-empty voxels from the occupancy bitmap. All give the same bits.
template <int VARIANT>
__global__ __launch_bounds__(kBlock) void fm_force_move(const gell_cell* __restrict__ sorted,
                                                        const float4* __restrict__ nb, const int2* __restrict__ se,
                                                        gell_cell* __restrict__ cells, int32_t n, int32_t N, float dx,
                                                        float dt, const uint32_t* __restrict__ occ,
                                                        const int32_t* __restrict__ np) {
    n = count_of(n, np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in_range = i < n;
    // The cell's own (x, y, z, r) and liveness come from its 16-byte entry of
    // the compact array (same values as its record); the 72-byte record is
    // read after the pair loop, so it holds no registers during the walk.
    const float4 me = in_range ? load_nb(nb_rsrc(nb), (int32_t)i) : make_float4(0.f, 0.f, 0.f, __uint_as_float(kDeadTag));
    const bool live = __float_as_uint(me.w) != kDeadTag;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (VARIANT == 2) {
        // every lane of the wave takes part in the candidate loop's vote
        cell_force_flat(nb, se, N, dx, live, live ? me.x : 0.f, live ? me.y : 0.f, live ? me.z : 0.f,
                        live ? me.w : 0.f, fx, fy, fz);
    } else if (VARIANT == 3) {
        if (live) cell_force_grouped(nb, se, N, dx, me.x, me.y, me.z, me.w, fx, fy, fz);
    } else if (VARIANT == 6) {
        if (live) cell_force_occ(nb, se, occ, N, dx, me.x, me.y, me.z, me.w, fx, fy, fz);
    } else if (live) {
        cell_force<VARIANT == 1>(nb, se, N, dx, me.x, me.y, me.z, me.w, fx, fy, fz);
    }
    if (!in_range) return;
    CellRegs c = load_cell(sorted, i);
    if (live) move_cell(c, fx, fy, fz, N, dx, dt);
    store_cell(cells, i, c);
}

// Variant 4: the grouped candidate walk of variant 3, with the cell's own
// (x, y, z, r) taken from the compact neighbour array. The block's records,
// contiguous in sorted order, are read as coalesced 16-byte chunks together
// with that first load and staged in LDS; after the force loop each thread
// applies the movement to its record there, and the block writes the records
// back as 16-byte chunks. Same bits as variants 0-3. Needs 16-byte aligned
// record arrays (the launcher falls back to variant 3 otherwise).
constexpr int kRecDw = (int)(sizeof(gell_cell) / 4);  // 18 dwords

__global__ __launch_bounds__(kBlock) void fm_force_move_staged(const gell_cell* __restrict__ sorted,
                                                               const float4* __restrict__ nb,
                                                               const int2* __restrict__ se,
                                                               gell_cell* __restrict__ cells, int32_t n, int32_t N,
                                                               float dx, float dt, const int32_t* __restrict__ np) {
    n = count_of(n, np);
    __shared__ uint32_t rec[kBlock * kRecDw];
    const int64_t b0 = (int64_t)blockIdx.x * kBlock;
    // records in this block (none in the blocks past a device-resident count)
    const int m = (int)max<int64_t>(0, min<int64_t>((int64_t)n - b0, kBlock));
    const int ndw = m * kRecDw;
    const int nq = ndw / 4;  // whole 16-byte chunks; an odd m leaves one 8-byte tail
    const uint4* src = reinterpret_cast<const uint4*>(sorted + b0);
    const int64_t i = b0 + threadIdx.x;
    const bool in_range = (int)threadIdx.x < m;
    const float4 me = in_range ? nb[i] : make_float4(0.f, 0.f, 0.f, __uint_as_float(kDeadTag));
    for (int q = threadIdx.x; q < nq; q += kBlock) {
        const uint4 v = src[q];
        rec[4 * q] = v.x;
        rec[4 * q + 1] = v.y;
        rec[4 * q + 2] = v.z;
        rec[4 * q + 3] = v.w;
    }
    if (threadIdx.x == 0 && (ndw & 3)) {
        const uint2 v = reinterpret_cast<const uint2*>(sorted + b0)[2 * nq];
        rec[4 * nq] = v.x;
        rec[4 * nq + 1] = v.y;
    }
    const bool live = in_range && __float_as_uint(me.w) != kDeadTag;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (live) cell_force_grouped(nb, se, N, dx, me.x, me.y, me.z, me.w, fx, fy, fz);
    __syncthreads();
    if (in_range) {
        CellRegs c;
#pragma unroll
        for (int w = 0; w < kRecDw; ++w) c.w[w] = rec[threadIdx.x * kRecDw + w];
        if (live) move_cell(c, fx, fy, fz, N, dx, dt);
#pragma unroll
        for (int w = 0; w < kRecDw; ++w) rec[threadIdx.x * kRecDw + w] = c.w[w];
    }
    __syncthreads();
    uint4* dst = reinterpret_cast<uint4*>(cells + b0);
    for (int q = threadIdx.x; q < nq; q += kBlock)
        dst[q] = make_uint4(rec[4 * q], rec[4 * q + 1], rec[4 * q + 2], rec[4 * q + 3]);
    if (threadIdx.x == 0 && (ndw & 3))
        reinterpret_cast<uint2*>(cells + b0)[2 * nq] = make_uint2(rec[4 * nq], rec[4 * nq + 1]);
}

// Workgroup -> tile order for the eight XCDs: the hardware hands workgroup i
// to XCD i mod 8, and each XCD has its own L2. Tile = XCD x's share in one
// contiguous run of the sorted cells (a contiguous stretch of the Morton
// curve), so the workgroups an XCD runs at the same time are spatial
// neighbours and share their stencils' ranges and candidates in that L2.
// A bijection of [0, G).
__device__ __forceinline__ int xcd_tile(int i, int G) {
    const int x = i & 7, q = i >> 3;
    const int per = G >> 3, rem = G & 7;
    return x * per + (x < rem ? x : rem) + q;
}

// Per-substrate uptake accumulation fused into the mechanics step (PreLOD's
// oxy_consumption_update, LODkernel.cuh:12-22, on the moved cells): each live
// cell inside the domain adds its Cell::O2_consume() rate of every substrate
// to the substrate's Morton-indexed int64 accumulator at its new mesh_idx, and
// its voxel's Morton code goes to `list` (-1 for none). Every lane of the
// wave calls it (lanes without a cell take nothing).
__device__ __forceinline__ void fused_uptake(bool in_range, const CellRegs& c, int64_t slot, int32_t N, float dx,
                                             const UptakeSet& u, int32_t* __restrict__ list) {
    const int64_t V = (int64_t)N * N * N;
    const int32_t v = in_range ? (int32_t)c.w[W_MESH] : -1;
    const bool take = in_range && csign(c) && v >= 0 && v < V;
    int32_t code;
    uptake_accumulate_wave(take, v, cf(c, W_V), (int32_t)c.w[W_PHASE], N, dx, u, code);
    if (in_range) list[slot] = code;
}

// Variants 7, 8 and 9: the occupancy walk of variant 6 (RANK: the walk of the
// rank-compacted list, 9) on XCD-ordered tiles. STAGED (7, 9): the tile's
// records are read and written as coalesced 16-byte chunks through LDS (as
// variant 4) instead of 8-byte per-lane accesses. With a list the moved
// cells' uptake rates are accumulated as well. Same force / movement bits as
// variants 0-6.
// One wave per workgroup: a tile's slot is freed as soon as its wave is done,
// without waiting at barriers for the tile's other waves (against 256-thread
// tiles, alternated on one machine: 10k cells at 64^3 16.3 -> 15.6 us, 1M
// cells at 256^3 53.2 -> 52.5 us, 900k cells at 100^3 / 25 um 286 -> 283 us).
constexpr int kForceTile = 64;
template <bool STAGED, bool RANK>
__global__ __launch_bounds__(kForceTile) void fm_force_move_xcd(const gell_cell* __restrict__ sorted,
                                                            const float4* __restrict__ nb,
                                                            const int2* __restrict__ se,
                                                            gell_cell* __restrict__ cells, int32_t n, int32_t N,
                                                            float dx, float dt, const uint32_t* __restrict__ occ,
                                                            const uint4* __restrict__ occr,
                                                            const int32_t* __restrict__ cs, UptakeSet u,
                                                            int32_t* __restrict__ list,
                                                            const int32_t* __restrict__ np) {
    n = count_of(n, np);
    __shared__ __align__(16) uint32_t rec[STAGED ? kForceTile * kRecDw : 1];
    const int64_t b0 = (int64_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * kForceTile;
    const int64_t i = b0 + threadIdx.x;
    const bool in_range = i < n;
    const float4 me = in_range ? load_nb(nb_rsrc(nb), (int32_t)i) : make_float4(0.f, 0.f, 0.f, __uint_as_float(kDeadTag));
    const bool live = __float_as_uint(me.w) != kDeadTag;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    __shared__ int2 wq[kForceTile];  // each wave's queue of accepted pairs
    // RANK with staged records: the staging LDS holds each wave's listed
    // stencil ranges during the walk (kBalCap ranges and keys in its quarter)
    static_assert(kBalCap * (sizeof(int2) + sizeof(uint16_t)) <= 64 * kRecDw * sizeof(uint32_t), "rq fits");
    int2* rq = (RANK && STAGED) ? reinterpret_cast<int2*>(rec + (threadIdx.x & ~63u) * kRecDw) : nullptr;
    cell_force_queued<RANK>(nb, se, occ, occr, cs, N, dx, live, me.x, me.y, me.z, me.w, wq + (threadIdx.x & ~63u), rq,
                            fx, fy, fz);
    if (!STAGED) {
        CellRegs c{};
        if (in_range) {
            c = load_cell(sorted, i);
            if (live) move_cell(c, fx, fy, fz, N, dx, dt);
            store_cell(cells, i, c);
        }
        if (list) fused_uptake(in_range, c, i, N, dx, u, list);
        return;
    }
    // the tile's records, coalesced, into LDS (after the walk, so that the
    // walk's registers are free again, and with RANK after every wave's
    // walk, whose ranges were kept there); none past a device-resident count
    if (RANK) __syncthreads();
    const int m = (int)max<int64_t>(0, min<int64_t>((int64_t)n - b0, kForceTile));
    const int ndw = m * kRecDw;
    const int nq = ndw / 4;  // whole 16-byte chunks; an odd m leaves one 8-byte tail
    const uint4* src = reinterpret_cast<const uint4*>(sorted + b0);
    for (int q = threadIdx.x; q < nq; q += kForceTile) {
        const uint4 v = src[q];
        rec[4 * q] = v.x;
        rec[4 * q + 1] = v.y;
        rec[4 * q + 2] = v.z;
        rec[4 * q + 3] = v.w;
    }
    if (threadIdx.x == 0 && (ndw & 3)) {
        const uint2 v = reinterpret_cast<const uint2*>(sorted + b0)[2 * nq];
        rec[4 * nq] = v.x;
        rec[4 * nq + 1] = v.y;
    }
    __syncthreads();
    CellRegs c{};
    if (in_range) {
#pragma unroll
        for (int w = 0; w < kRecDw; ++w) c.w[w] = rec[threadIdx.x * kRecDw + w];
        if (live) move_cell(c, fx, fy, fz, N, dx, dt);
#pragma unroll
        for (int w = 0; w < kRecDw; ++w) rec[threadIdx.x * kRecDw + w] = c.w[w];
    }
    if (list) fused_uptake(in_range, c, i, N, dx, u, list);
    __syncthreads();
    uint4* dst = reinterpret_cast<uint4*>(cells + b0);
    for (int q = threadIdx.x; q < nq; q += kForceTile)
        dst[q] = make_uint4(rec[4 * q], rec[4 * q + 1], rec[4 * q + 2], rec[4 * q + 3]);
    if (threadIdx.x == 0 && (ndw & 3))
        reinterpret_cast<uint2*>(cells + b0)[2 * nq] = make_uint2(rec[4 * nq], rec[4 * nq + 1]);
}

__global__ void blk_reset(int2* __restrict__ blk, int64_t nblk, int32_t* __restrict__ n_out) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblk) blk[b] = make_int2(0x7FFFFFFF, -1);
    if (b == 0) *n_out = 0;
}

// Variant 5: one workgroup per aligned 8^3 Morton block. The block's cells
// are contiguous in sorted order (first / last slot from the gather). The
// workgroup loads the ranges of the 10^3 halo voxels and then every candidate
// neighbour of the halo into LDS (each once, all loads independent); each cell
// then walks its 27 stencil voxels in the reference's order (x outer, z inner,
// ascending slot) from LDS, so the bits are those of variants 0-4. A block
// whose halo holds more than kCand candidates walks global memory instead
// (variant 0's loop). The outside cells, which sort first, are copied as they
// are by a grid-stride loop of the same launch.
constexpr int kHalo = 10, kHaloV = 1000, kCand = 2048;

__global__ __launch_bounds__(kBlock) void fm_force_move_blocks(const gell_cell* __restrict__ sorted,
                                                               const float4* __restrict__ nb,
                                                               const int2* __restrict__ se,
                                                               const int2* __restrict__ blk,
                                                               const int32_t* __restrict__ n_out,
                                                               gell_cell* __restrict__ cells, int32_t N, float dx,
                                                               float dt) {
    __shared__ int ho[kBlock * 4], hc[kBlock * 4];
    __shared__ float4 cand[kCand];
    __shared__ int lds4[kBlock / 64];
    const int32_t nout = *n_out;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nout; i += (int64_t)gridDim.x * kBlock)
        store_cell(cells, i, load_cell(sorted, i));
    const int2 range = blk[blockIdx.x];
    if (range.x > range.y) return;  // empty block (uniform across the workgroup)
    const uint32_t m0 = (uint32_t)blockIdx.x << 9;
    const int32_t bx = (int32_t)compact3(m0), by = (int32_t)compact3(m0 >> 1), bz = (int32_t)compact3(m0 >> 2);
    // halo ranges: thread t owns halo voxels 4t .. 4t + 3 (x fastest, 10 per row)
    int hsv[4], hcv[4], local = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int h = 4 * threadIdx.x + e;
        hsv[e] = 0;
        hcv[e] = 0;
        if (h < kHaloV) {
            const int32_t x = bx - 1 + h % kHalo, y = by - 1 + (h / kHalo) % kHalo, z = bz - 1 + h / (kHalo * kHalo);
            if (inside(x, y, z, N)) {
                const int2 r = se[morton3(x, y, z)];
                hsv[e] = r.x;
                hcv[e] = r.y >= r.x ? r.y - r.x + 1 : 0;
            }
        }
        local += hcv[e];
    }
    int total;
    int off = block_exclusive_scan(local, lds4, total);
    const bool fits = total <= kCand;  // uniform
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        ho[4 * threadIdx.x + e] = off;
        hc[4 * threadIdx.x + e] = hcv[e];
        if (fits)
            for (int k = 0; k < hcv[e]; ++k) cand[off + k] = nb[hsv[e] + k];
        off += hcv[e];
    }
    __syncthreads();
    for (int64_t i = range.x + threadIdx.x; i <= range.y; i += kBlock) {
        CellRegs c = load_cell(sorted, i);
        const float4 me = nb[i];
        if (__float_as_uint(me.w) != kDeadTag) {
            float fx = 0.f, fy = 0.f, fz = 0.f;
            if (fits) {
                // FMkernel.cuh:56-118 from the LDS copy
                const int32_t lx = trunc_index(me.x / dx) - bx, ly = trunc_index(me.y / dx) - by,
                              lz = trunc_index(me.z / dx) - bz;
                Cutoff cut;  // thres = Tumor_CC_Max_Dist, carried across pairs (FMkernel.cuh:72)
                for (int q = 0; q < 27; ++q) {
                    const int h = ((lz + q % 3) * kHalo + (ly + (q / 3) % 3)) * kHalo + (lx + q / 9);
                    const int k0 = ho[h], k1 = k0 + hc[h];
                    for (int k = k0; k < k1; ++k) force_pair(cand[k], me.x, me.y, me.z, me.w, cut, fx, fy, fz);
                }
            } else {
                cell_force<false>(nb, se, N, dx, me.x, me.y, me.z, me.w, fx, fy, fz);
            }
            move_cell(c, fx, fy, fz, N, dx, dt);
        }
        store_cell(cells, i, c);
    }
}

__global__ void occ_clear(uint32_t* occ, int64_t words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < words) occ[i] = 0u;
}

__global__ void se_fill(int2* se, int64_t V, int32_t s_empty) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) se[v] = make_int2(s_empty, -1);
}

// Rank-compacted list -> linear (x fastest) ranges of the reference: s is
// the voxel's first live slot, e the last live slot before the next occupied
// voxel's first (only dead cells lie in between).
__global__ void cs_to_linear(const uint4* __restrict__ occr, const int32_t* __restrict__ cs,
                             const float4* __restrict__ nb, int2* __restrict__ out, int32_t N, int32_t s_empty) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t V = (int64_t)N * N * N;
    if (v >= V) return;
    const int32_t i = (int32_t)(v % N), j = (int32_t)((v / N) % N), k = (int32_t)(v / ((int64_t)N * N));
    const uint32_t code = morton3(i, j, k);
    const uint4 rec = occr[code >> 6];
    const uint32_t half = (code & 32u) ? rec.y : rec.x;
    if (!((half >> (code & 31u)) & 1u)) {
        out[v] = make_int2(s_empty, -1);
        return;
    }
    const int32_t r = voxel_rank(rec, code);
    const int32_t s = cs[r];
    int32_t e = cs[r + 1] - 1;
    while (e > s && __float_as_uint(nb[e].w) == kDeadTag) --e;
    out[v] = make_int2(s, e);
}

// Morton-indexed ranges -> linear (x fastest) layout of the reference.
__global__ void se_to_linear(const int2* __restrict__ se, int2* __restrict__ out, int32_t N) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t V = (int64_t)N * N * N;
    if (v >= V) return;
    const int32_t i = (int32_t)(v % N), j = (int32_t)((v / N) % N), k = (int32_t)(v / ((int64_t)N * N));
    out[v] = se[morton3(i, j, k)];
}

}  // namespace

SortPlan plan_sort(int key_bits, int max_bits) {
    SortPlan p{};
    p.passes = key_bits <= max_bits ? 1 : (key_bits + max_bits - 1) / max_bits;
    int shift = 0;
    for (int q = 0; q < p.passes; ++q) {
        p.bits[q] = key_bits / p.passes + (q < key_bits % p.passes ? 1 : 0);
        p.shift[q] = shift;
        shift += p.bits[q];
    }
    return p;
}

hipError_t fm_init_ranges(gell_ctx* ctx) {
    const unsigned blocks = (unsigned)((ctx->KS + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(se_fill, dim3(blocks), dim3(kBlock), 0, ctx->stream, ctx->se, ctx->KS, ctx->P.capacity + 1);
    hipLaunchKernelGGL(occ_clear, dim3((unsigned)((ctx->occ_words + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, ctx->occ, ctx->occ_words);
    hipLaunchKernelGGL(occ_clear, dim3((unsigned)((4 * ctx->occr_words + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       ctx->stream, reinterpret_cast<uint32_t*>(ctx->occr), 4 * ctx->occr_words);
    ctx->se_list_n = 0;
    ctx->cl_mode = 0;
    return hipGetLastError();
}

hipError_t fm_ranges_linear(gell_ctx* ctx, int2* out) {
    const unsigned blocks = (unsigned)((ctx->V + kBlock - 1) / kBlock);
    if (ctx->cl_mode == 1)
        hipLaunchKernelGGL(cs_to_linear, dim3(blocks), dim3(kBlock), 0, ctx->stream, ctx->occr, ctx->cs, ctx->nb, out,
                           ctx->N, ctx->P.capacity + 1);
    else
        hipLaunchKernelGGL(se_to_linear, dim3(blocks), dim3(kBlock), 0, ctx->stream, ctx->se, out, ctx->N);
    return hipGetLastError();
}

// The variant a launch for n cells runs. 0-9 run as set; 10 (the default)
// picks the rank-compacted list (9) for large sparse tissue and 7 otherwise
// (same bits). The list costs four extra launches per step and pays only
// where most stencil voxels are empty and there are many cells: measured,
// 1M cells in 256^3 (1/16 cell per grid voxel) 92.0 -> 90.1 ms per biology
// step with it; 10k cells in 64^3, 131k-200k-cell replicates in 128^3-192^3
// and the reference's dense 100^3 example (up to ~1 cell per voxel) slower.
constexpr int32_t kRankMin = 524288;
int fm_variant(const gell_ctx* ctx, int32_t n) {
    const int v = ctx->tune[GELL_KNOB_FM_PREFETCH];
    if (v != 10) return v;
    const int64_t m = grid_count(ctx, n);
    return (m >= kRankMin && 8 * m <= (int64_t)ctx->N * ctx->N * ctx->N) ? 9 : 7;
}

hipError_t fm_keys_launch(gell_ctx* ctx, const gell_cell* cells, int32_t n) {
    const int32_t m = ctx->devn ? ctx->P.capacity : (n > ctx->se_list_n ? n : ctx->se_list_n);
    if (m == 0) return hipSuccess;
    const bool ranked = ctx->cl_mode == 1;  // what the previous gather wrote, cleared here
    hipLaunchKernelGGL(fm_keys, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, cells, n, ctx->N,
                       ctx->P.dx, ctx->keys[0], ctx->se_list, ctx->se_list_n, ctx->se, ctx->P.capacity + 1,
                       ranked ? reinterpret_cast<uint32_t*>(ctx->occr) : ctx->occ, (int)ranked, dev_count(ctx),
                       dev_count(ctx, DN_LIST));
    ctx->se_list_n = 0;
    return hipGetLastError();
}

hipError_t fm_sort_launch(gell_ctx* ctx, int32_t n) {
    const int32_t gn = grid_count(ctx, n);
    if (gn == 0) return hipSuccess;
    const int key_bits = 3 * ceil_log2(ctx->N) + 1;
    if (gn <= kSmallMax) {
        hipLaunchKernelGGL(radix_sort_small, dim3(1), dim3(kSmallThreads), 0, ctx->stream, ctx->keys[0], ctx->vals[0],
                           ctx->keys[1], ctx->vals[1], ctx->out_order, ctx->out_keys, n,
                           plan_sort(key_bits, kSmallBits), dev_count(ctx));
        return hipGetLastError();
    }
    const SortPlan plan = plan_sort(key_bits, 9);
    int cur = 0;
    for (int q = 0; q < plan.passes; ++q) {
        const bool last = q == plan.passes - 1;
        const int32_t* vin = q == 0 ? nullptr : ctx->vals[cur];
        int32_t* vout = last ? ctx->out_order : ctx->vals[1 - cur];
        hipError_t e = kPass[plan.bits[q]](ctx, ctx->keys[cur], vin, ctx->keys[1 - cur], vout,
                                           last ? ctx->out_keys : nullptr, n, plan.shift[q]);
        if (e != hipSuccess) return e;
        cur = 1 - cur;
    }
    return hipSuccess;
}

hipError_t fm_gather_launch(gell_ctx* ctx, const gell_cell* cells, int32_t n) {
    const int32_t gn = grid_count(ctx, n);
    if (gn == 0) return hipSuccess;
    const bool blocks = ctx->tune[GELL_KNOB_FM_PREFETCH] == 5;
    if (blocks)
        hipLaunchKernelGGL(blk_reset, dim3((unsigned)((ctx->nmblk + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, ctx->mblk, ctx->nmblk, ctx->n_out);
    const bool ranked = fm_variant(ctx, n) == 9;
    hipLaunchKernelGGL(fm_gather, dim3((gn + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, cells,
                       ctx->out_order, n, ctx->N, ctx->P.dx, ctx->sorted, ctx->nb, ranked ? nullptr : ctx->se,
                       ctx->se_list, ctx->out_keys, blocks ? ctx->mblk : nullptr, ctx->n_out,
                       ranked ? reinterpret_cast<uint32_t*>(ctx->occr) : ctx->occ, (int)ranked, dev_count(ctx),
                       dev_count_out(ctx, DN_LIST));
    ctx->se_list_n = n;
    ctx->cl_mode = ranked ? 1 : 0;
    if (ranked) {
        // ranks of the occupied voxels, then each one's first live slot
        const unsigned rb = (unsigned)ctx->rank_blocks;
        hipLaunchKernelGGL(rank_count, dim3(rb), dim3(kBlock), 0, ctx->stream, ctx->occr, ctx->occr_words,
                           ctx->rank_tot);
        hipLaunchKernelGGL(rank_scan, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->rank_tot, (int32_t)rb, ctx->cs, n,
                           dev_count(ctx));
        hipLaunchKernelGGL(rank_write, dim3(rb), dim3(kBlock), 0, ctx->stream, ctx->occr, ctx->occr_words,
                           ctx->rank_tot, ctx->cs);
        hipLaunchKernelGGL(fm_cs, dim3((gn + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, ctx->out_keys,
                           ctx->nb, n, ctx->occr, ctx->cs, dev_count(ctx));
    }
    return hipGetLastError();
}

bool fm_fuses_uptake(const gell_ctx* ctx) {
    const int v = ctx->tune[GELL_KNOB_FM_PREFETCH];
    return v >= 7 && v <= 10;
}

hipError_t fm_force_move_launch(gell_ctx* ctx, gell_cell* cells, int32_t n, int32_t* uptake_list) {
    const int32_t gn = grid_count(ctx, n);
    if (gn == 0) return hipSuccess;
    const dim3 grid((gn + kBlock - 1) / kBlock);
    const int32_t* np = dev_count(ctx);
    int v = fm_variant(ctx, n);
    UptakeSet ua{};
    if (uptake_list) {
        // only the XCD-ordered kernels accumulate (fm_fuses_uptake)
        if (v < 7 || v > 9) return hipErrorInvalidValue;
        ua = prelod_uptake_set(ctx);
    }
    if (v == 5) {
        hipLaunchKernelGGL(fm_force_move_blocks, dim3((unsigned)ctx->nmblk), dim3(kBlock), 0, ctx->stream,
                           ctx->sorted, ctx->nb, ctx->se, ctx->mblk, ctx->n_out, cells, ctx->N, ctx->P.dx,
                           ctx->P.dt_mech);
        return hipGetLastError();
    }
    const bool aligned16 = ((reinterpret_cast<uintptr_t>(cells) | reinterpret_cast<uintptr_t>(ctx->sorted)) & 15) == 0;
    if (v == 4 && !aligned16) v = 3;
    if (v == 7 && !aligned16) v = 8;
    if (v >= 7 && v <= 9) {
        const dim3 grid((gn + kForceTile - 1) / kForceTile);
        const gell_cell* so = ctx->sorted;
        if (v == 7)
            hipLaunchKernelGGL((fm_force_move_xcd<true, false>), grid, dim3(kForceTile), 0, ctx->stream, so, ctx->nb,
                               ctx->se, cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, ctx->occr, ctx->cs, ua, uptake_list, np);
        else if (v == 8)
            hipLaunchKernelGGL((fm_force_move_xcd<false, false>), grid, dim3(kForceTile), 0, ctx->stream, so, ctx->nb,
                               ctx->se, cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, ctx->occr, ctx->cs, ua, uptake_list, np);
        else if (aligned16)
            hipLaunchKernelGGL((fm_force_move_xcd<true, true>), grid, dim3(kForceTile), 0, ctx->stream, so, ctx->nb,
                               ctx->se, cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, ctx->occr, ctx->cs, ua, uptake_list, np);
        else
            hipLaunchKernelGGL((fm_force_move_xcd<false, true>), grid, dim3(kForceTile), 0, ctx->stream, so, ctx->nb,
                               ctx->se, cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, ctx->occr, ctx->cs, ua, uptake_list, np);
        return hipGetLastError();
    }
    if (v == 4)
        hipLaunchKernelGGL(fm_force_move_staged, grid, dim3(kBlock), 0, ctx->stream, ctx->sorted, ctx->nb, ctx->se,
                           cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, np);
    else if (v == 6)
        hipLaunchKernelGGL(fm_force_move<6>, grid, dim3(kBlock), 0, ctx->stream, ctx->sorted, ctx->nb, ctx->se,
                           cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, np);
    else if (v == 3)
        hipLaunchKernelGGL(fm_force_move<3>, grid, dim3(kBlock), 0, ctx->stream, ctx->sorted, ctx->nb, ctx->se,
                           cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, np);
    else if (v == 2)
        hipLaunchKernelGGL(fm_force_move<2>, grid, dim3(kBlock), 0, ctx->stream, ctx->sorted, ctx->nb, ctx->se,
                           cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, np);
    else if (v == 1)
        hipLaunchKernelGGL(fm_force_move<1>, grid, dim3(kBlock), 0, ctx->stream, ctx->sorted, ctx->nb, ctx->se,
                           cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, np);
    else
        hipLaunchKernelGGL(fm_force_move<0>, grid, dim3(kBlock), 0, ctx->stream, ctx->sorted, ctx->nb, ctx->se,
                           cells, n, ctx->N, ctx->P.dx, ctx->P.dt_mech, ctx->occ, np);
    return hipGetLastError();
}

}  // namespace gell
