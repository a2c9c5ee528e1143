// pd.hip -- proliferation and death on gfx950 (CellBirth_kernel /
// CellDeath_kernel, PDkernel.cuh:27-117).
//
// Birth: the reference draws from a per-slot XORWOW state and hands out
// daughter slots with atomicAdd, so which mother gets which slot depends on
// scheduling. Here the draws are Philox4x32-10(key = seed, counter = (slot,
// step)) and the slots come from an exclusive prefix count of the dividing
// mothers in slot order (wave ballots + a small scan of per-block counts);
// divisions beyond capacity - 1 are dropped in that same order.
// Death: stable stream compaction of live records by the same scan.
#include <algorithm>
#include <mutex>
#include <string>

#include "device_common.hpp"
#include "internal.hpp"

namespace gell {
namespace {

constexpr int kBlock = 256;
constexpr int kScanThreads = 1024;

// Rank of this thread's flag among the flagged threads of the block (all
// 256 threads call it); returns the block total through `total`.
__device__ __forceinline__ int block_rank(bool flag, int* lds4, int& total) {
    const uint64_t bal = __ballot(flag);
    const int w = threadIdx.x / 64;
    if (lane_id() == 0) lds4[w] = __popcll(bal);
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
    return off + __popcll(bal & lanemask_lt());
}

// In-place exclusive scan of a[0, m) by one workgroup; the sum goes to *total.
__global__ __launch_bounds__(kScanThreads) void scan_small(int32_t* __restrict__ a, int32_t m,
                                                           int32_t* __restrict__ total) {
    __shared__ int wsum[kScanThreads / 64];
    const int tid = threadIdx.x, lane = lane_id(), w = tid / 64;
    int carry = 0;
    for (int base = 0; base < m; base += kScanThreads) {
        const int idx = base + tid;
        const int x = idx < m ? a[idx] : 0;
        int inc = x;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(inc, d, 64);
            if (lane >= d) inc += y;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        int off = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < kScanThreads / 64; ++q) {
            const int s = wsum[q];
            if (q < w) off += s;
            tot += s;
        }
        __syncthreads();
        if (idx < m) a[idx] = carry + off + inc - x;
        carry += tot;
    }
    if (tid == 0) *total = carry;
}

__device__ __forceinline__ Bio bio_of(const CellRegs& c) {
    Bio b;
    b.V = cf(c, W_V);
    b.Vf = cf(c, W_VF);
    b.Vcs = cf(c, W_VCS);
    b.Vns = cf(c, W_VNS);
    b.r = cf(c, W_R);
    b.clock = cf(c, W_CLOCK);
    b.phase = (int32_t)c.w[W_PHASE];
    b.sign = csign(c);
    return b;
}

// Birthkernel's per-cell part (PDkernel.cuh:28-49): draw, O2 lookup,
// Volume_update, Phase_update_ki67adv.
// With a device-resident count (np) the step comes from stepp and a full
// array (n >= capacity) updates nothing, as the host-count path skips the
// launch then (PDkernel.cuh:31).
__device__ __forceinline__ void step_of(const uint32_t* stepp, uint32_t& s0, uint32_t& s1) {
    if (stepp) {
        s0 = stepp[0];
        s1 = stepp[1];
    }
}

__global__ __launch_bounds__(kBlock) void birth_update(gell_cell* __restrict__ cells, int32_t n,
                                                       const float* __restrict__ p0, int32_t N, float dx,
                                                       float dt_bio, uint32_t k0, uint32_t k1, uint32_t s0,
                                                       uint32_t s1, int32_t* __restrict__ flags,
                                                       int32_t* __restrict__ blk, const int32_t* __restrict__ np,
                                                       const uint32_t* __restrict__ stepp, int32_t capacity) {
    __shared__ int lds4[kBlock / 64];
    n = count_of(n, np);
    step_of(stepp, s0, s1);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool divide = false;
    if (i < n && n < capacity) {
        CellRegs c = load_cell(cells, i);
        if (csign(c)) {
            const uint4 r = philox4x32_10(make_uint4((uint32_t)i, s0, s1, 0u), k0, k1);
            const float u = u01(r.x);
            int32_t a = trunc_index(cf(c, W_PX) / dx), b = trunc_index(cf(c, W_PY) / dx),
                    k = trunc_index(cf(c, W_PZ) / dx);
            a = min(max(a, 0), N - 1);
            b = min(max(b, 0), N - 1);
            k = min(max(k, 0), N - 1);
            const float O2 = p0[lin(a, b, k, N)];
            Bio bio = bio_of(c);
            volume_update(bio, dt_bio);
            divide = phase_update(bio, u, O2, dt_bio);
            cset(c, W_V, bio.V);
            cset(c, W_VF, bio.Vf);
            cset(c, W_VCS, bio.Vcs);
            cset(c, W_VNS, bio.Vns);
            cset(c, W_R, bio.r);
            cset(c, W_CLOCK, bio.clock);
            c.w[W_PHASE] = (uint32_t)bio.phase;
            cset_sign(c, bio.sign);
            store_cell(cells, i, c);
        }
    }
    if (i < n) flags[i] = divide ? 1 : 0;
    int total;
    block_rank(divide, lds4, total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// Division (PDkernel.cuh:51-85) for the mothers whose slot fits.
__global__ __launch_bounds__(kBlock) void birth_divide(gell_cell* __restrict__ cells, int32_t n, int32_t capacity,
                                                       const int32_t* __restrict__ flags,
                                                       const int32_t* __restrict__ blk_off, uint32_t k0, uint32_t k1,
                                                       uint32_t s0, uint32_t s1, int32_t* __restrict__ slots,
                                                       const int32_t* __restrict__ np,
                                                       const uint32_t* __restrict__ stepp) {
    __shared__ int lds4[kBlock / 64];
    n = count_of(n, np);
    step_of(stepp, s0, s1);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool flag = i < n && flags[i] != 0;
    int total;
    const int rank = blk_off[blockIdx.x] + block_rank(flag, lds4, total);
    if (i >= n) return;
    int32_t slot = -1;
    if (flag && (int64_t)n + rank < (int64_t)capacity - 1) {
        slot = n + rank;
        const uint4 r = philox4x32_10(make_uint4((uint32_t)i, s0, s1, 0u), k0, k1);
        const float u = u01(r.y), v = u01(r.z);
        // randomvect, RandomFunctions.cuh:23-32
        const float theta = 2.f * u * 3.1415926f;
        const float phi = acosf(2.f * v - 1.f);
        const float dxr = sinf(theta) * sinf(phi);
        const float dyr = cosf(theta) * sinf(phi);
        const float dzr = cosf(phi);
        CellRegs m = load_cell(cells, i);
        float rm = cf(m, W_R);
        const float distance = 0.2063f * rm;
        rm *= 0.7937f;
        const float x = cf(m, W_PX), y = cf(m, W_PY), z = cf(m, W_PZ);
        cset(m, W_R, rm);
        cset(m, W_PX, x - dxr * distance);
        cset(m, W_PY, y - dyr * distance);
        cset(m, W_PZ, z - dzr * distance);
        m.w[W_PHASE] = 2u;
        cset(m, W_CLOCK, 0.f);
        cset(m, W_V, cf(m, W_V) / 2.f);
        cset(m, W_VF, cf(m, W_VF) / 2.f);
        cset(m, W_VCS, cf(m, W_VCS) / 2.f);
        cset(m, W_VNS, cf(m, W_VNS) / 2.f);
        store_cell(cells, i, m);
        CellRegs d;
#pragma unroll
        for (int q = 0; q < 18; ++q) d.w[q] = 0u;
        d.w[W_SIGN] = 1u;
        d.w[W_MESH] = 0xFFFFFFFFu;  // -1
        cset(d, W_PX, x + dxr * distance);
        cset(d, W_PY, y + dyr * distance);
        cset(d, W_PZ, z + dzr * distance);
        cset(d, W_R, rm);
        d.w[W_V] = m.w[W_V];
        d.w[W_VF] = m.w[W_VF];
        d.w[W_VCS] = m.w[W_VCS];
        d.w[W_VNS] = m.w[W_VNS];
        d.w[W_PHASE] = 2u;
        store_cell(cells, slot, d);
    }
    slots[i] = slot;
}

__global__ __launch_bounds__(kBlock) void death_count(const gell_cell* __restrict__ cells, int32_t n,
                                                      int32_t* __restrict__ blk, const int32_t* __restrict__ np) {
    __shared__ int lds4[kBlock / 64];
    n = count_of(n, np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n && (reinterpret_cast<const uint32_t*>(cells + i)[W_SIGN] & 0xFFu) != 0;
    int total;
    block_rank(live, lds4, total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

__global__ __launch_bounds__(kBlock) void death_scatter(const gell_cell* __restrict__ cells, int32_t n,
                                                        const int32_t* __restrict__ blk_off,
                                                        gell_cell* __restrict__ out, const int32_t* __restrict__ np) {
    __shared__ int lds4[kBlock / 64];
    n = count_of(n, np);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n && (reinterpret_cast<const uint32_t*>(cells + i)[W_SIGN] & 0xFFu) != 0;
    int total;
    const int rank = blk_off[blockIdx.x] + block_rank(live, lds4, total);
    if (live) store_cell(out, rank, load_cell(cells, i));
}

// Copies the first *count records of src to dst (count read on the device, so
// the compaction needs no host round trip before its copy-back). Grid-stride
// over 8-byte words; the grid is sized for the upper bound n.
__global__ __launch_bounds__(kBlock) void copy_counted(const uint2* __restrict__ src, uint2* __restrict__ dst,
                                                       const int32_t* __restrict__ count) {
    const int64_t words = (int64_t)(*count) * (int64_t)(sizeof(gell_cell) / sizeof(uint2));
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) dst[i] = src[i];
}

// The new device-resident count after birth: n plus the births that fit
// below capacity - 1 (PDkernel.cuh:51-54); unchanged for a full array.
__global__ void birth_count(int32_t* __restrict__ np, const int32_t* __restrict__ births, int32_t capacity) {
    const int32_t n = *np;
    if (n <= 0 || n >= capacity) return;
    const int64_t room = (int64_t)capacity - 1 - n;
    const int64_t b = *births < room ? *births : (room > 0 ? room : 0);
    *np = (int32_t)(n + b);
}

// The new device-resident count after death: the survivors.
__global__ void death_set_count(int32_t* __restrict__ np, const int32_t* __restrict__ live) { *np = *live; }

}  // namespace

hipError_t birth_begin(gell_ctx* ctx, gell_cell* cells, int32_t n, uint64_t step) {
    hipStream_t st = ctx->stream;
    const bool dev = ctx->devn;
    if (!dev && (n <= 0 || n >= ctx->P.capacity)) {
        // PDkernel.cuh:31: nothing happens once the array is full.
        if (n > 0) return hipMemsetAsync(ctx->slots, 0xFF, sizeof(int32_t) * (size_t)n, st);
        return hipSuccess;
    }
    // with device-resident counts every case runs on the device, sized for
    // the capacity, and the step is read from dn[DN_STEP]
    const int32_t gn = grid_count(ctx, n);
    const unsigned blocks = (unsigned)((gn + kBlock - 1) / kBlock);
    const uint32_t k0 = (uint32_t)ctx->P.seed, k1 = (uint32_t)(ctx->P.seed >> 32);
    const uint32_t s0 = (uint32_t)step, s1 = (uint32_t)(step >> 32);
    const int32_t* np = dev_count(ctx);
    const uint32_t* stepp = dev ? reinterpret_cast<const uint32_t*>(ctx->dn + DN_STEP) : nullptr;
    hipLaunchKernelGGL(birth_update, dim3(blocks), dim3(kBlock), 0, st, cells, n, ctx->field[0], ctx->N, ctx->P.dx,
                       ctx->P.dt_bio, k0, k1, s0, s1, ctx->flags, ctx->blk, np, stepp, ctx->P.capacity);
    hipLaunchKernelGGL(scan_small, dim3(1), dim3(kScanThreads), 0, st, ctx->blk, (int32_t)blocks, ctx->d_total);
    hipLaunchKernelGGL(birth_divide, dim3(blocks), dim3(kBlock), 0, st, cells, n, ctx->P.capacity, ctx->flags,
                       ctx->blk, k0, k1, s0, s1, ctx->slots, np, stepp);
    if (dev) {
        hipLaunchKernelGGL(birth_count, dim3(1), dim3(1), 0, st, ctx->dn + DN_CELLS, ctx->d_total, ctx->P.capacity);
        return hipGetLastError();
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(ctx->h_total, ctx->d_total, sizeof(int32_t), hipMemcpyDeviceToHost, st);
}

hipError_t birth_finish(gell_ctx* ctx, int32_t n, int32_t* n_out) {
    *n_out = n;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess || n <= 0 || n >= ctx->P.capacity) return e;
    const int64_t room = (int64_t)ctx->P.capacity - 1 - n;
    const int64_t births = *ctx->h_total < room ? *ctx->h_total : (room > 0 ? room : 0);
    *n_out = (int32_t)(n + births);
    return hipSuccess;
}

hipError_t birth_launch(gell_ctx* ctx, gell_cell* cells, int32_t n, uint64_t step, int32_t* n_out) {
    *n_out = n;
    hipError_t e = birth_begin(ctx, cells, n, step);
    return e != hipSuccess ? e : birth_finish(ctx, n, n_out);
}

hipError_t death_begin(gell_ctx* ctx, gell_cell* cells, int32_t n) {
    const bool dev = ctx->devn;
    if (!dev && n <= 0) return hipSuccess;
    hipStream_t st = ctx->stream;
    const int32_t gn = grid_count(ctx, n);
    const int32_t* np = dev_count(ctx);
    const unsigned blocks = (unsigned)((gn + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(death_count, dim3(blocks), dim3(kBlock), 0, st, cells, n, ctx->blk, np);
    hipLaunchKernelGGL(scan_small, dim3(1), dim3(kScanThreads), 0, st, ctx->blk, (int32_t)blocks, ctx->d_total);
    hipLaunchKernelGGL(death_scatter, dim3(blocks), dim3(kBlock), 0, st, cells, n, ctx->blk, ctx->compact, np);
    const int64_t words = (int64_t)gn * (int64_t)(sizeof(gell_cell) / sizeof(uint2));
    const int64_t cblocks = std::max<int64_t>(1, std::min<int64_t>((words + kBlock - 1) / kBlock, 8192));
    hipLaunchKernelGGL(copy_counted, dim3((unsigned)cblocks), dim3(kBlock), 0, st,
                       reinterpret_cast<const uint2*>(ctx->compact), reinterpret_cast<uint2*>(cells), ctx->d_total);
    if (dev) {
        hipLaunchKernelGGL(death_set_count, dim3(1), dim3(1), 0, st, ctx->dn + DN_CELLS, ctx->d_total);
        return hipGetLastError();
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(ctx->h_total, ctx->d_total, sizeof(int32_t), hipMemcpyDeviceToHost, st);
}

hipError_t death_finish(gell_ctx* ctx, int32_t n, int32_t* n_out) {
    *n_out = n;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess || n <= 0) return e;
    *n_out = *ctx->h_total;
    return hipSuccess;
}

hipError_t death_launch(gell_ctx* ctx, gell_cell* cells, int32_t n, int32_t* n_out) {
    *n_out = n;
    hipError_t e = death_begin(ctx, cells, n);
    return e != hipSuccess ? e : death_finish(ctx, n, n_out);
}

}  // namespace gell

// ---------------------------------------------------------------------------
// Death without allocation (gell_death_noalloc): the same four kernels as
// gell_death, with their scratch (per-block counts and offsets, the live
// count, the compacted records) in one file-scope device array instead of a
// context's buffers. Each device holds its own copy of the array, at its own
// address, until the process ends; host code uses only the address
// hipGetSymbolAddress returns for the current device, never the array's name.
// ---------------------------------------------------------------------------
namespace {

constexpr int32_t kNaMaxCells = GELL_DEATH_NOALLOC_MAX_CELLS;
constexpr int32_t kNaMaxBlocks = (kNaMaxCells + gell::kBlock - 1) / gell::kBlock;
constexpr size_t kNaBlkOff = 0;
constexpr size_t kNaTotalOff = ((sizeof(int32_t) * (size_t)kNaMaxBlocks + 255) / 256) * 256;
constexpr size_t kNaCellsOff = kNaTotalOff + 256;
constexpr size_t kNaBytes = kNaCellsOff + sizeof(gell_cell) * (size_t)kNaMaxCells;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
unsigned char* g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) on its first call.
hipError_t noalloc_scratch(unsigned char** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kNaMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_na_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_na_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int gell_death_noalloc(gell_cell* d_cells, int32_t n, void* stream, int32_t* n_out) {
    using namespace gell;
    if (!n_out) return fail_global(GELL_E_ARG, "gell_death_noalloc: null n_out");
    if (n < 0 || (n > 0 && !d_cells)) return fail_global(GELL_E_ARG, "gell_death_noalloc: bad cell array or count");
    if (n > kNaMaxCells)
        return fail_global(GELL_E_ARG,
                           "gell_death_noalloc: n exceeds GELL_DEATH_NOALLOC_MAX_CELLS (use gell_death for more cells)");
    *n_out = n;
    if (n == 0) return GELL_OK;
    unsigned char* base = nullptr;
    hipError_t e = noalloc_scratch(&base);
    if (e != hipSuccess) return fail_global(GELL_E_HIP, std::string("gell_death_noalloc: scratch address: ") +
                                                            hipGetErrorString(e));
    int32_t* blk = reinterpret_cast<int32_t*>(base + kNaBlkOff);
    int32_t* total = reinterpret_cast<int32_t*>(base + kNaTotalOff);
    gell_cell* compact = reinterpret_cast<gell_cell*>(base + kNaCellsOff);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    // every word the kernels read is written earlier in the same call: the
    // counts by death_count, offsets and total by scan_small, the records by
    // death_scatter; nothing has to start at zero
    hipLaunchKernelGGL(death_count, dim3(blocks), dim3(kBlock), 0, st, d_cells, n, blk, nullptr);
    hipLaunchKernelGGL(scan_small, dim3(1), dim3(kScanThreads), 0, st, blk, (int32_t)blocks, total);
    hipLaunchKernelGGL(death_scatter, dim3(blocks), dim3(kBlock), 0, st, d_cells, n, blk, compact, nullptr);
    const int64_t words = (int64_t)n * (int64_t)(sizeof(gell_cell) / sizeof(uint2));
    const int64_t cblocks = std::min<int64_t>((words + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(copy_counted, dim3((unsigned)cblocks), dim3(kBlock), 0, st,
                       reinterpret_cast<const uint2*>(compact), reinterpret_cast<uint2*>(d_cells), total);
    e = hipGetLastError();
    int32_t live = n;
    if (e == hipSuccess) e = hipMemcpyAsync(&live, total, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail_global(GELL_E_HIP, std::string("gell_death_noalloc: ") + hipGetErrorString(e));
    *n_out = live;
    return GELL_OK;
}
