// nn.hip -- kernels of the LOD surrogate network on gfx950: 3x3x3
// convolutions as implicit GEMMs on the BF16 matrix cores (FP32
// accumulation), their weight gradient, and the elementwise pieces (pooling,
// upsampling, skip adds, ReLU masks, bias gradients, loss, Adam).
//
// Tensors are channels-last: element (x, y, z, c) of an X x Y x Z grid with C
// channels sits at ((z Y + y) X + x) C + c. Activations are BF16 (uint16
// storage, round-to-nearest-even), parameters FP32 with a BF16 copy in the
// GEMM layout W[o][k], k = tap C_in + c, tap = (dz 3 + dy) 3 + dx, the input
// voxel of tap (dx, dy, dz) being the output voxel + (dx - 1, dy - 1, dz - 1),
// zero outside the grid. Rows are padded to whole 16 x 32 MFMA tiles with
// zeros.
//
// MFMA fragment use (v_mfma_f32_16x16x32_bf16): A = weights (16 output
// channels x 32 k), B = input (32 k x 16 voxels); lane l supplies row/column
// l % 16 and the 8 consecutive k of group l / 16 for both operands, and gets
// D[4 (l / 16) + i][l % 16], i = 0..3: four consecutive output channels of
// one voxel, stored as one 8-byte write.
#include <cstdint>

#include "internal.hpp"
#include "nn.hpp"

namespace gell {
namespace nn {
namespace {

typedef __attribute__((ext_vector_type(4))) float f4;
typedef __attribute__((ext_vector_type(8))) __bf16 bf8;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bf16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

typedef __attribute__((ext_vector_type(2))) float f2v;
typedef __attribute__((ext_vector_type(2))) __bf16 bf2v;

// Two floats -> packed BF16 pair (low half = a), round to nearest even: one
// v_cvt_pk_bf16_f32, the rounding f2bf spells out.
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v{a, b}, bf2v));
}

constexpr int kThreads = 256;  // 4 waves

// Stages a halo of HX x HY x HZ voxels x CH 16-byte channel chunks into LDS,
// chunk (hz, hy, hx, ch) at lds[((hz HY + hy) HX + hx) CH + ch], straight from
// memory without passing through registers. The halo's origin is voxel
// (x0, y0, z0) of the X x Y x Z tensor `base` (CIN channels, BF16); voxels
// outside it read zero (an offset past the buffer's end reads zero). A
// thread's (hy, hx, ch) is worked out once per plane slot; each z plane then
// costs it one bounds test and one address. All loads are in flight together;
// stage_wait() waits for them (before the barrier).
template <int HX, int HY, int HZ, int CH>
__device__ __forceinline__ void stage_halo(const uint16_t* base, int X, int Y, int Z, int x0, int y0, int z0,
                                           uint4* lds) {
    constexpr int PC = HX * HY * CH;  // chunks per z plane
    constexpr int CIN = CH * 8;
    const uint32_t num_bytes = (uint32_t)((int64_t)X * Y * Z * CIN * 2);
    const uint32_t plane = (uint32_t)X * (uint32_t)Y * CIN * 2u;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(base), (short)0, (int)num_bytes, 0x00020000);
    const int wbase = (int)(threadIdx.x & ~63u);
#pragma unroll
    for (int j = 0; j < (PC + kThreads - 1) / kThreads; ++j) {
        const int p = j * kThreads + (int)threadIdx.x;
        if (p < PC) {
            const int ch = p % CH, hv = p / CH;
            const int gx = x0 + hv % HX, gy = y0 + hv / HX;
            const bool okxy = gx >= 0 && gx < X && gy >= 0 && gy < Y;
            const uint32_t row = (((uint32_t)gy * (uint32_t)X + (uint32_t)gx) * CIN + (uint32_t)ch * 8u) * 2u;
#pragma unroll
            for (int hz = 0; hz < HZ; ++hz) {
                const int gz = z0 + hz;
                const uint32_t vo = (okxy && gz >= 0 && gz < Z) ? row + (uint32_t)gz * plane : num_bytes;
                // lane L of the wave lands at lds[hz PC + j kThreads + wbase + L]
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs, (__attribute__((address_space(3))) void*)(lds + hz * PC + j * kThreads + wbase), 16, vo, 0, 0,
                    0);
            }
        }
    }
}
__device__ __forceinline__ void stage_wait() { __builtin_amdgcn_s_waitcnt(0); }
constexpr int kTX = 8, kTY = 8;

// ---------------------------------------------------------------------------
// Forward convolution. Block = 8 x 8 x TZ output voxels and CG of the layer's
// 16-channel output tiles (blockIdx.z also walks the NCT / CG tile groups);
// its input halo (10 x 10 x (TZ + 2) voxels x CIN) is staged in LDS as
// 16-byte chunks. The block's voxels form B / 16 tiles of 16, split over the
// 4 waves; each wave computes its voxel tiles for the block's CG output tiles,
// so one LDS fragment feeds CG matrix-core operations. Grids of 256 blocks or
// more take every output tile in one block (CG = NCT); smaller ones (the
// 16^3 level) one tile per block (CG = 1), for more blocks and fewer weight
// loads per wave.
// Epilogue: + bias, ReLU (when relu), BF16 store.
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int TZ, int CG>
__global__ __launch_bounds__(kThreads) void conv3_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
                                                      const float* __restrict__ bias, uint16_t* __restrict__ y, int X,
                                                      int Y, int Z, int relu, uint16_t* __restrict__ pooled) {
    constexpr int HX = kTX + 2, HY = kTY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int CH = CIN / 8;
    constexpr int KP = kpad(CIN);
    constexpr int KSTEPS = KP / 32;
    constexpr int NCT = cpad(COUT) / 16;
    constexpr int NVT = kTX * kTY * TZ / 16;
    constexpr int VPW = NVT / 4;
    static_assert(CIN % 8 == 0 && NVT % 4 == 0 && NCT % CG == 0, "tile shape");
    __shared__ uint4 halo[HV * CH];
    const int zb = Z / TZ;
    const int bz = blockIdx.z % zb, ct0 = (blockIdx.z / zb) * CG;
    const int x0 = blockIdx.x * kTX - 1, y0 = blockIdx.y * kTY - 1, z0 = bz * TZ - 1;
    stage_halo<HX, HY, HZ, CH>(x, X, Y, Z, x0, y0, z0, halo);
    stage_wait();
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    f4 acc[VPW][CG];
#pragma unroll
    for (int a = 0; a < VPW; ++a)
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[a][c] = f4{0.f, 0.f, 0.f, 0.f};
    // In-block position of the voxel this lane feeds as a B column: tile a of
    // wave wv is the 16 voxels (wv + 4 a) 16 .. + 15, i.e. z slice a, rows
    // 2 wv and 2 wv + 1; so fragment a sits PLANE chunks after fragment 0 and
    // every LDS read below is this lane's base plus a compile-time offset.
    const int hbase0 = (2 * wv + (r >> 3)) * HX + (r & 7);
    // B chunk of fragment a at step s: lane base (its voxel, and its k group's
    // channel chunk and tap shift) + tap offset of the step + a PLANE.
    // CIN >= 32: the four k groups of a step are four channel chunks of one
    // tap. CIN = 16: groups 0, 1 are the channel halves of tap 2 s, groups 2,
    // 3 those of tap 2 s + 1, whose offset from tap 2 s is one of three.
    // chunk (ch, voxel v) at v CH + ch (a layout with each channel chunk's
    // voxels 16 bytes apart measured no faster)
    constexpr int PLANE = HX * HY * CH, VS = CH, CS = 1;
    const uint4* lb1 = halo + hbase0 * VS + (CIN >= 32 ? g * CS : (g & 1) * CS + (g >> 1) * VS);
    const uint4* lb2 = halo + hbase0 * VS + (g & 1) * CS + (g >> 1) * VS * (HX - 2);
    const uint4* lb3 = halo + hbase0 * VS + (g & 1) * CS + (g >> 1) * VS * (HX * HY - 2 * HX - 2);
    const uint16_t* wr = w + ((ct0 * 16) + r) * KP + g * 8;
    // weight fragments of step s + 1 are loaded while step s computes
    bf8 af[CG], an[CG];
#pragma unroll
    for (int c = 0; c < CG; ++c) an[c] = *reinterpret_cast<const bf8*>(wr + c * 16 * KP);
    // two steps per iteration: enough B reads in flight ahead of the products
    // without holding every step's fragments in registers
#pragma unroll 2
    for (int s = 0; s < KSTEPS; ++s) {
        const int k0 = s * 32 + g * 8;
        const int tap = k0 / CIN, c0 = k0 % CIN;
#pragma unroll
        for (int c = 0; c < CG; ++c) af[c] = an[c];
        if (s + 1 < KSTEPS) {
#pragma unroll
            for (int c = 0; c < CG; ++c) an[c] = *reinterpret_cast<const bf8*>(wr + c * 16 * KP + (s + 1) * 32);
        }
        // all of the step's B fragments are read before the first product,
        // so the LDS reads overlap one another and the matrix-core work
        bf8 bf[VPW];
        if constexpr (CIN >= 32 || CIN == 16) {
            const int kb = s * 32, t0 = kb / CIN, cb = kb % CIN;  // tap and chunk of k group 0
            const int toff = (((t0 / 9) * HY + (t0 / 3) % 3) * HX + t0 % 3) * VS + (cb / 8) * CS;
            const uint4* p = CIN >= 32 || t0 % 3 != 2 ? lb1 : ((t0 / 3) % 3 != 2 ? lb2 : lb3);
            const bool valid = CIN >= 32 || t0 + 1 < 27 || g < 2;  // the padding tap 27 reads zero
#pragma unroll
            for (int a = 0; a < VPW; ++a)
                bf[a] = __builtin_bit_cast(bf8, valid ? p[toff + a * PLANE] : make_uint4(0, 0, 0, 0));
        } else {
            const bool valid = tap < 27;
            const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
            const int off = (dz * HY + dy) * HX + dx;
#pragma unroll
            for (int a = 0; a < VPW; ++a)
                bf[a] = __builtin_bit_cast(
                    bf8, valid ? halo[(hbase0 + a * HX * HY + off) * VS + (c0 / 8) * CS] : make_uint4(0, 0, 0, 0));
        }
#pragma unroll
        for (int a = 0; a < VPW; ++a)
#pragma unroll
            for (int c = 0; c < CG; ++c)
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[c], bf[a], acc[a][c], 0, 0, 0);
    }
    uint2 hv[VPW][CG];  // the stored BF16 values, kept for the pooling below
#pragma unroll
    for (int a = 0; a < VPW; ++a) {
        const int v = (wv + 4 * a) * 16 + r;
        const int gx = blockIdx.x * kTX + v % kTX, gy = blockIdx.y * kTY + (v / kTX) % kTY,
                  gz = bz * TZ + v / (kTX * kTY);
        const int64_t vox = ((int64_t)gz * Y + gy) * X + gx;
#pragma unroll
        for (int c = 0; c < CG; ++c) {
            const int o = (ct0 + c) * 16 + 4 * g;
            hv[a][c] = make_uint2(0u, 0u);
            if (o >= COUT) continue;
            float t[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                t[i] = acc[a][c][i] + bias[o + i];
                if (relu) t[i] = fmaxf(t[i], 0.f);
            }
            hv[a][c] = make_uint2(pack_bf2(t[0], t[1]), pack_bf2(t[2], t[3]));
            *reinterpret_cast<uint2*>(y + vox * COUT + o) = hv[a][c];
        }
    }
    if constexpr (VPW % 2 == 0) {
        if (pooled) {
            // pool2 of the output, from registers: the 2 x 2 x 2 cell of an
            // even voxel is lanes r, r ^ 1 (x), r ^ 8 (y), r ^ 9 of tiles a
            // and a + 1 (z); summed in pool2's order (x fastest, then y, z).
            // The lane exchanges stay inside 16-lane rows: DPP moves (a
            // quad permutation for r ^ 1, a rotation by 8 for r ^ 8), not
            // LDS permutes, which this kernel's B reads keep busy.
            const auto xor1 = [](float v) {
                return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF,
                                                                             0xF, false));
            };
            const auto xor8 = [](float v) {
                return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x128, 0xF,
                                                                             0xF, false));
            };
            const bool writer = (r & 1) == 0 && r < 8;
#pragma unroll
            for (int a = 0; a < VPW; a += 2) {
#pragma unroll
                for (int c = 0; c < CG; ++c) {
                    const int o = (ct0 + c) * 16 + 4 * g;
                    float q[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const uint32_t w0 = i < 2 ? hv[a][c].x : hv[a][c].y, w1 = i < 2 ? hv[a + 1][c].x : hv[a + 1][c].y;
                        const float z0 = __uint_as_float((i & 1) ? (w0 & 0xffff0000u) : (w0 << 16));
                        const float z1 = __uint_as_float((i & 1) ? (w1 & 0xffff0000u) : (w1 << 16));
                        const float c0[4] = {z0, xor1(z0), xor8(z0), xor8(xor1(z0))};
                        const float c1[4] = {z1, xor1(z1), xor8(z1), xor8(xor1(z1))};
                        float sum = 0.f;
#pragma unroll
                        for (int d = 0; d < 4; ++d) sum += c0[d];
#pragma unroll
                        for (int d = 0; d < 4; ++d) sum += c1[d];
                        q[i] = sum * 0.125f;
                    }
                    const int v = (wv + 4 * a) * 16 + r;
                    const int px = (blockIdx.x * kTX + v % kTX) / 2, py = (blockIdx.y * kTY + (v / kTX) % kTY) / 2,
                              pz = (bz * TZ + v / (kTX * kTY)) / 2;
                    if (writer && o < COUT)
                        *reinterpret_cast<uint2*>(pooled + (((int64_t)pz * (Y / 2) + py) * (X / 2) + px) * COUT + o) =
                            make_uint2(pack_bf2(q[0], q[1]), pack_bf2(q[2], q[3]));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Forward convolution for the coarse levels (CIN a multiple of 32), split
// along K. Block = one 8 x 8 voxel z slice (4 tiles of 16 voxels: rows 2t,
// 2t + 1) and CG 16-channel output tiles; its halo, 10 x 10 x 3 voxels x CIN,
// sits in LDS. Wave w computes every tile of the block over its own quarter
// of the K steps, so a wave issues 4 CG products per weight load (the plain
// kernel, one wave per voxel tile, issued CG and waited on each weight load
// in turn: the coarse levels have few blocks and long K loops, and that
// latency bounded them). The weights come from memory in groups of U steps,
// the next group in flight while the current one computes. The four
// partial sums meet in LDS (over the halo, after a barrier); wave w adds up
// tile w and applies the epilogue of conv3_fwd.
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int CG>
__global__ __launch_bounds__(kThreads) void conv3_fwd_splitk(const uint16_t* __restrict__ x,
                                                             const uint16_t* __restrict__ w,
                                                             const float* __restrict__ bias, uint16_t* __restrict__ y,
                                                             int X, int Y, int Z, int relu) {
    constexpr int HX = kTX + 2, HY = kTY + 2, HZ = 3;
    constexpr int CH = CIN / 8;
    constexpr int KP = 27 * CIN;  // a multiple of 32: no padding
    constexpr int KSTEPS = KP / 32;
    constexpr int KW = (KSTEPS + 3) / 4;  // K steps per wave
    constexpr int U = 4;                  // weight steps per load group
    constexpr int NCT = cpad(COUT) / 16;
    constexpr int HALO = HX * HY * HZ * CH;       // uint4 chunks
    constexpr int RED = 4 * 4 * CG * 64;          // 4 waves x 4 tiles x CG x 64 lanes (float4 each)
    static_assert(CIN % 32 == 0 && NCT % CG == 0, "tile shape");
    __shared__ uint4 lds[HALO > RED ? HALO : RED];
    const int ng = NCT / CG;
    const int bz = blockIdx.z / ng, ct0 = (blockIdx.z % ng) * CG;
    stage_halo<HX, HY, HZ, CH>(x, X, Y, Z, blockIdx.x * kTX - 1, blockIdx.y * kTY - 1, bz - 1, lds);
    stage_wait();
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    // B chunk of tile t at step s: this lane's base (voxel row r >> 3,
    // column r & 7 of tile 0; channel chunk g) + tap offset + 2 t HX CH
    const uint4* lb = lds + ((r >> 3) * HX + (r & 7)) * CH + g;
    f4 acc[4][CG];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[t][c] = f4{0.f, 0.f, 0.f, 0.f};
    const int s0 = wv * KW, s1 = s0 + KW < KSTEPS ? s0 + KW : KSTEPS;
    const uint16_t* wr = w + ((ct0 * 16) + r) * KP + g * 8;
    bf8 A[U][CG], An[U][CG];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int c = 0; c < CG; ++c)
            if (s0 + u < s1) An[u][c] = *reinterpret_cast<const bf8*>(wr + c * 16 * KP + (s0 + u) * 32);
    for (int sg = s0; sg < s1; sg += U) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < CG; ++c) A[u][c] = An[u][c];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < CG; ++c)
                if (sg + U + u < s1) An[u][c] = *reinterpret_cast<const bf8*>(wr + c * 16 * KP + (sg + U + u) * 32);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int st = sg + u;
            if (st < s1) {
                const int kb = st * 32, t0 = kb / CIN, cb = kb % CIN;
                const uint4* pb = lb + (((t0 / 9) * HY + (t0 / 3) % 3) * HX + t0 % 3) * CH + cb / 8;
                bf8 bf[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) bf[t] = __builtin_bit_cast(bf8, pb[2 * t * HX * CH]);
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int c = 0; c < CG; ++c)
                        acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[u][c], bf[t], acc[t][c], 0, 0, 0);
            }
        }
    }
    // partial sums: wave wv's tile t, group c at red[((t CG + c) 4 + wv) 64 + lane]
    __syncthreads();  // every wave is done with the halo
    f4* red = reinterpret_cast<f4*>(lds);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int c = 0; c < CG; ++c) red[((t * CG + c) * 4 + wv) * 64 + lane] = acc[t][c];
    __syncthreads();
    const int t = wv;  // this wave finishes tile wv
    const int gx = blockIdx.x * kTX + (r & 7), gy = blockIdx.y * kTY + 2 * t + (r >> 3);
    const int64_t vox = ((int64_t)bz * Y + gy) * X + gx;
#pragma unroll
    for (int c = 0; c < CG; ++c) {
        const int o = (ct0 + c) * 16 + 4 * g;
        // fixed order over the four K quarters
        f4 sum = red[((t * CG + c) * 4 + 0) * 64 + lane];
#pragma unroll
        for (int q = 1; q < 4; ++q) sum += red[((t * CG + c) * 4 + q) * 64 + lane];
        if (o >= COUT) continue;
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            v[i] = sum[i] + bias[o + i];
            if (relu) v[i] = fmaxf(v[i], 0.f);
        }
        *reinterpret_cast<uint2*>(y + vox * COUT + o) = make_uint2(pack_bf2(v[0], v[1]), pack_bf2(v[2], v[3]));
    }
}

// ---------------------------------------------------------------------------
// Convolution of an upsampled tensor, on the coarse grid. y = conv3(up2(d))
// with up2 the nearest-neighbour 2x upsampling of the S^3 tensor d: along
// each axis fine voxel 2c + p (parity p) reads up2(d) at 2c + p + t - 1,
// t = 0, 1, 2, which is coarse voxel c - 1 + p + e with e = (p + t + 1) / 2 - p
// (0, 1, 1 for p = 0; 0, 0, 1 for p = 1), zero outside on both grids. So each
// of the 8 parities is a 2x2x2 convolution of d whose weights are sums of
// the original taps (combine_up_weights): 8 CIN products per output voxel
// instead of 27 CIN, and no upsampled tensor.
// Block = 4 x 4 x TZ coarse voxels (their halo, 6 x 6 x (TZ + 2) x CIN, in
// LDS) for all 8 parities; a 16-voxel tile is 4 x 4 coarse voxels at one z
// for one parity; wave w takes parities 2w and 2w + 1. K = 8 CIN, k = e3 CIN
// + c, e3 = (ez 2 + ey) 2 + ex. Weights W'[p][o][k] (combine_up_weights).
// Epilogue as conv3_fwd (bias, ReLU, BF16) at fine voxel 2c + p of the
// (2S)^3 output; with `skip` the stored value is BF16(BF16(that) + skip),
// the U-Net's skip connection rounded as a separate add would round it.
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int TZ, int CG>
__global__ __launch_bounds__(kThreads) void upconv_fwd(const uint16_t* __restrict__ d, const uint16_t* __restrict__ w,
                                                       const float* __restrict__ bias,
                                                       const uint16_t* __restrict__ skip, uint16_t* __restrict__ y,
                                                       int S, int relu) {
    constexpr int HX = 6, HY = 6, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int CH = CIN / 8;
    constexpr int KP = 8 * CIN;
    constexpr int KSTEPS = KP / 32;
    constexpr int NCT = cpad(COUT) / 16;
    constexpr int VPW = 2 * TZ;  // two parities x TZ z slices
    static_assert(CIN % 32 == 0 && NCT % CG == 0, "tile shape");
    // With every output channel in the block (CG = NCT) the block's output,
    // 8 x 8 x 2 TZ fine voxels, is staged in LDS and leaves in whole rows
    // (each lane's own voxels are 2 apart along x), the skip read the same way.
    constexpr bool STAGE = CG == NCT;
    constexpr int NF = 64 * 2 * TZ;  // fine voxels of the block
    __shared__ uint4 halo[HV * CH];
    __shared__ uint2 ostage[STAGE ? NF * COUT / 4 : 1];
    const int zb = S / TZ;
    const int bz = blockIdx.z % zb, ct0 = (blockIdx.z / zb) * CG;
    const int cx0 = blockIdx.x * 4, cy0 = blockIdx.y * 4, cz0 = bz * TZ;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int lx = r & 3, ly = r >> 2;
    const int F = 2 * S;
    // tile a: parity 2 wv + (a & 1), coarse z slice a >> 1; its output voxel
    // for this lane
    const auto vox_of = [&](int a) {
        const int p = 2 * wv + (a & 1);
        const int fx = 2 * (cx0 + lx) + (p & 1), fy = 2 * (cy0 + ly) + ((p >> 1) & 1),
                  fz = 2 * (cz0 + (a >> 1)) + (p >> 2);
        return ((int64_t)fz * F + fy) * F + fx;
    };
    stage_halo<HX, HY, HZ, CH>(d, S, S, S, cx0 - 1, cy0 - 1, cz0 - 1, halo);
    stage_wait();
    __syncthreads();
    f4 acc[VPW][CG];
#pragma unroll
    for (int a = 0; a < VPW; ++a)
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[a][c] = f4{0.f, 0.f, 0.f, 0.f};
    // The halo voxel of window offset 0 for this lane's coarse voxel in tile
    // a (parity p = 2 wv + (a & 1): px = a & 1, py = wv & 1, pz = wv >> 1) is
    // hb0 + (a >> 1) HX HY + (a & 1); with CIN a multiple of 32 the four k
    // groups of a step are four channel chunks of one window offset, so every
    // LDS read is this lane's base plus a compile-time offset.
    const int hb0 = (wv >> 1) * HX * HY + (ly + (wv & 1)) * HX + lx;
    const uint4* lb = halo + hb0 * CH + g;
    // A fragments of the wave's two parities, one step ahead
    const uint16_t* wr0 = w + ((int64_t)(2 * wv) * cpad(COUT) + ct0 * 16 + r) * KP + g * 8;
    const uint16_t* wr1 = wr0 + (int64_t)cpad(COUT) * KP;
    bf8 af[2][CG], an[2][CG];
#pragma unroll
    for (int c = 0; c < CG; ++c) {
        an[0][c] = *reinterpret_cast<const bf8*>(wr0 + c * 16 * KP);
        an[1][c] = *reinterpret_cast<const bf8*>(wr1 + c * 16 * KP);
    }
#pragma unroll 2
    for (int s = 0; s < KSTEPS; ++s) {
        const int kb = s * 32, e3 = kb / CIN;  // window offset and chunk of k group 0
        const uint4* pb = lb + (((e3 >> 2) * HY + ((e3 >> 1) & 1)) * HX + (e3 & 1)) * CH + (kb % CIN) / 8;
#pragma unroll
        for (int c = 0; c < CG; ++c) {
            af[0][c] = an[0][c];
            af[1][c] = an[1][c];
        }
        if (s + 1 < KSTEPS) {
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                an[0][c] = *reinterpret_cast<const bf8*>(wr0 + c * 16 * KP + (s + 1) * 32);
                an[1][c] = *reinterpret_cast<const bf8*>(wr1 + c * 16 * KP + (s + 1) * 32);
            }
        }
        bf8 bf[VPW];
#pragma unroll
        for (int a = 0; a < VPW; ++a) bf[a] = __builtin_bit_cast(bf8, pb[((a >> 1) * HX * HY + (a & 1)) * CH]);
#pragma unroll
        for (int a = 0; a < VPW; ++a)
#pragma unroll
            for (int c = 0; c < CG; ++c)
                acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a & 1][c], bf[a], acc[a][c], 0, 0, 0);
    }
    if constexpr (STAGE) {
        // BF16(act(y)) into the block's tile (fine voxel (lx, ly, lz) of the
        // block at (lz 8 + ly) 8 + lx, COUT channels each) ...
#pragma unroll
        for (int a = 0; a < VPW; ++a) {
            const int p = 2 * wv + (a & 1);
            const int lv = ((2 * (a >> 1) + (p >> 2)) * 8 + 2 * ly + ((p >> 1) & 1)) * 8 + 2 * lx + (p & 1);
#pragma unroll
            for (int c = 0; c < CG; ++c) {
                const int o = (ct0 + c) * 16 + 4 * g;
                float t[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    t[i] = acc[a][c][i] + bias[o + i];
                    if (relu) t[i] = fmaxf(t[i], 0.f);
                }
                ostage[(lv * COUT + o) / 4] = make_uint2(pack_bf2(t[0], t[1]), pack_bf2(t[2], t[3]));
            }
        }
        __syncthreads();
        // ... then out by 16-byte chunks, consecutive chunks along the fine
        // rows, with the skip added: BF16(BF16(y) + skip), the rounding of a
        // separate add
        constexpr int CPV = COUT / 8;  // chunks per voxel
        const uint4* st4 = reinterpret_cast<const uint4*>(ostage);
        for (int q = threadIdx.x; q < NF * CPV; q += kThreads) {
            const int lv = q / CPV, ch = q % CPV;
            const int fxl = lv & 7, fyl = (lv >> 3) & 7, fzl = lv >> 6;
            const int64_t gv = ((int64_t)(2 * cz0 + fzl) * F + 2 * cy0 + fyl) * F + 2 * cx0 + fxl;
            uint4 v = st4[q];
            if (skip) {
                const uint4 sk = *reinterpret_cast<const uint4*>(skip + gv * COUT + ch * 8);
                const uint32_t vw[4] = {v.x, v.y, v.z, v.w}, sw[4] = {sk.x, sk.y, sk.z, sk.w};
                uint32_t ow[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float lo = __uint_as_float(vw[k] << 16) + __uint_as_float(sw[k] << 16);
                    const float hi = __uint_as_float(vw[k] & 0xffff0000u) + __uint_as_float(sw[k] & 0xffff0000u);
                    ow[k] = pack_bf2(lo, hi);
                }
                v = make_uint4(ow[0], ow[1], ow[2], ow[3]);
            }
            *reinterpret_cast<uint4*>(y + gv * COUT + ch * 8) = v;
        }
        return;
    }
    // the skip values, all loads issued before the first is needed (here,
    // not before the halo: holding them through the K loop took 16 registers
    // and a wave per SIMD)
    uint2 skv[VPW][CG];
#pragma unroll
    for (int a = 0; a < VPW; ++a)
#pragma unroll
        for (int c = 0; c < CG; ++c) {
            const int o = (ct0 + c) * 16 + 4 * g;
            skv[a][c] = (skip && o < COUT) ? *reinterpret_cast<const uint2*>(skip + vox_of(a) * COUT + o)
                                           : make_uint2(0u, 0u);
        }
#pragma unroll
    for (int a = 0; a < VPW; ++a) {
#pragma unroll
        for (int c = 0; c < CG; ++c) {
            const int o = (ct0 + c) * 16 + 4 * g;
            if (o >= COUT) continue;
            float t[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                t[i] = acc[a][c][i] + bias[o + i];
                if (relu) t[i] = fmaxf(t[i], 0.f);
            }
            uint2 pk = make_uint2(pack_bf2(t[0], t[1]), pack_bf2(t[2], t[3]));
            if (skip) {
                // BF16(BF16(y) + skip): the rounding of a separate add
                const uint2 sk = skv[a][c];
                const float u0 = __uint_as_float(pk.x << 16) + __uint_as_float(sk.x << 16);
                const float u1 = __uint_as_float(pk.x & 0xffff0000u) + __uint_as_float(sk.x & 0xffff0000u);
                const float u2 = __uint_as_float(pk.y << 16) + __uint_as_float(sk.y << 16);
                const float u3 = __uint_as_float(pk.y & 0xffff0000u) + __uint_as_float(sk.y & 0xffff0000u);
                pk = make_uint2(pack_bf2(u0, u1), pack_bf2(u2, u3));
            }
            *reinterpret_cast<uint2*>(y + vox_of(a) * COUT + o) = pk;
        }
    }
}

// W'[p][o][e3 CIN + c] = BF16(sum of the FP32 master weights m[o][c][tap]
// over the taps t with e(t) = e3 on every axis), for the 8 parities p; rows
// o >= cout are zero.
__global__ void combine_up_weights(const float* __restrict__ m, int cin, int cout, uint16_t* __restrict__ w) {
    const int coutp = cpad(cout), kp = 8 * cin;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)8 * coutp * kp) return;
    const int p = (int)(i / ((int64_t)coutp * kp));
    const int o = (int)((i / kp) % coutp), k = (int)(i % kp);
    const int e3 = k / cin, c = k % cin;
    const int pp[3] = {p & 1, (p >> 1) & 1, p >> 2}, ee[3] = {e3 & 1, (e3 >> 1) & 1, e3 >> 2};
    float sum = 0.f;
    if (o < cout) {
        for (int tz = 0; tz < 3; ++tz)
            for (int ty = 0; ty < 3; ++ty)
                for (int tx = 0; tx < 3; ++tx) {
                    const int t[3] = {tx, ty, tz};
                    bool hit = true;
                    for (int a = 0; a < 3; ++a) hit = hit && ((pp[a] + t[a] + 1) / 2 - pp[a] == ee[a]);
                    if (hit) sum += m[((int64_t)o * cin + c) * 27 + (tz * 3 + ty) * 3 + tx];
                }
    }
    w[i] = f2bf(sum);
}

// ---------------------------------------------------------------------------
// Weight gradient dW[o][tap CIN + c] = sum_v dy[v][o] x[v + d(tap)][c] over
// one block of 8 x 8 x TZ voxels, as MFMAs with the voxels on the K axis:
// A = dy^T (16 channels x 32 voxels), B = shifted x (32 voxels x 16 k). The
// block's halo of x and its dy tile sit in LDS; fragments gather 8 voxels per
// lane. Each block writes its partial sums to part[block][COUTP][KP].
// ---------------------------------------------------------------------------
template <int CIN, int COUT, int TZ>
__global__ __launch_bounds__(kThreads) void conv3_wgrad(const uint16_t* __restrict__ x,
                                                        const uint16_t* __restrict__ dy, float* __restrict__ part,
                                                        int X, int Y, int Z) {
    constexpr int HX = kTX + 2, HY = kTY + 2, HZ = TZ + 2, HV = HX * HY * HZ;
    constexpr int KP = kpad(CIN);
    constexpr int COUTP = cpad(COUT);
    constexpr int NB = kTX * kTY * TZ;   // voxels in the block
    constexpr int NOT = COUTP / 16;      // output-channel tiles
    constexpr int NKT = KP / 16;         // k tiles
    constexpr int NT = NOT * NKT;        // (o, k) tiles
    constexpr int TPW = (NT + 3) / 4;    // tiles per wave
    __shared__ uint16_t hx_s[HV * CIN];
    __shared__ uint16_t dy_s[NB * COUTP];
    const int x0 = blockIdx.x * kTX - 1, y0 = blockIdx.y * kTY - 1, z0 = blockIdx.z * TZ - 1;
    for (int q = threadIdx.x; q < HV * (CIN / 8); q += kThreads) {
        const int ch = q % (CIN / 8), hv = q / (CIN / 8);
        const int hxx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
        const int gx = x0 + hxx, gy = y0 + hy, gz = z0 + hz;
        uint4 val = make_uint4(0, 0, 0, 0);
        if (gx >= 0 && gx < X && gy >= 0 && gy < Y && gz >= 0 && gz < Z)
            val = *reinterpret_cast<const uint4*>(x + (((int64_t)gz * Y + gy) * X + gx) * CIN + ch * 8);
        *reinterpret_cast<uint4*>(&hx_s[hv * CIN + ch * 8]) = val;
    }
    for (int q = threadIdx.x; q < NB * COUTP; q += kThreads) {
        const int o = q % COUTP, v = q / COUTP;
        const int gx = blockIdx.x * kTX + v % kTX, gy = blockIdx.y * kTY + (v / kTX) % kTY,
                  gz = blockIdx.z * TZ + v / (kTX * kTY);
        dy_s[q] = o < COUT ? dy[(((int64_t)gz * Y + gy) * X + gx) * COUT + o] : (uint16_t)0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    for (int tt = 0; tt < TPW; ++tt) {
        const int t = wv + 4 * tt;
        if (t >= NT) break;
        const int ot = t / NKT, kt = t % NKT;
        const int o = ot * 16 + r;
        const int k = kt * 16 + r;
        const int tap = k / CIN, c = k % CIN;
        const bool kvalid = tap < 27;
        const int dz = tap / 9, dyy = (tap / 3) % 3, dx = tap % 3;
        f4 acc = f4{0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < NB / 32; ++s) {
            uint16_t av[8], bv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int v = s * 32 + g * 8 + e;
                av[e] = dy_s[v * COUTP + o];
                const int lx = v % kTX, ly = (v / kTX) % kTY, lz = v / (kTX * kTY);
                bv[e] = kvalid ? hx_s[(((lz + dz) * HY + ly + dyy) * HX + lx + dx) * CIN + c] : (uint16_t)0;
            }
            bf8 af, bfv;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                af[e] = __builtin_bit_cast(__bf16, av[e]);
                bfv[e] = __builtin_bit_cast(__bf16, bv[e]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfv, acc, 0, 0, 0);
        }
        // acc[i] = dW[ot 16 + 4 g + i][kt 16 + r]
        const int64_t blk = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        float* dst = part + blk * COUTP * KP;
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[(ot * 16 + 4 * g + i) * KP + kt * 16 + r] = acc[i];
    }
}

// out[i] = sum over blocks of part[b][i] (fixed order: deterministic)
__global__ void reduce_blocks(const float* __restrict__ part, int64_t nblk, int64_t per, float* __restrict__ out,
                              int accumulate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per) return;
    float s = 0.f;
    for (int64_t b = 0; b < nblk; ++b) s += part[b * per + i];
    out[i] = accumulate ? out[i] + s : s;
}

// W'[c][tap' COUT + o] = W[o][(26 - tap') CIN + c]: the kernel of the input
// gradient (a convolution of dy with the flipped, transposed weights).
__global__ void flip_weights(const uint16_t* __restrict__ w, int cin, int cout, uint16_t* __restrict__ wf) {
    const int kp = kpad(cin), kpf = kpad(cout), coutp_f = cpad(cin);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)coutp_f * kpf) return;
    const int c = (int)(i / kpf), kf = (int)(i % kpf);
    const int tapf = kf / cout, o = kf % cout;
    uint16_t v = 0;
    if (c < cin && tapf < 27) v = w[o * kp + (26 - tapf) * cin + c];
    wf[i] = v;
}

// FP32 master weights (PyTorch-like [o][c][tap] order, see param layout) ->
// BF16 GEMM layout W[o][tap cin + c], zero padded.
__global__ void pack_weights(const float* __restrict__ m, int cin, int cout, uint16_t* __restrict__ w) {
    const int kp = kpad(cin), coutp = cpad(cout);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)coutp * kp) return;
    const int o = (int)(i / kp), k = (int)(i % kp);
    const int tap = k / cin, c = k % cin;
    uint16_t v = 0;
    if (o < cout && tap < 27) v = f2bf(m[((int64_t)o * cin + c) * 27 + tap]);
    w[i] = v;
}

// GEMM-layout gradient -> master layout ([o][c][tap])
__global__ void unpack_grad(const float* __restrict__ gk, int cin, int cout, float* __restrict__ gm, int accumulate) {
    const int kp = kpad(cin);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)cout * cin * 27) return;
    const int tap = (int)(i % 27), c = (int)((i / 27) % cin), o = (int)(i / (27 * cin));
    const float v = gk[o * kp + tap * cin + c];
    gm[i] = accumulate ? gm[i] + v : v;
}

// dy *= [y > 0] in place (ReLU backward from the layer's own output)
__global__ void relu_mask(uint16_t* __restrict__ dy, const uint16_t* __restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(bf2f(y[i]) > 0.f)) dy[i] = 0;
}

// db[o] = sum_v dy[v][o] (one block per channel, fixed order)
__global__ __launch_bounds__(256) void bias_grad(const uint16_t* __restrict__ dy, int64_t nvox, int cout,
                                                 float* __restrict__ db, int accumulate) {
    __shared__ float red[256];
    const int o = blockIdx.x;
    float s = 0.f;
    for (int64_t v = threadIdx.x; v < nvox; v += 256) s += bf2f(dy[v * cout + o]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[o] = accumulate ? db[o] + red[0] : red[0];
}

// 8 bf16 <-> 8 floats
__device__ __forceinline__ void unpack8(uint4 u, float (&f)[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        f[2 * q] = __uint_as_float(w[q] << 16);
        f[2 * q + 1] = __uint_as_float(w[q] & 0xffff0000u);
    }
}

__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = (uint32_t)f2bf(f[2 * q]) | ((uint32_t)f2bf(f[2 * q + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// 2 x 2 x 2 average pool (out grid X x Y x Z, input 2X x 2Y x 2Z); one thread
// per 8 channels of a coarse voxel
__global__ void pool2(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int X, int Y, int Z, int C) {
    const int C8 = C / 8;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)X * Y * Z * C8) return;
    const int c8 = (int)(i % C8);
    const int64_t v = i / C8;
    const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((int64_t)X * Y));
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int d = 0; d < 8; ++d) {
        const int xi = 2 * x + (d & 1), yi = 2 * y + ((d >> 1) & 1), zi = 2 * z + (d >> 2);
        float f[8];
        unpack8(*reinterpret_cast<const uint4*>(in + (((int64_t)zi * 2 * Y + yi) * 2 * X + xi) * C + c8 * 8), f);
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] += f[e];
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) s[e] *= 0.125f;
    *reinterpret_cast<uint4*>(out + v * C + c8 * 8) = pack8(s);
}

// backward of pool2: every fine voxel gets dy / 8 (one thread per 8 channels
// of a fine voxel)
__global__ void pool2_back(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx, int X, int Y, int Z, int C) {
    const int C8 = C / 8;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)8 * X * Y * Z * C8) return;
    const int c8 = (int)(i % C8);
    const int64_t v = i / C8;
    const int xi = (int)(v % (2 * X)), yi = (int)((v / (2 * X)) % (2 * Y)), zi = (int)(v / ((int64_t)4 * X * Y));
    float f[8];
    unpack8(*reinterpret_cast<const uint4*>(dy + (((int64_t)(zi / 2) * Y + yi / 2) * X + xi / 2) * C + c8 * 8), f);
#pragma unroll
    for (int e = 0; e < 8; ++e) f[e] *= 0.125f;
    *reinterpret_cast<uint4*>(dx + v * C + c8 * 8) = pack8(f);
}

// nearest x2 upsample (out grid 2X x 2Y x 2Z), 8 channels per thread
__global__ void up2(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int X, int Y, int Z, int C) {
    const int C8 = C / 8;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)8 * X * Y * Z * C8) return;
    const int c8 = (int)(i % C8);
    const int64_t v = i / C8;
    const int xi = (int)(v % (2 * X)), yi = (int)((v / (2 * X)) % (2 * Y)), zi = (int)(v / ((int64_t)4 * X * Y));
    *reinterpret_cast<uint4*>(out + v * C + c8 * 8) =
        *reinterpret_cast<const uint4*>(in + (((int64_t)(zi / 2) * Y + yi / 2) * X + xi / 2) * C + c8 * 8);
}

// backward of up2: sum of the 8 fine gradients, 8 channels per thread
__global__ void up2_back(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx, int X, int Y, int Z, int C) {
    const int C8 = C / 8;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)X * Y * Z * C8) return;
    const int c8 = (int)(i % C8);
    const int64_t v = i / C8;
    const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((int64_t)X * Y));
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int d = 0; d < 8; ++d) {
        const int xi = 2 * x + (d & 1), yi = 2 * y + ((d >> 1) & 1), zi = 2 * z + (d >> 2);
        float f[8];
        unpack8(*reinterpret_cast<const uint4*>(dy + (((int64_t)zi * 2 * Y + yi) * 2 * X + xi) * C + c8 * 8), f);
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] += f[e];
    }
    *reinterpret_cast<uint4*>(dx + v * C + c8 * 8) = pack8(s);
}

// out = a + b (bf16), 8 elements per thread (n a multiple of 8)
__global__ void add_bf(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b, uint16_t* __restrict__ out,
                       int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n / 8) return;
    float fa[8], fb[8];
    unpack8(reinterpret_cast<const uint4*>(a)[i], fa);
    unpack8(reinterpret_cast<const uint4*>(b)[i], fb);
#pragma unroll
    for (int e = 0; e < 8; ++e) fa[e] = fa[e] + fb[e];
    reinterpret_cast<uint4*>(out)[i] = pack8(fa);
}

// The increment r = p1 - p0 of one real substep (p1 = the field after one
// LOD substep from p0 with the uptake map held fixed) and its scale. Within
// one call the substep is affine in the field (Setboundary, the implicit
// uptake p / (1 - rate dt) and the three Thomas sweeps), so the field after
// K substeps is p0 + (1 + A + ... + A^(K-1)) r with A the substep's linear
// part: a linear map of r that depends only on the uptake map. The network
// sees r in units of its interior RMS s and predicts the K-substep increment
// in the same units, so its error scales with how far the field moves.
//
// incr_sums: per-block partial sums (double) of r^2 over the interior (the
// voxels two or more from every face) and of p0^2 over the grid. Each
// thread takes 4 consecutive voxels of an x line (N is a multiple of 128, so
// N^3 / 4 < 2^31 up to N = 1024).
__global__ __launch_bounds__(256) void incr_sums(const float* __restrict__ p0, const float* __restrict__ p1, int N,
                                                 double* __restrict__ part) {
    __shared__ double r0[256], r1[256];
    const uint32_t q = (uint32_t)N / 4u, nq = (uint32_t)N * (uint32_t)N * q;
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < nq; c += gridDim.x * 256u) {
        const uint32_t line = c / q, x0 = (c - line * q) * 4u;
        const uint32_t y = line % (uint32_t)N, z = line / (uint32_t)N;
        const float4 a = reinterpret_cast<const float4*>(p0)[c];
        const float4 b = reinterpret_cast<const float4*>(p1)[c];
        const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
        const bool row_in = y >= 2u && z >= 2u && y < (uint32_t)N - 2u && z < (uint32_t)N - 2u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double ad = (double)av[k];
            s1 += ad * ad;
            const uint32_t x = x0 + (uint32_t)k;
            if (row_in && x >= 2u && x < (uint32_t)N - 2u) {
                const double d = (double)bv[k] - ad;
                s0 += d * d;
            }
        }
    }
    r0[threadIdx.x] = s0;
    r1[threadIdx.x] = s1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r0[0];
        part[2 * blockIdx.x + 1] = r1[0];
    }
}

// One block, fixed order: norm[0] = s = sqrt(sum r^2 / interior count),
// norm[1] = 1 / s (0 when s = 0: then the prediction is p0 itself);
// sums[0..1] = the two sums.
__global__ __launch_bounds__(256) void incr_finalize(const double* __restrict__ part, int nblk, double inv_count,
                                                     float* __restrict__ norm, double* __restrict__ sums,
                                                     double* host_sums) {
    __shared__ double r0[256], r1[256];
    double s0 = 0.0, s1 = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        s0 += part[2 * b];
        s1 += part[2 * b + 1];
    }
    r0[threadIdx.x] = s0;
    r1[threadIdx.x] = s1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double s = sqrt(r0[0] * inv_count);
        norm[0] = (float)s;
        norm[1] = s > 0.0 ? (float)(1.0 / s) : 0.f;
        sums[0] = r0[0];
        sums[1] = r1[0];
        if (host_sums) {
            // the two sums straight into pinned host memory as well (the
            // host reads them after an event, without a copy in the stream)
            host_sums[0] = r0[0];
            host_sums[1] = r1[0];
        }
    }
}

// Input features on the half-resolution grid (X = N / 2), 16 channels per
// coarse voxel: q = (p1 - p0) / s at its 8 fine voxels rounded to BF16
// (channels 0..7), then the uptake rate times rate_scale at the same 8
// (8..15). Fine voxel d = (z & 1) 4 + (y & 1) 2 + (x & 1).
__global__ void features(const float* __restrict__ p0, const float* __restrict__ p1, const float* __restrict__ rate,
                         int N, const float* __restrict__ norm, float rate_scale, uint16_t* __restrict__ f) {
    const int X = N / 2;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (int64_t)X * X * X) return;
    const int x = (int)(v % X), y = (int)((v / X) % X), z = (int)(v / ((int64_t)X * X));
    const float inv_s = norm[1];
    // the fine voxels come in x pairs: four 8-byte loads per field
    float qv[8], rv[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int yi = 2 * y + (q & 1), zi = 2 * z + (q >> 1);
        const int64_t o = ((int64_t)zi * N + yi) * N + 2 * x;
        const float2 a = *reinterpret_cast<const float2*>(p0 + o);
        const float2 b = *reinterpret_cast<const float2*>(p1 + o);
        const float2 r = *reinterpret_cast<const float2*>(rate + o);
        qv[2 * q] = (b.x - a.x) * inv_s;
        qv[2 * q + 1] = (b.y - a.y) * inv_s;
        rv[2 * q] = r.x * rate_scale;
        rv[2 * q + 1] = r.y * rate_scale;
    }
    uint4 pk[2];
    pk[0] = make_uint4(pack_bf2(qv[0], qv[1]), pack_bf2(qv[2], qv[3]), pack_bf2(qv[4], qv[5]), pack_bf2(qv[6], qv[7]));
    pk[1] = make_uint4(pack_bf2(rv[0], rv[1]), pack_bf2(rv[2], rv[3]), pack_bf2(rv[4], rv[5]), pack_bf2(rv[6], rv[7]));
    reinterpret_cast<uint4*>(f + v * 16)[0] = pk[0];
    reinterpret_cast<uint4*>(f + v * 16)[1] = pk[1];
}

// Full-resolution output: out[fine] = p0[fine] + s out_scale head[coarse][sub],
// sub = (zi & 1) 4 + (yi & 1) 2 + (xi & 1), in the interior; the two outer
// layers, which every substep first resets to the boundary value, take p1
// (out may alias p0 or p1: each thread reads a voxel before writing it).
// One thread per coarse voxel: its 8 head values in one 16-byte load, its
// fine voxels as 4 x-pairs.
__global__ void head_out(const float* p0, const float* p1, const uint16_t* __restrict__ head, int N,
                         const float* __restrict__ norm, float out_scale, float* out) {
    const int X = N / 2;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)X * X * X) return;
    const int x = (int)(c % X), y = (int)((c / X) % X), z = (int)(c / ((int64_t)X * X));
    const float scale = norm[0] * out_scale;
    float h[8];
    unpack8(*reinterpret_cast<const uint4*>(head + c * 8), h);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int yi = 2 * y + (q & 1), zi = 2 * z + (q >> 1);
        const int64_t o = ((int64_t)zi * N + yi) * N + 2 * x;
        const bool row_edge = yi < 2 || zi < 2 || yi >= N - 2 || zi >= N - 2;
        const int x0 = 2 * x, x1 = 2 * x + 1;
        const bool e0 = row_edge || x0 < 2 || x0 >= N - 2, e1 = row_edge || x1 < 2 || x1 >= N - 2;
        // p1 only on the outer layers, p0 only inside: each read where used
        const float2 a = (e0 && e1) ? make_float2(0.f, 0.f) : *reinterpret_cast<const float2*>(p0 + o);
        const float2 b = (e0 || e1) ? *reinterpret_cast<const float2*>(p1 + o) : make_float2(0.f, 0.f);
        float2 r;
        r.x = e0 ? b.x : a.x + scale * h[q * 2];
        r.y = e1 ? b.y : a.y + scale * h[q * 2 + 1];
        *reinterpret_cast<float2*>(out + o) = r;
    }
}

// Loss = mean over interior fine voxels of (head - (target - p0) / (s
// out_scale))^2: the squared error of the predicted increment in units of s
// out_scale; writes dL/dhead (bf16, coarse x 8) and the per-block partial
// sums of the loss.
__global__ __launch_bounds__(256) void mse_head(const float* __restrict__ p0, const uint16_t* __restrict__ head,
                                                const float* __restrict__ target, int N,
                                                const float* __restrict__ norm, float out_scale, float inv_count,
                                                uint16_t* __restrict__ dhead, float* __restrict__ loss_part) {
    __shared__ float red[256];
    const int X = N / 2;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // coarse voxel
    const float inv = norm[1] / out_scale;
    float l = 0.f;
    if (c < (int64_t)X * X * X) {
        const int x = (int)(c % X), y = (int)((c / X) % X), z = (int)(c / ((int64_t)X * X));
        uint16_t gh[8];
        for (int sub = 0; sub < 8; ++sub) {
            const int xi = 2 * x + (sub & 1), yi = 2 * y + ((sub >> 1) & 1), zi = 2 * z + (sub >> 2);
            const int64_t o = ((int64_t)zi * N + yi) * N + xi;
            const bool edge = xi < 2 || yi < 2 || zi < 2 || xi >= N - 2 || yi >= N - 2 || zi >= N - 2;
            const float h = bf2f(head[c * 8 + sub]);
            const float e = edge ? 0.f : h - (target[o] - p0[o]) * inv;
            l += e * e;
            gh[sub] = f2bf(2.f * e * inv_count);
        }
        uint4 pk;
        pk.x = (uint32_t)gh[0] | ((uint32_t)gh[1] << 16);
        pk.y = (uint32_t)gh[2] | ((uint32_t)gh[3] << 16);
        pk.z = (uint32_t)gh[4] | ((uint32_t)gh[5] << 16);
        pk.w = (uint32_t)gh[6] | ((uint32_t)gh[7] << 16);
        *reinterpret_cast<uint4*>(dhead + c * 8) = pk;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * inv_count;
}

// Adam with bias correction (b1, b2, eps), step t >= 1.
__global__ void adam(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m1,
                     float* __restrict__ m2, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * gr;
    const float b = b2 * m2[i] + (1.f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void f32_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

__global__ void bf16_to_f32(const uint16_t* __restrict__ a, float* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = bf2f(a[i]);
}

inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Block depth per input width: halo LDS stays within ~77 KB.
constexpr int tz_for(int cin) { return cin <= 16 ? 8 : (cin <= 32 ? 4 : (cin == 64 ? 2 : 1)); }

template <int CIN, int COUT>
hipError_t conv_t(const uint16_t* x, const uint16_t* w, const float* b, uint16_t* y, int X, int Y, int Z, int relu,
                  hipStream_t st, uint16_t* pooled, bool* pooled_done) {
    constexpr int TZ = tz_for(CIN);
    constexpr int NCT = cpad(COUT) / 16;
    if (pooled_done) *pooled_done = false;
    if (X % kTX || Y % kTY || Z % TZ) return hipErrorInvalidValue;
    if constexpr (CIN % 32 == 0) {
        // coarse levels: K split over the waves up to 32^3 grids, and at 64^3
        // unless the output is to be pooled too (the plain kernel's blocks
        // pool their own output there: 34.8 + 6.0 -> 36 us, against the split
        // kernel's 31 us alone)
        if ((X <= 32 && Y <= 32) || (X <= 64 && Y <= 64 && !pooled)) {
            constexpr int CGS = NCT % 2 == 0 ? 2 : 1;
            hipLaunchKernelGGL((conv3_fwd_splitk<CIN, COUT, CGS>), dim3(X / kTX, Y / kTY, Z * (NCT / CGS)),
                               dim3(kThreads), 0, st, x, w, b, y, X, Y, Z, relu);
            return hipGetLastError();
        }
    }
    const int64_t base = (int64_t)(X / kTX) * (Y / kTY) * (Z / TZ);
    // blocks of two or more z slices can pool their own output
    uint16_t* const pool_to = TZ % 2 == 0 ? pooled : nullptr;
    if (base >= 256 || NCT == 1)
        hipLaunchKernelGGL((conv3_fwd<CIN, COUT, TZ, NCT>), dim3(X / kTX, Y / kTY, Z / TZ), dim3(kThreads), 0, st, x,
                           w, b, y, X, Y, Z, relu, pool_to);
    else
        hipLaunchKernelGGL((conv3_fwd<CIN, COUT, TZ, 1>), dim3(X / kTX, Y / kTY, (Z / TZ) * NCT), dim3(kThreads), 0,
                           st, x, w, b, y, X, Y, Z, relu, pool_to);
    if (pooled_done) *pooled_done = pool_to != nullptr;
    return hipGetLastError();
}

// Coarse z slices per up-convolution block: the halo stays within ~37 KB.
constexpr int up_tz(int cin) { return cin >= 128 ? 2 : 4; }

template <int CIN, int COUT>
hipError_t upconv_t(const uint16_t* d, const uint16_t* w, const float* b, const uint16_t* skip, uint16_t* y, int S,
                    int relu, hipStream_t st) {
    constexpr int TZ = up_tz(CIN);
    constexpr int NCT = cpad(COUT) / 16;
    if (S % 4 || S % TZ) return hipErrorInvalidValue;
    const int64_t base = (int64_t)(S / 4) * (S / 4) * (S / TZ);
    if (base >= 256 || NCT == 1)
        hipLaunchKernelGGL((upconv_fwd<CIN, COUT, TZ, NCT>), dim3(S / 4, S / 4, S / TZ), dim3(kThreads), 0, st, d, w, b,
                           skip, y, S, relu);
    else
        hipLaunchKernelGGL((upconv_fwd<CIN, COUT, TZ, 1>), dim3(S / 4, S / 4, (S / TZ) * NCT), dim3(kThreads), 0, st,
                           d, w, b, skip, y, S, relu);
    return hipGetLastError();
}

template <int CIN, int COUT>
hipError_t wgrad_t(const uint16_t* x, const uint16_t* dy, float* part, float* out, int X, int Y, int Z, int accumulate,
                   hipStream_t st) {
    constexpr int TZ = tz_for(CIN) < 2 ? 2 : tz_for(CIN);  // 64 voxels min: whole 32-voxel K steps
    if (X % kTX || Y % kTY || Z % TZ) return hipErrorInvalidValue;
    const dim3 grid(X / kTX, Y / kTY, Z / TZ);
    hipLaunchKernelGGL((conv3_wgrad<CIN, COUT, TZ>), grid, dim3(kThreads), 0, st, x, dy, part, X, Y, Z);
    const int64_t nblk = (int64_t)grid.x * grid.y * grid.z;
    const int64_t per = (int64_t)cpad(COUT) * kpad(CIN);
    hipLaunchKernelGGL(reduce_blocks, dim3(blocks_for(per)), dim3(256), 0, st, part, nblk, per, out, accumulate);
    return hipGetLastError();
}

}  // namespace

int64_t wgrad_scratch_floats(int cin, int cout, int X, int Y, int Z) {
    const int tz = tz_for(cin) < 2 ? 2 : tz_for(cin);
    return (int64_t)(X / kTX) * (Y / kTY) * (Z / tz) * cpad(cout) * kpad(cin);
}

hipError_t conv3(const uint16_t* x, const uint16_t* w, const float* b, uint16_t* y, int cin, int cout, int X, int Y,
                 int Z, int relu, hipStream_t st, uint16_t* pooled, bool* pooled_done) {
#define GELL_NN_CASE(a, c) \
    if (cin == a && cout == c) return conv_t<a, c>(x, w, b, y, X, Y, Z, relu, st, pooled, pooled_done);
    GELL_NN_CASE(16, 16) GELL_NN_CASE(16, 32) GELL_NN_CASE(32, 32)
    GELL_NN_CASE(32, 64) GELL_NN_CASE(64, 64) GELL_NN_CASE(64, 128) GELL_NN_CASE(128, 128) GELL_NN_CASE(128, 64)
    GELL_NN_CASE(64, 32) GELL_NN_CASE(32, 16) GELL_NN_CASE(16, 8) GELL_NN_CASE(8, 16)
#undef GELL_NN_CASE
    return hipErrorInvalidValue;
}

int64_t up_weight_count(int cin, int cout) { return (int64_t)8 * cpad(cout) * 8 * cin; }

hipError_t up_weights(const float* m, int cin, int cout, uint16_t* w8, hipStream_t st) {
    hipLaunchKernelGGL(combine_up_weights, dim3(blocks_for(up_weight_count(cin, cout))), dim3(256), 0, st, m, cin, cout,
                       w8);
    return hipGetLastError();
}

hipError_t upconv3(const uint16_t* d, const uint16_t* w8, const float* b, uint16_t* y, int cin, int cout, int S,
                   int relu, hipStream_t st, const uint16_t* skip) {
    if (cin == 128 && cout == 64) return upconv_t<128, 64>(d, w8, b, skip, y, S, relu, st);
    if (cin == 64 && cout == 32) return upconv_t<64, 32>(d, w8, b, skip, y, S, relu, st);
    if (cin == 32 && cout == 16) return upconv_t<32, 16>(d, w8, b, skip, y, S, relu, st);
    return hipErrorInvalidValue;
}

hipError_t conv3_weight_grad(const uint16_t* x, const uint16_t* dy, float* scratch, float* out, int cin, int cout, int X,
                             int Y, int Z, int accumulate, hipStream_t st) {
#define GELL_NN_CASE(a, c) \
    if (cin == a && cout == c) return wgrad_t<a, c>(x, dy, scratch, out, X, Y, Z, accumulate, st);
    GELL_NN_CASE(16, 16) GELL_NN_CASE(16, 32) GELL_NN_CASE(32, 32) GELL_NN_CASE(32, 64)
    GELL_NN_CASE(64, 64) GELL_NN_CASE(64, 128) GELL_NN_CASE(128, 128) GELL_NN_CASE(128, 64) GELL_NN_CASE(64, 32)
    GELL_NN_CASE(32, 16) GELL_NN_CASE(16, 8)
#undef GELL_NN_CASE
    return hipErrorInvalidValue;
}

hipError_t weights_flip(const uint16_t* w, int cin, int cout, uint16_t* wf, hipStream_t st) {
    const int64_t n = (int64_t)cpad(cin) * kpad(cout);
    hipLaunchKernelGGL(flip_weights, dim3(blocks_for(n)), dim3(256), 0, st, w, cin, cout, wf);
    return hipGetLastError();
}

hipError_t weights_pack(const float* m, int cin, int cout, uint16_t* w, hipStream_t st) {
    const int64_t n = (int64_t)cpad(cout) * kpad(cin);
    hipLaunchKernelGGL(pack_weights, dim3(blocks_for(n)), dim3(256), 0, st, m, cin, cout, w);
    return hipGetLastError();
}

hipError_t grad_unpack(const float* gk, int cin, int cout, float* gm, int accumulate, hipStream_t st) {
    const int64_t n = (int64_t)cout * cin * 27;
    hipLaunchKernelGGL(unpack_grad, dim3(blocks_for(n)), dim3(256), 0, st, gk, cin, cout, gm, accumulate);
    return hipGetLastError();
}

hipError_t relu_back(uint16_t* dy, const uint16_t* y, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(relu_mask, dim3(blocks_for(n)), dim3(256), 0, st, dy, y, n);
    return hipGetLastError();
}

hipError_t bias_back(const uint16_t* dy, int64_t nvox, int cout, float* db, int accumulate, hipStream_t st) {
    hipLaunchKernelGGL(bias_grad, dim3(cout), dim3(256), 0, st, dy, nvox, cout, db, accumulate);
    return hipGetLastError();
}

hipError_t pool(const uint16_t* in, uint16_t* out, int X, int Y, int Z, int C, hipStream_t st) {
    hipLaunchKernelGGL(pool2, dim3(blocks_for((int64_t)X * Y * Z * (C / 8))), dim3(256), 0, st, in, out, X, Y, Z, C);
    return hipGetLastError();
}

hipError_t pool_back(const uint16_t* dy, uint16_t* dx, int X, int Y, int Z, int C, hipStream_t st) {
    hipLaunchKernelGGL(pool2_back, dim3(blocks_for((int64_t)8 * X * Y * Z * (C / 8))), dim3(256), 0, st, dy, dx, X, Y, Z,
                       C);
    return hipGetLastError();
}

hipError_t upsample(const uint16_t* in, uint16_t* out, int X, int Y, int Z, int C, hipStream_t st) {
    hipLaunchKernelGGL(up2, dim3(blocks_for((int64_t)8 * X * Y * Z * (C / 8))), dim3(256), 0, st, in, out, X, Y, Z, C);
    return hipGetLastError();
}

hipError_t upsample_back(const uint16_t* dy, uint16_t* dx, int X, int Y, int Z, int C, hipStream_t st) {
    hipLaunchKernelGGL(up2_back, dim3(blocks_for((int64_t)X * Y * Z * (C / 8))), dim3(256), 0, st, dy, dx, X, Y, Z, C);
    return hipGetLastError();
}

hipError_t add(const uint16_t* a, const uint16_t* b, uint16_t* out, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(add_bf, dim3(blocks_for(n / 8)), dim3(256), 0, st, a, b, out, n);
    return hipGetLastError();
}

hipError_t increment_norms(const float* p0, const float* p1, int N, double* part, float* norm, double* sums,
                           hipStream_t st, double* host_sums) {
    const int64_t interior = (int64_t)(N - 4) * (N - 4) * (N - 4);
    hipLaunchKernelGGL(incr_sums, dim3(kNormBlocks), dim3(256), 0, st, p0, p1, N, part);
    hipLaunchKernelGGL(incr_finalize, dim3(1), dim3(256), 0, st, part, kNormBlocks, 1.0 / (double)interior, norm,
                       sums, host_sums);
    return hipGetLastError();
}

hipError_t finalize_norms(const double* part, int nblk, int N, float* norm, double* sums, hipStream_t st,
                          double* host_sums) {
    const int64_t interior = (int64_t)(N - 4) * (N - 4) * (N - 4);
    hipLaunchKernelGGL(incr_finalize, dim3(1), dim3(256), 0, st, part, nblk, 1.0 / (double)interior, norm, sums,
                       host_sums);
    return hipGetLastError();
}

hipError_t input_features(const float* p0, const float* p1, const float* rate, int N, const float* norm,
                          float rate_scale, uint16_t* f, hipStream_t st) {
    const int64_t n = (int64_t)(N / 2) * (N / 2) * (N / 2);
    hipLaunchKernelGGL(features, dim3(blocks_for(n)), dim3(256), 0, st, p0, p1, rate, N, norm, rate_scale, f);
    return hipGetLastError();
}

hipError_t output_field(const float* p0, const float* p1, const uint16_t* head, int N, const float* norm,
                        float out_scale, float* out, hipStream_t st) {
    hipLaunchKernelGGL(head_out, dim3(blocks_for((int64_t)(N / 2) * (N / 2) * (N / 2))), dim3(256), 0, st, p0, p1,
                       head, N, norm, out_scale, out);
    return hipGetLastError();
}

hipError_t loss_head(const float* p0, const uint16_t* head, const float* target, int N, const float* norm,
                     float out_scale, uint16_t* dhead, float* loss_part, int64_t* nparts, hipStream_t st) {
    const int64_t nc = (int64_t)(N / 2) * (N / 2) * (N / 2);
    const int64_t interior = (int64_t)(N - 4) * (N - 4) * (N - 4);
    *nparts = blocks_for(nc);
    hipLaunchKernelGGL(mse_head, dim3((unsigned)*nparts), dim3(256), 0, st, p0, head, target, N, norm, out_scale,
                       1.f / (float)interior, dhead, loss_part);
    return hipGetLastError();
}

hipError_t adam_step(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1, float b2, float eps,
                     int t, hipStream_t st) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam, dim3(blocks_for(n)), dim3(256), 0, st, w, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t to_bf16(const float* a, uint16_t* b, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(f32_to_bf16, dim3(blocks_for(n)), dim3(256), 0, st, a, b, n);
    return hipGetLastError();
}

hipError_t to_f32(const uint16_t* a, float* b, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(bf16_to_f32, dim3(blocks_for(n)), dim3(256), 0, st, a, b, n);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace gell
