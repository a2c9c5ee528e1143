// nn.hpp -- launchers of the surrogate network kernels (nn.hip). All tensors
// channels-last, BF16 activations as uint16_t, FP32 parameters; every call
// enqueues on `st` and returns the launch status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gell {
namespace nn {

// GEMM K extent of a 3x3x3 layer with `cin` inputs, padded to whole 32-wide
// MFMA steps; output channels padded to whole 16-row tiles.
constexpr int kpad(int cin) { return (27 * cin + 31) / 32 * 32; }
constexpr int cpad(int c) { return (c + 15) / 16 * 16; }

// y = act(conv3x3x3(x, w) + b); w in GEMM layout [cpad(cout)][kpad(cin)].
// With pooled, the kernel also writes pool2(y) there when its blocks can
// (*pooled_done then true; otherwise the caller pools y itself).
hipError_t conv3(const uint16_t* x, const uint16_t* w, const float* b, uint16_t* y, int cin, int cout, int X, int Y,
                 int Z, int relu, hipStream_t st, uint16_t* pooled = nullptr, bool* pooled_done = nullptr);
// y = act(conv3x3x3(up2(d), W) + b) (+ skip) without forming up2(d): d is
// S^3 x cin, y (2S)^3 x cout; w8 = up_weights of the layer's FP32 master
// weights ([8][cpad(cout)][8 cin], up_weight_count elements). (cin, cout) in
// {(128, 64), (64, 32), (32, 16)}.
hipError_t upconv3(const uint16_t* d, const uint16_t* w8, const float* b, uint16_t* y, int cin, int cout, int S,
                   int relu, hipStream_t st, const uint16_t* skip = nullptr);
hipError_t up_weights(const float* m, int cin, int cout, uint16_t* w8, hipStream_t st);
int64_t up_weight_count(int cin, int cout);
// dW (GEMM layout, FP32) = sum_v dy[v] x[v + tap]; `scratch` holds
// wgrad_scratch_floats(...) floats of per-block partial sums.
hipError_t conv3_weight_grad(const uint16_t* x, const uint16_t* dy, float* scratch, float* out, int cin, int cout,
                             int X, int Y, int Z, int accumulate, hipStream_t st);
int64_t wgrad_scratch_floats(int cin, int cout, int X, int Y, int Z);
// kernel of the input gradient: layout [cpad(cin)][kpad(cout)]
hipError_t weights_flip(const uint16_t* w, int cin, int cout, uint16_t* wf, hipStream_t st);
// master [cout][cin][27] FP32 -> GEMM BF16, and GEMM FP32 gradient -> master layout
hipError_t weights_pack(const float* m, int cin, int cout, uint16_t* w, hipStream_t st);
hipError_t grad_unpack(const float* gk, int cin, int cout, float* gm, int accumulate, hipStream_t st);

hipError_t relu_back(uint16_t* dy, const uint16_t* y, int64_t n, hipStream_t st);
hipError_t bias_back(const uint16_t* dy, int64_t nvox, int cout, float* db, int accumulate, hipStream_t st);
// grids given at the coarse side (X x Y x Z; the fine side is twice that)
hipError_t pool(const uint16_t* in, uint16_t* out, int X, int Y, int Z, int C, hipStream_t st);
hipError_t pool_back(const uint16_t* dy, uint16_t* dx, int X, int Y, int Z, int C, hipStream_t st);
hipError_t upsample(const uint16_t* in, uint16_t* out, int X, int Y, int Z, int C, hipStream_t st);
hipError_t upsample_back(const uint16_t* dy, uint16_t* dx, int X, int Y, int Z, int C, hipStream_t st);
hipError_t add(const uint16_t* a, const uint16_t* b, uint16_t* out, int64_t n, hipStream_t st);

// The one-substep increment r = p1 - p0: per-block partial sums (part: 2
// kNormBlocks doubles) reduced in fixed order to norm[0] = s, the RMS of r
// over the interior (the voxels two or more from every face), norm[1] = 1 / s
// (0 when s = 0), and sums[0] = sum of r^2 over the interior, sums[1] = sum
// of p0^2 over the grid.
constexpr int kNormBlocks = 1024;
// host_sums (optional): a device-mapped pointer to pinned host memory that
// receives the two sums as well.
hipError_t increment_norms(const float* p0, const float* p1, int N, double* part, float* norm, double* sums,
                           hipStream_t st, double* host_sums = nullptr);
// The same reduction from nblk partials summed elsewhere (an out-of-place
// substep's last kernel).
hipError_t finalize_norms(const double* part, int nblk, int N, float* norm, double* sums, hipStream_t st,
                          double* host_sums = nullptr);
// p0, p1 (N^3) + uptake map -> 16-channel features on the (N/2)^3 grid: r / s
// and rate * rate_scale, rounded to BF16 (8 fine voxels each)
hipError_t input_features(const float* p0, const float* p1, const float* rate, int N, const float* norm,
                          float rate_scale, uint16_t* f, hipStream_t st);
// (N/2)^3 x 8 head -> N^3 field: p0 + s out_scale head in the interior, p1 on
// the two outer layers (sub-voxel order z, y, x); out may alias p1
hipError_t output_field(const float* p0, const float* p1, const uint16_t* head, int N, const float* norm,
                        float out_scale, float* out, hipStream_t st);
// mean squared error of the increment in units of s out_scale over the
// interior; dL/dhead and per-block loss partial sums (*nparts of them)
hipError_t loss_head(const float* p0, const uint16_t* head, const float* target, int N, const float* norm,
                     float out_scale, uint16_t* dhead, float* loss_part, int64_t* nparts, hipStream_t st);
hipError_t adam_step(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1, float b2, float eps,
                     int t, hipStream_t st);
hipError_t to_bf16(const float* a, uint16_t* b, int64_t n, hipStream_t st);
hipError_t to_f32(const uint16_t* a, float* b, int64_t n, hipStream_t st);

}  // namespace nn
}  // namespace gell
