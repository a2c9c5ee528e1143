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
hipError_t conv3(const uint16_t* x, const uint16_t* w, const float* b, uint16_t* y, int cin, int cout, int X, int Y,
                 int Z, int relu, hipStream_t st);
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

// N^3 field + uptake map -> 24-channel features on the (N/2)^3 grid
hipError_t input_features(const float* p, const float* rate, int N, float bnd, float rate_scale, uint16_t* f,
                          hipStream_t st);
// (N/2)^3 x 8 head -> N^3 field: p + out_scale head (sub-voxel order z, y, x)
hipError_t output_field(const float* p, const uint16_t* head, int N, float out_scale, float bnd, float* out,
                        hipStream_t st);
// mean squared error in units of out_scale over the interior; dL/dhead and
// per-block loss partial sums (*nparts of them)
hipError_t loss_head(const float* p, const uint16_t* head, const float* target, int N, float out_scale,
                     uint16_t* dhead, float* loss_part, int64_t* nparts, hipStream_t st);
hipError_t adam_step(float* w, const float* g, float* m1, float* m2, int64_t n, float lr, float b1, float b2, float eps,
                     int t, hipStream_t st);
hipError_t to_bf16(const float* a, uint16_t* b, int64_t n, hipStream_t st);
hipError_t to_f32(const uint16_t* a, float* b, int64_t n, hipStream_t st);

}  // namespace nn
}  // namespace gell
