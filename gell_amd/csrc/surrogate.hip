// surrogate.hip -- a learned stand-in for K LOD substeps with a fixed uptake
// map (no counterpart in the reference: its time loop always runs every
// substep, Main.cu:121-123).
//
// What it predicts: with the uptake map fixed, one substep is affine in the
// field, p -> A p + c (Setboundary, the implicit uptake p / (1 - rate dt),
// the three Thomas sweeps; LODkernel.cuh:24-136). So the field after K
// substeps from p0 is p0 + (1 + A + ... + A^(K-1)) r, r = p1 - p0 being the
// increment of the first substep: a linear map of r that depends only on the
// uptake map. The network takes r in units of its interior RMS s together
// with the uptake map and returns the K-substep increment in units of
// s out_scale; the prediction is p0 + s out_scale head. Its error thus scales
// with how far the field actually moves: near a quasi-steady state, where the
// K substeps change little, so does the prediction.
//
// The network, on the half-resolution grid (N/2)^3, is a four-level U-Net of
// 3x3x3 convolutions (BF16 on the matrix cores, FP32 accumulation):
//
//   features(r / s, rate) 16 ch: r / s and rate (BF16) at the 8 fine
//                   voxels of each coarse one
//   level 1 (N/2):  L0 16->16 = e1                        pool
//   level 2 (N/4):  L1 16->32, L2 32->32 = e2             pool
//   level 3 (N/8):  L3 32->64, L4 64->64 = e3             pool
//   level 4 (N/16): L5 64->128, L6 128->128 = e4          up
//   level 3:        L7 128->64 (+ e3), L8 64->64 = d3     up
//   level 2:        L9 64->32 (+ e2), L10 32->32 = d2     up
//   level 1:        L11 32->16 (+ e1) = x1
//   head:           L12 16->8, no ReLU: the 2x2x2 fine voxels of each coarse
//                   voxel, added to p0 (in units of s out_scale); the two
//                   outer layers, which every substep resets first, take p1.
// Level 1 holds one convolution on each side (the other levels two): its
// 16-channel layers on the (N/2)^3 grid cost the most per layer.
//
// Every layer but the head applies ReLU; the "+ e" skips are added after it.
// Training (gradients + Adam) runs on the same kernels: input gradients are
// convolutions with the flipped weights, weight gradients MFMA reductions.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "internal.hpp"
#include "nn.hpp"

using namespace gell::nn;

namespace {

struct LayerSpec {
    int cin, cout, level;
    bool relu;
};

const LayerSpec kLayers[] = {
    {16, 16, 1, true},  {16, 32, 2, true},   {32, 32, 2, true},  {32, 64, 3, true}, {64, 64, 3, true},
    {64, 128, 4, true}, {128, 128, 4, true}, {128, 64, 3, true}, {64, 64, 3, true}, {64, 32, 2, true},
    {32, 32, 2, true},  {32, 16, 1, true},   {16, 8, 1, false},
};
constexpr int kNL = 13;

struct Tensor {
    uint16_t* d = nullptr;
    int S = 0, C = 0;  // grid side, channels
    int64_t n() const { return (int64_t)S * S * S * C; }
};

}  // namespace

struct gell_surrogate {
    gell_surrogate_config cfg{};
    int N = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // parameters: per layer W[cout][cin][27] then b[cout], FP32
    int64_t w_off[kNL] = {}, b_off[kNL] = {}, nparam = 0;
    float *param = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t* wb[kNL] = {};  // GEMM-layout BF16 weights
    uint16_t* wf[kNL] = {};  // flipped, for input gradients
    uint16_t* wu[kNL] = {};  // the layers after an upsampling: per-parity 2x2x2 weights
    bool packed = false;     // wb / wf current
    float* zero_bias = nullptr;
    // forward tensors
    Tensor f0, e1, p1, a2, e2, p2, a4, e3, p3, a6, e4, u3, c3, x3, d3, u2, c2, x2, d2, u1, c1, x1, h;
    // gradient scratch: two per level + head
    Tensor g1a, g1b, g1c, g2a, g2b, g2c, g3a, g3b, g3c, g4a, g4b, dh;
    float* wg_scratch = nullptr;
    int64_t wg_scratch_n = 0;
    float* wg_out = nullptr;
    float* loss_part = nullptr;
    float* h_red = nullptr;  // pinned
    float* pred = nullptr;   // N^3 scratch for advance: the prediction at a check
    float* p_first = nullptr;  // N^3 scratch for advance: the field after its first substep
    float* red_part = nullptr;
    double* norm_part = nullptr;  // increment_norms partials
    float* norm = nullptr;        // s, 1 / s (device)
    double* sums = nullptr;       // sum r^2 (interior), sum p0^2 (device)
    double* h_sums = nullptr;     // pinned, coherent copy of sums
    double* h_sums_dev = nullptr; // its device-mapped address (written by the norms' last kernel)
    hipEvent_t sums_ready = nullptr;  // recorded after the kernel that writes h_sums
    int adam_t = 0;
    int accumulate = 0;  // add this call's gradients to the stored ones
    // The in-loop gate (gell_surrogate_advance): the largest measured ratio
    // of a prediction's true error to the relative size of its first
    // substep's increment (< 0: none measured yet), the calls since the last
    // exact check, the error spent in the current budget window, and counts
    // of each kind of call.
    double ratio = -1.0;
    int since_check = 0;
    bool last_kept = false;  // the last call between checks kept its prediction
    double spent = 0.0;
    int64_t n_checked = 0, n_accepted = 0, n_fallback = 0;
    std::vector<void*> allocs;
};

namespace {

thread_local std::string g_sur_err;

int sfail(gell_surrogate* s, int code, const std::string& m) {
    if (s) s->err = m;
    g_sur_err = m;
    return code;
}

template <class T>
hipError_t salloc(gell_surrogate* s, T** p, int64_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), (size_t)(count > 0 ? count : 1) * sizeof(T));
    if (e == hipSuccess) s->allocs.push_back(*p);
    return e;
}

hipError_t tensor(gell_surrogate* s, Tensor& t, int S, int C) {
    t.S = S;
    t.C = C;
    return salloc(s, &t.d, t.n());
}

#define SCHK(x)                        \
    do {                               \
        hipError_t e_ = (x);           \
        if (e_ != hipSuccess) return e_; \
    } while (0)

// The layers whose input is an upsampled tensor (7, 9, 11).
bool after_up(int l) { return l == 7 || l == 9 || l == 11; }

hipError_t pack_all(gell_surrogate* s) {
    for (int l = 0; l < kNL; ++l) {
        const LayerSpec& L = kLayers[l];
        SCHK(weights_pack(s->param + s->w_off[l], L.cin, L.cout, s->wb[l], s->stream));
        SCHK(weights_flip(s->wb[l], L.cin, L.cout, s->wf[l], s->stream));
        if (after_up(l)) SCHK(up_weights(s->param + s->w_off[l], L.cin, L.cout, s->wu[l], s->stream));
    }
    s->packed = true;
    return hipSuccess;
}

hipError_t conv_layer(gell_surrogate* s, int l, const Tensor& in, Tensor& out) {
    const LayerSpec& L = kLayers[l];
    return conv3(in.d, s->wb[l], s->param + s->b_off[l], out.d, L.cin, L.cout, in.S, in.S, in.S, L.relu ? 1 : 0,
                 s->stream);
}

// Layer l, then pooled = pool2(out): in the convolution's own epilogue where
// its blocks allow, else by the pooling kernel (the same values either way).
hipError_t conv_pool_layer(gell_surrogate* s, int l, const Tensor& in, Tensor& out, Tensor& pooled) {
    const LayerSpec& L = kLayers[l];
    bool done = false;
    SCHK(conv3(in.d, s->wb[l], s->param + s->b_off[l], out.d, L.cin, L.cout, in.S, in.S, in.S, L.relu ? 1 : 0,
               s->stream, pooled.d, &done));
    if (!done) SCHK(pool(out.d, pooled.d, pooled.S, pooled.S, pooled.S, pooled.C, s->stream));
    return hipSuccess;
}

// out = layer l (one of 8, 10, 12) of up2(coarse) + skip, on the coarse grid.
hipError_t up_layer(gell_surrogate* s, int l, const Tensor& coarse, Tensor& out, const uint16_t* skip) {
    const LayerSpec& L = kLayers[l];
    return upconv3(coarse.d, s->wu[l], s->param + s->b_off[l], out.d, L.cin, L.cout, coarse.S, L.relu ? 1 : 0,
                   s->stream, skip);
}

// The increment's norms (s on the device) and the network's head from
// (p0, p1 = one substep from p0, rate). With train each upsampling is formed
// (u3, u2, u1) and the skip connections keep the convolution's output (c3,
// c2, c1: the backward pass needs its ReLU mask) apart from the sum; without,
// the layers after an upsampling run on the coarse grid (upconv3: its
// per-parity weights are sums of the layer's FP32 taps rounded to BF16 once,
// where the formed path rounds each tap, a difference of one BF16 rounding
// per weight) and store the sum themselves.
// norms_ready: the norms of this (p0, p1) are already on the device. With
// out (inference only) the predicted field is written there as well.
hipError_t forward(gell_surrogate* s, const float* p0, const float* p1, const float* rate, bool train,
                   bool norms_ready = false, float* out = nullptr) {
    if (!s->packed) SCHK(pack_all(s));
    hipStream_t st = s->stream;
    if (!norms_ready) SCHK(increment_norms(p0, p1, s->N, s->norm_part, s->norm, s->sums, st));
    SCHK(input_features(p0, p1, rate, s->N, s->norm, s->cfg.rate_scale, s->f0.d, st));
    SCHK(conv_pool_layer(s, 0, s->f0, s->e1, s->p1));
    SCHK(conv_layer(s, 1, s->p1, s->a2));
    SCHK(conv_pool_layer(s, 2, s->a2, s->e2, s->p2));
    SCHK(conv_layer(s, 3, s->p2, s->a4));
    SCHK(conv_pool_layer(s, 4, s->a4, s->e3, s->p3));
    SCHK(conv_layer(s, 5, s->p3, s->a6));
    SCHK(conv_layer(s, 6, s->a6, s->e4));
    if (train) {
        SCHK(upsample(s->e4.d, s->u3.d, s->e4.S, s->e4.S, s->e4.S, s->e4.C, st));
        SCHK(conv_layer(s, 7, s->u3, s->c3));
        SCHK(add(s->c3.d, s->e3.d, s->x3.d, s->x3.n(), st));
    } else {
        SCHK(up_layer(s, 7, s->e4, s->x3, s->e3.d));
    }
    SCHK(conv_layer(s, 8, s->x3, s->d3));
    if (train) {
        SCHK(upsample(s->d3.d, s->u2.d, s->d3.S, s->d3.S, s->d3.S, s->d3.C, st));
        SCHK(conv_layer(s, 9, s->u2, s->c2));
        SCHK(add(s->c2.d, s->e2.d, s->x2.d, s->x2.n(), st));
    } else {
        SCHK(up_layer(s, 9, s->d3, s->x2, s->e2.d));
    }
    SCHK(conv_layer(s, 10, s->x2, s->d2));
    if (train) {
        SCHK(upsample(s->d2.d, s->u1.d, s->d2.S, s->d2.S, s->d2.S, s->d2.C, st));
        SCHK(conv_layer(s, 11, s->u1, s->c1));
        SCHK(add(s->c1.d, s->e1.d, s->x1.d, s->x1.n(), st));
    } else {
        SCHK(up_layer(s, 11, s->d2, s->x1, s->e1.d));
    }
    SCHK(conv_layer(s, 12, s->x1, s->h));
    if (out && !train) SCHK(output_field(p0, p1, s->h.d, s->N, s->norm, s->cfg.out_scale, out, st));
    return hipSuccess;
}

// Gradients of layer l given dA (gradient at the layer's pre-activation, already
// masked by ReLU): bias, weights (into the master layout), and, when dx is
// given, the input gradient.
hipError_t layer_back(gell_surrogate* s, int l, const Tensor& in, const Tensor& dA, Tensor* dx) {
    const LayerSpec& L = kLayers[l];
    hipStream_t st = s->stream;
    const int S = in.S;
    SCHK(bias_back(dA.d, (int64_t)S * S * S, L.cout, s->grad + s->b_off[l], s->accumulate, st));
    SCHK(conv3_weight_grad(in.d, dA.d, s->wg_scratch, s->wg_out, L.cin, L.cout, S, S, S, 0, st));
    SCHK(grad_unpack(s->wg_out, L.cin, L.cout, s->grad + s->w_off[l], s->accumulate, st));
    if (dx) SCHK(conv3(dA.d, s->wf[l], s->zero_bias, dx->d, L.cout, L.cin, S, S, S, 0, st));
    return hipSuccess;
}

hipError_t backward(gell_surrogate* s) {
    hipStream_t st = s->stream;
    // head (no ReLU): dh holds dL/dh
    SCHK(layer_back(s, 12, s->x1, s->dh, &s->g1b));                 // g1b = dx1 (= dc1, part of de1)
    SCHK(hipMemcpyAsync(s->g1c.d, s->g1b.d, s->g1b.n() * 2, hipMemcpyDeviceToDevice, st));  // g1c = de1 so far
    SCHK(relu_back(s->g1b.d, s->c1.d, s->g1b.n(), st));
    // The input gradients of L11, L9, L7, L5, L3 and L1 have the shape of
    // the layer's own input (u1, u2, u3, p3, p2, p1), which the backward pass
    // no longer needs once that layer's weight gradient is taken: they are
    // written in place of it.
    SCHK(layer_back(s, 11, s->u1, s->g1b, &s->u1));                 // u1 <- du1
    SCHK(upsample_back(s->u1.d, s->g2a.d, s->d2.S, s->d2.S, s->d2.S, s->d2.C, st));  // g2a = dd2
    SCHK(relu_back(s->g2a.d, s->d2.d, s->g2a.n(), st));
    SCHK(layer_back(s, 10, s->x2, s->g2a, &s->g2b));                // g2b = dx2
    SCHK(hipMemcpyAsync(s->g2c.d, s->g2b.d, s->g2b.n() * 2, hipMemcpyDeviceToDevice, st));  // g2c = de2 so far
    SCHK(relu_back(s->g2b.d, s->c2.d, s->g2b.n(), st));
    SCHK(layer_back(s, 9, s->u2, s->g2b, &s->u2));                  // u2 <- du2 (64 ch, level 2)
    SCHK(upsample_back(s->u2.d, s->g3a.d, s->d3.S, s->d3.S, s->d3.S, s->d3.C, st));  // g3a = dd3
    SCHK(relu_back(s->g3a.d, s->d3.d, s->g3a.n(), st));
    SCHK(layer_back(s, 8, s->x3, s->g3a, &s->g3b));                 // g3b = dx3
    SCHK(hipMemcpyAsync(s->g3c.d, s->g3b.d, s->g3b.n() * 2, hipMemcpyDeviceToDevice, st));  // g3c = de3 so far
    SCHK(relu_back(s->g3b.d, s->c3.d, s->g3b.n(), st));
    SCHK(layer_back(s, 7, s->u3, s->g3b, &s->u3));                  // u3 <- du3 (128 ch, level 3)
    SCHK(upsample_back(s->u3.d, s->g4a.d, s->e4.S, s->e4.S, s->e4.S, s->e4.C, st));  // g4a = de4
    SCHK(relu_back(s->g4a.d, s->e4.d, s->g4a.n(), st));
    SCHK(layer_back(s, 6, s->a6, s->g4a, &s->g4b));                 // g4b = da6
    SCHK(relu_back(s->g4b.d, s->a6.d, s->g4b.n(), st));
    SCHK(layer_back(s, 5, s->p3, s->g4b, &s->p3));                  // p3 <- dp3 (64 ch, level 4)
    SCHK(pool_back(s->p3.d, s->g3a.d, s->p3.S, s->p3.S, s->p3.S, s->p3.C, st));  // g3a = pool part of de3
    SCHK(add(s->g3a.d, s->g3c.d, s->g3c.d, s->g3c.n(), st));        // g3c = de3
    SCHK(relu_back(s->g3c.d, s->e3.d, s->g3c.n(), st));
    SCHK(layer_back(s, 4, s->a4, s->g3c, &s->g3a));                 // g3a = da4
    SCHK(relu_back(s->g3a.d, s->a4.d, s->g3a.n(), st));
    SCHK(layer_back(s, 3, s->p2, s->g3a, &s->p2));                  // p2 <- dp2 (32 ch, level 3)
    SCHK(pool_back(s->p2.d, s->g2a.d, s->p2.S, s->p2.S, s->p2.S, s->p2.C, st));
    SCHK(add(s->g2a.d, s->g2c.d, s->g2c.d, s->g2c.n(), st));        // g2c = de2
    SCHK(relu_back(s->g2c.d, s->e2.d, s->g2c.n(), st));
    SCHK(layer_back(s, 2, s->a2, s->g2c, &s->g2a));                 // g2a = da2
    SCHK(relu_back(s->g2a.d, s->a2.d, s->g2a.n(), st));
    SCHK(layer_back(s, 1, s->p1, s->g2a, &s->p1));                  // p1 <- dp1 (16 ch, level 2)
    SCHK(pool_back(s->p1.d, s->g1a.d, s->p1.S, s->p1.S, s->p1.S, s->p1.C, st));
    SCHK(add(s->g1a.d, s->g1c.d, s->g1c.d, s->g1c.n(), st));        // g1c = de1
    SCHK(relu_back(s->g1c.d, s->e1.d, s->g1c.n(), st));
    SCHK(layer_back(s, 0, s->f0, s->g1c, nullptr));
    return hipSuccess;
}

// sum of squares of (a - b) and of a, in fixed-order per-block partials
__global__ __launch_bounds__(256) void diff_norms(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
                                                  float* __restrict__ part) {
    __shared__ double r0[256], r1[256];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double d = (double)a[i] - (double)b[i];
        s0 += d * d;
        s1 += (double)a[i] * (double)a[i];
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
        part[2 * blockIdx.x] = (float)r0[0];
        part[2 * blockIdx.x + 1] = (float)r1[0];
    }
}

constexpr int kRedBlocks = 1024;

}  // namespace

extern "C" {

void gell_surrogate_default_config(gell_surrogate_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->N = 256;
    c->K = 50;
    c->rate_scale = 0.1f;
    c->out_scale = 8.f;
    c->check_every = 6;
    c->seed = 7;
}

int gell_surrogate_create(const gell_surrogate_config* cfg, void* stream, gell_surrogate** out) {
    if (!out) return sfail(nullptr, GELL_E_ARG, "null output pointer");
    *out = nullptr;
    if (!cfg) return sfail(nullptr, GELL_E_ARG, "null config");
    if (cfg->N < 128 || cfg->N % 128 != 0 || cfg->N > 1024)
        return sfail(nullptr, GELL_E_UNSUPPORTED, "surrogate grid must be a multiple of 128 (128 .. 1024)");
    if (cfg->K < 1 || !(cfg->out_scale > 0.f))
        return sfail(nullptr, GELL_E_ARG, "bad surrogate config");
    gell_surrogate* s = new (std::nothrow) gell_surrogate;
    if (!s) return sfail(nullptr, GELL_E_NOMEM, "surrogate allocation");
    s->cfg = *cfg;
    s->N = cfg->N;
    s->stream = reinterpret_cast<hipStream_t>(stream);
    int64_t off = 0;
    for (int l = 0; l < kNL; ++l) {
        s->w_off[l] = off;
        off += (int64_t)kLayers[l].cout * kLayers[l].cin * 27;
        s->b_off[l] = off;
        off += kLayers[l].cout;
    }
    s->nparam = off;
    const int S1 = s->N / 2, S2 = S1 / 2, S3 = S2 / 2, S4 = S3 / 2;
    hipError_t e = hipSuccess;
#define SA(x) \
    if (e == hipSuccess) e = (x)
    SA(salloc(s, &s->param, s->nparam));
    SA(salloc(s, &s->grad, s->nparam));
    SA(salloc(s, &s->m1, s->nparam));
    SA(salloc(s, &s->m2, s->nparam));
    for (int l = 0; l < kNL; ++l) {
        SA(salloc(s, &s->wb[l], (int64_t)cpad(kLayers[l].cout) * kpad(kLayers[l].cin)));
        SA(salloc(s, &s->wf[l], (int64_t)cpad(kLayers[l].cin) * kpad(kLayers[l].cout)));
        if (after_up(l)) SA(salloc(s, &s->wu[l], up_weight_count(kLayers[l].cin, kLayers[l].cout)));
    }
    SA(salloc(s, &s->zero_bias, 128));
    SA(tensor(s, s->f0, S1, 16));
    SA(tensor(s, s->e1, S1, 16));
    SA(tensor(s, s->p1, S2, 16));
    SA(tensor(s, s->a2, S2, 32));
    SA(tensor(s, s->e2, S2, 32));
    SA(tensor(s, s->p2, S3, 32));
    SA(tensor(s, s->a4, S3, 64));
    SA(tensor(s, s->e3, S3, 64));
    SA(tensor(s, s->p3, S4, 64));
    SA(tensor(s, s->a6, S4, 128));
    SA(tensor(s, s->e4, S4, 128));
    SA(tensor(s, s->u3, S3, 128));
    SA(tensor(s, s->c3, S3, 64));
    SA(tensor(s, s->x3, S3, 64));
    SA(tensor(s, s->d3, S3, 64));
    SA(tensor(s, s->u2, S2, 64));
    SA(tensor(s, s->c2, S2, 32));
    SA(tensor(s, s->x2, S2, 32));
    SA(tensor(s, s->d2, S2, 32));
    SA(tensor(s, s->u1, S1, 32));
    SA(tensor(s, s->c1, S1, 16));
    SA(tensor(s, s->x1, S1, 16));
    SA(tensor(s, s->h, S1, 8));
    SA(tensor(s, s->g1a, S1, 16));
    SA(tensor(s, s->g1b, S1, 16));
    SA(tensor(s, s->g1c, S1, 16));
    SA(tensor(s, s->g2a, S2, 32));
    SA(tensor(s, s->g2b, S2, 32));
    SA(tensor(s, s->g2c, S2, 32));
    SA(tensor(s, s->g3a, S3, 64));
    SA(tensor(s, s->g3b, S3, 64));
    SA(tensor(s, s->g3c, S3, 64));
    SA(tensor(s, s->g4a, S4, 128));
    SA(tensor(s, s->g4b, S4, 128));
    SA(tensor(s, s->dh, S1, 8));
    int64_t wgn = 0, wgo = 0;
    for (int l = 0; l < kNL; ++l) {
        const int S = s->N >> kLayers[l].level;
        const int64_t a = wgrad_scratch_floats(kLayers[l].cin, kLayers[l].cout, S, S, S);
        wgn = a > wgn ? a : wgn;
        const int64_t b = (int64_t)cpad(kLayers[l].cout) * kpad(kLayers[l].cin);
        wgo = b > wgo ? b : wgo;
    }
    s->wg_scratch_n = wgn;
    SA(salloc(s, &s->wg_scratch, wgn));
    SA(salloc(s, &s->wg_out, wgo));
    SA(salloc(s, &s->loss_part, (int64_t)S1 * S1 * S1 / 256 + 1));
    SA(salloc(s, &s->pred, (int64_t)s->N * s->N * s->N));
    SA(salloc(s, &s->p_first, (int64_t)s->N * s->N * s->N));
    SA(salloc(s, &s->red_part, 2 * kRedBlocks));
    SA(salloc(s, &s->norm_part, 2 * std::max(kNormBlocks, gell::substep_norm_blocks(s->N))));
    SA(salloc(s, &s->norm, 2));
    SA(salloc(s, &s->sums, 2));
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&s->h_red), sizeof(float) * (2 * kRedBlocks + 8));
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&s->h_sums), sizeof(double) * 2,
                          hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->h_sums_dev), s->h_sums, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->sums_ready, hipEventDisableTiming);
#undef SA
    if (e == hipSuccess) {
        // He-normal initialisation from a seeded host generator; the head
        // starts at 1/100 of it so the first predictions stay near p.
        std::vector<float> h(s->nparam, 0.f);
        std::mt19937_64 rng(cfg->seed);
        for (int l = 0; l < kNL; ++l) {
            const double sd = std::sqrt(2.0 / (27.0 * kLayers[l].cin)) * (l == kNL - 1 ? 0.01 : 1.0);
            std::normal_distribution<double> nd(0.0, sd);
            for (int64_t i = 0; i < (int64_t)kLayers[l].cout * kLayers[l].cin * 27; ++i)
                h[s->w_off[l] + i] = (float)nd(rng);
        }
        e = hipMemcpy(s->param, h.data(), sizeof(float) * s->nparam, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(s->m1, 0, sizeof(float) * s->nparam);
        if (e == hipSuccess) e = hipMemset(s->m2, 0, sizeof(float) * s->nparam);
        if (e == hipSuccess) e = hipMemset(s->grad, 0, sizeof(float) * s->nparam);
        if (e == hipSuccess) e = hipMemset(s->zero_bias, 0, sizeof(float) * 128);
    }
    if (e != hipSuccess) {
        const std::string m = std::string("surrogate allocation: ") + hipGetErrorString(e);
        gell_surrogate_destroy(s);
        return sfail(nullptr, GELL_E_NOMEM, m);
    }
    *out = s;
    return GELL_OK;
}

void gell_surrogate_destroy(gell_surrogate* s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->stream);
    for (void* p : s->allocs) (void)hipFree(p);
    if (s->h_red) (void)hipHostFree(s->h_red);
    if (s->h_sums) (void)hipHostFree(s->h_sums);
    if (s->sums_ready) (void)hipEventDestroy(s->sums_ready);
    delete s;
}

const char* gell_surrogate_last_error(const gell_surrogate* s) { return s ? s->err.c_str() : g_sur_err.c_str(); }

int64_t gell_surrogate_param_count(const gell_surrogate* s) { return s ? s->nparam : -1; }

int gell_surrogate_get_params(gell_surrogate* s, float* host, int64_t n) {
    if (!s || !host || n != s->nparam) return sfail(s, GELL_E_ARG, "parameter buffer size");
    hipError_t e = hipMemcpyAsync(host, s->param, sizeof(float) * n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? GELL_OK : sfail(s, GELL_E_HIP, hipGetErrorString(e));
}

int gell_surrogate_set_params(gell_surrogate* s, const float* host, int64_t n) {
    if (!s || !host || n != s->nparam) return sfail(s, GELL_E_ARG, "parameter buffer size");
    hipError_t e = hipMemcpyAsync(s->param, host, sizeof(float) * n, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    s->packed = false;
    s->adam_t = 0;
    if (e == hipSuccess) e = hipMemsetAsync(s->m1, 0, sizeof(float) * n, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->m2, 0, sizeof(float) * n, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? GELL_OK : sfail(s, GELL_E_HIP, hipGetErrorString(e));
}

int gell_surrogate_predict(gell_surrogate* s, const float* field, const float* field1, const float* rate, float* out) {
    if (!s || !field || !field1 || !rate || !out) return sfail(s, GELL_E_ARG, "null argument");
    hipError_t e = forward(s, field, field1, rate, false, false, out);
    return e == hipSuccess ? GELL_OK : sfail(s, GELL_E_HIP, std::string("predict: ") + hipGetErrorString(e));
}

int gell_surrogate_gradients(gell_surrogate* s, const float* field, const float* field1, const float* rate,
                             const float* target, int32_t accumulate, float* loss, float* grad_host) {
    if (!s || !field || !field1 || !rate || !target) return sfail(s, GELL_E_ARG, "null argument");
    s->accumulate = accumulate ? 1 : 0;
    hipError_t e = forward(s, field, field1, rate, true);
    int64_t np = 0;
    if (e == hipSuccess)
        e = loss_head(field, s->h.d, target, s->N, s->norm, s->cfg.out_scale, s->dh.d, s->loss_part, &np,
                      s->stream);
    if (e == hipSuccess) e = backward(s);
    if (e == hipSuccess && loss) {
        std::vector<float> lp(np);
        e = hipMemcpyAsync(lp.data(), s->loss_part, sizeof(float) * np, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        double acc = 0.0;
        for (float v : lp) acc += v;
        *loss = (float)acc;
    }
    if (e == hipSuccess && grad_host) {
        e = hipMemcpyAsync(grad_host, s->grad, sizeof(float) * s->nparam, hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    }
    return e == hipSuccess ? GELL_OK : sfail(s, GELL_E_HIP, std::string("gradients: ") + hipGetErrorString(e));
}

int gell_surrogate_adam(gell_surrogate* s, float lr) {
    if (!s) return sfail(nullptr, GELL_E_ARG, "null surrogate");
    ++s->adam_t;
    hipError_t e = adam_step(s->param, s->grad, s->m1, s->m2, s->nparam, lr, 0.9f, 0.999f, 1e-8f, s->adam_t, s->stream);
    s->packed = false;
    return e == hipSuccess ? GELL_OK : sfail(s, GELL_E_HIP, std::string("adam: ") + hipGetErrorString(e));
}

namespace {

// ||a - b|| / ||a|| over n floats (fixed-order reduction); synchronises.
hipError_t rel_change(gell_surrogate* s, const float* a, const float* b, int64_t n, double* out) {
    hipLaunchKernelGGL(diff_norms, dim3(kRedBlocks), dim3(256), 0, s->stream, a, b, n, s->red_part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->h_red, s->red_part, sizeof(float) * 2 * kRedBlocks, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return e;
    double num = 0.0, den = 0.0;
    for (int b = 0; b < kRedBlocks; ++b) {
        num += s->h_red[2 * b];
        den += s->h_red[2 * b + 1];
    }
    *out = den > 0.0 ? std::sqrt(num / den) : 0.0;
    return hipSuccess;
}

}  // namespace

// The gate. Every call first runs one real substep (the first of the K) and
// measures the relative size of its increment, rho = ||p1 - p0|| (interior)
// / ||p0||. On the first call and every check_every-th call after, the
// network predicts and the other K - 1 real substeps run as well; the field
// keeps their exact result, and the prediction's true relative L2 error e
// against it calibrates ratio = max(e / rho) over the checks. Between checks
// the estimate 1.5 ratio rho decides before the network runs: the prediction
// is made and kept when that estimate, added to the estimates of the
// predictions already kept in the current window (gell_surrogate_reset_budget
// starts a new one), stays within tol; otherwise the other K - 1 substeps
// run and the network does not. The error a window adds to the field is thus
// bounded by tol as far as the calibrated estimate holds; tol = 0 makes every
// call exact. The first substep runs out of place (p0 stays in the field, p1
// in a scratch field, from which the second substep of an exact run writes
// back), so no copy of the field is made; on checks, and after a call that
// kept its prediction, the network is queued before the host reads rho.
int gell_surrogate_advance(gell_surrogate* s, gell_ctx* ctx, int32_t sub, float tol, int32_t* used, float* residual) {
    if (!s || !ctx || !used) return sfail(s, GELL_E_ARG, "null argument");
    if (ctx->N != s->N) return sfail(s, GELL_E_ARG, "grid size differs from the surrogate's");
    if (sub < 0 || sub >= ctx->P.n_substrates) return sfail(s, GELL_E_ARG, "substrate out of range");
    if (ctx->stream != s->stream) return sfail(s, GELL_E_ARG, "surrogate and context must share a stream");
    hipStream_t st = s->stream;
    const int64_t V = ctx->V;
    float* const field = ctx->field[sub];  // p0 until the call's end
    const float* const rate = ctx->rate[sub];
    *used = 0;
    const int every = s->cfg.check_every > 0 ? s->cfg.check_every : 1;
    const bool check = s->ratio < 0.0 || s->since_check + 1 >= every;
    // The network runs before the host has read rho when the call is a check
    // (it always predicts) or when the last call kept its prediction (it
    // likely will again); a call that then falls back has spent it for
    // nothing.
    const bool early = check || (tol > 0.f && s->last_kept);
    // the first substep for real, out of place: p0 stays in the field, p1
    // goes to the scratch; then its increment's norms, read back while the
    // network runs
    hipError_t e = gell::join_substrates(ctx);
    gell::SubstepNorms sn;
    sn.p0 = field;
    sn.part = s->norm_part;
    if (e == hipSuccess) e = gell::lod_substep_launch(ctx, sub, field, s->p_first, st, &sn);
    // the fused substep sums the norms in its z sweep; the other forms do not
    if (e == hipSuccess)
        e = sn.nblk ? finalize_norms(sn.part, sn.nblk, s->N, s->norm, s->sums, st, s->h_sums_dev)
                    : increment_norms(field, s->p_first, s->N, s->norm_part, s->norm, s->sums, st, s->h_sums_dev);
    if (e == hipSuccess) e = hipEventRecord(s->sums_ready, st);
    // (a check's prediction goes to the scratch; a kept one over the field,
    // which a fallback then rebuilds from p1 anyway)
    float* const pred_to = check ? s->pred : field;
    if (e == hipSuccess && early) e = forward(s, field, s->p_first, rate, false, true, pred_to);
    if (e == hipSuccess) e = hipEventSynchronize(s->sums_ready);
    if (e != hipSuccess) return sfail(s, GELL_E_HIP, std::string("advance: ") + hipGetErrorString(e));
    const double rho = s->h_sums[1] > 0.0 ? std::sqrt(s->h_sums[0] / s->h_sums[1]) : 0.0;
    if (residual) *residual = (float)rho;
    // the other K - 1 substeps: the second from p1 back into the field
    const auto exact_rest = [&]() {
        hipError_t r = gell::lod_substep_launch(ctx, sub, s->p_first, field, st);
        for (int k = 2; k < s->cfg.K && r == hipSuccess; ++k) r = gell::lod_substep_launch(ctx, sub, st);
        return r;
    };
    if (check) {
        // exact check: the prediction, then the other K - 1 real substeps,
        // kept; the prediction's error calibrates the estimate
        e = exact_rest();
        double err = 0.0;
        if (e == hipSuccess) e = rel_change(s, field, s->pred, V, &err);
        if (e != hipSuccess) return sfail(s, GELL_E_HIP, std::string("advance check: ") + hipGetErrorString(e));
        const double r = err / (rho > 1e-30 ? rho : 1e-30);
        s->ratio = r > s->ratio ? r : s->ratio;
        s->since_check = 0;
        ++s->n_checked;
        return GELL_OK;
    }
    ++s->since_check;
    const double est = 1.5 * s->ratio * rho;
    if (tol > 0.f && s->spent + est <= (double)tol) {
        s->spent += est;
        *used = 1;
        ++s->n_accepted;
        // in place: the interior from p0 and the head, the outer layers from p1
        if (!early) e = forward(s, field, s->p_first, rate, false, true, field);
    } else {
        ++s->n_fallback;
        e = exact_rest();
    }
    s->last_kept = *used != 0;
    return e == hipSuccess ? GELL_OK : sfail(s, GELL_E_HIP, std::string("advance: ") + hipGetErrorString(e));
}

int gell_surrogate_reset_budget(gell_surrogate* s) {
    if (!s) return sfail(nullptr, GELL_E_ARG, "null surrogate");
    s->spent = 0.0;
    return GELL_OK;
}

int gell_surrogate_gate_stats(const gell_surrogate* s, int64_t* checked, int64_t* accepted, int64_t* fallback,
                              double* ratio, double* spent) {
    if (!s) return sfail(nullptr, GELL_E_ARG, "null surrogate");
    if (checked) *checked = s->n_checked;
    if (accepted) *accepted = s->n_accepted;
    if (fallback) *fallback = s->n_fallback;
    if (ratio) *ratio = s->ratio;
    if (spent) *spent = s->spent;
    return GELL_OK;
}

// One layer after an upsampling, as inference runs it: y ((2S)^3 x cout) =
// act(conv3(up2(x)) + b) (+ skip when not NULL) computed on the coarse S^3
// grid of x. x, skip, y BF16 device tensors; w_master [cout][cin][27] and
// b[cout] FP32 on the host. Allocates; synchronises.
int gell_nn_upconv3(const uint16_t* x, const float* w_master, const float* b, const uint16_t* skip, uint16_t* y,
                    int32_t cin, int32_t cout, int32_t S, int32_t relu) {
    if (!x || !w_master || !b || !y || S < 4 || S % 4) return sfail(nullptr, GELL_E_ARG, "bad upconv arguments");
    const int64_t nw = (int64_t)cout * cin * 27;
    float *dm = nullptr, *db = nullptr;
    uint16_t* w8 = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dm), sizeof(float) * nw);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&db), sizeof(float) * cout);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&w8), 2 * (size_t)up_weight_count(cin, cout));
    if (e == hipSuccess) e = hipMemcpy(dm, w_master, sizeof(float) * nw, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(db, b, sizeof(float) * cout, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = up_weights(dm, cin, cout, w8, nullptr);
    if (e == hipSuccess) e = upconv3(x, w8, db, y, cin, cout, S, relu, nullptr, skip);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    for (void* p : {(void*)dm, (void*)db, (void*)w8})
        if (p) (void)hipFree(p);
    return e == hipSuccess ? GELL_OK : sfail(nullptr, GELL_E_HIP, std::string("upconv3: ") + hipGetErrorString(e));
}

// Single 3x3x3 layer on an S^3 grid, for checking the kernels one at a time.
// x, y BF16 device tensors; w_master FP32 [cout][cin][27] and b FP32 on the
// host. Allocates and frees its own packing buffers; synchronises.
namespace {
int nn_conv3_impl(const uint16_t* x, const float* w_master, const float* b, uint16_t* y, int32_t cin, int32_t cout,
                  int32_t S, int32_t relu, int32_t input_grad, uint16_t* pooled, int32_t* in_kernel) {
    if (!x || !w_master || !b || !y || S < 8 || S % 8) return sfail(nullptr, GELL_E_ARG, "bad conv arguments");
    const int64_t nw = (int64_t)cout * cin * 27;
    float *dm = nullptr, *db = nullptr;
    uint16_t *wb = nullptr, *wf = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&dm), sizeof(float) * nw);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&db), sizeof(float) * 128);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&wb), 2 * (size_t)cpad(cout) * kpad(cin));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&wf), 2 * (size_t)cpad(cin) * kpad(cout));
    if (e == hipSuccess) e = hipMemcpy(dm, w_master, sizeof(float) * nw, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(db, 0, sizeof(float) * 128);
    if (e == hipSuccess && !input_grad) e = hipMemcpy(db, b, sizeof(float) * cout, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = weights_pack(dm, cin, cout, wb, nullptr);
    if (e == hipSuccess && input_grad) e = weights_flip(wb, cin, cout, wf, nullptr);
    // input_grad: y (cin channels) = input gradient of the layer for x = dy (cout channels)
    bool done = false;
    if (e == hipSuccess)
        e = input_grad ? conv3(x, wf, db, y, cout, cin, S, S, S, 0, nullptr)
                       : conv3(x, wb, db, y, cin, cout, S, S, S, relu, nullptr, pooled, &done);
    if (e == hipSuccess && pooled && !done) e = pool(y, pooled, S / 2, S / 2, S / 2, cout, nullptr);
    if (in_kernel) *in_kernel = done ? 1 : 0;
    if (e == hipSuccess) e = hipDeviceSynchronize();
    for (void* p : {(void*)dm, (void*)db, (void*)wb, (void*)wf})
        if (p) (void)hipFree(p);
    return e == hipSuccess ? GELL_OK : sfail(nullptr, GELL_E_HIP, std::string("conv3: ") + hipGetErrorString(e));
}
}  // namespace

int gell_nn_conv3(const uint16_t* x, const float* w_master, const float* b, uint16_t* y, int32_t cin, int32_t cout,
                  int32_t S, int32_t relu, int32_t input_grad) {
    return nn_conv3_impl(x, w_master, b, y, cin, cout, S, relu, input_grad, nullptr, nullptr);
}

int gell_nn_conv3_pooled(const uint16_t* x, const float* w_master, const float* b, uint16_t* y, uint16_t* pooled,
                         int32_t cin, int32_t cout, int32_t S, int32_t relu, int32_t* in_kernel) {
    if (!pooled) return sfail(nullptr, GELL_E_ARG, "null pooled");
    return nn_conv3_impl(x, w_master, b, y, cin, cout, S, relu, 0, pooled, in_kernel);
}

// Weight gradient of one layer: dw_master (host, [cout][cin][27]) =
// sum_v dy[v][o] x[v + tap][c]. Allocates; synchronises.
int gell_nn_conv3_wgrad(const uint16_t* x, const uint16_t* dy, float* dw_master, int32_t cin, int32_t cout,
                        int32_t S) {
    if (!x || !dy || !dw_master || S < 8 || S % 8) return sfail(nullptr, GELL_E_ARG, "bad wgrad arguments");
    const int64_t ns = wgrad_scratch_floats(cin, cout, S, S, S);
    const int64_t nk = (int64_t)cpad(cout) * kpad(cin), nw = (int64_t)cout * cin * 27;
    float *scr = nullptr, *gk = nullptr, *gm = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&scr), sizeof(float) * ns);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&gk), sizeof(float) * nk);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&gm), sizeof(float) * nw);
    if (e == hipSuccess) e = conv3_weight_grad(x, dy, scr, gk, cin, cout, S, S, S, 0, nullptr);
    if (e == hipSuccess) e = grad_unpack(gk, cin, cout, gm, 0, nullptr);
    if (e == hipSuccess) e = hipMemcpy(dw_master, gm, sizeof(float) * nw, hipMemcpyDeviceToHost);
    for (void* p : {(void*)scr, (void*)gk, (void*)gm})
        if (p) (void)hipFree(p);
    return e == hipSuccess ? GELL_OK : sfail(nullptr, GELL_E_HIP, std::string("wgrad: ") + hipGetErrorString(e));
}

}  // extern "C"
