// util.hip -- a plain streaming copy kernel: 16-byte loads and stores, one
// pass over the data. It backs gell_stream_copy and is the known-size kernel
// that the HBM byte counters (FETCH_SIZE / WRITE_SIZE) are calibrated
// against.
#include "device_common.hpp"
#include "internal.hpp"

namespace gell {
namespace {

__global__ __launch_bounds__(256) void stream_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                     int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

__global__ void stream_copy1(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace
}  // namespace gell

extern "C" int gell_stream_copy(void* dst, const void* src, size_t bytes, void* stream) {
    if ((!dst || !src) && bytes) return GELL_E_ARG;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) == 0;
    const int64_t n16 = aligned ? (int64_t)(bytes / 16) : 0;
    if (n16 > 0) {
        const int64_t blocks = (n16 + 255) / 256 < 8192 ? (n16 + 255) / 256 : 8192;
        hipLaunchKernelGGL(gell::stream_copy16, dim3((unsigned)blocks), dim3(256), 0, st,
                           reinterpret_cast<uint4*>(dst), reinterpret_cast<const uint4*>(src), n16);
    }
    const int64_t rest = (int64_t)bytes - n16 * 16;
    if (rest > 0)
        hipLaunchKernelGGL(gell::stream_copy1, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, st,
                           reinterpret_cast<uint8_t*>(dst) + n16 * 16, reinterpret_cast<const uint8_t*>(src) + n16 * 16,
                           rest);
    return hipGetLastError() == hipSuccess ? GELL_OK : GELL_E_HIP;
}
