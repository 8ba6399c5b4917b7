// Image output: the sRGB tonemap of src/kernel/tonemap.cu:7-39 (ToneMapping /
// ToneMappingKernel / LinearToSrgb / ApplySrgbGamma).
//
// Per channel: c < 0.0031308 ? 12.92 c : 1.055 c^(1/2.4) - 0.055, selected as glm::lerp does
// (x (1 - t) + y t, so a NaN in the branch not taken still propagates), clamped to [0, 1] with
// glm's (x < y ? y : x) / (y < x ? y : x), scaled by 255 and truncated to a byte; NaN gives 0.
// The power is evaluated in double and rounded once to float, so the CPU restatement gets the
// same bits from its own double pow.
#include "cx_internal.h"

namespace cx {

__device__ __forceinline__ uint8_t srgb_byte(float c) {
    const bool cond = c < 0.0031308f;
    const float ifTrue = 12.92f * c;
    const float p = (float)pow((double)c, (double)(1.0f / 2.4f));
    const float ifFalse = 1.055f * p - 0.055f;
    const float t = cond ? 1.0f : 0.0f;
    float v = ifFalse * (1.0f - t) + ifTrue * t;
    v = (v < 0.0f) ? 0.0f : v;
    v = (1.0f < v) ? 1.0f : v;
    v = v * 255.0f;
    return isnan(v) ? (uint8_t)0 : (uint8_t)v;
}

__global__ void k_tonemap(const float* __restrict__ hdr, uint8_t* __restrict__ ldr, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ldr[i] = srgb_byte(hdr[i]);
}

hipError_t launch_tonemap(const float* hdr, uint8_t* ldr, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tonemap, dim3((n + 255) / 256), dim3(256), 0, s, hdr, ldr, n);
    return hipGetLastError();
}

}  // namespace cx
