// The renderer's float elementary functions (dmath.h) on caller arrays: on the
// device, as the shading kernels evaluate them, and on the host, from the same
// source, so their agreement with the C library's double functions can be
// checked with or without a GPU.
#include <hip/hip_runtime.h>

#include "azimuth_gpu.h"
#include "dmath.h"
#include "gpu_context.h"

namespace az {
namespace {

constexpr uint32_t kMathBlock = 256;
constexpr uint32_t kMathFunctions = 9;

__host__ __device__ inline float math_apply(uint32_t fn, float x, float y) {
    switch (fn) {
    case AZ_MATH_SIN: return SinF(x);
    case AZ_MATH_COS: return CosF(x);
    case AZ_MATH_ATAN2: return Atan2F(x, y);
    case AZ_MATH_POW: return PowF(x, y);
    case AZ_MATH_EXP: return Exp(x);
    case AZ_MATH_ASIN: return Asin(x);
    case AZ_MATH_ACOS: return Acos(x);
    case AZ_MATH_POW4: return PowIF<4>(x);
    default: return PowIF<5>(x);
    }
}

__global__ __launch_bounds__(kMathBlock) void k_math_eval(uint32_t fn, const float *__restrict__ x,
                                                         const float *__restrict__ y, uint32_t n,
                                                         float *__restrict__ out) {
    const uint32_t i = blockIdx.x * kMathBlock + threadIdx.x;
    if (i >= n) return;
    out[i] = math_apply(fn, x[i], y ? y[i] : 0.0f);
}

bool two_args(uint32_t fn) { return fn == AZ_MATH_ATAN2 || fn == AZ_MATH_POW; }

}  // namespace
}  // namespace az

using namespace az;

extern "C" az_status az_math_eval(u32 Fn, const f32 *X, const f32 *Y, u32 N, f32 *Out) {
    if (Fn >= kMathFunctions || (N && (!X || !Out || (two_args(Fn) && !Y)))) return AZ_ERR_INVALID_ARGUMENT;
    for (u32 i = 0; i < N; ++i) Out[i] = math_apply(Fn, X[i], Y ? Y[i] : 0.0f);
    return AZ_OK;
}

extern "C" az_status az_gpu_math_eval(az_gpu_ctx *C, u32 Fn, const f32 *d_X, const f32 *d_Y, u32 N, f32 *d_Out) {
    if (!C || Fn >= kMathFunctions || (N && (!d_X || !d_Out || (two_args(Fn) && !d_Y))))
        return AZ_ERR_INVALID_ARGUMENT;
    if (N == 0) return AZ_OK;
    hipError_t e = hipSetDevice(C->device);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_math_eval, dim3((N + kMathBlock - 1) / kMathBlock), dim3(kMathBlock), 0, C->stream, Fn,
                           d_X, two_args(Fn) ? d_Y : nullptr, N, d_Out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) {
        C->lastError = std::string("az_gpu_math_eval: ") + hipGetErrorString(e);
        return AZ_ERR_DEVICE;
    }
    return AZ_OK;
}
