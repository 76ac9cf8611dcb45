// Fused clamp + mean-squared-error loss of the fit step: the caller's
//   loss = mean((clamp(img, 0, 1) - target)^2)
// (gsplat2d caller: gaussianlig.py:95-103 with utils.py:21-22), in two
// passes instead of torch's clamp / mse / mse-backward / clamp-backward chain.
//   forward : per-block FP32 partial sums, then one FP64 reduction of the
//             partials (fixed order: deterministic);
//   backward: grad = (0 <= img <= 1) ? ((2/n) * (clamp(img) - target)) * g : 0,
//             the clamp mask inclusive at both bounds as torch's clamp backward.
#include "lig_device.h"

namespace lig {
namespace {

constexpr int kLossThreads = 256;
constexpr int kLossItems = 32;  // floats per thread per block
constexpr long long kLossChunk = (long long)kLossThreads * kLossItems;

__device__ __forceinline__ float sq_err(float x, float t) {
    const float c = fminf(fmaxf(x, 0.0f), 1.0f);
    const float d = c - t;
    return d * d;
}

__global__ __launch_bounds__(kLossThreads) void loss_partial_kernel(
    long long n, const float *__restrict__ img, const float *__restrict__ target,
    float *__restrict__ partial) {
    __shared__ float ws[kLossThreads / LIG_WAVE];
    const long long base = (long long)blockIdx.x * kLossChunk;
    float acc = 0.0f;
    if (base + kLossChunk <= n && ((reinterpret_cast<uintptr_t>(img) |
                                   reinterpret_cast<uintptr_t>(target)) & 15) == 0) {
        const float4 *x4 = reinterpret_cast<const float4 *>(img + base);
        const float4 *t4 = reinterpret_cast<const float4 *>(target + base);
#pragma unroll
        for (int k = 0; k < kLossItems / 4; ++k) {
            const int i = k * kLossThreads + threadIdx.x;
            const float4 x = x4[i], t = t4[i];
            acc += sq_err(x.x, t.x) + sq_err(x.y, t.y) + sq_err(x.z, t.z) + sq_err(x.w, t.w);
        }
    } else {
        for (int k = 0; k < kLossItems; ++k) {
            const long long i = base + (long long)k * kLossThreads + threadIdx.x;
            if (i < n) acc += sq_err(img[i], target[i]);
        }
    }
    for (int off = LIG_WAVE / 2; off > 0; off >>= 1) acc += __shfl_down(acc, off, LIG_WAVE);
    if ((threadIdx.x & (LIG_WAVE - 1)) == 0) ws[threadIdx.x / LIG_WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
#pragma unroll
        for (int w = 0; w < kLossThreads / LIG_WAVE; ++w) s += ws[w];
        partial[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kLossThreads) void loss_final_kernel(int nblocks,
                                                                 const float *__restrict__ partial,
                                                                 double inv_n,
                                                                 float *__restrict__ loss) {
    __shared__ double ws[kLossThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += kLossThreads) acc += partial[i];
    ws[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kLossThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) ws[threadIdx.x] += ws[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = (float)(ws[0] * inv_n);
}

__global__ __launch_bounds__(kLossThreads) void loss_grad_kernel(long long n,
                                                                const float *__restrict__ img,
                                                                const float *__restrict__ target,
                                                                const float *__restrict__ g,
                                                                float norm,
                                                                float *__restrict__ grad) {
    const float gs = g[0];
    const long long i0 = ((long long)blockIdx.x * kLossThreads + threadIdx.x) * 4;
    if (i0 + 4 <= n && ((reinterpret_cast<uintptr_t>(img) | reinterpret_cast<uintptr_t>(target) |
                         reinterpret_cast<uintptr_t>(grad)) & 15) == 0) {
        const float4 x = *reinterpret_cast<const float4 *>(img + i0);
        const float4 t = *reinterpret_cast<const float4 *>(target + i0);
        float4 r;
        r.x = (x.x >= 0.0f && x.x <= 1.0f) ? (norm * (x.x - t.x)) * gs : 0.0f;
        r.y = (x.y >= 0.0f && x.y <= 1.0f) ? (norm * (x.y - t.y)) * gs : 0.0f;
        r.z = (x.z >= 0.0f && x.z <= 1.0f) ? (norm * (x.z - t.z)) * gs : 0.0f;
        r.w = (x.w >= 0.0f && x.w <= 1.0f) ? (norm * (x.w - t.w)) * gs : 0.0f;
        *reinterpret_cast<float4 *>(grad + i0) = r;
    } else {
        for (long long i = i0; i < i0 + 4 && i < n; ++i) {
            const float x = img[i];
            grad[i] = (x >= 0.0f && x <= 1.0f) ? (norm * (x - target[i])) * gs : 0.0f;
        }
    }
}

}  // namespace
}  // namespace lig

using namespace lig;

extern "C" size_t lig_l2_loss_workspace_size(long long n) {
    const long long nb = n > 0 ? (n + kLossChunk - 1) / kLossChunk : 1;
    return (size_t)nb * sizeof(float) + 256;
}

extern "C" int lig_l2_clamped_loss_forward(long long n, const float *img, const float *target,
                                           float *loss, void *workspace,
                                           size_t workspace_bytes, void *stream) {
    if (n <= 0 || !img || !target || !loss) return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_l2_loss_workspace_size(n)) return LIG_ERR_WORKSPACE;
    const long long nb = (n + kLossChunk - 1) / kLossChunk;
    if (nb > INT32_MAX) return LIG_ERR_INVALID_ARGUMENT;
    hipStream_t st = as_stream(stream);
    float *partial = reinterpret_cast<float *>(workspace);
    hipLaunchKernelGGL(loss_partial_kernel, dim3((unsigned)nb), dim3(kLossThreads), 0, st, n, img,
                       target, partial);
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(kLossThreads), 0, st, (int)nb,
                       (const float *)partial, 1.0 / (double)n, loss);
    return launch_status();
}

extern "C" int lig_l2_clamped_loss_backward(long long n, const float *img, const float *target,
                                            const float *grad_loss, float *grad_img,
                                            void *stream) {
    if (n <= 0 || !img || !target || !grad_loss || !grad_img) return LIG_ERR_INVALID_ARGUMENT;
    const long long nthreads = (n + 3) / 4;
    const long long nb = (nthreads + kLossThreads - 1) / kLossThreads;
    if (nb > INT32_MAX) return LIG_ERR_INVALID_ARGUMENT;
    const float norm = (float)(2.0 / (double)n);
    hipLaunchKernelGGL(loss_grad_kernel, dim3((unsigned)nb), dim3(kLossThreads), 0,
                       as_stream(stream), n, img, target, grad_loss, norm, grad_img);
    return launch_status();
}
