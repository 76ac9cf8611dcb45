// Fused optimiser steps over one FP32 parameter tensor: Adam (torch.optim.Adam
// semantics, L2 weight decay added to the gradient) and Adan (the reference's
// optimizer.py:238-293 single-tensor algorithm).  One thread per element (4
// per thread when all pointers are 16-byte aligned): a step reads and writes
// param + state once, 5 (Adam) or 7 (Adan) floats read, 4 or 6 written.
#include <cmath>

#include "lig_device.h"

namespace lig {
namespace {

constexpr int kOptThreads = 256;

struct AdamArgs {
    float lr, b1, b2, eps, wd;
    float bc1, bc2_sqrt;  // 1 - b1^t, sqrt(1 - b2^t)
};

__device__ __forceinline__ void adam_elem(float &p, float g, float &m, float &v,
                                          const AdamArgs &a) {
    if (a.wd != 0.0f) g = __builtin_fmaf(a.wd, p, g);
    m = __builtin_fmaf(a.b1, m, (1.0f - a.b1) * g);
    v = __builtin_fmaf(a.b2, v, (1.0f - a.b2) * g * g);
    const float denom = __builtin_sqrtf(v) / a.bc2_sqrt + a.eps;
    p -= (a.lr / a.bc1) * m / denom;
}

__device__ __forceinline__ void adam_thread(long long n, float *__restrict__ p,
                                            const float *__restrict__ g, float *__restrict__ m,
                                            float *__restrict__ v, const AdamArgs &a, int vec4) {
    const long long t = (long long)blockIdx.x * kOptThreads + threadIdx.x;
    if (vec4) {
        const long long i = 4 * t;
        if (i + 4 <= n) {
            float4 P = reinterpret_cast<float4 *>(p)[t];
            const float4 G = reinterpret_cast<const float4 *>(g)[t];
            float4 M = reinterpret_cast<float4 *>(m)[t];
            float4 V = reinterpret_cast<float4 *>(v)[t];
            adam_elem(P.x, G.x, M.x, V.x, a);
            adam_elem(P.y, G.y, M.y, V.y, a);
            adam_elem(P.z, G.z, M.z, V.z, a);
            adam_elem(P.w, G.w, M.w, V.w, a);
            reinterpret_cast<float4 *>(p)[t] = P;
            reinterpret_cast<float4 *>(m)[t] = M;
            reinterpret_cast<float4 *>(v)[t] = V;
        } else {
            for (long long k = i; k < n; ++k) adam_elem(p[k], g[k], m[k], v[k], a);
        }
    } else if (t < n) {
        adam_elem(p[t], g[t], m[t], v[t], a);
    }
}

__global__ __launch_bounds__(kOptThreads) void adam_kernel(long long n, float *__restrict__ p,
                                                           const float *__restrict__ g,
                                                           float *__restrict__ m,
                                                           float *__restrict__ v, AdamArgs a,
                                                           int vec4) {
    adam_thread(n, p, g, m, v, a, vec4);
}

// Step control of a captured fit step: {step, skipped, max count, ok}.
struct StepControl {
    int32_t step, skipped, max_count, ok;
};

// Adam with the step number and its bias corrections taken from device
// memory (so one captured launch serves every step); does nothing when the
// step was not admitted.
__global__ __launch_bounds__(kOptThreads) void adam_device_kernel(
    long long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
    float *__restrict__ v, AdamArgs a, int vec4, const StepControl *__restrict__ ctl,
    const float2 *__restrict__ bias) {
    if (!ctl->ok) return;
    const float2 b = bias[ctl->step - 1];
    a.bc1 = b.x;
    a.bc2_sqrt = b.y;
    adam_thread(n, p, g, m, v, a, vec4);
}

// Admits a step when its intersection count fits the capacity and the bias
// table still has an entry for it.  One thread.
__global__ void step_begin_kernel(const int32_t *__restrict__ count, int capacity, int table_len,
                                  StepControl *__restrict__ ctl) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int c = *count;
    StepControl s = *ctl;
    s.ok = (c <= capacity && s.step < table_len) ? 1 : 0;
    if (s.ok)
        s.step += 1;
    else
        s.skipped += 1;
    s.max_count = max(s.max_count, c);
    *ctl = s;
}

// The bias corrections lig_adam_step uses for a step: 1 - beta1^t and
// sqrt(1 - beta2^t), formed in FP64 from the FP32 betas and rounded once.
void adam_bias(float beta1, float beta2, int step, float *bc1, float *bc2_sqrt) {
    *bc1 = (float)(1.0 - std::pow((double)beta1, (double)step));
    *bc2_sqrt = (float)std::sqrt(1.0 - std::pow((double)beta2, (double)step));
}

struct AdanArgs {
    float lr, b1, b2, b3, eps, wd, clip;
    float step_size, step_size_diff, bc3_sqrt;
    int no_prox, first;
};

// optimizer.py:264-293 per element; neg_pre holds -g of the previous step
// (set to -g * clip on the first step, so d = 0 there).
__device__ __forceinline__ void adan_elem(float &p, float g, float &m, float &v, float &diff,
                                          float &neg_pre, const AdanArgs &a) {
    g *= a.clip;
    if (a.first) neg_pre = -g;
    const float d = neg_pre + g;
    m = a.b1 * m + (1.0f - a.b1) * g;
    diff = a.b2 * diff + (1.0f - a.b2) * d;
    const float t = a.b2 * d + g;
    v = a.b3 * v + (1.0f - a.b3) * t * t;
    const float denom = __builtin_sqrtf(v) / a.bc3_sqrt + a.eps;
    if (a.no_prox) {
        p *= 1.0f - a.lr * a.wd;
        p -= a.step_size * m / denom;
        p -= a.step_size_diff * diff / denom;
    } else {
        p -= a.step_size * m / denom;
        p -= a.step_size_diff * diff / denom;
        p /= 1.0f + a.lr * a.wd;
    }
    neg_pre = -g;
}

__global__ __launch_bounds__(kOptThreads) void adan_kernel(long long n, float *__restrict__ p,
                                                           const float *__restrict__ g,
                                                           float *__restrict__ m,
                                                           float *__restrict__ v,
                                                           float *__restrict__ diff,
                                                           float *__restrict__ neg_pre,
                                                           AdanArgs a) {
    const long long t = (long long)blockIdx.x * kOptThreads + threadIdx.x;
    if (t < n) adan_elem(p[t], g[t], m[t], v[t], diff[t], neg_pre[t], a);
}

bool aligned16(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; }

}  // namespace
}  // namespace lig

using namespace lig;

extern "C" int lig_adam_step(long long n, float *param, const float *grad, float *exp_avg,
                             float *exp_avg_sq, float lr, float beta1, float beta2, float eps,
                             float weight_decay, int step, void *stream) {
    if (n < 0 || step < 1 || (n > 0 && (!param || !grad || !exp_avg || !exp_avg_sq)))
        return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    AdamArgs a;
    a.lr = lr;
    a.b1 = beta1;
    a.b2 = beta2;
    a.eps = eps;
    a.wd = weight_decay;
    adam_bias(beta1, beta2, step, &a.bc1, &a.bc2_sqrt);
    const int vec4 = aligned16(param) && aligned16(grad) && aligned16(exp_avg) &&
                     aligned16(exp_avg_sq);
    const long long threads = vec4 ? (n + 3) / 4 : n;
    const long long blocks = (threads + kOptThreads - 1) / kOptThreads;
    if (blocks > INT32_MAX) return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)blocks), dim3(kOptThreads), 0,
                       as_stream(stream), n, param, grad, exp_avg, exp_avg_sq, a, vec4);
    return launch_status();
}

extern "C" int lig_adan_step(long long n, float *param, const float *grad, float *exp_avg,
                             float *exp_avg_sq, float *exp_avg_diff, float *neg_pre_grad,
                             float lr, float beta1, float beta2, float beta3, float eps,
                             float weight_decay, int no_prox, float clip_global_grad_norm,
                             int step, void *stream) {
    if (n < 0 || step < 1 ||
        (n > 0 &&
         (!param || !grad || !exp_avg || !exp_avg_sq || !exp_avg_diff || !neg_pre_grad)))
        return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    AdanArgs a;
    a.lr = lr;
    a.b1 = beta1;
    a.b2 = beta2;
    a.b3 = beta3;
    a.eps = eps;
    a.wd = weight_decay;
    a.clip = clip_global_grad_norm;
    const double bc1 = 1.0 - std::pow((double)beta1, (double)step);
    const double bc2 = 1.0 - std::pow((double)beta2, (double)step);
    const double bc3 = 1.0 - std::pow((double)beta3, (double)step);
    a.step_size = (float)(lr / bc1);
    a.step_size_diff = (float)(lr * beta2 / bc2);
    a.bc3_sqrt = (float)std::sqrt(bc3);
    a.no_prox = no_prox != 0;
    a.first = step == 1;
    const long long blocks = (n + kOptThreads - 1) / kOptThreads;
    if (blocks > INT32_MAX) return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(adan_kernel, dim3((unsigned)blocks), dim3(kOptThreads), 0,
                       as_stream(stream), n, param, grad, exp_avg, exp_avg_sq, exp_avg_diff,
                       neg_pre_grad, a);
    return launch_status();
}

extern "C" int lig_adam_bias_corrections(float beta1, float beta2, int first_step, int count,
                                         float *out) {
    if (first_step < 1 || count < 0 || (count > 0 && !out)) return LIG_ERR_INVALID_ARGUMENT;
    for (int k = 0; k < count; ++k) adam_bias(beta1, beta2, first_step + k, &out[2 * k],
                                              &out[2 * k + 1]);
    return LIG_OK;
}

extern "C" int lig_fit_step_begin(const int32_t *count, int capacity, int table_len,
                                  void *control, void *stream) {
    if (!count || !control || capacity < 0 || table_len < 0) return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(64), 0, as_stream(stream), count,
                       capacity, table_len, reinterpret_cast<StepControl *>(control));
    return launch_status();
}

extern "C" int lig_adam_step_device(long long n, float *param, const float *grad, float *exp_avg,
                                    float *exp_avg_sq, float lr, float beta1, float beta2,
                                    float eps, float weight_decay, const void *control,
                                    const float *bias_table, void *stream) {
    if (n < 0 || !control || !bias_table ||
        (n > 0 && (!param || !grad || !exp_avg || !exp_avg_sq)))
        return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    AdamArgs a;
    a.lr = lr;
    a.b1 = beta1;
    a.b2 = beta2;
    a.eps = eps;
    a.wd = weight_decay;
    a.bc1 = 1.0f;
    a.bc2_sqrt = 1.0f;
    const int vec4 = aligned16(param) && aligned16(grad) && aligned16(exp_avg) &&
                     aligned16(exp_avg_sq);
    const long long threads = vec4 ? (n + 3) / 4 : n;
    const long long blocks = (threads + kOptThreads - 1) / kOptThreads;
    if (blocks > INT32_MAX) return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(adam_device_kernel, dim3((unsigned)blocks), dim3(kOptThreads), 0,
                       as_stream(stream), n, param, grad, exp_avg, exp_avg_sq, a, vec4,
                       reinterpret_cast<const StepControl *>(control),
                       reinterpret_cast<const float2 *>(bias_table));
    return launch_status();
}
