// Gradient-aware Hermite (bicubic) upscaler: forward and a deterministic
// backward.
//
// Forward: one thread per destination pixel maps it into the source ROI
// (half-pixel grid, ROI extent truncated to int), takes the four clamped
// corners and evaluates the bicubic Hermite patch from (f, fx, fy, fxy).
//
// Backward: the forward is separable -- the weight of source pixel (sy, sx)
// in destination pixel (dy, dx) is wx(sx, dx) * wy(sy, dy) with value
// weights h0/h1 and derivative weights h2/h3 -- so the transpose is computed
// as two gathers instead of the reference's 48 float atomics per destination
// pixel: pass X sums each destination row over the destination columns that
// touch source column sx (a contiguous range, because the map is monotone),
// pass Y does the same over rows.  Summation order is fixed, so the result
// is bit-reproducible.
//
// Reference behaviour followed (paths relative to the reference repository):
//   gsplat2d/gsplat2d/cuda/csrc/gradient_aware_upscaler.cu:137-371,
//   bindings.cu:508-605, gsplat2d/gsplat2d/upscale.py:50-82.
#include "lig_device.h"

namespace lig {
namespace {

struct Axis {
    int dst, src;
    float roi0, roi1;
};

// Source coordinate of destination index d along one axis (upscaler.cu:201-218).
__device__ __forceinline__ void axis_map(const Axis &ax, int d, int &c0, int &c1, float &t) {
#pragma clang fp contract(off)
    const int roi_w = f2i_sat(ax.roi1 - ax.roi0);
    const float s = ax.roi0 + div_rn(((float)d + 0.5f) * (float)roi_w, (float)ax.dst) - 0.5f;
    const int k = f2i_sat(floorf(s));
    t = s - (float)k;
    // The reference leaves c1 < 0 unclamped (a read before the image when the
    // ROI starts left of -0.5); it is clamped to 0 here, like c0.
    const int kc = min(k, ax.src - 1);
    c1 = max(0, min(kc + 1, ax.src - 1));
    c0 = max(0, kc);
}

// Hermite basis (C_inv^T [1, t, t^2, t^3]).
__device__ __forceinline__ void hermite(float t, float h[4]) {
#pragma clang fp contract(off)
    const float t2 = t * t, t3 = t2 * t;
    h[0] = 1.0f - 3.0f * t2 + 2.0f * t3;
    h[1] = 3.0f * t2 - 2.0f * t3;
    h[2] = t - 2.0f * t2 + t3;
    h[3] = t3 - t2;
}

// Destination pixel (x, y) of the Hermite patch upscale (upscaler.cu:181-246).
// This and bilinear_pixel evaluate without FP contraction, so every kernel
// that inlines them gets the same bits (contraction choices depend on the
// surrounding code).
__device__ __forceinline__ void hermite_pixel(const Axis &ax, const Axis &ay, int x, int y,
                                              const float *__restrict__ f,
                                              const float *__restrict__ fx,
                                              const float *__restrict__ fy,
                                              const float *__restrict__ fxy, float v[3]) {
#pragma clang fp contract(off)
    int x0, x1, y0, y1;
    float tx, ty, hx[4], hy[4];
    axis_map(ax, x, x0, x1, tx);
    axis_map(ay, y, y0, y1, ty);
    hermite(tx, hx);
    hermite(ty, hy);
    const int w = ax.src;
    const size_t i00 = 3 * ((size_t)y0 * w + x0), i01 = 3 * ((size_t)y0 * w + x1);
    const size_t i10 = 3 * ((size_t)y1 * w + x0), i11 = 3 * ((size_t)y1 * w + x1);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const float c0 = f[i00 + ch] * hy[0] + f[i10 + ch] * hy[1] + fy[i00 + ch] * hy[2] +
                         fy[i10 + ch] * hy[3];
        const float c1 = f[i01 + ch] * hy[0] + f[i11 + ch] * hy[1] + fy[i01 + ch] * hy[2] +
                         fy[i11 + ch] * hy[3];
        const float c2 = fx[i00 + ch] * hy[0] + fx[i10 + ch] * hy[1] + fxy[i00 + ch] * hy[2] +
                         fxy[i10 + ch] * hy[3];
        const float c3 = fx[i01 + ch] * hy[0] + fx[i11 + ch] * hy[1] + fxy[i01 + ch] * hy[2] +
                         fxy[i11 + ch] * hy[3];
        v[ch] = hx[0] * c0 + hx[1] * c1 + hx[2] * c2 + hx[3] * c3;
    }
}

__global__ __launch_bounds__(256) void upscale_fwd_kernel(Axis ax, Axis ay,
                                                          const float *__restrict__ f,
                                                          const float *__restrict__ fx,
                                                          const float *__restrict__ fy,
                                                          const float *__restrict__ fxy,
                                                          float *__restrict__ out) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= ax.dst || y >= ay.dst) return;
    float v[3];
    hermite_pixel(ax, ay, x, y, f, fx, fy, fxy, v);
    const size_t o = 3 * ((size_t)y * ax.dst + x);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) out[o + ch] = v[ch];
}

// For each source index s along an axis, the destination range [lo, hi] whose
// corner pair (c0, c1) contains s.  c0 <= c1 <= c0 + 1 and both are
// non-decreasing in d, so the range is {d : c0(d) <= s <= c1(d)}.
__global__ void axis_range_kernel(Axis ax, int2 *__restrict__ range) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ax.src) return;
    int lo = 0, hi = ax.dst;  // first d with c1(d) >= s
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        int c0, c1;
        float t;
        axis_map(ax, mid, c0, c1, t);
        if (c1 >= s) hi = mid; else lo = mid + 1;
    }
    const int first = lo;
    lo = 0;
    hi = ax.dst;  // first d with c0(d) > s
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        int c0, c1;
        float t;
        axis_map(ax, mid, c0, c1, t);
        if (c0 > s) hi = mid; else lo = mid + 1;
    }
    range[s] = make_int2(first, lo - 1);
}

// Weights of source index s in destination index d: value weight and
// derivative weight (both corners when they coincide at a border).
__device__ __forceinline__ void axis_weights(const Axis &ax, int d, int s, float &wv, float &wd) {
    int c0, c1;
    float t, h[4];
    axis_map(ax, d, c0, c1, t);
    hermite(t, h);
    wv = (s == c0 ? h[0] : 0.0f) + (s == c1 ? h[1] : 0.0f);
    wd = (s == c0 ? h[2] : 0.0f) + (s == c1 ? h[3] : 0.0f);
}

// Pass X: for destination row y and source column sx,
//   Tv[y][sx] = sum_dx g[y][dx] * wv_x,  Td[y][sx] = sum_dx g[y][dx] * wd_x.
__global__ __launch_bounds__(256) void upscale_bwd_x_kernel(Axis ax, int dst_h,
                                                            const int2 *__restrict__ xr,
                                                            const float *__restrict__ g,
                                                            float *__restrict__ tv,
                                                            float *__restrict__ td) {
    const int sx = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (sx >= ax.src || y >= dst_h) return;
    const int2 r = xr[sx];
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, d0 = 0.f, d1 = 0.f, d2 = 0.f;
    const float *row = g + 3 * (size_t)y * ax.dst;
    for (int dx = r.x; dx <= r.y; ++dx) {
        float wv, wd;
        axis_weights(ax, dx, sx, wv, wd);
        const float g0 = row[3 * dx], g1 = row[3 * dx + 1], g2 = row[3 * dx + 2];
        v0 += g0 * wv; v1 += g1 * wv; v2 += g2 * wv;
        d0 += g0 * wd; d1 += g1 * wd; d2 += g2 * wd;
    }
    const size_t o = 3 * ((size_t)y * ax.src + sx);
    tv[o] = v0; tv[o + 1] = v1; tv[o + 2] = v2;
    td[o] = d0; td[o + 1] = d1; td[o + 2] = d2;
}

// Pass Y: for source pixel (sy, sx), sum over destination rows dy in range:
//   grad_f   += Tv * wv_y    grad_fx  += Td * wv_y
//   grad_fy  += Tv * wd_y    grad_fxy += Td * wd_y
__global__ __launch_bounds__(256) void upscale_bwd_y_kernel(Axis ay, int src_w,
                                                            const int2 *__restrict__ yr,
                                                            const float *__restrict__ tv,
                                                            const float *__restrict__ td,
                                                            float *__restrict__ gf,
                                                            float *__restrict__ gfx,
                                                            float *__restrict__ gfy,
                                                            float *__restrict__ gfxy) {
    const int sx = blockIdx.x * 64 + (threadIdx.x & 63);
    const int sy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (sx >= src_w || sy >= ay.src) return;
    const int2 r = yr[sy];
    float f[3] = {0.f, 0.f, 0.f}, fxv[3] = {0.f, 0.f, 0.f};
    float fyv[3] = {0.f, 0.f, 0.f}, fxyv[3] = {0.f, 0.f, 0.f};
    for (int dy = r.x; dy <= r.y; ++dy) {
        float wv, wd;
        axis_weights(ay, dy, sy, wv, wd);
        const size_t i = 3 * ((size_t)dy * src_w + sx);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const float a = tv[i + ch], b = td[i + ch];
            f[ch] += a * wv;
            fxv[ch] += b * wv;
            fyv[ch] += a * wd;
            fxyv[ch] += b * wd;
        }
    }
    const size_t o = 3 * ((size_t)sy * src_w + sx);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        gf[o + ch] = f[ch];
        gfx[o + ch] = fxv[ch];
        gfy[o + ch] = fyv[ch];
        gfxy[o + ch] = fxyv[ch];
    }
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

bool roi_ok(int src_h, int src_w, int dst_h, int dst_w, float x1, float y1, float x2,
            float y2) {
    if (src_h <= 0 || src_w <= 0 || dst_h <= 0 || dst_w <= 0) return false;
    if (!(x2 >= x1) || !(y2 >= y1)) return false;  // also rejects NaN
    return true;
}

// Bilinear resize with the half-pixel grid of torch.nn.functional.interpolate
// (mode="bilinear", align_corners=False), the resampling train.py:103-105 and
// :193 apply between levels: destination index d of a full_h x full_w output
// samples source coordinate max(scale * (d + 0.5) - 0.5, 0), scale = src /
// full, between floor and floor + 1 (clamped to the last index).  The kernel
// writes the out_h x out_w window at (off_y, off_x) of that output, so a crop
// holds exactly the full output's values at its pixels.
// Pixel (gx, gy) of the full_h x full_w bilinear resize of src (scale =
// src / full per axis).
__device__ __forceinline__ void bilinear_pixel(int src_h, int src_w, float scale_y,
                                               float scale_x, int gy, int gx,
                                               const float *__restrict__ src, float v[3]) {
#pragma clang fp contract(off)
    const float fy = fmaxf(scale_y * ((float)gy + 0.5f) - 0.5f, 0.0f);
    const float fx = fmaxf(scale_x * ((float)gx + 0.5f) - 0.5f, 0.0f);
    const int y0 = min((int)fy, src_h - 1), x0 = min((int)fx, src_w - 1);
    const int y1 = y0 + (y0 < src_h - 1 ? 1 : 0), x1 = x0 + (x0 < src_w - 1 ? 1 : 0);
    const float ly = fy - (float)y0, lx = fx - (float)x0;
    const float hy = 1.0f - ly, hx = 1.0f - lx;
    const size_t i00 = 3 * ((size_t)y0 * src_w + x0), i01 = 3 * ((size_t)y0 * src_w + x1);
    const size_t i10 = 3 * ((size_t)y1 * src_w + x0), i11 = 3 * ((size_t)y1 * src_w + x1);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
        v[ch] = hy * (hx * src[i00 + ch] + lx * src[i01 + ch]) +
                ly * (hx * src[i10 + ch] + lx * src[i11 + ch]);
}

__global__ __launch_bounds__(256) void bilinear_kernel(int src_h, int src_w, float scale_y,
                                                       float scale_x, int off_y, int off_x,
                                                       int out_h, int out_w,
                                                       const float *__restrict__ src,
                                                       float *__restrict__ out) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= out_w || y >= out_h) return;
    float v[3];
    bilinear_pixel(src_h, src_w, scale_y, scale_x, y + off_y, x + off_x, src, v);
    const size_t o = 3 * ((size_t)y * out_w + x);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) out[o + ch] = v[ch];
}

// clamp(v, 0, 1) as torch.clamp computes it (NaN passes through).
__device__ __forceinline__ float clamp01(float v) {
    return v != v ? v : fminf(fmaxf(v, 0.0f), 1.0f);
}

// Batched crop composition: block b belongs to the crop whose block range
// holds it; its 256 threads take 256 consecutive pixels of the crop in
// row-major order.
__global__ __launch_bounds__(256) void compose_crops_kernel(int num_crops,
                                                            const LigCrop *__restrict__ crops) {
    const int b = blockIdx.x;
    int lo = 0, hi = num_crops - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (crops[mid].block_off <= b) lo = mid; else hi = mid - 1;
    }
    const LigCrop &c = crops[lo];
    LIG_CHECK(b >= c.block_off && (long long)(b - c.block_off) * 256 < (long long)c.w * c.h);
    const long long p = (long long)(b - c.block_off) * 256 + threadIdx.x;
    if (p >= (long long)c.w * c.h) return;
    const int y = (int)(p / c.w), x = (int)(p - (long long)y * c.w);
    LIG_CHECK(x + c.off_x < c.view_w && x >= 0 && y >= 0);
    float up[3];
    if (c.mode == 0) {
        const Axis ax{c.w, c.src_w, c.roi_x1, c.roi_x2};
        const Axis ay{c.h, c.src_h, c.roi_y1, c.roi_y2};
        hermite_pixel(ax, ay, x, y, c.f, c.fx, c.fy, c.fxy, up);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) up[ch] = clamp01(up[ch]);
    } else {
        const float sy = (float)c.src_h / (float)c.full_h, sx = (float)c.src_w / (float)c.full_w;
        bilinear_pixel(c.src_h, c.src_w, sy, sx, y + c.win_y, x + c.win_x, c.f, up);
    }
    const float *l2 = c.level2 ? c.level2 + 3 * ((size_t)(y + c.off_y) * c.view_w + x + c.off_x)
                               : nullptr;
    float *o = c.out + 3 * p;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
#pragma clang fp contract(off)
        const float a = clamp01(l2 ? l2[ch] : 1.0f);
        o[ch] = clamp01(((a * c.span + up[ch]) - 0.5f) + c.lo);
    }
}

}  // namespace
}  // namespace lig

using namespace lig;

extern "C" int lig_bilinear_resize(int src_h, int src_w, int full_h, int full_w, int off_y,
                                   int off_x, int out_h, int out_w, const float *src,
                                   float *output, void *stream) {
    if (src_h <= 0 || src_w <= 0 || full_h <= 0 || full_w <= 0 || out_h < 0 || out_w < 0 ||
        off_y < 0 || off_x < 0 || (long long)off_y + out_h > full_h ||
        (long long)off_x + out_w > full_w)
        return LIG_ERR_INVALID_ARGUMENT;
    if (out_h == 0 || out_w == 0) return LIG_OK;
    if (!src || !output) return LIG_ERR_INVALID_ARGUMENT;
    const float sy = (float)src_h / (float)full_h, sx = (float)src_w / (float)full_w;
    const dim3 grid(ceil_div(out_w, 64), ceil_div(out_h, 4));
    hipLaunchKernelGGL(bilinear_kernel, grid, dim3(256), 0, as_stream(stream), src_h, src_w, sy,
                       sx, off_y, off_x, out_h, out_w, src, output);
    return launch_status();
}

extern "C" int lig_compose_crops(int num_crops, const LigCrop *crops, int total_blocks,
                                  void *stream) {
    if (num_crops < 1 || total_blocks < 0 || !crops) return LIG_ERR_INVALID_ARGUMENT;
    if (total_blocks == 0) return LIG_OK;
    hipLaunchKernelGGL(compose_crops_kernel, dim3((unsigned)total_blocks), dim3(256), 0,
                       as_stream(stream), num_crops, crops);
    return launch_status();
}

extern "C" int lig_gradient_aware_upscale_forward(int src_h, int src_w, int dst_h, int dst_w,
                                                  float roi_x1, float roi_y1, float roi_x2,
                                                  float roi_y2, const float *render,
                                                  const float *dx, const float *dy,
                                                  const float *dxy, float *output,
                                                  void *stream) {
    if (!roi_ok(src_h, src_w, dst_h, dst_w, roi_x1, roi_y1, roi_x2, roi_y2))
        return LIG_ERR_INVALID_ARGUMENT;
    if (!render || !dx || !dy || !dxy || !output) return LIG_ERR_INVALID_ARGUMENT;
    const Axis ax{dst_w, src_w, roi_x1, roi_x2};
    const Axis ay{dst_h, src_h, roi_y1, roi_y2};
    const dim3 grid(ceil_div(dst_w, 64), ceil_div(dst_h, 4));
    hipLaunchKernelGGL(upscale_fwd_kernel, grid, dim3(256), 0, as_stream(stream), ax, ay, render,
                       dx, dy, dxy, output);
    return launch_status();
}

extern "C" size_t lig_upscale_backward_workspace_size(int src_h, int src_w, int dst_h,
                                                      int dst_w) {
    if (src_h <= 0 || src_w <= 0 || dst_h <= 0 || dst_w <= 0) return 0;
    return 2 * align_up((size_t)dst_h * src_w * 3 * 4) + align_up((size_t)src_w * 8) +
           align_up((size_t)src_h * 8);
}

extern "C" int lig_gradient_aware_upscale_backward(int src_h, int src_w, int dst_h, int dst_w,
                                                   float roi_x1, float roi_y1, float roi_x2,
                                                   float roi_y2, const float *grad_output,
                                                   float *grad_render, float *grad_dx,
                                                   float *grad_dy, float *grad_dxy,
                                                   void *workspace, size_t workspace_bytes,
                                                   void *stream) {
    if (!roi_ok(src_h, src_w, dst_h, dst_w, roi_x1, roi_y1, roi_x2, roi_y2))
        return LIG_ERR_INVALID_ARGUMENT;
    if (!grad_output || !grad_render || !grad_dx || !grad_dy || !grad_dxy)
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace ||
        workspace_bytes < lig_upscale_backward_workspace_size(src_h, src_w, dst_h, dst_w))
        return LIG_ERR_WORKSPACE;
    const Axis ax{dst_w, src_w, roi_x1, roi_x2};
    const Axis ay{dst_h, src_h, roi_y1, roi_y2};
    char *p = reinterpret_cast<char *>(workspace);
    float *tv = reinterpret_cast<float *>(p);
    p += align_up((size_t)dst_h * src_w * 3 * 4);
    float *td = reinterpret_cast<float *>(p);
    p += align_up((size_t)dst_h * src_w * 3 * 4);
    int2 *xr = reinterpret_cast<int2 *>(p);
    p += align_up((size_t)src_w * 8);
    int2 *yr = reinterpret_cast<int2 *>(p);
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(axis_range_kernel, dim3(ceil_div(src_w, 256)), dim3(256), 0, st, ax, xr);
    hipLaunchKernelGGL(axis_range_kernel, dim3(ceil_div(src_h, 256)), dim3(256), 0, st, ay, yr);
    hipLaunchKernelGGL(upscale_bwd_x_kernel, dim3(ceil_div(src_w, 64), ceil_div(dst_h, 4)),
                       dim3(256), 0, st, ax, dst_h, (const int2 *)xr, grad_output, tv, td);
    hipLaunchKernelGGL(upscale_bwd_y_kernel, dim3(ceil_div(src_w, 64), ceil_div(src_h, 4)),
                       dim3(256), 0, st, ay, src_w, (const int2 *)yr, (const float *)tv,
                       (const float *)td, grad_render, grad_dx, grad_dy, grad_dxy);
    return launch_status();
}
