// Projection of 2-D Gaussians: covariance (or Cholesky factor) -> conic,
// SnugBox extent and tile count; and the backward of both forms.
//
// Numerics contract: IEEE FP32 in the reference's operation order with FP
// contraction disabled, division and square root correctly rounded and the
// logarithm taken in FP64 and rounded once.  Results are therefore
// bit-reproducible on any IEEE host, which is how they are checked.
//
// Reference behaviour followed (paths relative to the reference repository):
//   gsplat2d/gsplat2d/cuda/csrc/helpers.cuh:47-93, forward.cu:13-105,
//   backward.cu:185-247, bindings.cu:19-66.
#include "lig_device.h"

namespace lig {
namespace {

constexpr int kThreads = 256;

// SnugBox bounds (helpers.cuh:47-81).  Returns false when det == 0.
__device__ __forceinline__ bool cov_bounds(float ca, float cb, float cc, float opacity,
                                           float &ka, float &kb, float &kc, float &ex,
                                           float &ey) {
#pragma clang fp contract(off)
    const float det = ca * cc - cb * cb;
    if (det == 0.0f) return false;
    const float inv_det = div_rn(1.0f, det);
    ka = cc * inv_det;
    kb = -cb * inv_det;
    kc = ca * inv_det;
    const float t = 2.0f * log_rd(255.0f * opacity);
    if (t <= 0.0f) {
        ex = 0.0f;
        ey = 0.0f;
        return true;
    }
    ex = ceilf(sqrt_rn(div_rn(t * kc, inv_det)));
    ey = ceilf(sqrt_rn(div_rn(t * ka, inv_det)));
    return true;
}

// One Gaussian of the forward projection; writes every output slot (zeros
// where the reference leaves its zero fill).
__device__ __forceinline__ bool project_one(int i, float ca, float cb, float cc, float mx,
                                            float my, float o, int tiles_x, int tiles_y,
                                            int block_width,
                                            float *__restrict__ xys,
                                            float *__restrict__ extents,
                                            float *__restrict__ conics,
                                            int32_t *__restrict__ num_tiles_hit) {
    float ka = 0.0f, kb = 0.0f, kc = 0.0f, ex = 0.0f, ey = 0.0f;
    const bool ok = cov_bounds(ca, cb, cc, o, ka, kb, kc, ex, ey);
    int area = 0;
    if (ok) {
        int x0, y0, x1, y1;
        tile_bbox(mx, my, ex, ey, tiles_x, tiles_y, block_width, x0, y0, x1, y1);
        area = (int)((unsigned)(x1 - x0) * (unsigned)(y1 - y0));
    }
    const bool hit = area > 0;
    conics[3 * i + 0] = ka;
    conics[3 * i + 1] = kb;
    conics[3 * i + 2] = kc;
    extents[2 * i + 0] = hit ? ex : 0.0f;
    extents[2 * i + 1] = hit ? ey : 0.0f;
    xys[2 * i + 0] = hit ? mx : 0.0f;
    xys[2 * i + 1] = hit ? my : 0.0f;
    num_tiles_hit[i] = hit ? area : 0;
    return hit;
}

__global__ __launch_bounds__(kThreads) void project_cov_kernel(
    int n, const float *__restrict__ cov2d, const float *__restrict__ means2d,
    const float *__restrict__ opacities, int tiles_x, int tiles_y, int block_width,
    float *__restrict__ xys, float *__restrict__ extents, float *__restrict__ conics,
    int32_t *__restrict__ num_tiles_hit) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    project_one(i, cov2d[3 * i], cov2d[3 * i + 1], cov2d[3 * i + 2], means2d[2 * i],
                means2d[2 * i + 1], opacities ? opacities[i] : 1.0f, tiles_x, tiles_y,
                block_width, xys, extents, conics, num_tiles_hit);
}

__global__ __launch_bounds__(kThreads) void project_chol_kernel(
    int n, const float *__restrict__ chol, const float *__restrict__ means2d,
    const float *__restrict__ opacities, int tiles_x, int tiles_y, int block_width,
    float *__restrict__ xys, float *__restrict__ extents, float *__restrict__ conics,
    int32_t *__restrict__ num_tiles_hit) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float l11 = chol[3 * i], l21 = chol[3 * i + 1], l22 = chol[3 * i + 2];
    const float ca = l11 * l11;
    const float cb = l11 * l21;
    const float cc = l21 * l21 + l22 * l22;
    project_one(i, ca, cb, cc, means2d[2 * i], means2d[2 * i + 1],
                opacities ? opacities[i] : 1.0f, tiles_x, tiles_y, block_width, xys, extents,
                conics, num_tiles_hit);
}

// Batched fit: Gaussian i belongs to the image of `tab` whose Gaussian range
// holds it and is projected onto that image's canvas (16 x 16 tiles).
__global__ __launch_bounds__(kThreads) void project_batched_chol_kernel(
    int n, const float *__restrict__ chol, const float *__restrict__ means2d,
    const BatchImage *__restrict__ tab, int nimg, float *__restrict__ xys,
    float *__restrict__ extents, float *__restrict__ conics,
    int32_t *__restrict__ num_tiles_hit) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const BatchImage &im = tab[batch_find<1>(tab, nimg, i)];
    const float l11 = chol[3 * i], l21 = chol[3 * i + 1], l22 = chol[3 * i + 2];
    const float ca = l11 * l11;
    const float cb = l11 * l21;
    const float cc = l21 * l21 + l22 * l22;
    project_one(i, ca, cb, cc, means2d[2 * i], means2d[2 * i + 1], 1.0f, im.tiles_x,
                im.tiles_y, 16, xys, extents, conics, num_tiles_hit);
}

// Projection into a view: the Gaussian seen in a window with origin (ox, oy)
// of the source canvas, sampled at (sx, sy) output pixels per source pixel:
//   mean' = (mean - o) * s,  L' = diag(s) L  (Cholesky: l11 sx, l21 sy, l22 sy),
//   cov' = S cov S  (a sx sx, b sx sy, c sy sy),
// the transform of the reference's scale_to (minimal_trainer_upscale.py:139-159)
// with an origin, in FP32 without contraction; then the projection above.
// Input j of (par, means2d) is written to output slot i.
__device__ __forceinline__ bool project_view_one(int i, int j, int cholesky,
                                                 const float *__restrict__ par,
                                                 const float *__restrict__ means2d, float o,
                                                 float ox, float oy, float sx, float sy,
                                                 int tiles_x, int tiles_y, int block_width,
                                                 float *__restrict__ xys,
                                                 float *__restrict__ extents,
                                                 float *__restrict__ conics,
                                                 int32_t *__restrict__ num_tiles_hit) {
#pragma clang fp contract(off)
    const float mx = (means2d[2 * j] - ox) * sx;
    const float my = (means2d[2 * j + 1] - oy) * sy;
    float ca, cb, cc;
    if (cholesky) {
        const float l11 = par[3 * j] * sx, l21 = par[3 * j + 1] * sy, l22 = par[3 * j + 2] * sy;
        ca = l11 * l11;
        cb = l11 * l21;
        cc = l21 * l21 + l22 * l22;
    } else {
        ca = (par[3 * j] * sx) * sx;
        cb = (par[3 * j + 1] * sx) * sy;
        cc = (par[3 * j + 2] * sy) * sy;
    }
    return project_one(i, ca, cb, cc, mx, my, o, tiles_x, tiles_y, block_width, xys, extents,
                       conics, num_tiles_hit);
}

__global__ __launch_bounds__(kThreads) void project_view_kernel(
    int n, int cholesky, const float *__restrict__ par, const float *__restrict__ means2d,
    const float *__restrict__ opacities, float ox, float oy, float sx, float sy, int tiles_x,
    int tiles_y, int block_width, float *__restrict__ xys, float *__restrict__ extents,
    float *__restrict__ conics, int32_t *__restrict__ num_tiles_hit) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    project_view_one(i, i, cholesky, par, means2d, opacities ? opacities[i] : 1.0f, ox, oy, sx,
                     sy, tiles_x, tiles_y, block_width, xys, extents, conics, num_tiles_hit);
}

// Several views in one launch: output slot i belongs to the view whose
// [gauss_off, gauss_off + n) range holds it (views sorted by gauss_off,
// ranges back to back from 0) and is that view's Gaussian i - gauss_off.
// A view with a non-positive size or scale emits nothing.  Colors are copied
// for the slots whose Gaussian hits a tile (no other slot is read later).
__device__ __forceinline__ int find_view(const LigView *__restrict__ views, int num_views,
                                         int key) {
    int lo = 0, hi = num_views - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].gauss_off <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void project_slot(const LigView &v, int i, int block_width,
                                             float *__restrict__ xys,
                                             float *__restrict__ extents,
                                             float *__restrict__ conics,
                                             int32_t *__restrict__ num_tiles_hit,
                                             float *__restrict__ colors) {
    const int j = i - v.gauss_off;
    LIG_CHECK(j >= 0);
    const bool ok = j < v.n && v.img_height > 0 && v.img_width > 0 && v.scale_x > 0.0f &&
                    v.scale_y > 0.0f;
    bool hit = false;
    if (ok) {
        const int tx = (v.img_width + block_width - 1) / block_width;
        const int ty = (v.img_height + block_width - 1) / block_width;
        hit = project_view_one(i, j, v.cholesky, v.par, v.means2d, 1.0f, v.origin_x, v.origin_y,
                               v.scale_x, v.scale_y, tx, ty, block_width, xys, extents, conics,
                               num_tiles_hit);
    } else {
        xys[2 * i] = xys[2 * i + 1] = 0.0f;
        extents[2 * i] = extents[2 * i + 1] = 0.0f;
        conics[3 * i] = conics[3 * i + 1] = conics[3 * i + 2] = 0.0f;
        num_tiles_hit[i] = 0;
    }
    if (colors && hit) {
        const float *c = v.rgbs + 3 * j;
        colors[3 * i] = c[0];
        colors[3 * i + 1] = c[1];
        colors[3 * i + 2] = c[2];
    }
}

__global__ __launch_bounds__(kThreads) void project_views_kernel(
    int num_views, const LigView *__restrict__ views, int n_total, int block_width,
    float *__restrict__ xys, float *__restrict__ extents, float *__restrict__ conics,
    int32_t *__restrict__ num_tiles_hit, float *__restrict__ colors) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    // the wave's first and last slots: when one view holds both (all but the
    // waves at a view boundary), its record is wave-uniform
    const int w0 = __builtin_amdgcn_readfirstlane(i - (int)(threadIdx.x & (LIG_WAVE - 1)));
    const int w1 = min(w0 + LIG_WAVE, n_total) - 1;
    if (w0 >= n_total) return;
    const int vf = find_view(views, num_views, w0);
    const int vl = find_view(views, num_views, w1);
    if (i >= n_total) return;
    if (vf == vl)
        project_slot(views[vf], i, block_width, xys, extents, conics, num_tiles_hit, colors);
    else
        project_slot(views[find_view(views, num_views, i)], i, block_width, xys, extents, conics,
                     num_tiles_hit, colors);
}

__global__ __launch_bounds__(kThreads) void cov2d_bounds_kernel(
    int n, const float *__restrict__ cov2d, const float *__restrict__ opacities,
    float *__restrict__ conics, float *__restrict__ extents) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    float ka = 0.0f, kb = 0.0f, kc = 0.0f, ex = 0.0f, ey = 0.0f;
    const float o = opacities ? opacities[i] : 1.0f;
    cov_bounds(cov2d[3 * i], cov2d[3 * i + 1], cov2d[3 * i + 2], o, ka, kb, kc, ex, ey);
    conics[3 * i + 0] = ka;
    conics[3 * i + 1] = kb;
    conics[3 * i + 2] = kc;
    extents[2 * i + 0] = ex;
    extents[2 * i + 1] = ey;
}

// v_Sigma = -X G X (helpers.cuh:84-93), 2x2 products in glm's column-major
// order, left to right.
__device__ __forceinline__ void conic_vjp(float a, float b, float c, float va, float vb,
                                          float vc, float &o0, float &o1, float &o2) {
#pragma clang fp contract(off)
    const float X[2][2] = {{a, b}, {b, c}};
    const float hb = vb / 2.0f;
    const float G[2][2] = {{va, hb}, {hb, vc}};
    const float N[2][2] = {{-a, -b}, {-b, -c}};
    float P[2][2], V[2][2];
#pragma unroll
    for (int col = 0; col < 2; ++col)
#pragma unroll
        for (int r = 0; r < 2; ++r) P[col][r] = N[0][r] * G[col][0] + N[1][r] * G[col][1];
#pragma unroll
    for (int col = 0; col < 2; ++col)
#pragma unroll
        for (int r = 0; r < 2; ++r) V[col][r] = P[0][r] * X[col][0] + P[1][r] * X[col][1];
    o0 = V[0][0];
    o1 = V[1][0] + V[0][1];
    o2 = V[1][1];
}

__global__ __launch_bounds__(kThreads) void project_bwd_kernel(
    int n, const float *__restrict__ extents, const float *__restrict__ conics,
    const float *__restrict__ v_xy, const float *__restrict__ v_conic,
    const float *__restrict__ chol, float *__restrict__ v_par, float *__restrict__ v_mean) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    float m0 = 0.0f, m1 = 0.0f, p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
    if (extents[2 * i] > 0.0f && extents[2 * i + 1] > 0.0f) {
        m0 = v_xy[2 * i];
        m1 = v_xy[2 * i + 1];
        float va, vb, vc;
        conic_vjp(conics[3 * i], conics[3 * i + 1], conics[3 * i + 2], v_conic[3 * i],
                  v_conic[3 * i + 1], v_conic[3 * i + 2], va, vb, vc);
        if (chol) {
            const float l11 = chol[3 * i], l21 = chol[3 * i + 1], l22 = chol[3 * i + 2];
            const float v11 = 2.0f * l11 * va + l21 * vb;
            const float v21 = l11 * vb + 2.0f * l21 * vc;
            const float v22 = 2.0f * l22 * vc;
            const bool fin = __builtin_isfinite(v11) && __builtin_isfinite(v21) &&
                             __builtin_isfinite(v22);
            p0 = fin ? v11 : 0.0f;
            p1 = fin ? v21 : 0.0f;
            p2 = fin ? v22 : 0.0f;
        } else {
            p0 = va;
            p1 = vb;
            p2 = vc;
        }
    }
    v_mean[2 * i] = m0;
    v_mean[2 * i + 1] = m1;
    v_par[3 * i] = p0;
    v_par[3 * i + 1] = p1;
    v_par[3 * i + 2] = p2;
}

bool dims_ok(int img_height, int img_width, int block_width) {
    return img_height > 0 && img_width > 0 && block_width >= 1 && block_width <= 16;
}

}  // namespace
}  // namespace lig

using namespace lig;

extern "C" int lig_compute_cov2d_bounds(int n, const float *cov2d, const float *opacities,
                                        float *conics, float *extents, void *stream) {
    if (n < 0 || (n > 0 && (!cov2d || !conics || !extents))) return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    hipLaunchKernelGGL(cov2d_bounds_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0,
                       as_stream(stream), n, cov2d, opacities, conics, extents);
    return launch_status();
}

static int project_fwd(bool cholesky, int n, const float *par, const float *means2d,
                       const float *opacities, int img_height, int img_width, int block_width,
                       float *xys, float *extents, float *conics, int32_t *num_tiles_hit,
                       void *stream) {
    if (n < 0 || !dims_ok(img_height, img_width, block_width)) return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    if (!par || !means2d || !xys || !extents || !conics || !num_tiles_hit)
        return LIG_ERR_INVALID_ARGUMENT;
    const int tx = (img_width + block_width - 1) / block_width;
    const int ty = (img_height + block_width - 1) / block_width;
    auto kern = cholesky ? project_chol_kernel : project_cov_kernel;
    hipLaunchKernelGGL(kern, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0, as_stream(stream),
                       n, par, means2d, opacities, tx, ty, block_width, xys, extents, conics,
                       num_tiles_hit);
    return launch_status();
}

extern "C" int lig_project_gaussians_forward(int n, const float *cov2d, const float *means2d,
                                             const float *opacities, int img_height,
                                             int img_width, int block_width, float *xys,
                                             float *extents, float *conics,
                                             int32_t *num_tiles_hit, void *stream) {
    return project_fwd(false, n, cov2d, means2d, opacities, img_height, img_width, block_width,
                       xys, extents, conics, num_tiles_hit, stream);
}

extern "C" int lig_project_gaussians_forward_cholesky(int n, const float *cholesky,
                                                      const float *means2d,
                                                      const float *opacities, int img_height,
                                                      int img_width, int block_width,
                                                      float *xys, float *extents,
                                                      float *conics, int32_t *num_tiles_hit,
                                                      void *stream) {
    return project_fwd(true, n, cholesky, means2d, opacities, img_height, img_width,
                       block_width, xys, extents, conics, num_tiles_hit, stream);
}

extern "C" int lig_project_gaussians_view(int n, const float *par, const float *means2d,
                                          const float *opacities, int cholesky, float origin_x,
                                          float origin_y, float scale_x, float scale_y,
                                          int img_height, int img_width, int block_width,
                                          float *xys, float *extents, float *conics,
                                          int32_t *num_tiles_hit, void *stream) {
    if (n < 0 || !dims_ok(img_height, img_width, block_width) || !(scale_x > 0.0f) ||
        !(scale_y > 0.0f))
        return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    if (!par || !means2d || !xys || !extents || !conics || !num_tiles_hit)
        return LIG_ERR_INVALID_ARGUMENT;
    const int tx = (img_width + block_width - 1) / block_width;
    const int ty = (img_height + block_width - 1) / block_width;
    hipLaunchKernelGGL(project_view_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0,
                       as_stream(stream), n, cholesky, par, means2d, opacities, origin_x,
                       origin_y, scale_x, scale_y, tx, ty, block_width, xys, extents, conics,
                       num_tiles_hit);
    return launch_status();
}

extern "C" int lig_project_views_batched(int num_views, const LigView *views, int n_total,
                                         int block_width, float *xys, float *extents,
                                         float *conics, int32_t *num_tiles_hit, float *colors,
                                         void *stream) {
    if (num_views < 1 || n_total < 0 || block_width < 1 || block_width > 16)
        return LIG_ERR_INVALID_ARGUMENT;
    if (n_total == 0) return LIG_OK;
    if (!views || !xys || !extents || !conics || !num_tiles_hit) return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(project_views_kernel, dim3(ceil_div(n_total, kThreads)), dim3(kThreads),
                       0, as_stream(stream), num_views, views, n_total, block_width, xys,
                       extents, conics, num_tiles_hit, colors);
    return launch_status();
}

extern "C" int lig_project_batched_cholesky(int n, const float *cholesky, const float *means2d,
                                            const void *batch_table, int num_images, float *xys,
                                            float *extents, float *conics,
                                            int32_t *num_tiles_hit, void *stream) {
    if (n < 0 || num_images < 1 || !batch_table) return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    if (!cholesky || !means2d || !xys || !extents || !conics || !num_tiles_hit)
        return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(project_batched_chol_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads),
                       0, as_stream(stream), n, cholesky, means2d,
                       reinterpret_cast<const BatchImage *>(batch_table), num_images, xys,
                       extents, conics, num_tiles_hit);
    return launch_status();
}

static int project_bwd(int n, const float *extents, const float *chol, const float *conics,
                       const float *v_xy, const float *v_conic, float *v_par, float *v_mean,
                       void *stream) {
    if (n < 0) return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    if (!extents || !conics || !v_xy || !v_conic || !v_par || !v_mean)
        return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(project_bwd_kernel, dim3(ceil_div(n, kThreads)), dim3(kThreads), 0,
                       as_stream(stream), n, extents, conics, v_xy, v_conic, chol, v_par,
                       v_mean);
    return launch_status();
}

extern "C" int lig_project_gaussians_backward(int n, const float *extents, const float *conics,
                                              const float *v_xy, const float *v_conic,
                                              float *v_cov2d, float *v_mean2d, void *stream) {
    return project_bwd(n, extents, nullptr, conics, v_xy, v_conic, v_cov2d, v_mean2d, stream);
}

extern "C" int lig_project_gaussians_backward_cholesky(int n, const float *extents,
                                                       const float *cholesky,
                                                       const float *conics, const float *v_xy,
                                                       const float *v_conic, float *v_cholesky,
                                                       float *v_mean2d, void *stream) {
    if (!cholesky && n > 0) return LIG_ERR_INVALID_ARGUMENT;
    return project_bwd(n, extents, cholesky, conics, v_xy, v_conic, v_cholesky, v_mean2d,
                       stream);
}
