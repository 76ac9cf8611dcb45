// SSIM and MS-SSIM of image pairs (the reference's evaluation metric,
// train.py:200 / utils.py:26-34, from the pytorch-msssim package: Gaussian
// window applied separably with no padding, per-plane means of the ssim and
// contrast-structure maps, 2 x 2 zero-padded average pools between scales).
//
// One pass per scale:
//   ssim_level_kernel<WIN>: one wave per 64 output columns of a band of
//     kSsimRows output rows of one plane.  Each input row of the band is
//     staged in the wave's LDS row buffer (64 + WIN - 1 floats per image),
//     each lane forms the five horizontal window sums of its column (x, y,
//     x^2, y^2, x y, about a sample of the lane's own windows) and folds them into a WIN-row register ring of vertical
//     sums; a row of the ring completes every input row and gives one output
//     pixel's ssim and cs, summed in FP64 per lane.  A block writes one FP64
//     partial pair; inputs are read once per scale (the band's WIN - 1 halo
//     rows and the strip's halo columns aside).
//   pool_kernel: the next scale's planes for both images, planar FP32.
//   ssim_finish_kernel: one wave per plane sums every scale's partials in a
//     fixed order and combines relu(cs_l)^w_l ... relu(ssim_last)^w_last.
// Every reduction has a fixed order, so results do not depend on timing.
#include <cmath>

#include "lig_device.h"

namespace lig {
namespace {

constexpr int kSsimWaves = 4;
constexpr int kSsimThreads = kSsimWaves * LIG_WAVE;
constexpr int kSsimRows = 64;  // output rows per block
constexpr int kSsimMaxWin = 11;
constexpr int kSsimMaxLevels = 8;
constexpr int kPoolThreads = 256;

struct PlaneView {
    const float *p;
    long long sn, sc, sh, sw;  // element strides
};

struct SsimLevelArgs {
    PlaneView x, y;
    int c, h, w;        // channels per image and the input plane size
    int ho, wo;         // output map size (h - WIN + 1, w - WIN + 1)
    float g[kSsimMaxWin];
    float c1, c2;
    double2 *partial;   // [planes][gridDim.y][gridDim.x]
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int WIN>
__global__ __launch_bounds__(kSsimThreads) void ssim_level_kernel(SsimLevelArgs a) {
    constexpr int kRow = LIG_WAVE + WIN - 1;
    __shared__ float rows[kSsimWaves][2][2][kRow];  // [wave][buffer][x|y][col]
    __shared__ double red[kSsimWaves][2];

    const int lane = threadIdx.x & (LIG_WAVE - 1);
    const int wv = threadIdx.x / LIG_WAVE;
    const int plane = blockIdx.z;
    const int n = plane / a.c, ch = plane - n * a.c;
    const float *px = a.x.p + n * a.x.sn + ch * a.x.sc;
    const float *py = a.y.p + n * a.y.sn + ch * a.y.sc;
    const int x0 = (blockIdx.x * kSsimWaves + wv) * LIG_WAVE;
    const int o0 = blockIdx.y * kSsimRows;
    const int rows_out = min(kSsimRows, a.ho - o0);

    float g[WIN];
#pragma unroll
    for (int k = 0; k < WIN; ++k) g[k] = in_vgpr(a.g[k]);

    double sum_ss = 0.0, sum_cs = 0.0;
    if (x0 < a.wo) {
        // columns x0 + lane and, for lane < WIN - 1, x0 + 64 + lane; columns
        // past the plane read 0 and only feed outputs past wo
        const int ca = x0 + lane, cb = x0 + LIG_WAVE + lane;
        const bool va = ca < a.w, vb = lane < WIN - 1 && cb < a.w;
        const bool out_ok = ca < a.wo;
        const int nin = rows_out + WIN - 1;

        auto load = [&](int r, float &xa, float &ya, float &xb, float &yb) {
            const long long ry = (long long)(o0 + r);
            xa = va ? px[ry * a.x.sh + ca * a.x.sw] : 0.0f;
            ya = va ? py[ry * a.y.sh + ca * a.y.sw] : 0.0f;
            xb = vb ? px[ry * a.x.sh + cb * a.x.sw] : 0.0f;
            yb = vb ? py[ry * a.y.sh + cb * a.y.sw] : 0.0f;
        };

        float acc[5][WIN];
#pragma unroll
        for (int q = 0; q < 5; ++q)
#pragma unroll
            for (int k = 0; k < WIN; ++k) acc[q][k] = 0.0f;

        float xa, ya, xb, yb;
        load(0, xa, ya, xb, yb);
        // Each lane forms its window moments about one sample of its own
        // windows (column x0 + lane + WIN/2 of the band's first row):
        // variances and covariance do not change, E[x'^2] - mu'^2 cancels far
        // less in FP32, and the means get the shift back.
        const int cc = x0 + lane + WIN / 2;
        const float sx = cc < a.w ? px[(long long)o0 * a.x.sh + cc * a.x.sw] : 0.0f;
        const float sy = cc < a.w ? py[(long long)o0 * a.y.sh + cc * a.y.sw] : 0.0f;
        for (int r0 = 0; r0 < nin; r0 += WIN) {
#pragma unroll
            for (int s = 0; s < WIN; ++s) {
                const int r = r0 + s;
                if (r < nin) {
                    // consecutive rows alternate buffers, so this write never
                    // meets the previous row's reads
                    float *bx = rows[wv][r & 1][0];
                    float *by = rows[wv][r & 1][1];
                    bx[lane] = xa;
                    by[lane] = ya;
                    if (lane < WIN - 1) {
                        bx[LIG_WAVE + lane] = xb;
                        by[LIG_WAVE + lane] = yb;
                    }
                    wave_lds_sync();
                    if (r + 1 < nin) load(r + 1, xa, ya, xb, yb);

                    float h0 = 0.0f, h1 = 0.0f, h2 = 0.0f, h3 = 0.0f, h4 = 0.0f;
#pragma unroll
                    for (int k = 0; k < WIN; ++k) {
                        const float u = bx[lane + k] - sx, v = by[lane + k] - sy;
                        const float gu = g[k] * u, gv = g[k] * v;
                        h0 += gu;
                        h1 += gv;
                        h2 = __builtin_fmaf(gu, u, h2);
                        h3 = __builtin_fmaf(gv, v, h3);
                        h4 = __builtin_fmaf(gu, v, h4);
                    }
                    const float hq[5] = {h0, h1, h2, h3, h4};
#pragma unroll
                    for (int k = 0; k < WIN; ++k) {
                        const int slot = (s - k + WIN) % WIN;
#pragma unroll
                        for (int q = 0; q < 5; ++q)
                            acc[q][slot] = __builtin_fmaf(g[k], hq[q], acc[q][slot]);
                    }
                    // output row r - (WIN - 1) took its last row; its slot
                    // is next used by output row r + 1
                    const int done = (s + 1) % WIN;
                    if (r >= WIN - 1 && out_ok) {
                        const float dx = acc[0][done], dy = acc[1][done];
                        const float sxx = acc[2][done] - dx * dx;
                        const float syy = acc[3][done] - dy * dy;
                        const float sxy = acc[4][done] - dx * dy;
                        const float mx = dx + sx, my = dy + sy;
                        const float mxy = mx * my;
                        const float cs = (2.0f * sxy + a.c2) / (sxx + syy + a.c2);
                        const float ss = (2.0f * mxy + a.c1) / (mx * mx + my * my + a.c1) * cs;
                        sum_ss += (double)ss;
                        sum_cs += (double)cs;
                    }
#pragma unroll
                    for (int q = 0; q < 5; ++q) acc[q][done] = 0.0f;
                }
            }
        }
    }
    for (int off = LIG_WAVE / 2; off > 0; off >>= 1) {
        sum_ss += __shfl_down(sum_ss, off, LIG_WAVE);
        sum_cs += __shfl_down(sum_cs, off, LIG_WAVE);
    }
    if (lane == 0) {
        red[wv][0] = sum_ss;
        red[wv][1] = sum_cs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int w = 0; w < kSsimWaves; ++w) {
            s0 += red[w][0];
            s1 += red[w][1];
        }
        const size_t idx = ((size_t)plane * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        a.partial[idx] = make_double2(s0, s1);
    }
}

// 2 x 2 average pool with stride 2 and (h mod 2, w mod 2) zero padding on
// each side, divided by 4 everywhere; sums rows then columns of the window.
__global__ __launch_bounds__(kPoolThreads) void pool_kernel(long long total, PlaneView x,
                                                           PlaneView y, int c, int h, int w,
                                                           int ho, int wo, float *__restrict__ ox,
                                                           float *__restrict__ oy) {
    const long long t = (long long)blockIdx.x * kPoolThreads + threadIdx.x;
    if (t >= total) return;
    const int j = (int)(t % wo);
    const long long t2 = t / wo;
    const int i = (int)(t2 % ho);
    const int plane = (int)(t2 / ho);
    const int n = plane / c, ch = plane - n * c;
    const int ph = h & 1, pw = w & 1;
    const int r0 = 2 * i - ph, c0 = 2 * j - pw;
    float sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int di = 0; di < 2; ++di) {
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const int r = r0 + di, cc = c0 + dj;
            if (r >= 0 && r < h && cc >= 0 && cc < w) {
                sx += x.p[n * x.sn + ch * x.sc + r * x.sh + cc * x.sw];
                sy += y.p[n * y.sn + ch * y.sc + r * y.sh + cc * y.sw];
            }
        }
    }
    ox[t] = sx * 0.25f;
    oy[t] = sy * 0.25f;
}

struct FinishArgs {
    int levels, planes;
    int nblocks[kSsimMaxLevels];      // partials per plane at each scale
    size_t part_off[kSsimMaxLevels];  // first partial of each scale (double2 units)
    double count[kSsimMaxLevels];     // output pixels per plane at each scale
    double weight[kSsimMaxLevels];
};

__global__ __launch_bounds__(LIG_WAVE) void ssim_finish_kernel(FinishArgs a,
                                                              const double2 *__restrict__ part,
                                                              float *__restrict__ out_value,
                                                              float *__restrict__ out_ssim,
                                                              float *__restrict__ out_cs) {
    const int plane = blockIdx.x;
    const int lane = threadIdx.x;
    double prod = 1.0;
    for (int l = 0; l < a.levels; ++l) {
        const double2 *p = part + a.part_off[l] + (size_t)plane * a.nblocks[l];
        double s0 = 0.0, s1 = 0.0;
        for (int b = lane; b < a.nblocks[l]; b += LIG_WAVE) {
            s0 += p[b].x;
            s1 += p[b].y;
        }
        for (int off = LIG_WAVE / 2; off > 0; off >>= 1) {
            s0 += __shfl_down(s0, off, LIG_WAVE);
            s1 += __shfl_down(s1, off, LIG_WAVE);
        }
        const double ss = s0 / a.count[l], cs = s1 / a.count[l];
        if (lane == 0) {
            if (out_ssim) out_ssim[(size_t)l * a.planes + plane] = (float)ss;
            if (out_cs) out_cs[(size_t)l * a.planes + plane] = (float)cs;
            const double term = l + 1 < a.levels ? fmax(cs, 0.0) : fmax(ss, 0.0);
            prod *= pow(term, a.weight[l]);
        }
    }
    if (lane == 0 && out_value) out_value[plane] = (float)prod;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct SsimPlan {
    int levels, win;
    int h[kSsimMaxLevels], w[kSsimMaxLevels];
    unsigned gx[kSsimMaxLevels], gy[kSsimMaxLevels];
    size_t pool_off[kSsimMaxLevels];  // byte offset of scale l's pooled x (l >= 1)
    size_t part_off;                  // byte offset of the partials
    size_t part_count[kSsimMaxLevels];
    size_t bytes;
};

// Scale sizes and the workspace layout; false when a scale is smaller than
// the window or the launch grid would not fit.
bool plan_ssim(int n, int c, int h, int w, int win, int levels, SsimPlan &pl) {
    if (n < 1 || c < 1 || h < 1 || w < 1 || levels < 1 || levels > kSsimMaxLevels) return false;
    if (win < 3 || win > kSsimMaxWin || (win & 1) == 0) return false;
    if ((long long)n * c > 65535) return false;
    const size_t planes = (size_t)n * c;
    pl.levels = levels;
    pl.win = win;
    size_t off = 0;
    for (int l = 0; l < levels; ++l) {
        if (l == 0) {
            pl.h[0] = h;
            pl.w[0] = w;
        } else {
            const int ph = pl.h[l - 1], pw = pl.w[l - 1];
            pl.h[l] = (ph + 2 * (ph & 1) - 2) / 2 + 1;
            pl.w[l] = (pw + 2 * (pw & 1) - 2) / 2 + 1;
            pl.pool_off[l] = off;
            off += 2 * align256(planes * pl.h[l] * pl.w[l] * sizeof(float));
        }
        if (pl.h[l] < win || pl.w[l] < win) return false;
        const int ho = pl.h[l] - win + 1, wo = pl.w[l] - win + 1;
        pl.gx[l] = ceil_div(wo, (size_t)kSsimWaves * LIG_WAVE);
        pl.gy[l] = ceil_div(ho, kSsimRows);
        if (pl.gy[l] > 65535) return false;
        pl.part_count[l] = planes * pl.gx[l] * pl.gy[l];
    }
    pl.part_off = off;
    for (int l = 0; l < levels; ++l) off += pl.part_count[l] * sizeof(double2);
    pl.bytes = align256(off);
    return true;
}

template <int WIN>
void launch_level(const SsimLevelArgs &a, dim3 grid, hipStream_t s) {
    hipLaunchKernelGGL(ssim_level_kernel<WIN>, grid, dim3(kSsimThreads), 0, s, a);
}

}  // namespace
}  // namespace lig

using namespace lig;

extern "C" size_t lig_ssim_workspace_size(int n, int c, int h, int w, int win_size,
                                          int levels) {
    SsimPlan pl;
    return plan_ssim(n, c, h, w, win_size, levels, pl) ? pl.bytes : 0;
}

extern "C" int lig_ssim_multiscale(int n, int c, int h, int w, const float *x,
                                   const long long *x_strides, const float *y,
                                   const long long *y_strides, float data_range, int win_size,
                                   const float *win, float k1, float k2, int levels,
                                   const float *weights, float *out_value, float *out_ssim,
                                   float *out_cs, void *ws, size_t ws_bytes, void *stream) {
    SsimPlan pl;
    if (!plan_ssim(n, c, h, w, win_size, levels, pl)) return LIG_ERR_INVALID_ARGUMENT;
    if (!x || !y || !x_strides || !y_strides || !win || !weights) return LIG_ERR_INVALID_ARGUMENT;
    if (!out_value && !out_ssim && !out_cs) return LIG_ERR_INVALID_ARGUMENT;
    if (!ws || ws_bytes < pl.bytes) return LIG_ERR_WORKSPACE;
    hipStream_t s = as_stream(stream);
    char *base = static_cast<char *>(ws);
    const int planes = n * c;

    SsimLevelArgs a;
    a.c = c;
    for (int k = 0; k < kSsimMaxWin; ++k) a.g[k] = k < win_size ? win[k] : 0.0f;
    a.c1 = (k1 * data_range) * (k1 * data_range);
    a.c2 = (k2 * data_range) * (k2 * data_range);
    PlaneView vx{x, x_strides[0], x_strides[1], x_strides[2], x_strides[3]};
    PlaneView vy{y, y_strides[0], y_strides[1], y_strides[2], y_strides[3]};

    FinishArgs f;
    f.levels = levels;
    f.planes = planes;
    size_t part_elems = 0;
    for (int l = 0; l < levels; ++l) {
        a.x = vx;
        a.y = vy;
        a.h = pl.h[l];
        a.w = pl.w[l];
        a.ho = pl.h[l] - win_size + 1;
        a.wo = pl.w[l] - win_size + 1;
        a.partial = reinterpret_cast<double2 *>(base + pl.part_off) + part_elems;
        const dim3 grid(pl.gx[l], pl.gy[l], planes);
        switch (win_size) {
            case 3: launch_level<3>(a, grid, s); break;
            case 5: launch_level<5>(a, grid, s); break;
            case 7: launch_level<7>(a, grid, s); break;
            case 9: launch_level<9>(a, grid, s); break;
            default: launch_level<11>(a, grid, s); break;
        }
        f.nblocks[l] = (int)(pl.gx[l] * pl.gy[l]);
        f.part_off[l] = part_elems;
        f.count[l] = (double)a.ho * (double)a.wo;
        f.weight[l] = (double)weights[l];
        part_elems += pl.part_count[l];
        if (l + 1 < levels) {
            const int nh = pl.h[l + 1], nw = pl.w[l + 1];
            const size_t plane_elems = (size_t)nh * nw;
            float *ox = reinterpret_cast<float *>(base + pl.pool_off[l + 1]);
            float *oy = reinterpret_cast<float *>(base + pl.pool_off[l + 1] +
                                                  align256(planes * plane_elems * sizeof(float)));
            const long long total = (long long)planes * plane_elems;
            hipLaunchKernelGGL(pool_kernel, dim3(ceil_div(total, kPoolThreads)),
                               dim3(kPoolThreads), 0, s, total, vx, vy, c, pl.h[l], pl.w[l], nh,
                               nw, ox, oy);
            const long long pe = (long long)plane_elems;
            vx = PlaneView{ox, (long long)c * pe, pe, (long long)nw, 1};
            vy = PlaneView{oy, (long long)c * pe, pe, (long long)nw, 1};
        }
    }
    hipLaunchKernelGGL(ssim_finish_kernel, dim3(planes), dim3(LIG_WAVE), 0, s, f,
                       reinterpret_cast<const double2 *>(base + pl.part_off), out_value, out_ssim,
                       out_cs);
    return launch_status();
}
