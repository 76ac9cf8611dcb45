// Tile-based sum rasteriser: forward (plain and gradient-aware) and backward.
//
// Work decomposition (one 64-lane wave per tile, workgroup = one wave):
//   * forward: lane = pixel.  A tile of bw x bw pixels is spread over the 64
//     lanes, PX = ceil(bw*bw/64) pixels per lane (4 at bw = 16).  The tile's
//     Gaussians are fetched 64 at a time (one per lane) into LDS and every
//     lane walks the batch, so each LDS read feeds PX pixels.  FP32
//     accumulation in bin order, as the reference.
//   * backward: lane = Gaussian.  Each lane takes one Gaussian of the tile's
//     bin and walks the tile's pixels (their upstream gradients staged in LDS
//     and read as wave-uniform values), accumulating that Gaussian's gradient
//     in registers.  No cross-lane reduction is needed; each Gaussian then
//     receives one set of atomics per tile.
//   * tiles are assigned XCD-contiguously (see xcd_remap).
//
// Pair evaluation: sigma*log2(e) from the pre-scaled conic with explicit FMAs
// (pair_s) and alpha = min(clamp, o * 2^-(sigma*log2 e)); a pair contributes
// unless sigma < 0 or alpha < 1/255.  Clamps: 0.999 in both forwards and the
// gradient-aware backward, 0.99 in the plain backward; no backward
// differentiates the clamp.
//
// Reference behaviour followed (paths relative to the reference repository):
//   gsplat2d/gsplat2d/cuda/csrc/forward.cu:173-279, backward.cu:8-183,
//   gradient_aware_rasterizer.cu:118-495, bindings.cu:303-506.
#include <atomic>
#include <cmath>

#include "lig_device.h"

namespace lig {
namespace {

constexpr int kThreads = LIG_WAVE;
constexpr int kMaxPix = 256;  // 16 x 16
constexpr int kMaxBwdSplit = 8;  // shares per tile of the 16x16 backward

// (clamp(x, 0, 1) - t)^2: the caller's loss term per channel (gaussianlig.py:95-103,
// utils.py:21-22).
__device__ __forceinline__ float clamped_sq_err(float x, float t) {
    const float d = fminf(fmaxf(x, 0.0f), 1.0f) - t;
    return d * d;
}

// dL/dI of mean((clamp(I, 0, 1) - t)^2) for upstream gradient g: zero outside
// [0, 1] (the clamp's mask, inclusive at both bounds), else (norm * (I - t)) * g
// with norm = 2 / numel.  Same expression as loss.hip's gradient kernel.
__device__ __forceinline__ float l2_grad(float x, float t, float norm, float g) {
    return (x >= 0.0f && x <= 1.0f) ? (norm * (x - t)) * g : 0.0f;
}

// With `extents` (fit step: 16 x 16 tiles, no phantom intersections) p3.zw
// also carry the Gaussian's record addressing: its intersection with tile
// (tx, ty) has emission slot  base + ty * w + tx,  where w is the width of its
// tile rectangle [x0, x1) x [y0, y1) (emission is row-major over it, from slot
// cum[g-1]) and base = cum[g-1] - y0 * w - x0.
// SKIP_EMPTY (render-only callers, cum given): Gaussians with no tile
// intersections are left unwritten -- no bin names them -- except record 0,
// which a zero-extent slot's (tile 0, gid 0) entry may name.
template <bool SKIP_EMPTY = false>
__global__ __launch_bounds__(256) void pack_kernel(int n, const float *__restrict__ xys,
                                                   const float *__restrict__ conics,
                                                   const float *__restrict__ colors,
                                                   const float *__restrict__ opacities,
                                                   const float *__restrict__ extents,
                                                   const int32_t *__restrict__ cum, int tiles_x,
                                                   int tiles_y, const BatchImage *__restrict__ tab,
                                                   int nimg, GaussPack *__restrict__ pack) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (SKIP_EMPTY && i > 0 && cum[i] == cum[i - 1]) return;
    const float a = conics[3 * i], b = conics[3 * i + 1], c = conics[3 * i + 2];
    const float o = opacities ? opacities[i] : 1.0f;
    GaussPack g;
    g.p0 = make_float4(xys[2 * i], xys[2 * i + 1], a * kHalfLog2e, (b * kHalfLog2e) * 2.0f);
    g.p1 = make_float4(c * kHalfLog2e, o, colors[3 * i], colors[3 * i + 1]);
    g.p2 = make_float4(colors[3 * i + 2], a, b, c);
    // Cull box: for a positive definite conic, alpha >= 1/255 needs
    // d^T Q d <= t = 2 ln(255 o), whose x-extent is sqrt(t * Sigma_xx) with
    // Sigma_xx = c / det.  Half a pixel of margin covers FP32 rounding of the
    // pair evaluation many times over, so culled pairs are exactly the ones the
    // alpha test would reject.  Indefinite conics are never culled (sigma < 0
    // there is a rejection rule of its own and the region is unbounded).
    const float det = a * c - b * b;
    float cx = __builtin_inff(), cy = __builtin_inff();
    if (a > 0.0f && c > 0.0f && det > 0.0f && __builtin_isfinite(det) && o == o) {
        const float t = 2.0f * __logf(255.0f * o);
        if (!(t > 0.0f)) {
            cx = cy = -1.0f;
        } else {
            cx = __builtin_sqrtf(t * c / det) + 0.5f;
            cy = __builtin_sqrtf(t * a / det) + 0.5f;
        }
    }
    int base = 0, w = 1;
    if (extents) {
        const float ex = extents[2 * i], ey = extents[2 * i + 1];
        if (ex > 0.0f && ey > 0.0f) {
            int txs = tiles_x, tys = tiles_y;
            if (tab) {  // batched: the Gaussian's own image grid, local tile coordinates
                const BatchImage &im = tab[batch_find<1>(tab, nimg, i)];
                txs = im.tiles_x;
                tys = im.tiles_y;
            }
            int x0, y0, x1, y1;
            tile_bbox(xys[2 * i], xys[2 * i + 1], ex, ey, txs, tys, 16, x0, y0, x1, y1);
            w = max(x1 - x0, 1);
            base = (i == 0 ? 0 : cum[i - 1]) - y0 * w - x0;
        }
    }
    g.p3 = make_float4(cx, cy, __int_as_float(base), __int_as_float(w));
    pack[i] = g;
}

template <int PX, bool GA>
__global__ __launch_bounds__(kThreads) void raster_fwd_kernel(
    int tiles_x, int bw, int img_w, int img_h, const int32_t *__restrict__ gids,
    const int2 *__restrict__ bins, const GaussPack *__restrict__ pack,
    float *__restrict__ out_img, float *__restrict__ out_wsum, int32_t *__restrict__ final_idx,
    float *__restrict__ out_dx, float *__restrict__ out_dy, float *__restrict__ out_dxy,
    int has_opac) {
    __shared__ float4 s0[kThreads];
    __shared__ float4 s1[kThreads];
    __shared__ float4 s2[kThreads];

    const int tile = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int npix = bw * bw;

    float px[PX], py[PX];
    float acc_r[PX], acc_g[PX], acc_b[PX], acc_w[PX];
    float dxr[GA ? PX : 1], dxg[GA ? PX : 1], dxb[GA ? PX : 1];
    float dyr[GA ? PX : 1], dyg[GA ? PX : 1], dyb[GA ? PX : 1];
    float dzr[GA ? PX : 1], dzg[GA ? PX : 1], dzb[GA ? PX : 1];
    int fin[PX];
#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int p = lane + LIG_WAVE * k;
        const int ly = p / bw;
        const int lx = p - ly * bw;
        px[k] = (float)(tx * bw + lx) + 0.5f;
        py[k] = (float)(ty * bw + ly) + 0.5f;
        acc_r[k] = acc_g[k] = acc_b[k] = acc_w[k] = 0.0f;
        fin[k] = 0;
        if constexpr (GA) {
            dxr[k] = dxg[k] = dxb[k] = 0.0f;
            dyr[k] = dyg[k] = dyb[k] = 0.0f;
            dzr[k] = dzg[k] = dzb[k] = 0.0f;
        }
    }

    const int2 range = bins[tile];
    for (int base = range.x; base < range.y; base += kThreads) {
        const int cnt = min(kThreads, range.y - base);
        if (lane < cnt) {
            const GaussPack gp = pack[gids[base + lane]];
            s0[lane] = gp.p0;
            s1[lane] = gp.p1;
            s2[lane] = gp.p2;
        }
        __syncthreads();
        for (int t = 0; t < cnt; ++t) {
            const float4 a0 = s0[t];
            const float4 a1 = s1[t];
            const float4 a2 = s2[t];
            const int kidx = base + t;
#pragma unroll
            for (int k = 0; k < PX; ++k) {
                // pixels of one lane share a column when bw == 16 (px[k] equal),
                // so the compiler can share dx, a'dx and b'dx between them
                const float dx = a0.x - px[k];
                const float dy = a0.y - py[k];
                const float s = pair_s(a0.z, row_terms(a0.w, a1.x, dy), dx);
                const float e = fast_exp2(-s);
                float al = fminf(a1.y * e, 0.999f);
                const bool ok = has_opac ? !(s < 0.0f) && !(al < kAlphaMin) : accept_s(s);
                al = ok ? al : 0.0f;
                acc_r[k] = __builtin_fmaf(al, a1.z, acc_r[k]);
                acc_g[k] = __builtin_fmaf(al, a1.w, acc_g[k]);
                acc_b[k] = __builtin_fmaf(al, a2.x, acc_b[k]);
                acc_w[k] += al;
                fin[k] = ok ? kidx : fin[k];
                if constexpr (GA) {
                    // d' = p - mean = -(dx, dy): sigma_x = -u, sigma_y = -v
                    const float u = __builtin_fmaf(a2.y, dx, a2.z * dy);
                    const float v = __builtin_fmaf(a2.z, dx, a2.w * dy);
                    const float ax = al * u;
                    const float ay = al * v;
                    const float axy = al * __builtin_fmaf(u, v, -a2.z);
                    dxr[k] = __builtin_fmaf(a1.z, ax, dxr[k]);
                    dxg[k] = __builtin_fmaf(a1.w, ax, dxg[k]);
                    dxb[k] = __builtin_fmaf(a2.x, ax, dxb[k]);
                    dyr[k] = __builtin_fmaf(a1.z, ay, dyr[k]);
                    dyg[k] = __builtin_fmaf(a1.w, ay, dyg[k]);
                    dyb[k] = __builtin_fmaf(a2.x, ay, dyb[k]);
                    dzr[k] = __builtin_fmaf(a1.z, axy, dzr[k]);
                    dzg[k] = __builtin_fmaf(a1.w, axy, dzg[k]);
                    dzb[k] = __builtin_fmaf(a2.x, axy, dzb[k]);
                }
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int k = 0; k < PX; ++k) {
        const int p = lane + LIG_WAVE * k;
        const int ly = p / bw;
        const int lx = p - ly * bw;
        const int i = ty * bw + ly;
        const int j = tx * bw + lx;
        if (p < npix && i < img_h && j < img_w) {
            const size_t pid = (size_t)i * img_w + j;
            out_img[3 * pid + 0] = acc_r[k];
            out_img[3 * pid + 1] = acc_g[k];
            out_img[3 * pid + 2] = acc_b[k];
            out_wsum[pid] = acc_w[k];
            final_idx[pid] = fin[k];
            if constexpr (GA) {
                out_dx[3 * pid + 0] = dxr[k];
                out_dx[3 * pid + 1] = dxg[k];
                out_dx[3 * pid + 2] = dxb[k];
                out_dy[3 * pid + 0] = dyr[k];
                out_dy[3 * pid + 1] = dyg[k];
                out_dy[3 * pid + 2] = dyb[k];
                out_dxy[3 * pid + 0] = dzr[k];
                out_dxy[3 * pid + 1] = dzg[k];
                out_dxy[3 * pid + 2] = dzb[k];
            }
        }
    }
}

// Interval of offsets u along one axis over which the region {q <= lim} of
// q(u, v) = Au u^2 + 2 B u v + Av v^2 (positive definite, det = Au Av - B^2)
// meets the band v0 <= v <= v1 of the other axis: false when the band misses
// the region.  The region's extreme u = +-uE, uE = sqrt(Av lim / det), lies at
// v = -+B uE / Av; clamped into the band that v gives the band's extreme u,
// from q(u, vc) = lim: u = (-B vc +- sqrt(Au lim - det vc^2)) / Au (equal to
// +-uE when no clamp applies).  Reciprocals and square roots are the fast
// hardware forms; callers widen the interval (their lim carries a margin).
__device__ __forceinline__ bool reach_span(float Au, float Av, float B, float det, float lim,
                                           float v0, float v1, float &ulo, float &uhi) {
    const float idet = __builtin_amdgcn_rcpf(det);
    const float aul = Au * lim;
    const float vE = __builtin_sqrtf(aul * idet);  // the region's half-extent along v
    if (v1 < -vE || v0 > vE) return false;
    const float uE = __builtin_sqrtf(Av * lim * idet);
    const float vt = -B * uE * __builtin_amdgcn_rcpf(Av);
    const float vc = fminf(fmaxf(vt, v0), v1);
    const float vd = fminf(fmaxf(-vt, v0), v1);
    const float hc = __builtin_sqrtf(fmaxf(0.0f, __builtin_fmaf(-det * vc, vc, aul)));
    const float hd = __builtin_sqrtf(fmaxf(0.0f, __builtin_fmaf(-det * vd, vd, aul)));
    const float iau = __builtin_amdgcn_rcpf(Au);
    uhi = (hc - B * vc) * iau;
    ulo = (-hd - B * vd) * iau;
    return true;
}

// Column strips of a 16x16 tile (bit k: columns 4k .. 4k + 3, all 16 rows)
// that the accept region of a Gaussian can reach.  In the forward's own terms
// the region is {u : s(u) <= T}, s(u) = a' u_x^2 + b2' u_x u_y + c' u_y^2 with
// the staged conic (a', b2', c') = (a, 2b, c) log2(e) / 2, u = pixel centre -
// mean, and T = log2(255 o); (bx0, by0) is the tile's first pixel centre.  The
// columns whose centres the region {s <= lim} reaches within the tile's rows
// form one run (reach_span along x over the band of row centres), widened by
// 1/64 of a column each way; a strip is kept when it holds a column of that
// run.  lim = T + 0.36 + 2^-10 T (0.5 + 2^-10 t in units of the unscaled
// d^T Q d <= t).  Where the conic's condition number is below 1e4
// (tr^2 <= 1e4 det), a' + c' <= 46 (a + c <= 64) and the mean lies within
// 65536 px, the FP32 error of the run's ends and of the per-pixel s together
// stay well inside that margin, so a dropped strip holds no pixel the accept
// test passes.  Other Gaussians keep all four strips.
__device__ __forceinline__ int reach_strips(float mx, float my, float ap, float bp2, float cp,
                                            float T, float bx0, float by0) {
    const float bp = 0.5f * bp2;
    const float det = ap * cp - bp * bp, tr = ap + cp;
    if (!(ap > 0.0f && cp > 0.0f && det > 0.0f && tr <= 46.0f && tr * tr <= 1e4f * det &&
          T > 0.0f && T < 1e4f && fabsf(mx) <= 65536.0f && fabsf(my) <= 65536.0f))
        return 0xF;
    const float lim = __builtin_fmaf(T, 0x1p-10f, T) + 0.36f;
    const float x0 = bx0 - mx, y0 = by0 - my;
    float xlo, xhi;
    if (!reach_span(ap, cp, bp, det, lim, y0, y0 + 15.0f, xlo, xhi)) return 0;
    // clamped in float first: the conversions stay in range
    const int c0 = (int)fminf(fmaxf(ceilf(xlo - x0 - 0x1p-6f), 0.0f), 16.0f);
    const int c1 = (int)fmaxf(fminf(floorf(xhi - x0 + 0x1p-6f), 15.0f), -1.0f);
    if (c0 > c1) return 0;
    return (int)(((2u << (c1 >> 2)) - 1u) & ~((1u << (c0 >> 2)) - 1u));
}

// Rows of a 16 x 16 tile that a Gaussian can reach, as the run [r0, r1]
// (false: none).  A row r is kept when the form a'x^2 + 2b'xy + c'y^2 of
// (pixel centre - mean) can come within the accept limit anywhere on it
// (centres bx0 .. bx0 + 15 at height by0 + r): the region {form <= lim} met
// with the tile's column band is one interval of heights (reach_span), widened
// by 1/64 of a row each way, lim = T (1 + 2^-10) + 0.36 (the margin of
// reach_strips, above the FP32 error of the per-pixel s).  All rows when the
// guards of reach_strips fail.
__device__ __forceinline__ bool reach_row_span(float mx, float my, float ap, float bp2, float cp,
                                               float T, float bx0, float by0, int &r0, int &r1) {
    const float bp = 0.5f * bp2;
    const float det = ap * cp - bp * bp, tr = ap + cp;
    r0 = 0;
    r1 = 15;
    if (!(ap > 0.0f && cp > 0.0f && det > 0.0f && tr <= 46.0f && tr * tr <= 1e4f * det &&
          T > 0.0f && T < 1e4f && fabsf(mx) <= 65536.0f && fabsf(my) <= 65536.0f))
        return true;
    const float lim = __builtin_fmaf(T, 0x1p-10f, T) + 0.36f;
    const float x0 = bx0 - mx, y0 = by0 - my;
    float ylo, yhi;
    if (!reach_span(cp, ap, bp, det, lim, x0, x0 + 15.0f, ylo, yhi)) return false;
    // clamped in float first: the conversions stay in range
    r0 = (int)fminf(fmaxf(ceilf(ylo - y0 - 0x1p-6f), 0.0f), 16.0f);
    r1 = (int)fmaxf(fminf(floorf(yhi - y0 + 0x1p-6f), 15.0f), -1.0f);
    return r0 <= r1;
}

// Forward for 16x16 tiles.  Lane l owns the four pixels (l % 4 + 4k, l / 4),
// k = 0..3, of one row: one row's dy, b'dy and c'dy serve all four, and pixel
// k of every lane lies in column strip k (columns 4k .. 4k + 3).  A Gaussian's
// cull box and accept region are tested against each strip once, in the
// staging lane; strips it cannot reach are skipped by a wave-uniform branch.
// Pair evaluation and accumulation order are those of raster_fwd_kernel, so
// the outputs are identical.
//
// TRAIN (fit step and render-only): no wsum / final index; with a target the
// tile's clamp + L2 sum, over its pixels and channels of (clamp(I, 0, 1) - t)^2,
// is written to loss_partial[tile] (FP32, fixed lane order).
//
// DEPTH (64, 128, 256): Gaussians staged in LDS per batch, DEPTH / 64 per lane;
// deeper batches keep more gathers in flight per wave and halve or quarter the
// barriers (chosen per workload by lig_amd.autotune).
// FIN (API mode): write the final index; without it (a backward that needs
// none) the forward writes the image and wsum only.
template <bool HAS_OPAC, bool TRAIN, int DEPTH, bool FIN = true>
__device__ __forceinline__ void fwd16_tile(
    int gtile, int tiles_x, int img_w, int img_h, const int32_t *__restrict__ gids,
    const int2 *__restrict__ bins, const GaussPack *__restrict__ pack,
    float *__restrict__ out_img, float *__restrict__ out_wsum,
    int32_t *__restrict__ final_idx, const float *__restrict__ target,
    float *__restrict__ loss_partial, const BatchImage *__restrict__ tab, int nimg,
    const int32_t *__restrict__ cum, float4 *s0, float4 *s1, float *s2) {
    // gtile: bins / partials index
    int tile = gtile;  // tile of its own image
    float init = 0.0f;
    if (tab) {  // batched fit: the tile's image, its canvas and pixel buffers
        const int b = batch_find<0>(tab, nimg, gtile);
        const BatchImage &im = tab[b];
        if (cum) {
            // an image none of whose Gaussians meets a tile renders ones, as
            // the same image fitted alone (rasterize.py:80-90)
            const int g0 = im.gauss_off, g1 = tab[b + 1].gauss_off;
            if ((g1 > 0 ? cum[g1 - 1] : 0) == (g0 > 0 ? cum[g0 - 1] : 0)) init = 1.0f;
        }
        tile = gtile - im.tile_off;
        tiles_x = im.tiles_x;
        img_w = im.w;
        img_h = im.h;
        out_img += 3 * im.pix_off;
        if (target) target += 3 * im.pix_off;
    }
    const int lane = threadIdx.x;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int x_org = tx * 16, y_org = ty * 16;
    const int px0 = lane & 3, py = lane >> 2;
    // centres of the lane's pixels (exact: integers + 0.5)
    float X[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) X[k] = (float)(x_org + px0 + 4 * k) + 0.5f;
    const float Y = (float)(y_org + py) + 0.5f;

    // the clamp is a literal operand: fminf against a VGPR the compiler cannot
    // prove canonical would add a quieting v_max before every min
    const float amin = HAS_OPAC ? in_vgpr(kAlphaMin) : kAlphaMin;
    float ar[4], ag[4], ab[4], aw[4];
    int fin[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        ar[q] = ag[q] = ab[q] = init;
        aw[q] = 0.0f;
        fin[q] = 0;
    }

    // the tile's range is wave-uniform: batch bounds and loop control stay scalar
    const int2 range = bins[gtile];
    const int r0 = __builtin_amdgcn_readfirstlane(range.x);
    const int r1 = __builtin_amdgcn_readfirstlane(range.y);
    constexpr int kSub = DEPTH / kThreads;
    for (int base = r0; base < r1; base += DEPTH) {
        const int cnt = min(DEPTH, r1 - base);
        int gq[kSub];
#pragma unroll
        for (int k = 0; k < kSub; ++k) {
            const int j = k * kThreads + lane;
            gq[k] = j < cnt ? gids[base + j] : -1;
        }
        // qbits[k][q]: ballot over the k-th 64 staged Gaussians of "reaches
        // strip q", so the loop below finds each Gaussian's strips with scalar
        // bit tests and skips Gaussians that reach none
        uint64_t qbits[kSub][4];
        int box[kSub];
#pragma unroll
        for (int k = 0; k < kSub; ++k) {
            const int j = k * kThreads + lane;
            box[k] = 0;
            if (gq[k] >= 0) {
                const GaussPack gp = pack[gq[k]];
                s0[j] = gp.p0;
                // without opacities the opacity slot carries blue
                s1[j] = HAS_OPAC ? gp.p1 : make_float4(gp.p1.x, gp.p2.x, gp.p1.z, gp.p1.w);
                if constexpr (HAS_OPAC) s2[j] = gp.p2.x;
                // strip k spans pixel centres x in [org + 4k + 0.5, org + 4k + 3.5]
                const float xl = gp.p0.x - gp.p3.x, xh = gp.p0.x + gp.p3.x;
                const float yl = gp.p0.y - gp.p3.y, yh = gp.p0.y + gp.p3.y;
                const float bx0 = (float)x_org + 0.5f, by0 = (float)y_org + 0.5f;
                // negated tests keep NaN boxes (NaN means) unculled, as the generic kernel
                const bool cy = !(yl > by0 + 15.0f) && !(yh < by0);
                int b = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    b |= (cy && !(xl > bx0 + (float)(4 * q + 3)) && !(xh < bx0 + (float)(4 * q)))
                             ? 1 << q
                             : 0;
                box[k] = b;
            }
        }
        // then the region itself, not only its box (about a fifth of the
        // strips the box reaches hold no accepted pixel), from the lane's own
        // staged entries
#pragma unroll
        for (int k = 0; k < kSub; ++k) {
            const int j = k * kThreads + lane;
            int m = box[k];
            if (m) {
                const float4 a0 = s0[j];
                const float4 a1 = s1[j];
                const float T = HAS_OPAC ? __log2f(255.0f * a1.y) : kAcceptLog2;
                m &= reach_strips(a0.x, a0.y, a0.z, a0.w, a1.x, T, (float)x_org + 0.5f,
                                  (float)y_org + 0.5f);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) qbits[k][q] = __ballot((m >> q) & 1);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSub; ++k) {
            uint64_t todo = qbits[k][0] | qbits[k][1] | qbits[k][2] | qbits[k][3];
            while (todo) {
                const int bit = __builtin_ctzll(todo);
                todo &= todo - 1;
                const int t = k * kThreads + bit;
                const float4 a0 = s0[t];
                const float4 a1 = s1[t];
                const float blue = HAS_OPAC ? s2[t] : a1.y;
                const int kidx = base + t;
                const RowTerms rw = row_terms(a0.w, a1.x, a0.y - Y);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (!((qbits[k][q] >> bit) & 1)) continue;
                    const float s = pair_s(a0.z, rw, a0.x - X[q]);
                    if constexpr (TRAIN && !HAS_OPAC) {
                        // accumulation under the lane mask of accepted pairs (a
                        // rejected pair adds nothing), skipped when no lane accepts
                        if (accept_s(s)) {
                            const float al = fminf(fast_exp2(-s), 0.999f);
                            ar[q] = __builtin_fmaf(al, a1.z, ar[q]);
                            ag[q] = __builtin_fmaf(al, a1.w, ag[q]);
                            ab[q] = __builtin_fmaf(al, blue, ab[q]);
                        }
                        continue;
                    }
                    const float e = fast_exp2(-s);
                    float al = fminf(HAS_OPAC ? a1.y * e : e, 0.999f);
                    const bool ok = HAS_OPAC ? !(s < 0.0f) && !(al < amin) : accept_s(s);
                    al = ok ? al : 0.0f;
                    ar[q] = __builtin_fmaf(al, a1.z, ar[q]);
                    ag[q] = __builtin_fmaf(al, a1.w, ag[q]);
                    ab[q] = __builtin_fmaf(al, blue, ab[q]);
                    if constexpr (!TRAIN) {
                        aw[q] += al;
                        if constexpr (FIN) fin[q] = ok ? kidx : fin[q];
                    }
                }
            }
        }
        __syncthreads();
    }

    float lsum = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = y_org + py;
        const int j = x_org + px0 + 4 * q;
        if (i < img_h && j < img_w) {
            const size_t pid = (size_t)i * img_w + j;
            out_img[3 * pid + 0] = ar[q];
            out_img[3 * pid + 1] = ag[q];
            out_img[3 * pid + 2] = ab[q];
            if constexpr (TRAIN) {
                if (target)
                    lsum += clamped_sq_err(ar[q], target[3 * pid + 0]) +
                            clamped_sq_err(ag[q], target[3 * pid + 1]) +
                            clamped_sq_err(ab[q], target[3 * pid + 2]);
            } else {
                out_wsum[pid] = aw[q];
                if constexpr (FIN) final_idx[pid] = fin[q];
            }
        }
    }
    if constexpr (TRAIN) {
        if (loss_partial) {
            for (int off = LIG_WAVE / 2; off > 0; off >>= 1)
                lsum += __shfl_down(lsum, off, LIG_WAVE);
            if (lane == 0) loss_partial[gtile] = lsum;
        }
    }
}

// Tile of the u-th launch slot when no launch order is given: on a grid of
// whole 8 x 8 super-tiles (both tile counts multiples of 8) slots run through
// each super-tile, super-tiles row-major, so the tiles an XCD has in flight at
// once (a contiguous run of slots, xcd_remap) form a compact patch whose
// Gaussians its L2 holds, instead of a sliver of one tile row; other grids
// row-major.  Images of a batched launch keep their own order (one_image
// false).
__device__ __forceinline__ int tile_of_slot(int u, int tiles_x, int tiles_y, bool one_image) {
    if (!one_image || (tiles_x & 7) || (tiles_y & 7)) return u;
    const int w = u & 63, su = u >> 6;
    const int sxn = tiles_x >> 3;
    const int sy = su / sxn, sx = su - sy * sxn;
    return ((sy << 3) + (w >> 3)) * tiles_x + (sx << 3) + (w & 7);
}

// One tile per block: the tile order[blockIdx.x], or without an order (none
// given, or order[0] = -1) the tile xcd_remap(blockIdx.x).
template <bool HAS_OPAC, bool TRAIN, int DEPTH = kThreads, bool FIN = true>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(TRAIN && DEPTH <= 128 ? 8 : 1))) void raster_fwd16_kernel(
    int tiles_x, int img_w, int img_h, const int32_t *__restrict__ gids,
    const int2 *__restrict__ bins, const GaussPack *__restrict__ pack,
    float *__restrict__ out_img, float *__restrict__ out_wsum,
    int32_t *__restrict__ final_idx, const float *__restrict__ target,
    float *__restrict__ loss_partial, const BatchImage *__restrict__ tab, int nimg,
    const int32_t *__restrict__ cum, const int32_t *__restrict__ order) {
    __shared__ float4 s0[DEPTH];
    __shared__ float4 s1[DEPTH];
    __shared__ float s2[HAS_OPAC ? DEPTH : 1];  // blue (with opacities)
    const int gtile = order && order[0] >= 0
                          ? order[blockIdx.x]
                          : tile_of_slot((int)xcd_remap(blockIdx.x, gridDim.x), tiles_x,
                                         (img_h + 15) >> 4, tab == nullptr);
    fwd16_tile<HAS_OPAC, TRAIN, DEPTH, FIN>(gtile, tiles_x, img_w, img_h, gids, bins, pack,
                                            out_img, out_wsum, final_idx, target, loss_partial,
                                            tab, nimg, cum, s0, s1, s2);
}

// Persistent form: each wave takes tiles from the per-XCD queue until none is
// left.
template <bool HAS_OPAC, bool TRAIN, int DEPTH = kThreads>
__global__ __launch_bounds__(kThreads) void raster_fwd16_queue_kernel(
    int tiles_x, int img_w, int img_h, const int32_t *__restrict__ gids,
    const int2 *__restrict__ bins, const GaussPack *__restrict__ pack,
    float *__restrict__ out_img, float *__restrict__ out_wsum,
    int32_t *__restrict__ final_idx, const float *__restrict__ target,
    float *__restrict__ loss_partial, const BatchImage *__restrict__ tab, int nimg,
    const int32_t *__restrict__ cum, TileQueue queue) {
    __shared__ float4 s0[DEPTH];
    __shared__ float4 s1[DEPTH];
    __shared__ float s2[HAS_OPAC ? DEPTH : 1];
    int cnt, x;
    int i = queue_resolve(queue, queue_issue(queue), cnt, x);
    while (i >= 0) {
        const int pending = queue_issue(queue);
        for (int k = 0; k < cnt; ++k)
            fwd16_tile<HAS_OPAC, TRAIN, DEPTH>(queue_tile(queue, x, i + k), tiles_x, img_w, img_h,
                                               gids, bins, pack, out_img, out_wsum, final_idx,
                                               target, loss_partial, tab, nimg, cum, s0, s1, s2);
        i = queue_resolve(queue, pending, cnt, x);
    }
}

// Stage one tile's upstream gradients in LDS: sv = (dL/dI rgb, dL/dW), sfin =
// final index (INT32_MIN outside the image).  GA adds dL/d(dx, dy, dxy).
template <bool GA>
__device__ __forceinline__ void stage_tile_grads(
    int bw, int x_org, int y_org, int img_w, int img_h, const int32_t *__restrict__ final_idx,
    const float *__restrict__ v_out, const float *__restrict__ v_wsum,
    const float *__restrict__ v_dx, const float *__restrict__ v_dy,
    const float *__restrict__ v_dxy, float4 *sv, int *sfin, float4 *sa, float4 *sb,
    float *sc) {
    const int npix = bw * bw;
    for (int p = threadIdx.x; p < npix; p += kThreads) {
        const int ly = p / bw;
        const int lx = p - ly * bw;
        const int i = y_org + ly;
        const int j = x_org + lx;
        if (i < img_h && j < img_w) {
            const size_t pid = (size_t)i * img_w + j;
            sv[p] = make_float4(v_out[3 * pid], v_out[3 * pid + 1], v_out[3 * pid + 2],
                                v_wsum ? v_wsum[pid] : 0.0f);
            if (sfin) sfin[p] = final_idx[pid];
            if constexpr (GA) {
                sa[p] = make_float4(v_dx[3 * pid], v_dx[3 * pid + 1], v_dx[3 * pid + 2],
                                    v_dy[3 * pid]);
                sb[p] = make_float4(v_dy[3 * pid + 1], v_dy[3 * pid + 2], v_dxy[3 * pid],
                                    v_dxy[3 * pid + 1]);
                sc[p] = v_dxy[3 * pid + 2];
            }
        } else {
            sv[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (sfin) sfin[p] = INT32_MIN;
            if constexpr (GA) {
                sa[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                sb[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                sc[p] = 0.0f;
            }
        }
    }
    __syncthreads();
}

// Stage the loss gradient of a 16 x 16 tile in LDS for the fit step:
// sv = (dL/dI rgb, 0) with dL/dI = l2_grad(img, target) and upstream gradient
// *grad_loss.  Pixels outside the image get zeros (they are never evaluated).
__device__ __forceinline__ void stage_tile_loss_grads(int x_org, int y_org, int img_w, int img_h,
                                                      const float *__restrict__ img,
                                                      const float *__restrict__ target,
                                                      float g, float norm, float4 *sv) {
    for (int p = threadIdx.x; p < kMaxPix; p += kThreads) {
        const int i = y_org + (p >> 4);
        const int j = x_org + (p & 15);
        float4 w = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i < img_h && j < img_w) {
            const size_t k = 3 * ((size_t)i * img_w + j);
            w.x = l2_grad(img[k + 0], target[k + 0], norm, g);
            w.y = l2_grad(img[k + 1], target[k + 1], norm, g);
            w.z = l2_grad(img[k + 2], target[k + 2], norm, g);
        }
        sv[p] = w;
    }
    __syncthreads();
}

// Per-intersection gradient record (deterministic path): 48 bytes,
//   r0 = (v_r, v_g, v_b, v_conic.a)  r1 = (v_conic.b, v_conic.c, v_x, v_y)
//   r2 = (|v_x|, |v_y|, v_opacity, gid bits)
// written at the intersection's emission slot, so that each Gaussian's
// records sit contiguously at [cum[g-1], cum[g]) in tile row-major order.
struct __align__(16) GradRec {
    float4 r0, r1, r2;
};

// Fit-step gradient record (32 bytes): r0 = (v_r, v_g, v_b, v_conic.a),
// r1 = (v_conic.b, v_conic.c, v_x, v_y); no v_xy_abs, no opacity.
struct __align__(16) FitRec {
    float4 r0, r1;
};

// Emission slot of the intersection at sorted position k: from the sorted
// emission ids when given, else from the pack's record addressing (16 x 16
// tiles, no opacities): base + ty * width + tx.
__device__ __forceinline__ int record_slot(const int32_t *__restrict__ slots, int k,
                                           const GaussPack &gp, int tx, int ty) {
    return slots ? slots[k] : __float_as_int(gp.p3.z) + ty * __float_as_int(gp.p3.w) + tx;
}

__device__ __forceinline__ void emit_grads(bool det, GradRec *__restrict__ rec, int slot, int g,
                                           float vr, float vg, float vb, float c0, float c1,
                                           float c2, float gx, float gy, float ax, float ay,
                                           float vo, bool has_opac, float *__restrict__ g_xy,
                                           float *__restrict__ g_xy_abs,
                                           float *__restrict__ g_conic,
                                           float *__restrict__ g_rgb,
                                           float *__restrict__ g_opac) {
    if (det) {
        GradRec r;
        r.r0 = make_float4(vr, vg, vb, c0);
        r.r1 = make_float4(c1, c2, gx, gy);
        r.r2 = make_float4(ax, ay, vo, __int_as_float(g));
        LIG_CHECK(slot >= 0);
        rec[slot] = r;
    } else {
        unsafeAtomicAdd(&g_rgb[3 * g + 0], vr);
        unsafeAtomicAdd(&g_rgb[3 * g + 1], vg);
        unsafeAtomicAdd(&g_rgb[3 * g + 2], vb);
        unsafeAtomicAdd(&g_conic[3 * g + 0], c0);
        unsafeAtomicAdd(&g_conic[3 * g + 1], c1);
        unsafeAtomicAdd(&g_conic[3 * g + 2], c2);
        unsafeAtomicAdd(&g_xy[2 * g + 0], gx);
        unsafeAtomicAdd(&g_xy[2 * g + 1], gy);
        unsafeAtomicAdd(&g_xy_abs[2 * g + 0], ax);
        unsafeAtomicAdd(&g_xy_abs[2 * g + 1], ay);
        if (has_opac) unsafeAtomicAdd(&g_opac[g], vo);
    }
}

// Slots of a zero-extent Gaussian hold records of Gaussian 0 (the reference
// bins them as (tile 0, gid 0)); their sums go to `phantom` and are added to
// Gaussian 0 by add_phantom_kernel.
// Record reductions.  Each Gaussian's records are its contiguous slot range
// [cum[g-1], cum[g]); a 16-lane group sums them in chunks of 16 counted from
// the range's own start: the group loads a chunk (coalesced), adds it by a
// fixed butterfly (xor 8, 4, 2, 1) whose result is taken from the group's
// first lane, and adds the chunk sums in chunk order.  The order depends only
// on the Gaussian's own records, so the API and fit paths, and a batched and
// a single-image fit, give the same bits.  The chunk loop runs to the wave's
// longest range (all lanes active for the shuffles; missing records read 0).
constexpr int kSegLanes = 16;
constexpr int kSegThreads = 256;
#ifndef LIG_SEG_UNROLL
#define LIG_SEG_UNROLL 2
#endif

// Sum over a 16-lane DPP row, valid in the row's first lane: quad swaps
// (xor 1, xor 2), then the half-row and row mirrors; a fixed pairing.
__device__ __forceinline__ float dpp_f(float v, int ctrl_sel) {
    const int x = __float_as_int(v);
    int y;
    switch (ctrl_sel) {
        case 0: y = __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false); break;
        case 1: y = __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false); break;
        case 2: y = __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false); break;
        default: y = __builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, false); break;
    }
    return __int_as_float(y);
}
__device__ __forceinline__ float row16_sum(float v) {
    v += dpp_f(v, 0);
    v += dpp_f(v, 1);
    v += dpp_f(v, 2);
    v += dpp_f(v, 3);
    return v;
}

// kSegUnroll chunks are loaded before any is reduced (more loads in flight);
// their sums are still added in chunk order, and chunks past a group's range
// (as those past the wave's longest range) add zeros.
constexpr int kSegUnroll = LIG_SEG_UNROLL;

template <int NF>
__device__ __forceinline__ void seg_sum(const float4 *__restrict__ src, int lo, int hi,
                                        float (&s)[NF]) {
    const int sub = threadIdx.x & (kSegLanes - 1);
    int nchunk = hi > lo ? (hi - lo + kSegLanes - 1) / kSegLanes : 0;
#pragma unroll
    for (int off = LIG_WAVE / 2; off > 0; off >>= 1)
        nchunk = max(nchunk, __shfl_xor(nchunk, off, LIG_WAVE));
#pragma unroll
    for (int k = 0; k < NF; ++k) s[k] = 0.0f;
    for (int c = 0; c < nchunk; c += kSegUnroll) {
        float v[kSegUnroll][NF];
#pragma unroll
        for (int u = 0; u < kSegUnroll; ++u) {
            const int r = lo + (c + u) * kSegLanes + sub;
            if (r < hi) {
#pragma unroll
                for (int q = 0; q < NF / 4; ++q) {
                    const float4 x = src[(size_t)r * (NF / 4) + q];
                    v[u][4 * q] = x.x;
                    v[u][4 * q + 1] = x.y;
                    v[u][4 * q + 2] = x.z;
                    v[u][4 * q + 3] = x.w;
                }
            } else {
#pragma unroll
                for (int k = 0; k < NF; ++k) v[u][k] = 0.0f;
            }
        }
#pragma unroll
        for (int u = 0; u < kSegUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < NF; ++k) s[k] += row16_sum(v[u][k]);
        }
    }
}

__global__ __launch_bounds__(kSegThreads) void reduce_records_kernel(
    int n, int m, const int32_t *__restrict__ cum, const GradRec *__restrict__ rec,
    int has_opac, float *__restrict__ g_xy, float *__restrict__ g_xy_abs,
    float *__restrict__ g_conic, float *__restrict__ g_rgb, float *__restrict__ g_opac,
    float *__restrict__ phantom) {
    const int g = (int)(((size_t)blockIdx.x * kSegThreads + threadIdx.x) / kSegLanes);
    int lo = 0, hi = 0;
    if (g < n) {
        lo = max(0, g == 0 ? 0 : cum[g - 1]);
        hi = min(cum[g], m);
    }
    float s[12];
    seg_sum<12>(reinterpret_cast<const float4 *>(rec), lo, hi, s);
    if (g >= n || (threadIdx.x & (kSegLanes - 1)) != 0) return;
    // records of zero-extent slots carry gid 0: their sums go to Gaussian 0
    if (hi > lo && __float_as_int(rec[lo].r2.w) != g) {
#pragma unroll
        for (int q = 0; q < 11; ++q) unsafeAtomicAdd(&phantom[q], s[q]);
#pragma unroll
        for (int q = 0; q < 11; ++q) s[q] = 0.0f;
    }
    g_rgb[3 * g + 0] = s[0];
    g_rgb[3 * g + 1] = s[1];
    g_rgb[3 * g + 2] = s[2];
    g_conic[3 * g + 0] = s[3];
    g_conic[3 * g + 1] = s[4];
    g_conic[3 * g + 2] = s[5];
    g_xy[2 * g + 0] = s[6];
    g_xy[2 * g + 1] = s[7];
    g_xy_abs[2 * g + 0] = s[8];
    g_xy_abs[2 * g + 1] = s[9];
    if (has_opac) g_opac[g] = s[10];
}

// Fit-step reduction (FitRec: no phantom slots exist without opacities).
__global__ __launch_bounds__(kSegThreads) void reduce_fit_records_kernel(
    int n, int m, const int32_t *__restrict__ cum, const FitRec *__restrict__ rec,
    float *__restrict__ g_xy, float *__restrict__ g_conic, float *__restrict__ g_rgb) {
    const int g = (int)(((size_t)blockIdx.x * kSegThreads + threadIdx.x) / kSegLanes);
    int lo = 0, hi = 0;
    if (g < n) {
        lo = max(0, g == 0 ? 0 : cum[g - 1]);
        hi = min(cum[g], m);
    }
    float s[8];
    seg_sum<8>(reinterpret_cast<const float4 *>(rec), lo, hi, s);
    if (g >= n || (threadIdx.x & (kSegLanes - 1)) != 0) return;
    g_rgb[3 * g + 0] = s[0];
    g_rgb[3 * g + 1] = s[1];
    g_rgb[3 * g + 2] = s[2];
    g_conic[3 * g + 0] = s[3];
    g_conic[3 * g + 1] = s[4];
    g_conic[3 * g + 2] = s[5];
    g_xy[2 * g + 0] = s[6];
    g_xy[2 * g + 1] = s[7];
}

__global__ void add_phantom_kernel(const float *__restrict__ phantom, int has_opac,
                                   float *__restrict__ g_xy, float *__restrict__ g_xy_abs,
                                   float *__restrict__ g_conic, float *__restrict__ g_rgb,
                                   float *__restrict__ g_opac) {
    if (threadIdx.x != 0) return;
    g_rgb[0] += phantom[0];
    g_rgb[1] += phantom[1];
    g_rgb[2] += phantom[2];
    g_conic[0] += phantom[3];
    g_conic[1] += phantom[4];
    g_conic[2] += phantom[5];
    g_xy[0] += phantom[6];
    g_xy[1] += phantom[7];
    g_xy_abs[0] += phantom[8];
    g_xy_abs[1] += phantom[9];
    if (has_opac) g_opac[0] += phantom[10];
}

// Plain backward (backward.cu:101-150 per pair).  Per lane and pixel:
//   v_rgb += alpha * dL/dI,  v_alpha = c . dL/dI + dL/dW,  v_sigma = -alpha v_alpha,
//   v_conic += (v_sigma dx^2 / 2, v_sigma dx dy, v_sigma dy^2 / 2),
//   v_xy += v_sigma (a dx + b dy, b dx + c dy),  v_xy_abs += |that|,
//   v_opacity += vis * v_alpha.
// Moments: within a pixel row dy is fixed and dx = D - lx, so per pixel only
// q_k = sum v_sigma lx^k (k = 1, 2; lx the column relative to the tile's
// centre column c0 = bw / 2, a constant of the unrolled pixel loop, and D =
// mean.x - (x_org + c0 + 0.5): about the centre the terms D^2 q0, 2 D q1 and
// q2 that make up sum v_sigma dx^2 stay small next to it, which keeps FP32
// cancellation down) are accumulated, beside the row's sums of alpha *
// dL/dI per channel (and of alpha * dL/dW when that gradient can be nonzero,
// WS).  At the end of the row q0 = sum v_sigma = -(c . those sums + the dL/dW
// sum) gives the dy moments sum v_sigma dy (my) and dy^2 (syy) and joins its
// total over rows; after the last row the totals give sum v_sigma dx (mx),
// dx^2 (sxx) and dx dy (sxy); the row's colour sums join v_rgb.  v_xy is
// formed once at the end as (a mx + b my, b mx + c my); v_xy_abs accumulates
// |v_sigma| |u'| with the scaled conic's u' = K (a dx + b dy) and is divided
// by K once at the end.
struct PlainAcc {
    float r, g, b, sxx, sxy, syy, mx, my, ax, ay, o;
    float rr, rg, rb, rw;  // current row: sum alpha * dL/dI (rgb), alpha * dL/dW
    float q1;              // current row: sum v_sigma * lx
    float q2;              // all rows: sum v_sigma * lx^2
    float Q0, Q1, Y1;      // sums over rows of q0, q1 and dy * q1
};

__device__ __forceinline__ void plain_acc_zero(PlainAcc &A) {
    A.r = A.g = A.b = A.sxx = A.sxy = A.syy = A.mx = A.my = A.ax = A.ay = A.o = 0.0f;
    A.rr = A.rg = A.rb = A.rw = A.q1 = A.q2 = A.Q0 = A.Q1 = A.Y1 = 0.0f;
}

// End of a pixel row: q0 = sum v_sigma from the row's sums (v_sigma = -alpha
// (c . dL/dI + dL/dW)); the dy moments take the row's dy once (sum v_sigma dy
// and dy^2 in my and syy, sum v_sigma lx dy in Y1), the dx moments only need
// totals over rows (plain_fold).
template <bool WS>
__device__ __forceinline__ void plain_row_end(PlainAcc &A, float dy, float cr, float cg,
                                              float cb) {
    const float cz = WS ? __builtin_fmaf(cb, A.rb, A.rw) : cb * A.rb;
    const float q0 = -__builtin_fmaf(cr, A.rr, __builtin_fmaf(cg, A.rg, cz));
    A.r += A.rr;
    A.g += A.rg;
    A.b += A.rb;
    A.Q0 += q0;
    A.Q1 += A.q1;
    A.my = __builtin_fmaf(dy, q0, A.my);
    A.syy = __builtin_fmaf(dy * q0, dy, A.syy);
    A.Y1 = __builtin_fmaf(dy, A.q1, A.Y1);
    A.rr = A.rg = A.rb = A.q1 = 0.0f;
    if constexpr (WS) A.rw = 0.0f;
}

// After the last row: with dx = D - lx (D = mean.x - (x_org + c0 + 0.5), lx the
// pixel's column relative to c0), sum v_sigma dx = D Q0 - Q1,
// sum v_sigma dx^2 = D (D Q0 - 2 Q1) + q2 and sum v_sigma dx dy = D my - Y1.
__device__ __forceinline__ void plain_fold(PlainAcc &A, float D) {
    A.mx = __builtin_fmaf(D, A.Q0, -A.Q1);
    A.sxx = __builtin_fmaf(D, __builtin_fmaf(D, A.Q0, -2.0f * A.Q1), A.q2);
    A.sxy = __builtin_fmaf(D, A.my, -A.Y1);
}

// One pair of the plain backward.  Rejected pairs (lane test, the accept
// test) leave the accumulators untouched, as the reference's `continue`: the
// accumulation runs under the lane mask of accepted pairs and is skipped when
// no lane accepts.  Without opacities the test is accept_s on s alone, and the
// exponential is evaluated inside the branch.  amin = 1/255 comes in a VGPR;
// clampv = 0.99 is passed as a literal (a VGPR clamp the compiler cannot prove
// canonical costs a quieting v_max before the min).  WS: dL/dW (w.w) may be
// nonzero (the API backward); the fit step's is zero.
template <bool HAS_OPAC, bool ABS = true, bool WS = true>
__device__ __forceinline__ void plain_pair(PlainAcc &A, bool lane_ok, float ap, float bp,
                                           const RowTerms &rt, float op, float cr, float cg,
                                           float cb, float dx, float lx, float amin,
                                           float clampv, const float4 *__restrict__ wp,
                                           bool col0 = false) {
    const float s = pair_s(ap, rt, dx);
    float vis = 0.0f, al = 0.0f;
    bool ok;
    if constexpr (HAS_OPAC) {
        vis = fast_exp2(-s);
        al = fminf(op * vis, clampv);
        ok = lane_ok && !(s < 0.0f) && !(al < amin);
    } else {
        ok = lane_ok && accept_s(s);
    }
    if (ok) {
        if constexpr (!HAS_OPAC) {
            vis = fast_exp2(-s);
            al = fminf(vis, clampv);
        }
        const float4 w = *wp;
        const float va =
            __builtin_fmaf(cr, w.x, __builtin_fmaf(cg, w.y, __builtin_fmaf(cb, w.z, w.w)));
        A.rr = __builtin_fmaf(al, w.x, A.rr);
        A.rg = __builtin_fmaf(al, w.y, A.rg);
        A.rb = __builtin_fmaf(al, w.z, A.rb);
        if constexpr (WS) A.rw = __builtin_fmaf(al, w.w, A.rw);
        if constexpr (HAS_OPAC) A.o = __builtin_fmaf(vis, va, A.o);
        const float vsig = -al * va;
        if (!col0) {  // col0: the centre column (known at compile time), no lx moments
            A.q1 = __builtin_fmaf(vsig, lx, A.q1);
            A.q2 = __builtin_fmaf(vsig, lx * lx, A.q2);
        }
        if constexpr (ABS) {
            const float u = __builtin_fmaf(ap, dx, rt.bdy);
            const float v = __builtin_fmaf(bp, dx, rt.cdy);
            A.ax = __builtin_fmaf(fabsf(vsig), fabsf(u), A.ax);
            A.ay = __builtin_fmaf(fabsf(vsig), fabsf(v), A.ay);
        }
    }
}


template <bool HAS_OPAC, bool DET>
__device__ __forceinline__ void plain_emit(const PlainAcc &A, const GaussPack &gp, bool act,
                                           int slot, int g, GradRec *__restrict__ rec,
                                           float *__restrict__ g_xy, float *__restrict__ g_xy_abs,
                                           float *__restrict__ g_conic, float *__restrict__ g_rgb,
                                           float *__restrict__ g_opac) {
    constexpr float kInvK = 1.0f / kHalfLog2e;
    const float a = gp.p2.y, b = gp.p2.z, c = gp.p2.w;
    const float gx = __builtin_fmaf(a, A.mx, b * A.my);
    const float gy = __builtin_fmaf(b, A.mx, c * A.my);
    if (act)
        emit_grads(DET, rec, slot, g, A.r, A.g, A.b, 0.5f * A.sxx, A.sxy, 0.5f * A.syy, gx, gy,
                   A.ax * kInvK, A.ay * kInvK, A.o, HAS_OPAC, g_xy, g_xy_abs, g_conic, g_rgb,
                   g_opac);
}

// Generic tile width (bw <= 16): pixel loop over the bw x bw tile.
template <bool HAS_OPAC, bool DET>
__global__ __launch_bounds__(kThreads) void raster_bwd_plain_kernel(
    int tiles_x, int bw, int img_w, int img_h, const int32_t *__restrict__ gids,
    const int32_t *__restrict__ slots, const int2 *__restrict__ bins,
    const GaussPack *__restrict__ pack, const int32_t *__restrict__ final_idx,
    const float *__restrict__ v_out, const float *__restrict__ v_wsum,
    GradRec *__restrict__ rec, float *__restrict__ g_xy, float *__restrict__ g_xy_abs,
    float *__restrict__ g_conic, float *__restrict__ g_rgb, float *__restrict__ g_opac) {
    __shared__ float4 sv[kMaxPix];
    __shared__ int sfin[kMaxPix];

    const int tile = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x;
    const int2 range = bins[tile];
    if (range.x >= range.y) return;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int x_org = tx * bw, y_org = ty * bw;
    stage_tile_grads<false>(bw, x_org, y_org, img_w, img_h, final_idx, v_out, v_wsum, nullptr,
                            nullptr, nullptr, sv, sfin, nullptr, nullptr, nullptr);

    const float amin = HAS_OPAC ? in_vgpr(kAlphaMin) : kAlphaMin, clampv = 0.99f;
    for (int base = range.x; base < range.y; base += kThreads) {
        const bool act = base + lane < range.y;
        const int kidx = act ? base + lane : INT32_MAX;
        const int g = act ? gids[base + lane] : 0;
        const GaussPack gp = pack[g];
        PlainAcc A;
        plain_acc_zero(A);
        const int c0 = bw >> 1;  // moments about the tile's centre column
        const float D = gp.p0.x - ((float)(x_org + c0) + 0.5f);
        for (int ly = 0; ly < bw; ++ly) {
            const float dy = gp.p0.y - ((float)(y_org + ly) + 0.5f);
            const RowTerms rt = row_terms(gp.p0.w, gp.p1.x, dy);
            float pxf = (float)x_org + 0.5f;  // exact: integer + 0.5, incremented by 1
            for (int lx = 0; lx < bw; ++lx, pxf += 1.0f) {
                const int p = ly * bw + lx;
                const int fp = sfin[p];
                if (fp < base) continue;  // wave-uniform: nothing in this batch reaches p
                plain_pair<HAS_OPAC>(A, kidx <= fp, gp.p0.z, 0.5f * gp.p0.w, rt, gp.p1.y, gp.p1.z,
                                     gp.p1.w, gp.p2.x, gp.p0.x - pxf, (float)(lx - c0), amin,
                                     clampv, sv + p);
            }
            plain_row_end<true>(A, dy, gp.p1.z, gp.p1.w, gp.p2.x);
        }
        plain_fold(A, D);
        plain_emit<HAS_OPAC, DET>(A, gp, act, DET && act ? slots[base + lane] : 0, g, rec, g_xy,
                                  g_xy_abs, g_conic, g_rgb, g_opac);
    }
}

// Fit-step backward, row grouping.  A lane's record depends only on its own
// Gaussian and the tile's pixels (the batch it shares changes nothing), so a
// chunk of up to kFitChunk of the tile's Gaussians is reordered before it is
// batched: Gaussians that reach no row of the tile get their all-zero record
// at once (the accumulators start at +0 and a rejected pair leaves them
// untouched), the rest are sorted, stably, by (first, last) reachable row, and
// each batch walks only the rows some member reaches.  Rows are still taken
// in increasing order, so every record keeps its bits.
constexpr int kFitChunk = 512;            // Gaussians sorted together (8 batches)
constexpr int kFitCntWords = 128;         // 256 16-bit class counters, two per word
constexpr int kFitLdsInts = kFitCntWords + kFitChunk / 2;  // counters + uint16 order
constexpr uint32_t kNoRows = 0xF0;        // class of a Gaussian reaching no row

// Sorts entries c0 .. c0 + cn - 1 of the tile's list into ord (chunk-local
// indices) and returns how many were sorted; cnt[j] ends holding the rows that
// sorted batch j (ord[64j .. 64j + 63]) walks: the union of the row runs of
// the classes it overlaps.  One wave; its lanes hand LDS data to each other
// through wave_lds_sync.  Kept out of line: inlined, its scheduling region raises the
// register count of the whole kernel past the seven-wave budget.
template <bool HAS_OPAC>
__device__ __attribute__((noinline)) int fit_chunk_sort(int c0, int cn, const int32_t *__restrict__ gids,
                                              const GaussPack *__restrict__ pack, int x_org,
                                              int y_org, int tx, int ty,
                                              FitRec *__restrict__ rec,
                                              GradRec *__restrict__ grec,
                                              const int32_t *__restrict__ slots, uint32_t *cnt,
                                              uint16_t *ord) {
    const int lane = threadIdx.x;
    for (int w = lane; w < kFitCntWords; w += kThreads) cnt[w] = 0u;
    wave_lds_sync();
    constexpr int kRounds = kFitChunk / kThreads;
    static_assert(kRounds <= 8, "one class byte per round in a 64-bit word");
    uint64_t kw = 0;  // the lane's class per round, one byte each
#pragma unroll 2
    for (int k = 0; k < kRounds; ++k) {
        const int i = lane + kThreads * k;
        uint32_t key = kNoRows;
        if (i < cn) {
            const int g = gids[c0 + i];
            const float4 p0 = pack[g].p0, p1 = pack[g].p1;
            const float T = HAS_OPAC ? __log2f(255.0f * p1.y) : kAcceptLog2;
            int rf, rl;
            if (reach_row_span(p0.x, p0.y, p0.z, p0.w, p1.x, T, (float)x_org + 0.5f,
                               (float)y_org + 0.5f, rf, rl)) {
                key = ((uint32_t)rf << 4) | (uint32_t)rl;
                LIG_CHECK((key >> 1) < (uint32_t)kFitCntWords && (key >> 4) <= (key & 15u));
                atomicAdd(&cnt[key >> 1], 1u << (16 * (key & 1u)));
            } else if (rec || grec) {
                // the Gaussian's record for this tile, all zeros (the API
                // record carries its gid: a zero-extent slot's is gid 0)
                const float4 p3 = pack[g].p3;
                const int slot = slots ? slots[c0 + i]
                                       : __float_as_int(p3.z) + ty * __float_as_int(p3.w) + tx;
                LIG_CHECK(slot >= 0);
                const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (rec) {
                    FitRec z;
                    z.r0 = z4;
                    z.r1 = z4;
                    rec[slot] = z;
                } else {
                    GradRec z;
                    z.r0 = z4;
                    z.r1 = z4;
                    z.r2 = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(g));
                    grec[slot] = z;
                }
            }
        }
        kw |= (uint64_t)key << (8 * k);
    }
    wave_lds_sync();
    // exclusive scan of the 256 counts: lane l holds classes 4l .. 4l + 3
    const uint32_t w0 = cnt[2 * lane], w1 = cnt[2 * lane + 1];
    const int n0 = (int)(w0 & 0xFFFFu), n1 = (int)(w0 >> 16), n2 = (int)(w1 & 0xFFFFu),
              n3 = (int)(w1 >> 16);
    const int local = n0 + n1 + n2 + n3;
    int incl = local;
#pragma unroll
    for (int off = 1; off < LIG_WAVE; off <<= 1) {
        const int y = __shfl_up(incl, off, LIG_WAVE);
        if (lane >= off) incl += y;
    }
    const int total = __shfl(incl, LIG_WAVE - 1, LIG_WAVE);
    const uint32_t s0 = (uint32_t)(incl - local), s1 = s0 + n0, s2 = s1 + n1, s3 = s2 + n2;
    cnt[2 * lane] = s0 | (s1 << 16);
    cnt[2 * lane + 1] = s2 | (s3 << 16);
    wave_lds_sync();
    // stable ranking, round by round in index order: lanes of one class are
    // matched by ballots, the class's first lane advances its counter
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (LIG_WAVE - lane));
#pragma unroll
    for (int k = 0; k < kRounds; ++k) {
        if (kThreads * k >= cn) break;  // wave-uniform
        const uint32_t key = (uint32_t)(kw >> (8 * k)) & 0xFFu;
        const bool valid = key != kNoRows;
        uint64_t mm = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (key >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            mm &= bit ? bb : ~bb;
        }
        const int rank = __popcll(mm & lt);
        const int before = (int)((cnt[key >> 1] >> (16 * (key & 1u))) & 0xFFFFu);
        LIG_CHECK(!valid || (before + rank >= 0 && before + rank < cn));
        if (valid) ord[before + rank] = (uint16_t)(lane + kThreads * k);
        if (valid && rank == 0)
            atomicAdd(&cnt[key >> 1], (uint32_t)__popcll(mm) << (16 * (key & 1u)));
        wave_lds_sync();
    }
    // batch row masks from the class ends: lane l holds classes 4l .. 4l + 3
    const uint32_t e01 = cnt[2 * lane], e23 = cnt[2 * lane + 1];
    const int prev = lane == 0 ? 0 : (int)(cnt[2 * lane - 1] >> 16);
    wave_lds_sync();
    if (lane < kRounds) cnt[lane] = 0u;
    wave_lds_sync();
    const int ends[4] = {(int)(e01 & 0xFFFFu), (int)(e01 >> 16), (int)(e23 & 0xFFFFu),
                         (int)(e23 >> 16)};
    int start = prev;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (start < ends[q]) {
            const int key = 4 * lane + q, r0 = key >> 4, r1 = key & 15;
            const uint32_t rows = ((2u << r1) - 1u) & ~((1u << r0) - 1u);
            for (int j = start / kThreads; j <= (ends[q] - 1) / kThreads; ++j) {
                LIG_CHECK(j >= 0 && j < kRounds);
                atomicOr(&cnt[j], rows);
            }
        }
        start = ends[q];
    }
    wave_lds_sync();
    return total;
}

// 16 x 16 tiles.  Per batch and pixel row, two ballots over the row's final
// indices give wave-uniform bit masks: pixels some lane of the batch reaches
// (fp >= base) and pixels every lane reaches (fp >= base + 63).  Only pixels
// of the first kind are evaluated.  A row whose reached pixels are all of the
// second kind runs without the per-lane test kidx <= fp; any other row makes
// the test for every pixel, with fp read from the row's lane by v_readlane.
//
// TRAIN (fit step): the upstream gradient is the clamp + L2 loss gradient,
// formed while staging from (img, target, *grad_loss), there is no final
// index, and each record's emission slot comes from the pack (base + ty * w +
// tx) instead of an emission-id array.  In sum mode the final index only skips pairs that the accept test
// rejects anyway (the forward and this kernel evaluate the same test on the
// same FP32 values), so every in-image pixel is evaluated and the records are
// those of the gated path.
// GROUP (API backward without a final index, !TRAIN): the pairs the forward
// accepted are exactly the ones the accept test accepts here -- in sum mode
// the forward's final index only marks the last accepted pair, and both
// kernels evaluate the same test on the same FP32 values -- so no pixel is
// gated, and each chunk of the tile's Gaussians is grouped by the rows they
// reach, as in the fit step; every record keeps its bits.
// WSUM: the caller passes a wsum gradient (API mode); without one its terms
// are dropped from the pair and row-end arithmetic (they would add zeros).
template <bool HAS_OPAC, bool DET, bool TRAIN, bool GROUP = false, bool WSUM = true>
__device__ __forceinline__ void bwd16_tile(
    int gtile, int part, int nparts, int tiles_x, int img_w, int img_h,
    const int32_t *__restrict__ gids,
    const int32_t *__restrict__ slots, const int2 *__restrict__ bins,
    const GaussPack *__restrict__ pack, const int32_t *__restrict__ final_idx,
    const float *__restrict__ v_out, const float *__restrict__ v_wsum,
    const float *__restrict__ img, const float *__restrict__ target,
    const float *__restrict__ grad_loss, float norm, GradRec *__restrict__ rec,
    float *__restrict__ g_xy, float *__restrict__ g_xy_abs, float *__restrict__ g_conic,
    float *__restrict__ g_rgb, float *__restrict__ g_opac, const BatchImage *__restrict__ tab,
    int nimg, float4 *sv, int *sfin) {
    // gtile: bins index; this block takes part `part` of `nparts` equal shares
    // of the tile's 64-Gaussian batches (each record has its own slot, so the
    // split changes no result)
    const int lane = threadIdx.x;
    int2 range = bins[gtile];
    if (nparts > 1) {
        const int nb = (range.y - range.x + kThreads - 1) / kThreads;
        const int b0 = (int)((long long)part * nb / nparts);
        const int b1 = (int)((long long)(part + 1) * nb / nparts);
        range.y = min(range.y, range.x + b1 * kThreads);
        range.x += b0 * kThreads;
    }
    if (range.x >= range.y) return;
    int tile = gtile;  // tile of its own image
    if (TRAIN && tab) {  // batched fit: the tile's image, canvas, buffers and loss scale
        const BatchImage &im = tab[batch_find<0>(tab, nimg, gtile)];
        tile = gtile - im.tile_off;
        tiles_x = im.tiles_x;
        img_w = im.w;
        img_h = im.h;
        img += 3 * im.pix_off;
        target += 3 * im.pix_off;
        norm = (float)(2.0 / (3.0 * (double)im.h * (double)im.w));
    }
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int x_org = tx * 16, y_org = ty * 16;
    if constexpr (TRAIN)
        stage_tile_loss_grads(x_org, y_org, img_w, img_h, img, target, grad_loss[0], norm, sv);
    else
        stage_tile_grads<false>(16, x_org, y_org, img_w, img_h, final_idx, v_out, v_wsum,
                                nullptr, nullptr, nullptr, sv, GROUP ? nullptr : sfin, nullptr,
                                nullptr, nullptr);
    // in-image rows and columns of the tile (TRAIN, GROUP; the final index
    // excludes the others)
    const int nrows = (TRAIN || GROUP) ? min(16, img_h - y_org) : 16;
    const uint32_t cols = img_w - x_org >= 16 ? 0xffffu : ((1u << (img_w - x_org)) - 1u);

    // one batch: lanes act (Gaussian g, list position kidx) over the rows in
    // `rows` (bit r: tile row r), in increasing order
    auto run_batch = [&](int base, bool act, int kidx, int g, uint32_t rows) {
        const GaussPack gp = pack[g];
        const float mx = gp.p0.x, my = gp.p0.y, ap = gp.p0.z, bp2 = gp.p0.w, bp = 0.5f * bp2;
        const float cp = gp.p1.x, op = gp.p1.y, cr = gp.p1.z, cg = gp.p1.w, cb = gp.p2.x;
        // exact: x_org + lx + 0.5 is representable, one rounding in the difference.
        // The fit step holds the 16 differences in VGPRs; the API path (more
        // accumulators, the final-index gating) keeps the 16 column centres
        // wave-uniform in SGPRs and forms each difference where it is used,
        // the same expression, to stay within 72 VGPRs.
        float dxs[TRAIN ? 16 : 1];
        float xcs[TRAIN ? 1 : 16];
#pragma unroll
        for (int lx = 0; lx < 16; ++lx) {
            if constexpr (TRAIN)
                dxs[lx] = mx - ((float)(x_org + lx) + 0.5f);
            else
                xcs[lx] = __int_as_float(__builtin_amdgcn_readfirstlane(
                    __float_as_int((float)(x_org + lx) + 0.5f)));
        }
        auto dx_of = [&](int lx) {
            if constexpr (TRAIN)
                return dxs[lx];
            else
                return mx - xcs[lx];
        };
        PlainAcc A;
        plain_acc_zero(A);
        const int full_from = base + kThreads - 1;
        const float amin = HAS_OPAC ? in_vgpr(kAlphaMin) : kAlphaMin, clampv = 0.99f;
        while (rows) {
            const int ly = __builtin_ctz(rows);
            rows &= rows - 1u;
            int fcol = 0;
            uint32_t any = cols, full = cols;
            if constexpr (!TRAIN && !GROUP) {
                fcol = sfin[ly * 16 + (lane & 15)];
                any = (uint32_t)__ballot(fcol >= base) & 0xffffu;
                if (any == 0) continue;
                full = (uint32_t)__ballot(fcol >= full_from) & 0xffffu;
            }
            const float dy = my - ((float)(y_org + ly) + 0.5f);
            const RowTerms rt = row_terms(bp2, cp, dy);
            // the row's LDS address as one byte offset (immediate offsets per pixel)
            const float4 *row = reinterpret_cast<const float4 *>(
                reinterpret_cast<const char *>(sv) + in_vgpr(ly * 16 * (int)sizeof(float4)));
            if (full == 0xffffu) {
#pragma unroll
                for (int lx = 0; lx < 16; ++lx)
                    plain_pair<HAS_OPAC, !TRAIN, !TRAIN && WSUM>(A, true, ap, bp, rt, op, cr, cg, cb, dx_of(lx),
                                                 (float)(lx - 8), amin, clampv, row + lx, lx == 8);
            } else if (full == any) {
#pragma unroll
                for (int lx = 0; lx < 16; ++lx)
                    if (any & (1u << lx))
                        plain_pair<HAS_OPAC, !TRAIN, !TRAIN && WSUM>(A, true, ap, bp, rt, op, cr, cg, cb,
                                                     dx_of(lx), (float)(lx - 8), amin, clampv,
                                                     row + lx, lx == 8);
            } else {
#pragma unroll
                for (int lx = 0; lx < 16; ++lx)
                    if (any & (1u << lx))
                        plain_pair<HAS_OPAC, true, WSUM>(A, kidx <= __builtin_amdgcn_readlane(fcol, lx), ap,
                                             bp, rt, op, cr, cg, cb, dx_of(lx), (float)(lx - 8),
                                             amin, clampv, row + lx, lx == 8);
            }
            plain_row_end<!TRAIN && WSUM>(A, dy, cr, cg, cb);
        }
        plain_fold(A, dx_of(8));
        if constexpr (TRAIN) {
            if (act) {
                const int slot = __float_as_int(gp.p3.z) + ty * __float_as_int(gp.p3.w) + tx;
                const float a = gp.p2.y, b = gp.p2.z, c = gp.p2.w;
                FitRec r;
                r.r0 = make_float4(A.r, A.g, A.b, 0.5f * A.sxx);
                r.r1 = make_float4(A.sxy, 0.5f * A.syy, __builtin_fmaf(a, A.mx, b * A.my),
                                   __builtin_fmaf(b, A.mx, c * A.my));
                LIG_CHECK(slot >= 0);
                reinterpret_cast<FitRec *>(rec)[slot] = r;
            }
        } else {
            // kidx: the lane's list position (its emission slot's index)
            const int slot = DET && act ? record_slot(slots, kidx, gp, tx, ty) : 0;
            plain_emit<HAS_OPAC, DET>(A, gp, act, slot, g, rec, g_xy, g_xy_abs, g_conic, g_rgb,
                                      g_opac);
        }
    };
    const uint32_t all_rows = (1u << nrows) - 1u;
    if constexpr (TRAIN || GROUP) {
        uint32_t *cnt = reinterpret_cast<uint32_t *>(sfin);
        uint16_t *ord = reinterpret_cast<uint16_t *>(sfin + kFitCntWords);
        for (int c0 = range.x; c0 < range.y; c0 += kFitChunk) {
            const int cn = min(kFitChunk, range.y - c0);
            const int m = fit_chunk_sort<HAS_OPAC>(
                c0, cn, gids, pack, x_org, y_org, tx, ty,
                TRAIN ? reinterpret_cast<FitRec *>(rec) : nullptr,
                !TRAIN && DET ? rec : nullptr, TRAIN ? nullptr : slots, cnt, ord);
            for (int b = 0; b < m; b += kThreads) {
                const bool act = b + lane < m;
                const int k = act ? c0 + (int)ord[b + lane] : 0;
                const int g = act ? gids[k] : 0;
                const uint32_t rows = (uint32_t)__builtin_amdgcn_readfirstlane(
                                          (int)cnt[b / kThreads]) & all_rows;
                run_batch(b, act, TRAIN ? INT32_MAX : k, g, rows);
            }
            wave_lds_sync();  // the next chunk rewrites cnt and ord
        }
    } else {
        for (int base = range.x; base < range.y; base += kThreads) {
            const bool act = base + lane < range.y;
            run_batch(base, act, act ? base + lane : INT32_MAX, act ? gids[base + lane] : 0,
                      all_rows);
        }
    }
}

// One share of a tile's batches per block (split = 1: the whole tile): block
// b takes share b % split of tile order[b / split], or without an order (none
// given, or order[0] = -1), with u = xcd_remap(b), share u % split of tile
// u / split.
template <bool HAS_OPAC, bool DET, bool TRAIN, bool GROUP = false, bool WSUM = true>
__global__ __launch_bounds__(kThreads) void raster_bwd16_kernel(
    int tiles_x, int img_w, int img_h, const int32_t *__restrict__ gids,
    const int32_t *__restrict__ slots, const int2 *__restrict__ bins,
    const GaussPack *__restrict__ pack, const int32_t *__restrict__ final_idx,
    const float *__restrict__ v_out, const float *__restrict__ v_wsum,
    const float *__restrict__ img, const float *__restrict__ target,
    const float *__restrict__ grad_loss, float norm, GradRec *__restrict__ rec,
    float *__restrict__ g_xy, float *__restrict__ g_xy_abs, float *__restrict__ g_conic,
    float *__restrict__ g_rgb, float *__restrict__ g_opac, const BatchImage *__restrict__ tab,
    int nimg, int split, const int32_t *__restrict__ order) {
    __shared__ float4 sv[kMaxPix];
    __shared__ int sfin[(TRAIN || GROUP) ? kFitLdsInts : kMaxPix];  // row-group sort
    int gtile, part;
    if (order && order[0] >= 0) {  // the shares of a tile are consecutive blocks
        const int k = (int)blockIdx.x / split;
        part = (int)blockIdx.x - k * split;
        gtile = order[k];
    } else {
        const int u = (int)xcd_remap(blockIdx.x, gridDim.x);
        const int k = u / split;
        part = u - k * split;
        gtile = tile_of_slot(k, tiles_x, (img_h + 15) >> 4, tab == nullptr);
    }
    bwd16_tile<HAS_OPAC, DET, TRAIN, GROUP, WSUM>(gtile, part, split, tiles_x, img_w, img_h,
                                     gids, slots, bins, pack, final_idx, v_out, v_wsum, img,
                                     target, grad_loss, norm, rec, g_xy, g_xy_abs, g_conic, g_rgb,
                                     g_opac, tab, nimg, sv, sfin);
}

// Persistent form of the fit-step backward (batched): each wave takes tiles
// from the per-XCD queue until none is left.
__global__ __launch_bounds__(kThreads) void raster_bwd16_fit_queue_kernel(
    const int32_t *__restrict__ gids, const int2 *__restrict__ bins,
    const GaussPack *__restrict__ pack, const float *__restrict__ img,
    const float *__restrict__ target, const float *__restrict__ grad_loss,
    GradRec *__restrict__ rec, const BatchImage *__restrict__ tab, int nimg, TileQueue queue) {
    __shared__ float4 sv[kMaxPix];
    __shared__ int sfin[kFitLdsInts];  // row-group sort
    int cnt, x;
    int i = queue_resolve(queue, queue_issue(queue), cnt, x);
    while (i >= 0) {
        const int pending = queue_issue(queue);
        for (int k = 0; k < cnt; ++k) {
            bwd16_tile<false, true, true>(queue_tile(queue, x, i + k), 0, 1, 1, 1, 1, gids,
                                          nullptr, bins, pack, nullptr, nullptr, nullptr, img,
                                          target, grad_loss, 0.0f, rec, nullptr, nullptr, nullptr,
                                          nullptr, nullptr, tab, nimg, sv, sfin);
            __syncthreads();  // the next tile restages sv
        }
        i = queue_resolve(queue, pending, cnt, x);
    }
}

// Gradient-aware backward (gradient_aware_rasterizer.cu:201-315 per pair),
// d' = pixel - mean, sigma_x = a d'x + b d'y, sigma_y = b d'x + c d'y.
// xys.absgrad follows the reference's reduction: v_xy is summed over each
// 32-thread warp of the bw x bw block (thread rank p = ly * bw + lx) and the
// absolute value of each warp's sum is added (gradient_aware_rasterizer.cu:
// 170, 320, 339-341).  A lane walks its Gaussian's pixels in rank order, so a
// group is a run of 32 consecutive p; pixels outside the image add 0.
template <bool DET>
__global__ __launch_bounds__(kThreads) void raster_bwd_ga_kernel(
    int tiles_x, int bw, int img_w, int img_h, const int32_t *__restrict__ gids,
    const int32_t *__restrict__ slots, const int2 *__restrict__ bins,
    const GaussPack *__restrict__ pack,
    const int32_t *__restrict__ final_idx, const float *__restrict__ v_out,
    const float *__restrict__ v_wsum, const float *__restrict__ v_dx,
    const float *__restrict__ v_dy, const float *__restrict__ v_dxy, int has_opac,
    GradRec *__restrict__ rec, float *__restrict__ g_xy, float *__restrict__ g_xy_abs,
    float *__restrict__ g_conic, float *__restrict__ g_rgb, float *__restrict__ g_opac) {
    __shared__ float4 sv[kMaxPix];
    __shared__ int sfin[kMaxPix];
    __shared__ float4 sa[kMaxPix];  // (vdx r, g, b, vdy r)
    __shared__ float4 sb[kMaxPix];  // (vdy g, b, vdxy r, g)
    __shared__ float sc[kMaxPix];   // vdxy b

    const int tile = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x;
    const int2 range = bins[tile];
    if (range.x >= range.y) return;
    const int ty = tile / tiles_x;
    const int tx = tile - ty * tiles_x;
    const int x_org = tx * bw, y_org = ty * bw;
    stage_tile_grads<true>(bw, x_org, y_org, img_w, img_h, final_idx, v_out, v_wsum, v_dx,
                           v_dy, v_dxy, sv, sfin, sa, sb, sc);

    constexpr float kClamp = 0.999f;
    for (int base = range.x; base < range.y; base += kThreads) {
        const bool act = base + lane < range.y;
        const int kidx = act ? base + lane : INT32_MAX;
        const int g = act ? gids[base + lane] : 0;
        const GaussPack gp = pack[g];
        const float mx = gp.p0.x, my = gp.p0.y, ap = gp.p0.z, bp2 = gp.p0.w;
        const float cp = gp.p1.x, op = gp.p1.y, cr = gp.p1.z, cg = gp.p1.w;
        const float cb = gp.p2.x, ka = gp.p2.y, kb = gp.p2.z, kc = gp.p2.w;

        float a_r = 0.f, a_g = 0.f, a_b = 0.f;
        float a_c0 = 0.f, a_c1 = 0.f, a_c2 = 0.f;
        float a_x = 0.f, a_y = 0.f, a_ax = 0.f, a_ay = 0.f, a_o = 0.f;
        float w_x = 0.f, w_y = 0.f;  // v_xy of the current 32-rank group

        for (int ly = 0; ly < bw; ++ly) {
            const float dy = my - ((float)(y_org + ly) + 0.5f);
            const RowTerms rt = row_terms(bp2, cp, dy);
            float pxf = (float)x_org + 0.5f;
            for (int lx = 0; lx < bw; ++lx, pxf += 1.0f) {
                const int p = ly * bw + lx;
                if ((p & 31) == 0) {
                    a_ax += fabsf(w_x);
                    a_ay += fabsf(w_y);
                    w_x = 0.f;
                    w_y = 0.f;
                }
                const int fp = sfin[p];
                if (fp < base) continue;
                const float dx = mx - pxf;
                const float s = pair_s(ap, rt, dx);
                const float vis = fast_exp2(-s);
                const float al0 = fminf(op * vis, kClamp);
                const bool ok =
                    kidx <= fp && (has_opac ? !(s < 0.0f) && !(al0 < kAlphaMin) : accept_s(s));
                const float al = ok ? al0 : 0.0f;
                const float vs_ = ok ? vis : 0.0f;
                const float4 w = sv[p];
                const float4 wa = sa[p];
                const float4 wb = sb[p];
                const float wc = sc[p];
                // upstream: vI = w.xyz, vW = w.w, vdx = wa.xyz, vdy = (wa.w, wb.x, wb.y),
                // vdxy = (wb.z, wb.w, wc).
                const float ex = -dx, ey = -dy;
                const float sx = __builtin_fmaf(ka, ex, kb * ey);
                const float sy = __builtin_fmaf(kb, ex, kc * ey);
                const float alx = -al * sx;
                const float aly = -al * sy;
                const float P = __builtin_fmaf(sx, sy, -kb);
                const float alxy = al * P;
                a_r += w.x * al + wa.x * alx + wa.w * aly + wb.z * alxy;
                a_g += w.y * al + wa.y * alx + wb.x * aly + wb.w * alxy;
                a_b += w.z * al + wa.z * alx + wb.y * aly + wc * alxy;
                const float va = __builtin_fmaf(w.x, cr, __builtin_fmaf(w.y, cg, __builtin_fmaf(w.z, cb, w.w)));
                const float vax = __builtin_fmaf(wa.x, cr, __builtin_fmaf(wa.y, cg, wa.z * cb));
                const float vay = __builtin_fmaf(wa.w, cr, __builtin_fmaf(wb.x, cg, wb.y * cb));
                const float vaxy = __builtin_fmaf(wb.z, cr, __builtin_fmaf(wb.w, cg, wc * cb));
                const float sx2 = sx * sx, sy2 = sy * sy, sxsy = sx * sy;
                const float gx = al * (va * sx + vax * (ka - sx2) + vay * (kb - sxsy) +
                                       vaxy * (sx2 * sy - 2.0f * kb * sx - ka * sy));
                const float gy = al * (va * sy + vax * (kb - sxsy) + vay * (kc - sy2) +
                                       vaxy * (sx * sy2 - 2.0f * kb * sy - kc * sx));
                const float ex2 = ex * ex, ey2 = ey * ey, exy = ex * ey;
                a_c0 += al * (va * (-0.5f * ex2) + vax * (0.5f * ex2 * sx - ex) +
                              vay * (0.5f * ex2 * sy) + vaxy * (-0.5f * ex2 * P + ex * sy));
                a_c1 += al * (va * (-exy) + vax * (exy * sx - ey) + vay * (exy * sy - ex) +
                              vaxy * (-exy * P + ey * sy + ex * sx - 1.0f));
                a_c2 += al * (va * (-0.5f * ey2) + vax * (0.5f * ey2 * sx) +
                              vay * (0.5f * ey2 * sy - ey) + vaxy * (-0.5f * ey2 * P + ey * sx));
                a_x += gx;
                a_y += gy;
                w_x += gx;
                w_y += gy;
                a_o = __builtin_fmaf(vs_, va, a_o);
            }
        }
        a_ax += fabsf(w_x);
        a_ay += fabsf(w_y);
        if (act)
            emit_grads(DET, rec, DET ? record_slot(slots, base + lane, gp, tx, ty) : 0, g, a_r,
                       a_g, a_b, a_c0, a_c1, a_c2, a_x, a_y, a_ax, a_ay, a_o, has_opac != 0,
                       g_xy, g_xy_abs, g_conic, g_rgb, g_opac);
    }
}

// Deterministic FP64 sum of `count` FP32 partials in two launches:
// sum_partials_kernel: block b sums partial[b*kSumChunk, (b+1)*kSumChunk) into
// block_sum[b]; sum_blocks_kernel: out = inv_n * sum of the block sums.  Fixed
// order throughout.
constexpr int kSumThreads = 256;
constexpr int kSumChunk = kSumThreads * 16;

template <int THREADS = kSumThreads>
__device__ __forceinline__ double block_sum_f64(double v, double *ws) {
    ws[threadIdx.x] = v;
    __syncthreads();
    for (int s = THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) ws[threadIdx.x] += ws[threadIdx.x + s];
        __syncthreads();
    }
    return ws[0];
}

__global__ __launch_bounds__(kSumThreads) void sum_partials_kernel(
    int count, const float *__restrict__ partial, double *__restrict__ block_sum) {
    __shared__ double ws[kSumThreads];
    const int base = blockIdx.x * kSumChunk;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kSumChunk / kSumThreads; ++k) {
        const int i = base + k * kSumThreads + (int)threadIdx.x;
        if (i < count) acc += partial[i];
    }
    const double t = block_sum_f64(acc, ws);
    if (threadIdx.x == 0) block_sum[blockIdx.x] = t;
}

__global__ __launch_bounds__(kSumThreads) void sum_blocks_kernel(int nblocks,
                                                                const double *__restrict__ bsum,
                                                                double inv_n,
                                                                float *__restrict__ out) {
    __shared__ double ws[kSumThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += kSumThreads) acc += bsum[i];
    const double t = block_sum_f64(acc, ws);
    if (threadIdx.x == 0) out[0] = (float)(t * inv_n);
}

// Batched fit losses: image b's mean over its 3 w h values of the per-tile
// clamp + L2 sums of its tile range (FP64, fixed order), then their total.
// One block per image.  Thread i sums the image's tile partials i, i + 1024,
// ... into eight FP64 accumulators (k mod 8), eight loads in flight per step,
// and adds them in a fixed order; then a fixed block tree.  A 262 K-tile
// image takes tens of microseconds instead of one latency-bound chain.
constexpr int kImgLossThreads = 1024;
constexpr int kImgLossLanes = 8;

__global__ __launch_bounds__(kImgLossThreads) void image_loss_kernel(
    const BatchImage *__restrict__ tab, const float *__restrict__ partial,
    float *__restrict__ loss_img) {
    __shared__ double ws[kImgLossThreads];
    const BatchImage im = tab[blockIdx.x];
    const int t1 = tab[blockIdx.x + 1].tile_off;
    double acc[kImgLossLanes];
#pragma unroll
    for (int j = 0; j < kImgLossLanes; ++j) acc[j] = 0.0;
    constexpr int kStep = kImgLossThreads * kImgLossLanes;
    for (int t0 = im.tile_off + (int)threadIdx.x; t0 < t1; t0 += kStep) {
        float v[kImgLossLanes];
#pragma unroll
        for (int j = 0; j < kImgLossLanes; ++j) {
            const int t = t0 + j * kImgLossThreads;
            v[j] = t < t1 ? partial[t] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < kImgLossLanes; ++j) acc[j] += (double)v[j];
    }
    const double mine = ((acc[0] + acc[1]) + (acc[2] + acc[3])) +
                        ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    const double sum = block_sum_f64<kImgLossThreads>(mine, ws);
    if (threadIdx.x == 0) loss_img[blockIdx.x] = (float)(sum / (3.0 * (double)im.h * im.w));
}

__global__ __launch_bounds__(kSumThreads) void total_loss_kernel(int nimg,
                                                                const float *__restrict__ loss_img,
                                                                float *__restrict__ total) {
    __shared__ double ws[kSumThreads];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nimg; b += kSumThreads) acc += loss_img[b];
    const double sum = block_sum_f64(acc, ws);
    if (threadIdx.x == 0) total[0] = (float)sum;
}

// Launch order of a tile grid.  Heaviest first: a tile's class falls with
// its pair count, four classes per octave (128 in all); order_hist_kernel
// counts the classes and order_scatter_kernel writes each tile at its class's
// next free place (tiles of one class in no fixed order, which changes no
// result: each tile's work and outputs are its own).  Launched so (block b
// takes tile order[b]; the hardware deals consecutive blocks to the eight
// XCDs in turn), long tiles start first and every XCD gets a like mix.  The
// default mapping, xcd_remap, gives each XCD a contiguous band of tiles, which
// keeps a band's Gaussians in that XCD's L2 but leaves the XCDs that own a
// dense region running long after the others.  So the heaviest-first order is
// taken when the caller asks for it (a launch of few rounds, whose tail it
// shortens) or when the bands' loads (pairs + kTileCost per tile) differ by
// more than a tenth; otherwise order[0] = -1 and the kernels keep xcd_remap.
constexpr int kOrderClasses = 128;
constexpr int kOrderThreads = 256;
constexpr int kOrderPer = 8;    // tiles per thread
constexpr int kTileCost = 16;   // a tile's fixed cost, in pairs

__device__ __forceinline__ int order_class(int2 bin) {
    const unsigned c = (unsigned)max(bin.y - bin.x, 0) + 1u;
    const int lg = 31 - __clz(c);
    const unsigned frac = lg >= 2 ? (c >> (lg - 2)) & 3u : (c << (2 - lg)) & 3u;
    return kOrderClasses - 1 - (4 * lg + (int)frac);
}

// xcd_remap's band (XCD share) of tile t of nt
__device__ __forceinline__ int xcd_band(int t, int nt) {
    const int q = nt >> 3, r = nt & 7;
    return t < r * (q + 1) ? t / (q + 1) : r + (t - r * (q + 1)) / max(q, 1);
}

__global__ __launch_bounds__(kOrderThreads) void order_hist_kernel(int ntiles,
                                                                   const int2 *__restrict__ bins,
                                                                   int *__restrict__ hist,
                                                                   long long *__restrict__ band) {
    __shared__ int lh[kOrderClasses];
    __shared__ long long lb[8];
    for (int i = threadIdx.x; i < kOrderClasses; i += kOrderThreads) lh[i] = 0;
    if (threadIdx.x < 8) lb[threadIdx.x] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kOrderPer; ++k) {
        const int t = (blockIdx.x * kOrderPer + k) * kOrderThreads + threadIdx.x;
        if (t < ntiles) {
            const int2 b = bins[t];
            atomicAdd(&lh[order_class(b)], 1);
            atomicAdd((unsigned long long *)&lb[xcd_band(t, ntiles)],
                      (unsigned long long)(max(b.y - b.x, 0) + kTileCost));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kOrderClasses; i += kOrderThreads)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
    if (threadIdx.x < 8 && lb[threadIdx.x])
        atomicAdd((unsigned long long *)&band[threadIdx.x], (unsigned long long)lb[threadIdx.x]);
}

__global__ __launch_bounds__(kOrderThreads) void order_scatter_kernel(
    int ntiles, const int2 *__restrict__ bins, const int *__restrict__ hist,
    const long long *__restrict__ band, int force, int *__restrict__ cursor,
    int32_t *__restrict__ order) {
    if (!force) {
        long long mx = 0, sum = 0;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            mx = max(mx, band[x]);
            sum += band[x];
        }
        if (mx * 80 <= sum * 11) {  // max <= 1.1 x mean: keep the bands
            if (blockIdx.x == 0 && threadIdx.x == 0) order[0] = -1;
            return;
        }
    }
    __shared__ int pre[kOrderClasses];  // exclusive start of each class
    __shared__ int lcnt[kOrderClasses];
    __shared__ int lbase[kOrderClasses];
    if (threadIdx.x == 0) {
        int run = 0;
        for (int i = 0; i < kOrderClasses; ++i) {
            pre[i] = run;
            run += hist[i];
        }
    }
    for (int i = threadIdx.x; i < kOrderClasses; i += kOrderThreads) lcnt[i] = 0;
    __syncthreads();
    int cls[kOrderPer], rank[kOrderPer];
#pragma unroll
    for (int k = 0; k < kOrderPer; ++k) {
        const int t = (blockIdx.x * kOrderPer + k) * kOrderThreads + threadIdx.x;
        cls[k] = t < ntiles ? order_class(bins[t]) : -1;
        rank[k] = cls[k] >= 0 ? atomicAdd(&lcnt[cls[k]], 1) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kOrderClasses; i += kOrderThreads)
        lbase[i] = lcnt[i] ? atomicAdd(&cursor[i], lcnt[i]) : 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kOrderPer; ++k) {
        const int t = (blockIdx.x * kOrderPer + k) * kOrderThreads + threadIdx.x;
        if (cls[k] >= 0) {
            LIG_CHECK(cls[k] < kOrderClasses && pre[cls[k]] + lbase[cls[k]] + rank[k] < ntiles);
            order[pre[cls[k]] + lbase[cls[k]] + rank[k]] = t;
        }
    }
}

__global__ void zero_f32_kernel(float *__restrict__ p, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0f;
}

// Workload counters for the roofline: E = pairs whose pixel centre lies in the
// Gaussian's clipped SnugBox box |p - mean| <= extent, C = pairs of E with
// sigma >= 0 and alpha >= 1/255 (FP64, clamp 0.999).  One thread per Gaussian.
__global__ __launch_bounds__(256) void count_pairs_kernel(
    int n, int img_h, int img_w, const float *__restrict__ xys, const float *__restrict__ extents,
    const float *__restrict__ conics, const float *__restrict__ opacities,
    const int32_t *__restrict__ ntiles, unsigned long long *__restrict__ out) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    unsigned long long e = 0, c = 0;
    if (g < n && ntiles[g] > 0) {
        const double mx = xys[2 * g], my = xys[2 * g + 1];
        const double ex = extents[2 * g], ey = extents[2 * g + 1];
        int j0 = (int)ceil(mx - ex - 0.5), j1 = (int)floor(mx + ex - 0.5);
        int i0 = (int)ceil(my - ey - 0.5), i1 = (int)floor(my + ey - 0.5);
        j0 = max(j0, 0);
        i0 = max(i0, 0);
        j1 = min(j1, img_w - 1);
        i1 = min(i1, img_h - 1);
        const double a = conics[3 * g], b = conics[3 * g + 1], cc = conics[3 * g + 2];
        const double o = opacities ? opacities[g] : 1.0;
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) {
                const double dx = mx - (j + 0.5), dy = my - (i + 0.5);
                const double sg = 0.5 * (a * dx * dx + cc * dy * dy) + b * dx * dy;
                double al = o * exp(-sg);
                al = al > 0.999 ? 0.999 : al;
                ++e;
                c += (sg < 0.0 || al < 1.0 / 255.0) ? 0 : 1;
            }
    }
    // wave reduction, then one atomic per wave
    for (int off = LIG_WAVE / 2; off > 0; off >>= 1) {
        e += __shfl_down(e, off, LIG_WAVE);
        c += __shfl_down(c, off, LIG_WAVE);
    }
    if ((threadIdx.x & (LIG_WAVE - 1)) == 0) {
        atomicAdd(&out[0], e);
        atomicAdd(&out[1], c);
    }
}

__global__ void zero_u64_kernel(unsigned long long *p, int count) {
    if ((int)threadIdx.x < count) p[threadIdx.x] = 0ull;
}

template <bool GA>
void launch_fwd(int px, dim3 grid, hipStream_t st, int tiles_x, int bw, int img_w, int img_h,
                const int32_t *gids, const int2 *bins, const GaussPack *pack, float *img,
                float *wsum, int32_t *fin, float *dx, float *dy, float *dxy, int has_opac) {
    switch (px) {
        case 1:
            hipLaunchKernelGGL((raster_fwd_kernel<1, GA>), grid, dim3(kThreads), 0, st, tiles_x,
                               bw, img_w, img_h, gids, bins, pack, img, wsum, fin, dx, dy, dxy,
                               has_opac);
            break;
        case 2:
            hipLaunchKernelGGL((raster_fwd_kernel<2, GA>), grid, dim3(kThreads), 0, st, tiles_x,
                               bw, img_w, img_h, gids, bins, pack, img, wsum, fin, dx, dy, dxy,
                               has_opac);
            break;
        case 3:
            hipLaunchKernelGGL((raster_fwd_kernel<3, GA>), grid, dim3(kThreads), 0, st, tiles_x,
                               bw, img_w, img_h, gids, bins, pack, img, wsum, fin, dx, dy, dxy,
                               has_opac);
            break;
        default:
            hipLaunchKernelGGL((raster_fwd_kernel<4, GA>), grid, dim3(kThreads), 0, st, tiles_x,
                               bw, img_w, img_h, gids, bins, pack, img, wsum, fin, dx, dy, dxy,
                               has_opac);
            break;
    }
}

}  // namespace
}  // namespace lig

using namespace lig;

extern "C" size_t lig_packed_gaussians_bytes(int n) {
    return n > 0 ? (size_t)n * sizeof(GaussPack) : 0;
}

extern "C" int lig_pack_gaussians(int n, const float *xys, const float *conics,
                                  const float *colors, const float *opacities, void *packed,
                                  void *stream) {
    if (n < 0 || (n > 0 && (!xys || !conics || !colors || !packed)))
        return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    hipLaunchKernelGGL(pack_kernel<>, dim3(ceil_div(n, 256)), dim3(256), 0, as_stream(stream), n,
                       xys, conics, colors, opacities, nullptr, nullptr, 0, 0, nullptr, 0,
                       reinterpret_cast<GaussPack *>(packed));
    return launch_status();
}


namespace lig {
namespace {
// LDS batch depths of the fit forward; 0 selects the default, 64.
bool depth_ok(int d) { return d == 0 || d == 64 || d == 128 || d == 256; }

// Grid of a persistent launch of `kernel` over `ntiles` tiles: the waves
// that fit on the device at once (a multiple of 8, so every XCD gets the same
// number), at most one per tile.
template <typename KernelT>
unsigned persistent_grid(KernelT kernel, int ntiles) {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0);
    long long g = (long long)(cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    if (g >= 8) g &= ~7ll;
    return (unsigned)(g < ntiles ? g : ntiles);
}

// Shares per tile for the 16x16 backward (any number gives the same
// records).  With R = the blocks resident at once:
// * T * s <= R for some s >= 2 (a small level: every share runs at once, so
//   the deepest tile sets the time): the largest such s up to 8 (512^2 x
//   20 K Gaussians of 700 pairs per tile: 0.35 -> 0.11 ms);
// * else a launch of T * s blocks runs in about ceil(T * s / R) rounds of
//   (mean batches per tile / s + c) batch times, c being a share's fixed cost
//   (staging the tile's pixel gradients; about 0.75 batch times, measured);
//   the s that minimises this, keeping at least 1.5 batches per share, when
//   it saves 3% (2048^2 x 1 M: s = 3, 10% faster; 8192^2 x 2 M: s = 1).
std::atomic<int> g_bwd_shares{0};  // lig_set_backward_shares: 0 = chosen per launch

template <typename KernelT>
int bwd16_split(KernelT kernel, int ntiles, long long nisect) {
    if (const int forced = g_bwd_shares.load(std::memory_order_relaxed)) return forced;
    if (ntiles <= 0 || nisect <= 0) return 1;
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0);
    const double resident = (double)(cus > 0 ? cus : 1) * (per > 0 ? per : 1);
    const int one_round = (int)std::min<double>(kMaxBwdSplit, std::floor(resident / ntiles));
    if (one_round >= 2) return one_round;
    const double batches = (double)nisect / ((double)ntiles * kThreads);
    constexpr double kShareCost = 0.75;
    auto cost = [&](int sh) {
        return std::ceil((double)ntiles * sh / resident) * (batches / sh + kShareCost);
    };
    int best = 1;
    double best_cost = cost(1);
    for (int sh = 2; sh <= kMaxBwdSplit && batches / sh >= 1.5; ++sh) {
        if ((long long)ntiles * sh > INT32_MAX) break;
        const double c = cost(sh);
        if (c < 0.97 * best_cost) {
            best = sh;
            best_cost = c;
        }
    }
    return best;
}

// Workspace of launch_tile_order: the order, then band loads, class counts
// and cursors.
size_t order_ws_bytes(int ntiles) {
    return (((size_t)ntiles * sizeof(int32_t) + 255) & ~size_t(255)) + 8 * sizeof(long long) +
           2 * kOrderClasses * sizeof(int);
}

// Blocks of `kernel` resident on the device at once.
template <typename KernelT>
long long resident_blocks(KernelT kernel) {
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0);
    return (long long)(cus > 0 ? cus : 1) * (per > 0 ? per : 1);
}

// Launch order of `ntiles` tiles (order_hist_kernel / order_scatter_kernel)
// in `ws`, order_ws_bytes(ntiles) bytes, 8-byte aligned.  `blocks` blocks
// will take the tiles and `resident` of them run at once: under 8 rounds the
// heaviest-first order is always taken; within one round every block starts
// at once and the order cannot matter, so none is computed (null).
const int32_t *launch_tile_order(int ntiles, const int2 *bins, long long blocks,
                                 long long resident, void *ws, hipStream_t st) {
    if (ntiles <= 0 || !ws || blocks <= resident) return nullptr;
    int32_t *order = reinterpret_cast<int32_t *>(ws);
    long long *band = reinterpret_cast<long long *>(
        reinterpret_cast<char *>(ws) + (((size_t)ntiles * sizeof(int32_t) + 255) & ~size_t(255)));
    int *hist = reinterpret_cast<int *>(band + 8);
    int *cursor = hist + kOrderClasses;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(512), 0, st, reinterpret_cast<float *>(band),
                       (size_t)(16 + 2 * kOrderClasses));
    const dim3 grid((unsigned)ceil_div(ntiles, kOrderThreads * kOrderPer));
    const int force = blocks < 8 * resident ? 1 : 0;
    hipLaunchKernelGGL(order_hist_kernel, grid, dim3(kOrderThreads), 0, st, ntiles, bins, hist,
                       band);
    hipLaunchKernelGGL(order_scatter_kernel, grid, dim3(kOrderThreads), 0, st, ntiles, bins,
                       (const int *)hist, (const long long *)band, force, cursor, order);
    return order;
}

// The order in the slack of a fit-backward workspace after its 32-byte
// records (null when it does not fit: the launch then takes xcd_remap order).
const int32_t *order_after_fit_records(int ntiles, int split, const int2 *bins, void *ws,
                                       size_t ws_bytes, long long nisect, hipStream_t st) {
    const size_t off = (256 + (size_t)nisect * 32 + 255) & ~size_t(255);
    if (off + order_ws_bytes(ntiles) > ws_bytes) return nullptr;
    return launch_tile_order(ntiles, bins, (long long)ntiles * split,
                             resident_blocks(raster_bwd16_kernel<false, true, true>),
                             reinterpret_cast<char *>(ws) + off, st);
}

// Fit-mode (image + loss partials) forward at LDS batch depth `depth`; with
// queue.ctr set, a persistent grid over queue.ntiles tiles (the counters
// must be zero).
void launch_fit_forward(int depth, dim3 grid, hipStream_t st, int tiles_x, int img_w, int img_h,
                        const int32_t *gids, const int2 *bins, const GaussPack *pk,
                        float *out_img, const float *target, float *partial,
                        const BatchImage *tab, int nimg, const int32_t *cum = nullptr,
                        TileQueue queue = TileQueue{nullptr, 0, nullptr},
                        const int32_t *order = nullptr) {
#define LIG_FIT_FWD(D)                                                                        \
    if (queue.ctr) {                                                                          \
        auto kern = raster_fwd16_queue_kernel<false, true, D>;                                \
        hipLaunchKernelGGL(kern, dim3(persistent_grid(kern, queue.ntiles)), dim3(kThreads), 0, \
                           st, tiles_x, img_w, img_h, gids, bins, pk, out_img, nullptr,        \
                           nullptr, target, partial, tab, nimg, cum, queue);                  \
    } else {                                                                                  \
        hipLaunchKernelGGL((raster_fwd16_kernel<false, true, D>), grid, dim3(kThreads), 0, st, \
                           tiles_x, img_w, img_h, gids, bins, pk, out_img, nullptr, nullptr,   \
                           target, partial, tab, nimg, cum, order);                            \
    }
    if (depth == 256) {
        LIG_FIT_FWD(256)
    } else if (depth == 128) {
        LIG_FIT_FWD(128)
    } else {
        LIG_FIT_FWD(64)
    }
#undef LIG_FIT_FWD
}

// Resident blocks of the fit-mode forward at LDS batch depth `depth`.
long long fit_forward_resident(int depth) {
    if (depth == 256) return resident_blocks(raster_fwd16_kernel<false, true, 256>);
    if (depth == 128) return resident_blocks(raster_fwd16_kernel<false, true, 128>);
    return resident_blocks(raster_fwd16_kernel<false, true, 64>);
}
int rasterize_forward_impl(int n, int img_height, int img_width, int block_width,
                                     const int32_t *gaussian_ids_sorted,
                                     const int32_t *tile_bins, const float *xys,
                                     const float *conics, const float *colors,
                                     const float *opacities, const float *extents,
                                     const int32_t *cum_tiles_hit, void *packed, float *out_img,
                                     float *out_wsum, int32_t *final_idx, float *out_dx,
                                     float *out_dy, float *out_dxy, void *workspace,
                                     size_t workspace_bytes, void *stream) {
    if (n < 0 || img_height <= 0 || img_width <= 0 || block_width < 1 || block_width > 16)
        return LIG_ERR_INVALID_ARGUMENT;
    const bool ga = out_dx != nullptr;
    // image only (no wsum / final_idx): 16 x 16 tiles, plain forward; wsum
    // without the final index: 16 x 16, plain
    const bool image_only = !out_wsum && !final_idx;
    const bool no_final = out_wsum && !final_idx;
    if (!tile_bins || !out_img || (!out_wsum && final_idx) ||
        ((image_only || no_final) && (ga || block_width != 16)))
        return LIG_ERR_INVALID_ARGUMENT;
    if (ga && (!out_dy || !out_dxy)) return LIG_ERR_INVALID_ARGUMENT;
    if (n > 0 && (!gaussian_ids_sorted || !xys || !conics || !colors || !packed))
        return LIG_ERR_INVALID_ARGUMENT;
    // record addressing for the backward without emission ids: 16 x 16 tiles,
    // no opacities (so no zero-extent phantom slots)
    const bool slots_in_pack = extents != nullptr;
    if (slots_in_pack && (!cum_tiles_hit || block_width != 16 || opacities))
        return LIG_ERR_INVALID_ARGUMENT;
    hipStream_t st = as_stream(stream);
    GaussPack *pk = reinterpret_cast<GaussPack *>(packed);
    const int tiles_x = (img_width + block_width - 1) / block_width;
    const int tiles_y = (img_height + block_width - 1) / block_width;
    if (n > 0)
        hipLaunchKernelGGL(pack_kernel<>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, xys, conics,
                           colors, opacities, extents, cum_tiles_hit, tiles_x, tiles_y, nullptr,
                           0, pk);
    const dim3 grid((unsigned)(tiles_x * tiles_y));
    const int px = (block_width * block_width + LIG_WAVE - 1) / LIG_WAVE;
    const int2 *bins = reinterpret_cast<const int2 *>(tile_bins);
    // 16 x 16 plain forward: launch order in the workspace when one is given
    const int ntiles = tiles_x * tiles_y;
    const int32_t *order = nullptr;
    if (!ga && block_width == 16 && workspace && workspace_bytes >= order_ws_bytes(ntiles)) {
        const long long resident =
            image_only ? (opacities ? resident_blocks(raster_fwd16_kernel<true, true>)
                                    : resident_blocks(raster_fwd16_kernel<false, true>))
                       : (opacities ? resident_blocks(raster_fwd16_kernel<true, false>)
                                    : resident_blocks(raster_fwd16_kernel<false, false>));
        order = launch_tile_order(ntiles, bins, ntiles, resident, workspace, st);
    }
    if (ga)
        launch_fwd<true>(px, grid, st, tiles_x, block_width, img_width, img_height,
                         gaussian_ids_sorted, bins, pk, out_img, out_wsum, final_idx, out_dx,
                         out_dy, out_dxy, opacities ? 1 : 0);
    else if (image_only && opacities)
        hipLaunchKernelGGL((raster_fwd16_kernel<true, true>), grid, dim3(kThreads), 0, st,
                           tiles_x, img_width, img_height, gaussian_ids_sorted, bins, pk, out_img,
                           nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, order);
    else if (image_only)
        hipLaunchKernelGGL((raster_fwd16_kernel<false, true>), grid, dim3(kThreads), 0, st,
                           tiles_x, img_width, img_height, gaussian_ids_sorted, bins, pk, out_img,
                           nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, order);
    else if (no_final && opacities)
        hipLaunchKernelGGL((raster_fwd16_kernel<true, false, kThreads, false>), grid,
                           dim3(kThreads), 0, st, tiles_x, img_width, img_height,
                           gaussian_ids_sorted, bins, pk, out_img, out_wsum, nullptr, nullptr,
                           nullptr, nullptr, 0, nullptr, order);
    else if (no_final)
        hipLaunchKernelGGL((raster_fwd16_kernel<false, false, kThreads, false>), grid,
                           dim3(kThreads), 0, st, tiles_x, img_width, img_height,
                           gaussian_ids_sorted, bins, pk, out_img, out_wsum, nullptr, nullptr,
                           nullptr, nullptr, 0, nullptr, order);
    else if (block_width == 16 && opacities)
        hipLaunchKernelGGL((raster_fwd16_kernel<true, false>), grid, dim3(kThreads), 0, st,
                           tiles_x, img_width, img_height, gaussian_ids_sorted, bins, pk, out_img,
                           out_wsum, final_idx, nullptr, nullptr, nullptr, 0, nullptr, order);
    else if (block_width == 16)
        hipLaunchKernelGGL((raster_fwd16_kernel<false, false>), grid, dim3(kThreads), 0, st,
                           tiles_x, img_width, img_height, gaussian_ids_sorted, bins, pk, out_img,
                           out_wsum, final_idx, nullptr, nullptr, nullptr, 0, nullptr, order);
    else
        launch_fwd<false>(px, grid, st, tiles_x, block_width, img_width, img_height,
                          gaussian_ids_sorted, bins, pk, out_img, out_wsum, final_idx, nullptr,
                          nullptr, nullptr, opacities ? 1 : 0);
    return launch_status();
}

}  // namespace
}  // namespace lig

extern "C" size_t lig_tile_order_workspace_size(int num_tiles) {
    return num_tiles > 0 ? order_ws_bytes(num_tiles) : 0;
}

extern "C" int lig_rasterize_forward(int n, int img_height, int img_width, int block_width,
                                     const int32_t *gaussian_ids_sorted,
                                     const int32_t *tile_bins, const float *xys,
                                     const float *conics, const float *colors,
                                     const float *opacities, const float *extents,
                                     const int32_t *cum_tiles_hit, void *packed, float *out_img,
                                     float *out_wsum, int32_t *final_idx, float *out_dx,
                                     float *out_dy, float *out_dxy, void *stream) {
    return rasterize_forward_impl(n, img_height, img_width, block_width, gaussian_ids_sorted,
                                  tile_bins, xys, conics, colors, opacities, extents,
                                  cum_tiles_hit, packed, out_img, out_wsum, final_idx, out_dx,
                                  out_dy, out_dxy, nullptr, 0, stream);
}

extern "C" int lig_rasterize_forward_ws(int n, int img_height, int img_width, int block_width,
                                        const int32_t *gaussian_ids_sorted,
                                        const int32_t *tile_bins, const float *xys,
                                        const float *conics, const float *colors,
                                        const float *opacities, const float *extents,
                                        const int32_t *cum_tiles_hit, void *packed,
                                        float *out_img, float *out_wsum, int32_t *final_idx,
                                        float *out_dx, float *out_dy, float *out_dxy,
                                        void *workspace, size_t workspace_bytes, void *stream) {
    return rasterize_forward_impl(n, img_height, img_width, block_width, gaussian_ids_sorted,
                                  tile_bins, xys, conics, colors, opacities, extents,
                                  cum_tiles_hit, packed, out_img, out_wsum, final_idx, out_dx,
                                  out_dy, out_dxy, workspace, workspace_bytes, stream);
}

extern "C" size_t lig_fit_render_workspace_size(int img_height, int img_width) {
    if (img_height <= 0 || img_width <= 0) return 256;
    const size_t tiles = (size_t)((img_height + 15) / 16) * (size_t)((img_width + 15) / 16);
    const size_t blocks = (tiles + kSumChunk - 1) / kSumChunk;
    // [block sums, FP64][tile partials, FP32][launch order at a 256-byte boundary]
    return ((blocks * sizeof(double) + tiles * sizeof(float) + 255) & ~size_t(255)) +
           order_ws_bytes((int)tiles);
}

extern "C" int lig_fit_render_l2(int n, int img_height, int img_width,
                                 const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                                 const float *xys, const float *extents,
                                 const int32_t *cum_tiles_hit, const float *conics,
                                 const float *colors, void *packed, const float *target,
                                 float *out_img, float *loss, int batch_depth, void *workspace,
                                 size_t workspace_bytes, void *stream) {
    if (n <= 0 || img_height <= 0 || img_width <= 0 || !gaussian_ids_sorted || !tile_bins ||
        !xys || !extents || !cum_tiles_hit || !conics || !colors || !packed || !target ||
        !out_img || !loss || !depth_ok(batch_depth))
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_fit_render_workspace_size(img_height, img_width))
        return LIG_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    GaussPack *pk = reinterpret_cast<GaussPack *>(packed);
    const int tiles_x = (img_width + 15) / 16;
    const int tiles_y = (img_height + 15) / 16;
    const int tiles = tiles_x * tiles_y;
    hipLaunchKernelGGL(pack_kernel<>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, xys, conics,
                       colors, nullptr, extents, cum_tiles_hit, tiles_x, tiles_y, nullptr, 0, pk);
    const int2 *bins = reinterpret_cast<const int2 *>(tile_bins);
    const int nsum = (tiles + kSumChunk - 1) / kSumChunk;
    double *bsum = reinterpret_cast<double *>(workspace);
    float *partial = reinterpret_cast<float *>(bsum + nsum);
    const size_t ooff = ((size_t)nsum * sizeof(double) + (size_t)tiles * sizeof(float) + 255) &
                        ~size_t(255);
    const int32_t *order = launch_tile_order(tiles, bins, tiles, fit_forward_resident(batch_depth),
                                             reinterpret_cast<char *>(workspace) + ooff, st);
    launch_fit_forward(batch_depth, dim3((unsigned)tiles), st, tiles_x, img_width, img_height,
                       gaussian_ids_sorted, bins, pk, out_img, target, partial, nullptr, 0,
                       nullptr, TileQueue{nullptr, 0, nullptr}, order);
    const double numel = 3.0 * (double)img_height * (double)img_width;
    hipLaunchKernelGGL(sum_partials_kernel, dim3((unsigned)nsum), dim3(kSumThreads), 0, st, tiles,
                       (const float *)partial, bsum);
    hipLaunchKernelGGL(sum_blocks_kernel, dim3(1), dim3(kSumThreads), 0, st, nsum,
                       (const double *)bsum, 1.0 / numel, loss);
    return launch_status();
}

extern "C" size_t lig_fit_batched_workspace_size(int num_tiles) {
    // [tile partials, FP32][tile-queue counters at a 256-byte boundary][launch order]
    const size_t t = num_tiles > 0 ? (size_t)num_tiles : 0;
    return ((t * sizeof(float) + 255) & ~size_t(255)) + kQueueInts * sizeof(int) +
           order_ws_bytes((int)t);
}

namespace lig {
namespace {

int fit_render_batched(int n, const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                       const float *xys, const float *extents, const int32_t *cum_tiles_hit,
                       const float *conics, const float *colors, void *packed,
                       const float *target, const void *batch_table, int num_images,
                       int num_tiles, float *out_img, float *loss_per_image, float *loss,
                       int batch_depth, void *workspace, size_t workspace_bytes, void *stream,
                       bool persistent) {
    if (n <= 0 || num_images < 1 || num_tiles < 1 || !gaussian_ids_sorted || !tile_bins ||
        !xys || !extents || !cum_tiles_hit || !conics || !colors || !packed || !target ||
        !batch_table || !out_img || !loss_per_image || !loss || !depth_ok(batch_depth))
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_fit_batched_workspace_size(num_tiles))
        return LIG_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    const BatchImage *tab = reinterpret_cast<const BatchImage *>(batch_table);
    GaussPack *pk = reinterpret_cast<GaussPack *>(packed);
    hipLaunchKernelGGL(pack_kernel<>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, xys, conics,
                       colors, nullptr, extents, cum_tiles_hit, 0, 0, tab, num_images, pk);
    float *partial = reinterpret_cast<float *>(workspace);
    char *after_partials =
        reinterpret_cast<char *>(workspace) + (((size_t)num_tiles * sizeof(float) + 255) & ~size_t(255));
    TileQueue q{nullptr, num_tiles, nullptr};
    // launch order (heaviest tiles first where it matters): the grid's block
    // order, or the order in which the queue deals tiles to the XCDs
    const int32_t *order =
        launch_tile_order(num_tiles, reinterpret_cast<const int2 *>(tile_bins), num_tiles,
                          fit_forward_resident(batch_depth),
                          after_partials + kQueueInts * sizeof(int), st);
    if (persistent) {
        q.ctr = reinterpret_cast<int *>(after_partials);
        q.order = order;
        order = nullptr;
        hipLaunchKernelGGL(zero_f32_kernel, dim3(ceil_div(kQueueInts, 256)), dim3(256), 0, st,
                           reinterpret_cast<float *>(q.ctr), (size_t)kQueueInts);
    }
    launch_fit_forward(batch_depth, dim3((unsigned)num_tiles), st, 1, 1, 1, gaussian_ids_sorted,
                       reinterpret_cast<const int2 *>(tile_bins), pk, out_img, target, partial,
                       tab, num_images, cum_tiles_hit, q, order);
    hipLaunchKernelGGL(image_loss_kernel, dim3((unsigned)num_images), dim3(kImgLossThreads), 0, st,
                       tab, (const float *)partial, loss_per_image);
    hipLaunchKernelGGL(total_loss_kernel, dim3(1), dim3(kSumThreads), 0, st, num_images,
                       (const float *)loss_per_image, loss);
    return launch_status();
}

}  // namespace
}  // namespace lig

extern "C" int lig_fit_render_l2_batched(int n, const int32_t *gaussian_ids_sorted,
                                         const int32_t *tile_bins, const float *xys,
                                         const float *extents, const int32_t *cum_tiles_hit,
                                         const float *conics, const float *colors, void *packed,
                                         const float *target, const void *batch_table,
                                         int num_images, int num_tiles, float *out_img,
                                         float *loss_per_image, float *loss, int batch_depth,
                                         void *workspace, size_t workspace_bytes,
                                         void *stream) {
    return fit_render_batched(n, gaussian_ids_sorted, tile_bins, xys, extents, cum_tiles_hit,
                              conics, colors, packed, target, batch_table, num_images, num_tiles,
                              out_img, loss_per_image, loss, batch_depth, workspace,
                              workspace_bytes, stream, false);
}

extern "C" int lig_fit_render_l2_batched_persistent(
    int n, const int32_t *gaussian_ids_sorted, const int32_t *tile_bins, const float *xys,
    const float *extents, const int32_t *cum_tiles_hit, const float *conics, const float *colors,
    void *packed, const float *target, const void *batch_table, int num_images, int num_tiles,
    float *out_img, float *loss_per_image, float *loss, int batch_depth, void *workspace,
    size_t workspace_bytes, void *stream) {
    return fit_render_batched(n, gaussian_ids_sorted, tile_bins, xys, extents, cum_tiles_hit,
                              conics, colors, packed, target, batch_table, num_images, num_tiles,
                              out_img, loss_per_image, loss, batch_depth, workspace,
                              workspace_bytes, stream, true);
}

// Image-only render of a batch of views (crop serving): the fit-mode forward
// with no target and no loss partials, launched over the batch's tiles in the
// heaviest-first order when the launch is short or uneven.  Record addressing
// is packed as for the fit (extents and cum given) but not used.
extern "C" int lig_render_batched(int n, const int32_t *gaussian_ids_sorted,
                                  const int32_t *tile_bins, const float *xys,
                                  const float *extents, const int32_t *cum_tiles_hit,
                                  const float *conics, const float *colors, void *packed,
                                  const void *batch_table, int num_images, int num_tiles,
                                  float *out_img, int batch_depth, void *workspace,
                                  size_t workspace_bytes, void *stream) {
    if (n <= 0 || num_images < 1 || num_tiles < 1 || !gaussian_ids_sorted || !tile_bins ||
        !xys || !extents || !cum_tiles_hit || !conics || !colors || !packed || !batch_table ||
        !out_img || !depth_ok(batch_depth))
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_fit_batched_workspace_size(num_tiles))
        return LIG_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    const BatchImage *tab = reinterpret_cast<const BatchImage *>(batch_table);
    GaussPack *pk = reinterpret_cast<GaussPack *>(packed);
    hipLaunchKernelGGL(pack_kernel<true>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, xys,
                       conics, colors, nullptr, extents, cum_tiles_hit, 0, 0, tab, num_images,
                       pk);
    char *after_partials = reinterpret_cast<char *>(workspace) +
                           (((size_t)num_tiles * sizeof(float) + 255) & ~size_t(255));
    const int2 *bins = reinterpret_cast<const int2 *>(tile_bins);
    const int32_t *order =
        launch_tile_order(num_tiles, bins, num_tiles, fit_forward_resident(batch_depth),
                          after_partials + kQueueInts * sizeof(int), st);
    launch_fit_forward(batch_depth, dim3((unsigned)num_tiles), st, 1, 1, 1, gaussian_ids_sorted,
                       bins, pk, out_img, nullptr, nullptr, tab, num_images, cum_tiles_hit,
                       TileQueue{nullptr, num_tiles, nullptr}, order);
    return launch_status();
}

extern "C" int lig_set_backward_shares(int shares) {
    if (shares < 0 || shares > kMaxBwdSplit) return LIG_ERR_INVALID_ARGUMENT;
    return g_bwd_shares.exchange(shares);
}

extern "C" size_t lig_rasterize_backward_workspace_size(int num_intersects) {
    return num_intersects > 0 ? (size_t)num_intersects * sizeof(GradRec) + 256 : 256;
}

namespace lig {
namespace {

// bwd16_split for the API backward's instantiation.
int bwd16_split_for(int has_opacities, bool det, bool group, int ntiles, long long nisect) {
    if (group) {
        if (has_opacities)
            return det ? bwd16_split(raster_bwd16_kernel<true, true, false, true>, ntiles, nisect)
                       : bwd16_split(raster_bwd16_kernel<true, false, false, true>, ntiles,
                                     nisect);
        return det ? bwd16_split(raster_bwd16_kernel<false, true, false, true>, ntiles, nisect)
                   : bwd16_split(raster_bwd16_kernel<false, false, false, true>, ntiles, nisect);
    }
    if (has_opacities)
        return det ? bwd16_split(raster_bwd16_kernel<true, true, false>, ntiles, nisect)
                   : bwd16_split(raster_bwd16_kernel<true, false, false>, ntiles, nisect);
    return det ? bwd16_split(raster_bwd16_kernel<false, true, false>, ntiles, nisect)
               : bwd16_split(raster_bwd16_kernel<false, false, false>, ntiles, nisect);
}

// Launches the tile kernel of the backward: records (det) or atomics.
int launch_backward_tiles(int n, int img_height, int img_width, int block_width,
                          const int32_t *gids, const int32_t *tile_bins, const void *packed,
                          const int32_t *final_idx, const float *v_output, const float *v_wsum,
                          const float *v_output_dx, const float *v_output_dy,
                          const float *v_output_dxy, int has_opacities, const int32_t *slots,
                          GradRec *rec, float *v_xy, float *v_xy_abs, float *v_conic,
                          float *v_colors, float *v_opacity, long long num_intersects,
                          hipStream_t st, void *order_ws = nullptr, size_t order_ws_len = 0) {
    const bool ga = v_output_dx != nullptr;
    const bool det = rec != nullptr;
    // no final index (16 x 16, plain): row-grouped, ungated
    const bool group = final_idx == nullptr;
    const int tiles_x = (img_width + block_width - 1) / block_width;
    const int tiles_y = (img_height + block_width - 1) / block_width;
    const dim3 grid((unsigned)(tiles_x * tiles_y));
    const int split = block_width == 16 && !ga
                          ? bwd16_split_for(has_opacities, det, group, tiles_x * tiles_y,
                                            num_intersects)
                          : 1;
    const dim3 grid16((unsigned)(tiles_x * tiles_y * split));
    // 16 x 16 plain backward: launch order in the caller's spare workspace
    const int32_t *order = nullptr;
    if (block_width == 16 && !ga && order_ws &&
        order_ws_len >= order_ws_bytes(tiles_x * tiles_y)) {
        const long long resident =
            group ? (has_opacities
                         ? (det ? resident_blocks(raster_bwd16_kernel<true, true, false, true>)
                                : resident_blocks(raster_bwd16_kernel<true, false, false, true>))
                         : (det ? resident_blocks(raster_bwd16_kernel<false, true, false, true>)
                                : resident_blocks(raster_bwd16_kernel<false, false, false, true>)))
            : has_opacities ? (det ? resident_blocks(raster_bwd16_kernel<true, true, false>)
                                   : resident_blocks(raster_bwd16_kernel<true, false, false>))
                            : (det ? resident_blocks(raster_bwd16_kernel<false, true, false>)
                                   : resident_blocks(raster_bwd16_kernel<false, false, false>));
        order = launch_tile_order(tiles_x * tiles_y, reinterpret_cast<const int2 *>(tile_bins),
                                  (long long)tiles_x * tiles_y * split, resident, order_ws, st);
    }
    const int2 *bins = reinterpret_cast<const int2 *>(tile_bins);
    const GaussPack *pk = reinterpret_cast<const GaussPack *>(packed);
#define LIG_BWD_ARGS_PLAIN                                                                    \
    tiles_x, block_width, img_width, img_height, gids, slots, bins, pk, final_idx, v_output,    \
        v_wsum, rec, v_xy, v_xy_abs, v_conic, v_colors, v_opacity
#define LIG_BWD_ARGS_16                                                                      \
    tiles_x, img_width, img_height, gids, slots, bins, pk, final_idx, v_output, v_wsum,         \
        nullptr, nullptr, nullptr, 0.0f, rec, v_xy, v_xy_abs, v_conic, v_colors, v_opacity,      \
        nullptr, 0, split, order
#define LIG_BWD_ARGS_GA                                                                      \
    tiles_x, block_width, img_width, img_height, gids, slots, bins, pk, final_idx, v_output,    \
        v_wsum, v_output_dx, v_output_dy, v_output_dxy, has_opacities, rec, v_xy, v_xy_abs,    \
        v_conic, v_colors, v_opacity
    if (ga) {
        if (det)
            hipLaunchKernelGGL(raster_bwd_ga_kernel<true>, grid, dim3(kThreads), 0, st,
                               LIG_BWD_ARGS_GA);
        else
            hipLaunchKernelGGL(raster_bwd_ga_kernel<false>, grid, dim3(kThreads), 0, st,
                               LIG_BWD_ARGS_GA);
    } else if (group) {
        // 16 x 16 without a final index: row-grouped, ungated; without a wsum
        // gradient its terms are left out
        const int sel = (v_wsum ? 4 : 0) | (has_opacities ? 2 : 0) | (det ? 1 : 0);
        switch (sel) {
            case 0: hipLaunchKernelGGL((raster_bwd16_kernel<false, false, false, true, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 1: hipLaunchKernelGGL((raster_bwd16_kernel<false, true, false, true, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 2: hipLaunchKernelGGL((raster_bwd16_kernel<true, false, false, true, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 3: hipLaunchKernelGGL((raster_bwd16_kernel<true, true, false, true, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 4: hipLaunchKernelGGL((raster_bwd16_kernel<false, false, false, true>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 5: hipLaunchKernelGGL((raster_bwd16_kernel<false, true, false, true>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 6: hipLaunchKernelGGL((raster_bwd16_kernel<true, false, false, true>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            default: hipLaunchKernelGGL((raster_bwd16_kernel<true, true, false, true>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
        }
    } else {
        // plain: (opacities, deterministic, 16x16) -> one of eight instantiations
        const int sel = (has_opacities ? 4 : 0) | (det ? 2 : 0) | (block_width == 16 ? 1 : 0);
        switch (sel) {
            case 0: hipLaunchKernelGGL((raster_bwd_plain_kernel<false, false>), grid, dim3(kThreads), 0, st, LIG_BWD_ARGS_PLAIN); break;
            case 1: hipLaunchKernelGGL((raster_bwd16_kernel<false, false, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 2: hipLaunchKernelGGL((raster_bwd_plain_kernel<false, true>), grid, dim3(kThreads), 0, st, LIG_BWD_ARGS_PLAIN); break;
            case 3: hipLaunchKernelGGL((raster_bwd16_kernel<false, true, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 4: hipLaunchKernelGGL((raster_bwd_plain_kernel<true, false>), grid, dim3(kThreads), 0, st, LIG_BWD_ARGS_PLAIN); break;
            case 5: hipLaunchKernelGGL((raster_bwd16_kernel<true, false, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
            case 6: hipLaunchKernelGGL((raster_bwd_plain_kernel<true, true>), grid, dim3(kThreads), 0, st, LIG_BWD_ARGS_PLAIN); break;
            default: hipLaunchKernelGGL((raster_bwd16_kernel<true, true, false>), grid16, dim3(kThreads), 0, st, LIG_BWD_ARGS_16); break;
        }
    }
#undef LIG_BWD_ARGS_PLAIN
#undef LIG_BWD_ARGS_16
#undef LIG_BWD_ARGS_GA
    return launch_status();
}

bool bwd_args_ok(int n, int img_height, int img_width, int block_width, const int32_t *gids,
                 const int32_t *tile_bins, const void *packed, const int32_t *final_idx,
                 const float *v_output, const float *v_output_dx, const float *v_output_dy,
                 const float *v_output_dxy) {
    if (n < 0 || img_height <= 0 || img_width <= 0 || block_width < 1 || block_width > 16)
        return false;
    if (v_output_dx && (!v_output_dy || !v_output_dxy)) return false;
    if (n > 0 && (!gids || !tile_bins || !packed || !v_output)) return false;
    // no final index: 16 x 16 plain only
    if (n > 0 && !final_idx && (block_width != 16 || v_output_dx)) return false;
    return true;
}

}  // namespace
}  // namespace lig

extern "C" int lig_rasterize_backward_records(int n, int img_height, int img_width,
                                              int block_width, const int32_t *gaussian_ids_sorted,
                                              const int32_t *tile_bins, const void *packed,
                                              const int32_t *final_idx, const float *v_output,
                                              const float *v_wsum, const float *v_output_dx,
                                              const float *v_output_dy,
                                              const float *v_output_dxy, int has_opacities,
                                              const int32_t *emission_ids_sorted,
                                              int num_intersects, void *workspace,
                                              size_t workspace_bytes, void *stream) {
    if (!bwd_args_ok(n, img_height, img_width, block_width, gaussian_ids_sorted, tile_bins,
                     packed, final_idx, v_output, v_output_dx, v_output_dy, v_output_dxy) ||
        num_intersects < 0 ||
        (n > 0 && !emission_ids_sorted && (block_width != 16 || has_opacities)))
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_rasterize_backward_workspace_size(num_intersects))
        return LIG_ERR_WORKSPACE;
    if (n == 0) return LIG_OK;
    GradRec *rec = reinterpret_cast<GradRec *>(reinterpret_cast<char *>(workspace) + 256);
    // room after the records (a workspace sized with lig_tile_order_workspace_size
    // on top) holds the launch order
    const size_t off = (256 + (size_t)num_intersects * sizeof(GradRec) + 255) & ~size_t(255);
    void *ows = workspace_bytes > off ? reinterpret_cast<char *>(workspace) + off : nullptr;
    return launch_backward_tiles(n, img_height, img_width, block_width, gaussian_ids_sorted,
                                 tile_bins, packed, final_idx, v_output, v_wsum, v_output_dx,
                                 v_output_dy, v_output_dxy, has_opacities, emission_ids_sorted,
                                 rec, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 num_intersects, as_stream(stream), ows,
                                 ows ? workspace_bytes - off : 0);
}

extern "C" int lig_fit_backward_records_l2(int n, int img_height, int img_width,
                                           const int32_t *gaussian_ids_sorted,
                                           const int32_t *tile_bins, const void *packed,
                                           const float *img, const float *target,
                                           const float *grad_loss, int num_intersects,
                                           void *workspace, size_t workspace_bytes,
                                           void *stream) {
    if (n <= 0 || img_height <= 0 || img_width <= 0 || num_intersects < 0 ||
        !gaussian_ids_sorted || !tile_bins || !packed || !img || !target || !grad_loss)
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_rasterize_backward_workspace_size(num_intersects))
        return LIG_ERR_WORKSPACE;
    GradRec *rec = reinterpret_cast<GradRec *>(reinterpret_cast<char *>(workspace) + 256);
    const int tiles_x = (img_width + 15) / 16;
    const int ntiles = tiles_x * ((img_height + 15) / 16);
    const int split = bwd16_split(raster_bwd16_kernel<false, true, true>, ntiles, num_intersects);
    const int2 *bins = reinterpret_cast<const int2 *>(tile_bins);
    const GaussPack *pk = reinterpret_cast<const GaussPack *>(packed);
    const float norm = (float)(2.0 / (3.0 * (double)img_height * (double)img_width));
    hipStream_t st = as_stream(stream);
    const int32_t *order = order_after_fit_records(ntiles, split, bins, workspace,
                                                   workspace_bytes, num_intersects, st);
    hipLaunchKernelGGL((raster_bwd16_kernel<false, true, true>), dim3((unsigned)(ntiles * split)),
                       dim3(kThreads), 0, st, tiles_x, img_width, img_height,
                       gaussian_ids_sorted, nullptr, bins, pk, nullptr, nullptr, nullptr, img,
                       target, grad_loss, norm, rec, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr, 0, split, order);
    return launch_status();
}

namespace lig {
namespace {

int fit_backward_batched(int n, const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                         const void *packed, const float *img, const float *target,
                         const float *grad_loss, const void *batch_table, int num_images,
                         int num_tiles, int num_intersects, void *workspace,
                         size_t workspace_bytes, void *stream, bool persistent) {
    if (n <= 0 || num_images < 1 || num_tiles < 1 || num_intersects < 0 ||
        !gaussian_ids_sorted || !tile_bins || !packed || !img || !target || !grad_loss ||
        !batch_table)
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_rasterize_backward_workspace_size(num_intersects))
        return LIG_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    GradRec *rec = reinterpret_cast<GradRec *>(reinterpret_cast<char *>(workspace) + 256);
    const int2 *bins = reinterpret_cast<const int2 *>(tile_bins);
    const GaussPack *pk = reinterpret_cast<const GaussPack *>(packed);
    const BatchImage *tab = reinterpret_cast<const BatchImage *>(batch_table);
    // fit records (32 bytes) start 256 bytes in; the workspace is sized for
    // 48-byte records, and the tile-queue counters go after the fit records
    // when that slack holds them (else the one-tile-per-block launch runs)
    const size_t qoff = (256 + (size_t)num_intersects * sizeof(FitRec) + 255) & ~size_t(255);
    if (persistent && qoff + kQueueInts * sizeof(int) <= workspace_bytes) {
        char *ctr = reinterpret_cast<char *>(workspace) + qoff;
        TileQueue q{reinterpret_cast<int *>(ctr), num_tiles, nullptr};
        // the queue's dealing order after the counters, when the slack holds it
        const size_t ooff = qoff + kQueueInts * sizeof(int);
        if (ooff + order_ws_bytes(num_tiles) <= workspace_bytes)
            q.order = launch_tile_order(num_tiles, bins, num_tiles,
                                        resident_blocks(raster_bwd16_fit_queue_kernel),
                                        ctr + kQueueInts * sizeof(int), st);
        hipLaunchKernelGGL(zero_f32_kernel, dim3(ceil_div(kQueueInts, 256)), dim3(256), 0, st,
                           reinterpret_cast<float *>(q.ctr), (size_t)kQueueInts);
        hipLaunchKernelGGL(raster_bwd16_fit_queue_kernel,
                           dim3(persistent_grid(raster_bwd16_fit_queue_kernel, num_tiles)),
                           dim3(kThreads), 0, st, gaussian_ids_sorted, bins, pk, img, target,
                           grad_loss, rec, tab, num_images, q);
    } else {
        const int split =
            bwd16_split(raster_bwd16_kernel<false, true, true>, num_tiles, num_intersects);
        const int32_t *order = order_after_fit_records(num_tiles, split, bins, workspace,
                                                       workspace_bytes, num_intersects, st);
        hipLaunchKernelGGL((raster_bwd16_kernel<false, true, true>),
                           dim3((unsigned)(num_tiles * split)), dim3(kThreads), 0, st, 1, 1, 1,
                           gaussian_ids_sorted, nullptr, bins, pk, nullptr, nullptr, nullptr, img,
                           target, grad_loss, 0.0f, rec, nullptr, nullptr, nullptr, nullptr,
                           nullptr, tab, num_images, split, order);
    }
    return launch_status();
}

}  // namespace
}  // namespace lig

extern "C" int lig_fit_backward_records_l2_batched(
    int n, const int32_t *gaussian_ids_sorted, const int32_t *tile_bins, const void *packed,
    const float *img, const float *target, const float *grad_loss, const void *batch_table,
    int num_images, int num_tiles, int num_intersects, void *workspace, size_t workspace_bytes,
    void *stream) {
    return fit_backward_batched(n, gaussian_ids_sorted, tile_bins, packed, img, target,
                                grad_loss, batch_table, num_images, num_tiles, num_intersects,
                                workspace, workspace_bytes, stream, false);
}

extern "C" int lig_fit_backward_records_l2_batched_persistent(
    int n, const int32_t *gaussian_ids_sorted, const int32_t *tile_bins, const void *packed,
    const float *img, const float *target, const float *grad_loss, const void *batch_table,
    int num_images, int num_tiles, int num_intersects, void *workspace, size_t workspace_bytes,
    void *stream) {
    return fit_backward_batched(n, gaussian_ids_sorted, tile_bins, packed, img, target,
                                grad_loss, batch_table, num_images, num_tiles, num_intersects,
                                workspace, workspace_bytes, stream, true);
}

extern "C" int lig_fit_reduce_records(int n, int num_intersects, const int32_t *cum_tiles_hit,
                                      const void *workspace, size_t workspace_bytes,
                                      float *v_xy, float *v_conic, float *v_colors,
                                      void *stream) {
    if (n <= 0 || num_intersects < 0 || !cum_tiles_hit || !v_xy || !v_conic || !v_colors)
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_rasterize_backward_workspace_size(num_intersects))
        return LIG_ERR_WORKSPACE;
    const FitRec *rec = reinterpret_cast<const FitRec *>(
        reinterpret_cast<const char *>(workspace) + 256);
    hipLaunchKernelGGL(reduce_fit_records_kernel,
                       dim3(ceil_div((size_t)n * kSegLanes, kSegThreads)), dim3(kSegThreads), 0,
                       as_stream(stream), n, num_intersects, cum_tiles_hit, rec, v_xy, v_conic,
                       v_colors);
    return launch_status();
}

extern "C" int lig_reduce_gradient_records(int n, int num_intersects,
                                           const int32_t *cum_tiles_hit, int has_opacities,
                                           const void *workspace, size_t workspace_bytes,
                                           float *v_xy, float *v_xy_abs, float *v_conic,
                                           float *v_colors, float *v_opacity, void *stream) {
    if (n < 0 || num_intersects < 0) return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    if (!cum_tiles_hit || !v_xy || !v_xy_abs || !v_conic || !v_colors ||
        (has_opacities && !v_opacity))
        return LIG_ERR_INVALID_ARGUMENT;
    if (!workspace || workspace_bytes < lig_rasterize_backward_workspace_size(num_intersects))
        return LIG_ERR_WORKSPACE;
    hipStream_t st = as_stream(stream);
    // the first 256 bytes of the workspace hold the zero-extent-slot accumulator
    float *phantom = reinterpret_cast<float *>(const_cast<void *>(workspace));
    const GradRec *rec =
        reinterpret_cast<const GradRec *>(reinterpret_cast<const char *>(workspace) + 256);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, st, phantom, (size_t)16);
    hipLaunchKernelGGL(reduce_records_kernel, dim3(ceil_div((size_t)n * kSegLanes, kSegThreads)),
                       dim3(kSegThreads), 0, st, n, num_intersects, cum_tiles_hit, rec, has_opacities, v_xy,
                       v_xy_abs, v_conic, v_colors, v_opacity, phantom);
    hipLaunchKernelGGL(add_phantom_kernel, dim3(1), dim3(64), 0, st, (const float *)phantom,
                       has_opacities, v_xy, v_xy_abs, v_conic, v_colors, v_opacity);
    return launch_status();
}

extern "C" int lig_rasterize_backward(int n, int img_height, int img_width, int block_width,
                                      const int32_t *gaussian_ids_sorted,
                                      const int32_t *tile_bins, const void *packed,
                                      const int32_t *final_idx, const float *v_output,
                                      const float *v_wsum, const float *v_output_dx,
                                      const float *v_output_dy, const float *v_output_dxy,
                                      int has_opacities, const int32_t *emission_ids_sorted,
                                      const int32_t *cum_tiles_hit, int num_intersects,
                                      void *workspace, size_t workspace_bytes, float *v_xy,
                                      float *v_xy_abs, float *v_conic, float *v_colors,
                                      float *v_opacity, void *stream) {
    if (!bwd_args_ok(n, img_height, img_width, block_width, gaussian_ids_sorted, tile_bins,
                     packed, final_idx, v_output, v_output_dx, v_output_dy, v_output_dxy))
        return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    if (!v_xy || !v_xy_abs || !v_conic || !v_colors || (has_opacities && !v_opacity))
        return LIG_ERR_INVALID_ARGUMENT;
    if (emission_ids_sorted) {
        int rc = lig_rasterize_backward_records(
            n, img_height, img_width, block_width, gaussian_ids_sorted, tile_bins, packed,
            final_idx, v_output, v_wsum, v_output_dx, v_output_dy, v_output_dxy, has_opacities,
            emission_ids_sorted, num_intersects, workspace, workspace_bytes, stream);
        if (rc != LIG_OK) return rc;
        return lig_reduce_gradient_records(n, num_intersects, cum_tiles_hit, has_opacities,
                                           workspace, workspace_bytes, v_xy, v_xy_abs, v_conic,
                                           v_colors, v_opacity, stream);
    }
    hipStream_t st = as_stream(stream);
    const size_t nn = (size_t)n;
    hipLaunchKernelGGL(zero_f32_kernel, dim3(ceil_div(2 * nn, 256)), dim3(256), 0, st, v_xy, 2 * nn);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(ceil_div(2 * nn, 256)), dim3(256), 0, st, v_xy_abs,
                       2 * nn);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(ceil_div(3 * nn, 256)), dim3(256), 0, st, v_conic,
                       3 * nn);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(ceil_div(3 * nn, 256)), dim3(256), 0, st, v_colors,
                       3 * nn);
    if (has_opacities)
        hipLaunchKernelGGL(zero_f32_kernel, dim3(ceil_div(nn, 256)), dim3(256), 0, st, v_opacity,
                           nn);
    return launch_backward_tiles(n, img_height, img_width, block_width, gaussian_ids_sorted,
                                 tile_bins, packed, final_idx, v_output, v_wsum, v_output_dx,
                                 v_output_dy, v_output_dxy, has_opacities, nullptr, nullptr, v_xy,
                                 v_xy_abs, v_conic, v_colors, v_opacity, num_intersects, st);
}

extern "C" int lig_count_pairs(int n, int img_height, int img_width, const float *xys,
                               const float *extents, const float *conics,
                               const float *opacities, const int32_t *num_tiles_hit,
                               unsigned long long *counts, void *stream) {
    if (n < 0 || img_height <= 0 || img_width <= 0 || !counts) return LIG_ERR_INVALID_ARGUMENT;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, counts, 2);
    if (n == 0) return launch_status();
    if (!xys || !extents || !conics || !num_tiles_hit) return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(count_pairs_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n,
                       img_height, img_width, xys, extents, conics, opacities, num_tiles_hit,
                       counts);
    return launch_status();
}
