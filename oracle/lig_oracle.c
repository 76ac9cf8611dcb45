/*
 * CPU oracle for the gsplat2d 2-D Gaussian-splatting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (lig_amd/) links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do, and only as the checker.
 *
 * Precision contract (each function says which it follows):
 *   * Projection (K1/K2/K3/K10/K11) is restated in IEEE FP32 with the operation
 *     order of the reference, no FP contraction, correctly rounded division and
 *     square root, and log evaluated in FP64 then rounded to FP32.  The GPU
 *     kernels follow the same contract, so projection parity is bit-exact.
 *   * Binning (K4/K5 + the sort) is integer work and bit-exact: a stable sort
 *     of the gid-major emission order by signed 64-bit key.
 *   * Rasterisation forward/backward (K6..K9) is evaluated in FP64.  Pairs that
 *     sit within a relative 1e-5 of the alpha >= 1/255 cut, or whose quadratic
 *     form is within rounding of zero, are reported as ambiguous so that tests
 *     can exclude them (FP32 rounding on the device may decide them either way).
 *     A second mode ("ref32") evaluates every pair in FP32 with the reference's
 *     expression order and libm expf, accumulating in FP64: that is the
 *     restatement of the reference gradient.
 *   * Upscaler (K12/K13) computes the source coordinate in FP32 exactly as the
 *     reference does, then the Hermite weights and sums in FP64.
 *
 * Reference files followed (paths relative to the reference repository):
 *   gsplat2d/gsplat2d/cuda/csrc/helpers.cuh:12-93
 *   gsplat2d/gsplat2d/cuda/csrc/forward.cu:13-279
 *   gsplat2d/gsplat2d/cuda/csrc/backward.cu:8-247
 *   gsplat2d/gsplat2d/cuda/csrc/gradient_aware_rasterizer.cu:118-495
 *   gsplat2d/gsplat2d/cuda/csrc/gradient_aware_upscaler.cu:137-371
 *   gsplat2d/gsplat2d/cuda/csrc/bindings.cu:19-66 (compute_cov2d_bounds_kernel)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* FP contraction is disabled by the build (-ffp-contract=off). */

/* ------------------------------------------------------------------------- */
/* FP32 helpers with a fixed, compiler-independent meaning.                    */
/* ------------------------------------------------------------------------- */

/* float -> int32 truncation that saturates and maps NaN to 0 (the behaviour of
 * the GPU conversion instruction the reference's (int) casts compile to). */
static int f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int)x;
}

static float div_rn(float a, float b) { return (float)((double)a / (double)b); }
static float sqrt_rn(float a) { return (float)sqrt((double)a); }
static float log_rd(float a) { return (float)log((double)a); }

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* helpers.cuh:12-26 get_bbox: inclusive min, exclusive max, clamped to bounds */
static void get_bbox(float cx, float cy, float dx, float dy, int bx, int by,
                     int *x0, int *y0, int *x1, int *y1) {
    *x0 = imin(imax(0, f2i_sat(cx - dx)), bx);
    *x1 = imin(imax(0, f2i_sat(cx + dx + 1.0f)), bx);
    *y0 = imin(imax(0, f2i_sat(cy - dy)), by);
    *y1 = imin(imax(0, f2i_sat(cy + dy + 1.0f)), by);
}

/* helpers.cuh:28-45 get_tile_bbox */
static void get_tile_bbox(float px, float py, float ex, float ey, int tx, int ty,
                          int bw, int *x0, int *y0, int *x1, int *y1) {
    float fb = (float)bw;
    get_bbox(div_rn(px, fb), div_rn(py, fb), div_rn(ex, fb), div_rn(ey, fb),
             tx, ty, x0, y0, x1, y1);
}

/* helpers.cuh:47-81 compute_cov2d_bounds (SnugBox).  Returns 0 when det == 0. */
static int cov2d_bounds(const float *cov, float opacity, float *conic, float *extent) {
    float det = cov[0] * cov[2] - cov[1] * cov[1];
    if (det == 0.0f) return 0;
    float inv_det = div_rn(1.0f, det);
    float a = cov[2] * inv_det;
    float b = -cov[1] * inv_det;
    float c = cov[0] * inv_det;
    conic[0] = a; conic[1] = b; conic[2] = c;
    float t = 2.0f * log_rd(255.0f * opacity);
    if (t <= 0.0f) {
        extent[0] = 0.0f; extent[1] = 0.0f;
        return 1;
    }
    extent[0] = ceilf(sqrt_rn(div_rn(t * c, inv_det)));
    extent[1] = ceilf(sqrt_rn(div_rn(t * a, inv_det)));
    return 1;
}

/* bindings.cu:19-39 compute_cov2d_bounds_kernel (outputs written even when
 * det == 0, with whatever the helper left: the caller zero-fills). */
void orc_compute_cov2d_bounds(int n, const float *cov, const float *opac,
                              float *conics, float *extents) {
    for (int i = 0; i < n; ++i) {
        float conic[3] = {0, 0, 0}, ext[2] = {0, 0};
        float o = opac ? opac[i] : 1.0f;
        /* the reference kernel stores uninitialised locals when det == 0;
         * the build defines them as zero */
        cov2d_bounds(cov + 3 * i, o, conic, ext);
        conics[3 * i + 0] = conic[0];
        conics[3 * i + 1] = conic[1];
        conics[3 * i + 2] = conic[2];
        extents[2 * i + 0] = ext[0];
        extents[2 * i + 1] = ext[1];
    }
}

static void project_one(int i, const float *cov, const float *means, const float *opac,
                        int tx, int ty, int bw, float *xys, float *extents, float *conics,
                        int *ntiles) {
    extents[2 * i] = 0.0f; extents[2 * i + 1] = 0.0f;
    ntiles[i] = 0;
    float conic[3], ext[2];
    float o = opac ? opac[i] : 1.0f;
    if (!cov2d_bounds(cov, o, conic, ext)) return;
    conics[3 * i + 0] = conic[0];
    conics[3 * i + 1] = conic[1];
    conics[3 * i + 2] = conic[2];
    float cx = means[2 * i], cy = means[2 * i + 1];
    int x0, y0, x1, y1;
    get_tile_bbox(cx, cy, ext[0], ext[1], tx, ty, bw, &x0, &y0, &x1, &y1);
    int area = (int)((unsigned)(x1 - x0) * (unsigned)(y1 - y0));
    if (area <= 0) return;
    ntiles[i] = area;
    extents[2 * i] = ext[0]; extents[2 * i + 1] = ext[1];
    xys[2 * i] = cx; xys[2 * i + 1] = cy;
}

/* forward.cu:13-53 project_gaussians_forward_kernel.  Outputs must be
 * zero-filled by the caller (bindings.cu:89-97). */
void orc_project_fwd(int n, const float *cov2d, const float *means, const float *opac,
                     int img_h, int img_w, int bw, float *xys, float *extents,
                     float *conics, int *ntiles) {
    int tx = (img_w + bw - 1) / bw, ty = (img_h + bw - 1) / bw;
    for (int i = 0; i < n; ++i)
        project_one(i, cov2d + 3 * i, means, opac, tx, ty, bw, xys, extents, conics, ntiles);
}

/* forward.cu:55-105 project_gaussians_forward_kernel_cholesky */
void orc_project_fwd_chol(int n, const float *chol, const float *means, const float *opac,
                          int img_h, int img_w, int bw, float *xys, float *extents,
                          float *conics, int *ntiles) {
    int tx = (img_w + bw - 1) / bw, ty = (img_h + bw - 1) / bw;
    for (int i = 0; i < n; ++i) {
        float l11 = chol[3 * i], l21 = chol[3 * i + 1], l22 = chol[3 * i + 2];
        float cov[3];
        cov[0] = l11 * l11;
        cov[1] = l11 * l21;
        cov[2] = l21 * l21 + l22 * l22;
        project_one(i, cov, means, opac, tx, ty, bw, xys, extents, conics, ntiles);
    }
}

/* helpers.cuh:84-93 cov2d_to_conic_vjp: v_Sigma = -X G X with glm's
 * column-major 2x2 products evaluated left to right. */
static void conic_vjp(const float *conic, const float *v_conic, float *v_cov) {
    float X[2][2] = {{conic[0], conic[1]}, {conic[1], conic[2]}};
    float hb = v_conic[1] / 2.0f;
    float G[2][2] = {{v_conic[0], hb}, {hb, v_conic[2]}};
    float N[2][2] = {{-X[0][0], -X[0][1]}, {-X[1][0], -X[1][1]}};
    float P[2][2], V[2][2];
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 2; ++r)
            P[c][r] = N[0][r] * G[c][0] + N[1][r] * G[c][1];
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 2; ++r)
            V[c][r] = P[0][r] * X[c][0] + P[1][r] * X[c][1];
    v_cov[0] = V[0][0];
    v_cov[1] = V[1][0] + V[0][1];
    v_cov[2] = V[1][1];
}

/* backward.cu:185-204 project_gaussians_backward_kernel (zero-filled outputs) */
void orc_project_bwd(int n, const float *extents, const float *conics, const float *v_xy,
                     const float *v_conic, float *v_cov, float *v_mean) {
    for (int i = 0; i < n; ++i) {
        if (extents[2 * i] <= 0.0f || extents[2 * i + 1] <= 0.0f) continue;
        v_mean[2 * i] = v_xy[2 * i];
        v_mean[2 * i + 1] = v_xy[2 * i + 1];
        conic_vjp(conics + 3 * i, v_conic + 3 * i, v_cov + 3 * i);
    }
}

/* backward.cu:206-247 project_gaussians_backward_kernel_cholesky */
void orc_project_bwd_chol(int n, const float *extents, const float *chol, const float *conics,
                          const float *v_xy, const float *v_conic, float *v_chol,
                          float *v_mean) {
    for (int i = 0; i < n; ++i) {
        if (extents[2 * i] <= 0.0f || extents[2 * i + 1] <= 0.0f) continue;
        v_mean[2 * i] = v_xy[2 * i];
        v_mean[2 * i + 1] = v_xy[2 * i + 1];
        float vc[3];
        conic_vjp(conics + 3 * i, v_conic + 3 * i, vc);
        float l11 = chol[3 * i], l21 = chol[3 * i + 1], l22 = chol[3 * i + 2];
        float v11 = 2.0f * l11 * vc[0] + l21 * vc[1];
        float v21 = l11 * vc[1] + 2.0f * l21 * vc[2];
        float v22 = 2.0f * l22 * vc[2];
        int fin = isfinite(v11) && isfinite(v21) && isfinite(v22);
        v_chol[3 * i + 0] = fin ? v11 : 0.0f;
        v_chol[3 * i + 1] = fin ? v21 : 0.0f;
        v_chol[3 * i + 2] = fin ? v22 : 0.0f;
    }
}

/* ------------------------------------------------------------------------- */
/* Binning                                                                     */
/* ------------------------------------------------------------------------- */

/* utils.py:61-67: inclusive int32 cumulative sum */
void orc_cumsum_i32(int n, const int *x, int *out) {
    int s = 0;
    for (int i = 0; i < n; ++i) { s += x[i]; out[i] = s; }
}

/* forward.cu:109-143 map_gaussian_to_intersects.  Slots of Gaussians with a
 * zero extent are left as the caller's zero fill (bindings.cu:260-263). */
void orc_map_intersects(int n, const float *xys, const float *depths, const float *extents,
                        const int *cum, int tx, int ty, int bw, int64_t *isect_ids,
                        int32_t *gids) {
    for (int i = 0; i < n; ++i) {
        if (extents[2 * i] <= 0.0f || extents[2 * i + 1] <= 0.0f) continue;
        int x0, y0, x1, y1;
        get_tile_bbox(xys[2 * i], xys[2 * i + 1], extents[2 * i], extents[2 * i + 1],
                      tx, ty, bw, &x0, &y0, &x1, &y1);
        int cur = i == 0 ? 0 : cum[i - 1];
        int32_t dbits;
        memcpy(&dbits, &depths[i], 4);
        int64_t depth_id = (int64_t)dbits;  /* sign-extended, as the reference */
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) {
                int64_t tile = (int64_t)y * tx + x;
                isect_ids[cur] = (int64_t)((uint64_t)tile << 32) | depth_id;
                gids[cur] = i;
                ++cur;
            }
    }
}

typedef struct { int64_t key; int32_t pos; int32_t val; } kv_t;
static int kv_cmp(const void *a, const void *b) {
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* utils.py:97-98 torch.sort + gather, specified as a STABLE sort by signed key */
void orc_sort_pairs(int m, const int64_t *keys, const int32_t *vals, int64_t *keys_out,
                    int32_t *vals_out) {
    kv_t *t = (kv_t *)malloc(sizeof(kv_t) * (size_t)(m > 0 ? m : 1));
    for (int i = 0; i < m; ++i) { t[i].key = keys[i]; t[i].pos = i; t[i].val = vals[i]; }
    qsort(t, (size_t)m, sizeof(kv_t), kv_cmp);
    for (int i = 0; i < m; ++i) { keys_out[i] = t[i].key; vals_out[i] = t[i].val; }
    free(t);
}

/* forward.cu:148-170 get_tile_bin_edges on top of a zero-filled bins array */
void orc_tile_bin_edges(int m, const int64_t *sorted, int *bins) {
    for (int i = 0; i < m; ++i) {
        int cur = (int)(sorted[i] >> 32);
        if (i == 0) bins[2 * cur] = 0;
        if (i == m - 1) bins[2 * cur + 1] = m;
        if (i == 0) continue;
        int prev = (int)(sorted[i - 1] >> 32);
        if (prev != cur) { bins[2 * prev + 1] = i; bins[2 * cur] = i; }
    }
}

/* ------------------------------------------------------------------------- */
/* Rasterisation                                                               */
/* ------------------------------------------------------------------------- */

#define AMBIG_REL 1e-5

typedef struct {
    int img_h, img_w, bw, tx, ty;
    const int32_t *gids;   /* sorted gaussian ids */
    const int32_t *bins;   /* [T,2] */
    const float *xys, *conics, *colors, *opac;
} scene_t;

/* One pair evaluated in FP64.  d = mu - p (forward.cu:252-258). */
typedef struct { double sigma, vis, alpha, dx, dy; int valid, ambiguous; } pair64_t;

static pair64_t eval_pair64(const scene_t *s, int g, double px, double py, double clamp) {
    pair64_t r;
    double a = s->conics[3 * g], b = s->conics[3 * g + 1], c = s->conics[3 * g + 2];
    double o = s->opac ? (double)s->opac[g] : 1.0;
    r.dx = (double)s->xys[2 * g] - px;
    r.dy = (double)s->xys[2 * g + 1] - py;
    double qa = a * r.dx * r.dx, qc = c * r.dy * r.dy, qb = b * r.dx * r.dy;
    r.sigma = 0.5 * (qa + qc) + qb;
    r.vis = exp(-r.sigma);
    double raw = o * r.vis;
    r.alpha = raw < clamp ? raw : clamp;
    r.valid = !(r.sigma < 0.0 || r.alpha < 1.0 / 255.0);
    double mag = 0.5 * (fabs(qa) + fabs(qc)) + fabs(qb);
    r.ambiguous = fabs(r.alpha * 255.0 - 1.0) < AMBIG_REL ||
                  fabs(r.sigma) <= AMBIG_REL * mag + 1e-30;
    return r;
}

/* One pair in FP32 with the reference's expression order and libm expf
 * (backward.cu:105-111, gradient_aware_rasterizer.cu:204-216). */
typedef struct { float sigma, vis, alpha, dx, dy; int valid; } pair32_t;

static pair32_t eval_pair32(const scene_t *s, int g, float px, float py, float clamp,
                            int d_is_p_minus_mu) {
    pair32_t r;
    float a = s->conics[3 * g], b = s->conics[3 * g + 1], c = s->conics[3 * g + 2];
    float o = s->opac ? s->opac[g] : 1.0f;
    if (d_is_p_minus_mu) {
        r.dx = px - s->xys[2 * g];
        r.dy = py - s->xys[2 * g + 1];
    } else {
        r.dx = s->xys[2 * g] - px;
        r.dy = s->xys[2 * g + 1] - py;
    }
    r.sigma = 0.5f * (a * r.dx * r.dx + c * r.dy * r.dy) + b * r.dx * r.dy;
    r.vis = expf(-r.sigma);
    float raw = o * r.vis;
    r.alpha = raw < clamp ? raw : clamp;
    r.valid = !(r.sigma < 0.0f || r.alpha < 1.0f / 255.0f);
    return r;
}

/* Pixel loop shared by the forward passes: calls back per valid pair. */
typedef struct {
    double img[3], dx[3], dy[3], dxy[3], wsum;
    int final_idx, ambiguous;
} pixacc_t;

/* forward.cu:173-279 (plain) and gradient_aware_rasterizer.cu:351-495 (GA):
 * FP64 sums over the tile's bin in bin order.  out_* are HWC doubles; dx/dy/dxy
 * may be NULL.  ambig_px (may be NULL) counts ambiguous pairs per pixel. */
void orc_rasterize_fwd(int img_h, int img_w, int bw, const int32_t *gids, const int32_t *bins,
                       const float *xys, const float *conics, const float *colors,
                       const float *opac, double *out_img, double *out_wsum,
                       double *out_dx, double *out_dy, double *out_dxy, int32_t *final_idx,
                       int32_t *ambig_px) {
    scene_t s = {img_h, img_w, bw, (img_w + bw - 1) / bw, (img_h + bw - 1) / bw,
                 gids, bins, xys, conics, colors, opac};
    int ga = out_dx != NULL;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < img_h; ++i) {
        for (int j = 0; j < img_w; ++j) {
            int tile = (i / bw) * s.tx + (j / bw);
            int lo = bins[2 * tile], hi = bins[2 * tile + 1];
            double px = j + 0.5, py = i + 0.5;
            pixacc_t acc;
            memset(&acc, 0, sizeof acc);
            for (int k = lo; k < hi; ++k) {
                int g = gids[k];
                pair64_t p = eval_pair64(&s, g, px, py, 0.999);
                acc.ambiguous += p.ambiguous;
                if (!p.valid) continue;
                const float *col = colors + 3 * g;
                for (int ch = 0; ch < 3; ++ch) acc.img[ch] += col[ch] * p.alpha;
                acc.wsum += p.alpha;
                acc.final_idx = k;
                if (ga) {
                    double a = conics[3 * g], b = conics[3 * g + 1], c = conics[3 * g + 2];
                    /* GA kernel uses d = p - mu */
                    double ddx = -p.dx, ddy = -p.dy;
                    double sx = a * ddx + b * ddy, sy = b * ddx + c * ddy;
                    double ax = -p.alpha * sx, ay = -p.alpha * sy;
                    double axy = -p.alpha * (b - sx * sy);
                    for (int ch = 0; ch < 3; ++ch) {
                        acc.dx[ch] += col[ch] * ax;
                        acc.dy[ch] += col[ch] * ay;
                        acc.dxy[ch] += col[ch] * axy;
                    }
                }
            }
            size_t pid = (size_t)i * img_w + j;
            for (int ch = 0; ch < 3; ++ch) {
                out_img[3 * pid + ch] = acc.img[ch];
                if (ga) {
                    out_dx[3 * pid + ch] = acc.dx[ch];
                    out_dy[3 * pid + ch] = acc.dy[ch];
                    out_dxy[3 * pid + ch] = acc.dxy[ch];
                }
            }
            out_wsum[pid] = acc.wsum;
            final_idx[pid] = acc.final_idx;
            if (ambig_px) ambig_px[pid] = acc.ambiguous;
        }
    }
}

/* Per-Gaussian gradient slots, FP64: rgb[3], conic[3], xy[2], xy_abs[2], opacity */
#define NG 11

static void add_plain_pair(double *acc, double alpha, double vis, double dx, double dy,
                           const double *conic, const double *col, const double *v_out,
                           double v_w) {
    double v_alpha = col[0] * v_out[0] + col[1] * v_out[1] + col[2] * v_out[2] + v_w;
    for (int ch = 0; ch < 3; ++ch) acc[ch] += alpha * v_out[ch];
    double v_sigma = -alpha * v_alpha;
    acc[3] += 0.5 * v_sigma * dx * dx;
    acc[4] += v_sigma * dx * dy;
    acc[5] += 0.5 * v_sigma * dy * dy;
    double gx = v_sigma * (conic[0] * dx + conic[1] * dy);
    double gy = v_sigma * (conic[1] * dx + conic[2] * dy);
    acc[6] += gx; acc[7] += gy;
    acc[8] += fabs(gx); acc[9] += fabs(gy);
    acc[10] += vis * v_alpha;
}

/* gradient_aware_rasterizer.cu:228-315, d = p - mu.  The position gradient
 * is added to acc[6..7] and to grp[0..1], the running sum of the pixel's
 * 32-thread group: the reference adds |group sum| to xys.absgrad per group
 * (gradient_aware_rasterizer.cu:320,339-341), not |v_xy| per pixel. */
static void add_ga_pair(double *acc, double *grp, double alpha, double vis, double dx,
                        double dy, const double *conic, const double *c, const double *vI,
                        const double *vdx, const double *vdy, const double *vdxy, double vW) {
    double a = conic[0], b = conic[1], cc = conic[2];
    double sx = a * dx + b * dy, sy = b * dx + cc * dy;
    double ax = -alpha * sx, ay = -alpha * sy, axy = alpha * (sx * sy - b);
    for (int ch = 0; ch < 3; ++ch)
        acc[ch] += vI[ch] * alpha + vdx[ch] * ax + vdy[ch] * ay + vdxy[ch] * axy;
    double va = vI[0] * c[0] + vI[1] * c[1] + vI[2] * c[2] + vW;
    double vax = vdx[0] * c[0] + vdx[1] * c[1] + vdx[2] * c[2];
    double vay = vdy[0] * c[0] + vdy[1] * c[1] + vdy[2] * c[2];
    double vaxy = vdxy[0] * c[0] + vdxy[1] * c[1] + vdxy[2] * c[2];
    double sx2 = sx * sx, sy2 = sy * sy, sxsy = sx * sy;
    double gx = alpha * (va * sx + vax * (a - sx2) + vay * (b - sxsy) +
                         vaxy * (sx2 * sy - 2.0 * b * sx - a * sy));
    double gy = alpha * (va * sy + vax * (b - sxsy) + vay * (cc - sy2) +
                         vaxy * (sx * sy2 - 2.0 * b * sy - cc * sx));
    double dx2 = dx * dx, dy2 = dy * dy, dxdy = dx * dy, P = sxsy - b;
    acc[3] += alpha * (va * (-0.5 * dx2) + vax * (0.5 * dx2 * sx - dx) + vay * (0.5 * dx2 * sy) +
                       vaxy * (-0.5 * dx2 * P + dx * sy));
    acc[4] += alpha * (va * (-dxdy) + vax * (dxdy * sx - dy) + vay * (dxdy * sy - dx) +
                       vaxy * (-dxdy * P + dy * sy + dx * sx - 1.0));
    acc[5] += alpha * (va * (-0.5 * dy2) + vax * (0.5 * dy2 * sx) + vay * (0.5 * dy2 * sy - dy) +
                       vaxy * (-0.5 * dy2 * P + dy * sx));
    acc[6] += gx; acc[7] += gy;
    grp[0] += gx; grp[1] += gy;
    acc[10] += vis * va;
}

/* backward.cu:8-183 (plain) and gradient_aware_rasterizer.cu:118-349 (GA).
 * mode 0: FP64 pair math; mode 1: FP32 pair math in the reference's expression
 * order (the "reference gradient"), FP64 accumulation.  v_dx == NULL selects
 * the plain kernel.  grads: [N, 11] doubles laid out as
 * rgb(3) conic(3) xy(2) xy_abs(2) opacity(1), zero-filled by this function.
 * xy_abs: the plain kernel sums |v_xy| per pixel (backward.cu:144-150); the GA
 * kernel sums v_xy over each 32-thread warp of the bw x bw block (thread rank
 * ty*bw + tx, backward.cu:67 / gradient_aware_rasterizer.cu:170) and adds the
 * absolute value of that sum per warp (gradient_aware_rasterizer.cu:320,
 * 339-341); threads outside the image contribute 0.
 * ambig_g (may be NULL) counts ambiguous pairs per Gaussian.
 * Accumulation is deterministic: threads own disjoint tiles and are reduced
 * in a fixed order. */
void orc_rasterize_bwd(int num_points, int img_h, int img_w, int bw, const int32_t *gids,
                       const int32_t *bins, const float *xys, const float *conics,
                       const float *colors, const float *opac, const int32_t *final_idx,
                       const float *v_out, const float *v_wsum, const float *v_dx,
                       const float *v_dy, const float *v_dxy, int mode, double *grads,
                       int32_t *ambig_g) {
    scene_t s = {img_h, img_w, bw, (img_w + bw - 1) / bw, (img_h + bw - 1) / bw,
                 gids, bins, xys, conics, colors, opac};
    int ga = v_dx != NULL;
    double clampv = ga ? 0.999 : 0.99;
    int ntiles = s.tx * s.ty;
    int npix = bw * bw;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    size_t per = (size_t)num_points * NG;
    double *bufs = (double *)calloc(per * (size_t)nthr, sizeof(double));
    int32_t *amb = (int32_t *)calloc((size_t)num_points * (size_t)nthr, sizeof(int32_t));
#pragma omp parallel num_threads(nthr)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *acc = bufs + per * (size_t)tid;
        int32_t *am = amb + (size_t)num_points * (size_t)tid;
#pragma omp for schedule(dynamic, 4)
        for (int tile = 0; tile < ntiles; ++tile) {
            int lo = bins[2 * tile], hi = bins[2 * tile + 1];
            if (lo >= hi) continue;
            int i0 = (tile / s.tx) * bw, j0 = (tile % s.tx) * bw;
            /* 32-thread groups of the block, by thread rank r = ly * bw + lx */
            for (int r0 = 0; r0 < npix; r0 += 32) {
                int r1 = imin(r0 + 32, npix);
                /* back to front as the reference; order only matters for FP sums */
                for (int k = hi - 1; k >= lo; --k) {
                    int g = gids[k];
                    double *gacc = acc + (size_t)g * NG;
                    double grp[2] = {0.0, 0.0};
                    for (int r = r0; r < r1; ++r) {
                        int i = i0 + r / bw, j = j0 + r % bw;
                        if (i >= img_h || j >= img_w) continue;
                        size_t pid = (size_t)i * img_w + j;
                        if (k > final_idx[pid]) continue;
                        double vo[3] = {v_out[3 * pid], v_out[3 * pid + 1], v_out[3 * pid + 2]};
                        double vw = v_wsum[pid];
                        double conic[3] = {conics[3 * g], conics[3 * g + 1], conics[3 * g + 2]};
                        double col[3] = {colors[3 * g], colors[3 * g + 1], colors[3 * g + 2]};
                        double alpha, vis, dx, dy;
                        int valid;
                        if (mode == 0) {
                            pair64_t p = eval_pair64(&s, g, j + 0.5, i + 0.5, clampv);
                            am[g] += p.ambiguous;
                            valid = p.valid; alpha = p.alpha; vis = p.vis;
                            dx = p.dx; dy = p.dy;
                        } else {
                            pair32_t p = eval_pair32(&s, g, (float)j + 0.5f, (float)i + 0.5f,
                                                     (float)clampv, ga);
                            valid = p.valid; alpha = p.alpha; vis = p.vis;
                            dx = p.dx; dy = p.dy;
                        }
                        if (!valid) continue;
                        if (!ga) {
                            if (mode == 1) {
                                /* backward.cu:127-150 in FP32 */
                                float fa = (float)alpha, fv = (float)vis;
                                float fdx = (float)dx, fdy = (float)dy;
                                float fo[3] = {v_out[3 * pid], v_out[3 * pid + 1],
                                               v_out[3 * pid + 2]};
                                const float *c = colors + 3 * g;
                                const float *cn = conics + 3 * g;
                                float v_alpha = 0.0f;
                                v_alpha += c[0] * fo[0];
                                v_alpha += c[1] * fo[1];
                                v_alpha += c[2] * fo[2];
                                v_alpha += v_wsum[pid];
                                float v_op = fv * v_alpha;
                                float v_sigma = -fa * v_alpha;
                                float vc0 = 0.5f * v_sigma * fdx * fdx;
                                float vc1 = v_sigma * fdx * fdy;
                                float vc2 = 0.5f * v_sigma * fdy * fdy;
                                float gx = v_sigma * (cn[0] * fdx + cn[1] * fdy);
                                float gy = v_sigma * (cn[1] * fdx + cn[2] * fdy);
                                for (int ch = 0; ch < 3; ++ch) gacc[ch] += (double)(fa * fo[ch]);
                                gacc[3] += vc0; gacc[4] += vc1; gacc[5] += vc2;
                                gacc[6] += gx; gacc[7] += gy;
                                gacc[8] += fabsf(gx); gacc[9] += fabsf(gy);
                                gacc[10] += v_op;
                            } else {
                                add_plain_pair(gacc, alpha, vis, dx, dy, conic, col, vo, vw);
                            }
                        } else {
                            double vx[3], vy[3], vxy[3];
                            for (int ch = 0; ch < 3; ++ch) {
                                vx[ch] = v_dx[3 * pid + ch];
                                vy[ch] = v_dy[3 * pid + ch];
                                vxy[ch] = v_dxy[3 * pid + ch];
                            }
                            /* GA: d = p - mu */
                            if (mode == 0) { dx = -dx; dy = -dy; }
                            add_ga_pair(gacc, grp, alpha, vis, dx, dy, conic, col, vo, vx, vy,
                                        vxy, vw);
                        }
                    }
                    if (ga) {
                        gacc[8] += fabs(grp[0]);
                        gacc[9] += fabs(grp[1]);
                    }
                }
            }
        }
    }
    memset(grads, 0, sizeof(double) * per);
    if (ambig_g) memset(ambig_g, 0, sizeof(int32_t) * (size_t)num_points);
    for (int t = 0; t < nthr; ++t) {
        const double *src = bufs + per * (size_t)t;
        for (size_t q = 0; q < per; ++q) grads[q] += src[q];
        if (ambig_g)
            for (int g = 0; g < num_points; ++g) ambig_g[g] += amb[(size_t)num_points * t + g];
    }
    free(bufs);
    free(amb);
}

/* ------------------------------------------------------------------------- */
/* Gradient-aware Hermite upscaler                                             */
/* ------------------------------------------------------------------------- */

/* gradient_aware_upscaler.cu:145-154: Hermite basis at t */
static void hermite(double t, double *h) {
    double t2 = t * t, t3 = t2 * t;
    h[0] = 1.0 - 3.0 * t2 + 2.0 * t3;
    h[1] = 3.0 * t2 - 2.0 * t3;
    h[2] = t - 2.0 * t2 + t3;
    h[3] = t3 - t2;
}

/* gradient_aware_upscaler.cu:201-218: FP32 source coordinate, floor, clamp */
static void src_coord(int d, int dst, int src, float roi0, float roi1, int *c0, int *c1,
                      double *t) {
    int roi_w = f2i_sat(roi1 - roi0);
    float s = roi0 + div_rn(((float)d + 0.5f) * (float)roi_w, (float)dst) - 0.5f;
    int k = f2i_sat(floorf(s));
    *t = (double)(s - (float)k);
    int kc = imin(k, src - 1);
    *c1 = imax(0, imin(kc + 1, src - 1)); /* c1 < 0 clamped to 0 like c0 */
    *c0 = imax(0, kc);
}

/* gradient_aware_upscaler.cu:181-246 forward, FP64.  Inputs HWC [src_h,src_w,3]. */
void orc_upscale_fwd(int src_h, int src_w, int dst_h, int dst_w, const float *roi,
                     const float *f, const float *fx, const float *fy, const float *fxy,
                     double *out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < dst_h; ++y) {
        int y0, y1; double ty; double hy[4];
        src_coord(y, dst_h, src_h, roi[1], roi[3], &y0, &y1, &ty);
        hermite(ty, hy);
        for (int x = 0; x < dst_w; ++x) {
            int x0, x1; double tx; double hx[4];
            src_coord(x, dst_w, src_w, roi[0], roi[2], &x0, &x1, &tx);
            hermite(tx, hx);
            size_t i00 = (size_t)y0 * src_w + x0, i01 = (size_t)y0 * src_w + x1;
            size_t i10 = (size_t)y1 * src_w + x0, i11 = (size_t)y1 * src_w + x1;
            for (int ch = 0; ch < 3; ++ch) {
#define V(a, i) ((double)(a)[3 * (i) + ch])
                double c0 = V(f, i00) * hy[0] + V(f, i10) * hy[1] + V(fy, i00) * hy[2] + V(fy, i10) * hy[3];
                double c1 = V(f, i01) * hy[0] + V(f, i11) * hy[1] + V(fy, i01) * hy[2] + V(fy, i11) * hy[3];
                double c2 = V(fx, i00) * hy[0] + V(fx, i10) * hy[1] + V(fxy, i00) * hy[2] + V(fxy, i10) * hy[3];
                double c3 = V(fx, i01) * hy[0] + V(fx, i11) * hy[1] + V(fxy, i01) * hy[2] + V(fxy, i11) * hy[3];
#undef V
                out[3 * ((size_t)y * dst_w + x) + ch] = hx[0] * c0 + hx[1] * c1 + hx[2] * c2 + hx[3] * c3;
            }
        }
    }
}

/* gradient_aware_upscaler.cu:248-371 backward (the exact transpose), FP64,
 * serial scatter.  Outputs [src_h,src_w,3] doubles, zero-filled here. */
void orc_upscale_bwd(int src_h, int src_w, int dst_h, int dst_w, const float *roi,
                     const float *g, double *gf, double *gfx, double *gfy, double *gfxy) {
    size_t ns = (size_t)src_h * src_w * 3;
    memset(gf, 0, ns * sizeof(double));
    memset(gfx, 0, ns * sizeof(double));
    memset(gfy, 0, ns * sizeof(double));
    memset(gfxy, 0, ns * sizeof(double));
    for (int y = 0; y < dst_h; ++y) {
        int y0, y1; double ty; double hy[4];
        src_coord(y, dst_h, src_h, roi[1], roi[3], &y0, &y1, &ty);
        hermite(ty, hy);
        for (int x = 0; x < dst_w; ++x) {
            int x0, x1; double tx; double hx[4];
            src_coord(x, dst_w, src_w, roi[0], roi[2], &x0, &x1, &tx);
            hermite(tx, hx);
            size_t i00 = (size_t)y0 * src_w + x0, i01 = (size_t)y0 * src_w + x1;
            size_t i10 = (size_t)y1 * src_w + x0, i11 = (size_t)y1 * src_w + x1;
            for (int ch = 0; ch < 3; ++ch) {
                double gv = g[3 * ((size_t)y * dst_w + x) + ch];
                gf[3 * i00 + ch] += gv * hx[0] * hy[0];
                gf[3 * i01 + ch] += gv * hx[1] * hy[0];
                gf[3 * i10 + ch] += gv * hx[0] * hy[1];
                gf[3 * i11 + ch] += gv * hx[1] * hy[1];
                gfx[3 * i00 + ch] += gv * hx[2] * hy[0];
                gfx[3 * i01 + ch] += gv * hx[3] * hy[0];
                gfx[3 * i10 + ch] += gv * hx[2] * hy[1];
                gfx[3 * i11 + ch] += gv * hx[3] * hy[1];
                gfy[3 * i00 + ch] += gv * hx[0] * hy[2];
                gfy[3 * i01 + ch] += gv * hx[1] * hy[2];
                gfy[3 * i10 + ch] += gv * hx[0] * hy[3];
                gfy[3 * i11 + ch] += gv * hx[1] * hy[3];
                gfxy[3 * i00 + ch] += gv * hx[2] * hy[2];
                gfxy[3 * i01 + ch] += gv * hx[3] * hy[2];
                gfxy[3 * i10 + ch] += gv * hx[2] * hy[3];
                gfxy[3 * i11 + ch] += gv * hx[3] * hy[3];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Workload counters for the roofline (tile-size independent).                 */
/* ------------------------------------------------------------------------- */

/* E = pixel/Gaussian pairs whose pixel centre lies inside the Gaussian's
 * clipped SnugBox box |p - mu| <= extent;  C = pairs of E with alpha >= 1/255
 * and sigma >= 0 (FP64).  Only Gaussians with num_tiles_hit > 0 count. */
void orc_count_pairs(int n, int img_h, int img_w, const float *xys, const float *extents,
                     const float *conics, const float *opac, const int *ntiles,
                     int64_t *E_out, int64_t *C_out) {
    int64_t E = 0, C = 0;
#pragma omp parallel for reduction(+ : E, C) schedule(dynamic, 256)
    for (int g = 0; g < n; ++g) {
        if (ntiles[g] <= 0) continue;
        double mx = xys[2 * g], my = xys[2 * g + 1];
        double ex = extents[2 * g], ey = extents[2 * g + 1];
        int j0 = (int)ceil(mx - ex - 0.5), j1 = (int)floor(mx + ex - 0.5);
        int i0 = (int)ceil(my - ey - 0.5), i1 = (int)floor(my + ey - 0.5);
        if (j0 < 0) j0 = 0;
        if (i0 < 0) i0 = 0;
        if (j1 > img_w - 1) j1 = img_w - 1;
        if (i1 > img_h - 1) i1 = img_h - 1;
        if (j1 < j0 || i1 < i0) continue;
        double a = conics[3 * g], b = conics[3 * g + 1], c = conics[3 * g + 2];
        double o = opac ? opac[g] : 1.0;
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) {
                double dx = mx - (j + 0.5), dy = my - (i + 0.5);
                double sg = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy;
                double al = o * exp(-sg);
                if (al > 0.999) al = 0.999;
                E += 1;
                if (!(sg < 0.0 || al < 1.0 / 255.0)) C += 1;
            }
    }
    *E_out = E;
    *C_out = C;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
