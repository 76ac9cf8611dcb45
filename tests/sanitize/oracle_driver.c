/* Runs every function of the C oracle (oracle/lig_oracle.c) on seeded random
 * scenes -- projection (covariance and Cholesky forms), scan, intersection
 * map, pair sort, bin edges, the FP64 and FP32-order rasterisers (plain and
 * gradient-aware, with and without opacities, tile widths 16, 8 and 5), the
 * upscaler with full and partial ROIs, the pair counter and the
 * backward projections -- so that a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer checks every array access they make.  Prints
 * "ok" and exits 0 when nothing fired (the sanitizers abort otherwise). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_compute_cov2d_bounds(int n, const float *cov, const float *opac, float *conics,
                              float *extents);
void orc_project_fwd(int n, const float *cov2d, const float *means, const float *opac,
                     int img_h, int img_w, int bw, float *xys, float *extents, float *conics,
                     int *ntiles);
void orc_project_fwd_chol(int n, const float *chol, const float *means, const float *opac,
                          int img_h, int img_w, int bw, float *xys, float *extents,
                          float *conics, int *ntiles);
void orc_project_bwd(int n, const float *extents, const float *conics, const float *v_xy,
                     const float *v_conic, float *v_cov, float *v_mean);
void orc_project_bwd_chol(int n, const float *extents, const float *chol, const float *conics,
                          const float *v_xy, const float *v_conic, float *v_chol,
                          float *v_mean);
void orc_cumsum_i32(int n, const int *x, int *out);
void orc_map_intersects(int n, const float *xys, const float *depths, const float *extents,
                        const int *cum, int tx, int ty, int bw, int64_t *isect_ids,
                        int32_t *gids);
void orc_sort_pairs(int m, const int64_t *keys, const int32_t *vals, int64_t *keys_out,
                    int32_t *vals_out);
void orc_tile_bin_edges(int m, const int64_t *sorted, int *bins);
void orc_rasterize_fwd(int img_h, int img_w, int bw, const int32_t *gids, const int32_t *bins,
                       const float *xys, const float *conics, const float *colors,
                       const float *opac, double *out_img, double *out_wsum, double *out_dx,
                       double *out_dy, double *out_dxy, int32_t *final_idx, int32_t *ambig_px);
void orc_rasterize_bwd(int num_points, int img_h, int img_w, int bw, const int32_t *gids,
                       const int32_t *bins, const float *xys, const float *conics,
                       const float *colors, const float *opac, const int32_t *final_idx,
                       const float *v_out, const float *v_wsum, const float *v_dx,
                       const float *v_dy, const float *v_dxy, int mode, double *grads,
                       int32_t *ambig_g);
void orc_upscale_fwd(int src_h, int src_w, int dst_h, int dst_w, const float *roi,
                     const float *f, const float *fx, const float *fy, const float *fxy,
                     double *out);
void orc_upscale_bwd(int src_h, int src_w, int dst_h, int dst_w, const float *roi,
                     const float *g, double *gf, double *gfx, double *gfy, double *gfxy);
void orc_count_pairs(int n, int img_h, int img_w, const float *xys, const float *extents,
                     const float *conics, const float *opac, const int *ntiles, int64_t *E_out,
                     int64_t *C_out);

static uint64_t rng_state;
static double urand(void) { /* xorshift64*, uniform in [0, 1) */
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (double)((rng_state * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}

/* Exactly-sized heap arrays, so that any access past an end is reported. */
static void *alloc(size_t count, size_t size) {
    void *p = calloc(count ? count : 1, size);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static void frand(float *x, size_t n, double lo, double hi) {
    for (size_t i = 0; i < n; ++i) x[i] = (float)(lo + (hi - lo) * urand());
}

static void scene(int n, int H, int W, int bw, int chol, int with_opac, int ga) {
    float *means = alloc(2 * n, 4), *par = alloc(3 * n, 4), *opac = alloc(n, 4);
    float *colors = alloc(3 * n, 4), *depths = alloc(n, 4);
    for (int i = 0; i < n; ++i) {
        means[2 * i] = (float)(-8.0 + (W + 16.0) * urand()); /* some off the canvas */
        means[2 * i + 1] = (float)(-8.0 + (H + 16.0) * urand());
        if (chol) {
            par[3 * i] = (float)(0.3 + 3.0 * urand());
            par[3 * i + 1] = (float)(-1.0 + 2.0 * urand());
            par[3 * i + 2] = (float)(0.3 + 3.0 * urand());
        } else { /* possibly indefinite, as the reference's U(0,1)^3 init */
            frand(par + 3 * i, 3, 0.0, 4.0);
        }
    }
    frand(opac, n, 0.0, 1.0);
    frand(colors, 3 * (size_t)n, 0.0, 1.0);
    const float *op = with_opac ? opac : NULL;
    float *xys = alloc(2 * n, 4), *ext = alloc(2 * n, 4), *con = alloc(3 * n, 4);
    int *nt = alloc(n, 4), *cum = alloc(n, 4);
    if (chol)
        orc_project_fwd_chol(n, par, means, op, H, W, bw, xys, ext, con, nt);
    else
        orc_project_fwd(n, par, means, op, H, W, bw, xys, ext, con, nt);
    orc_cumsum_i32(n, nt, cum);
    const int m = n ? cum[n - 1] : 0;
    const int tx = (W + bw - 1) / bw, ty = (H + bw - 1) / bw;
    int64_t *keys = alloc(m, 8), *keys_s = alloc(m, 8);
    int32_t *gids = alloc(m, 4), *gids_s = alloc(m, 4);
    orc_map_intersects(n, xys, depths, ext, cum, tx, ty, bw, keys, gids);
    orc_sort_pairs(m, keys, gids, keys_s, gids_s);
    int *bins = alloc(2 * (size_t)tx * ty, 4);
    orc_tile_bin_edges(m, keys_s, bins);
    const size_t P = (size_t)H * W;
    double *img = alloc(3 * P, 8), *wsum = alloc(P, 8);
    double *dx = ga ? alloc(3 * P, 8) : NULL, *dy = ga ? alloc(3 * P, 8) : NULL;
    double *dxy = ga ? alloc(3 * P, 8) : NULL;
    int32_t *fin = alloc(P, 4), *amb = alloc(P, 4);
    orc_rasterize_fwd(H, W, bw, gids_s, bins, xys, con, colors, op, img, wsum, dx, dy, dxy, fin,
                      amb);
    float *v_out = alloc(3 * P, 4), *v_wsum = alloc(P, 4);
    float *v_dx = ga ? alloc(3 * P, 4) : NULL, *v_dy = ga ? alloc(3 * P, 4) : NULL;
    float *v_dxy = ga ? alloc(3 * P, 4) : NULL;
    frand(v_out, 3 * P, -1.0, 1.0);
    frand(v_wsum, P, -1.0, 1.0);
    if (ga) {
        frand(v_dx, 3 * P, -1.0, 1.0);
        frand(v_dy, 3 * P, -1.0, 1.0);
        frand(v_dxy, 3 * P, -1.0, 1.0);
    }
    double *grads = alloc(11 * (size_t)n, 8);
    int32_t *ambg = alloc(n, 4);
    for (int mode = 0; mode < 2; ++mode)
        orc_rasterize_bwd(n, H, W, bw, gids_s, bins, xys, con, colors, op, fin, v_out, v_wsum,
                          v_dx, v_dy, v_dxy, mode, grads, ambg);
    int64_t E = 0, C = 0;
    orc_count_pairs(n, H, W, xys, ext, con, op, nt, &E, &C);
    float *v_xy = alloc(2 * n, 4), *v_con = alloc(3 * n, 4);
    float *v_par = alloc(3 * n, 4), *v_mean = alloc(2 * n, 4);
    frand(v_xy, 2 * (size_t)n, -1.0, 1.0);
    frand(v_con, 3 * (size_t)n, -1.0, 1.0);
    if (chol)
        orc_project_bwd_chol(n, ext, par, con, v_xy, v_con, v_par, v_mean);
    else
        orc_project_bwd(n, ext, con, v_xy, v_con, v_par, v_mean);
    float *con2 = alloc(3 * n, 4), *ext2 = alloc(2 * n, 4);
    orc_compute_cov2d_bounds(n, par, op, con2, ext2);
    printf("scene n=%d %dx%d bw=%d chol=%d opac=%d ga=%d: %d intersections, E=%lld C=%lld\n",
           n, H, W, bw, chol, with_opac, ga, m, (long long)E, (long long)C);
    void *all[] = {means, par, opac, colors, depths, xys, ext, con, nt, cum, keys, keys_s,
                   gids, gids_s, bins, img, wsum, dx, dy, dxy, fin, amb, v_out, v_wsum, v_dx,
                   v_dy, v_dxy, grads, ambg, v_xy, v_con, v_par, v_mean, con2, ext2};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]);
}

static void upscale(int sh, int sw, int dh, int dw, const float roi[4]) {
    const size_t ns = (size_t)sh * sw * 3, nd = (size_t)dh * dw * 3;
    float *f = alloc(ns, 4), *fx = alloc(ns, 4), *fy = alloc(ns, 4), *fxy = alloc(ns, 4);
    float *g = alloc(nd, 4);
    frand(f, ns, 0, 1);
    frand(fx, ns, -1, 1);
    frand(fy, ns, -1, 1);
    frand(fxy, ns, -1, 1);
    frand(g, nd, -1, 1);
    double *out = alloc(nd, 8), *gf = alloc(ns, 8), *gfx = alloc(ns, 8), *gfy = alloc(ns, 8);
    double *gfxy = alloc(ns, 8);
    orc_upscale_fwd(sh, sw, dh, dw, roi, f, fx, fy, fxy, out);
    orc_upscale_bwd(sh, sw, dh, dw, roi, g, gf, gfx, gfy, gfxy);
    printf("upscale %dx%d -> %dx%d roi (%g, %g, %g, %g)\n", sh, sw, dh, dw, roi[0], roi[1],
           roi[2], roi[3]);
    void *all[] = {f, fx, fy, fxy, g, out, gf, gfx, gfy, gfxy};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) free(all[i]);
}

int main(void) {
    rng_state = 0x9E3779B97F4A7C15ull;
    const int bws[3] = {16, 8, 5};
    for (int k = 0; k < 3; ++k)
        for (int chol = 0; chol < 2; ++chol)
            for (int op = 0; op < 2; ++op)
                for (int ga = 0; ga < 2; ++ga) {
                    const int H = 17 + (int)(60 * urand()), W = 23 + (int)(60 * urand());
                    scene(40 + (int)(200 * urand()), H, W, bws[k], chol, op, ga);
                }
    scene(0, 16, 16, 16, 1, 0, 0); /* no Gaussians */
    scene(1, 1, 1, 16, 0, 1, 1);   /* one pixel */
    const float full[4] = {0.0f, 0.0f, 13.0f, 9.0f};
    const float part[4] = {2.5f, 1.25f, 9.0f, 7.75f};
    const float left[4] = {-1.5f, -0.75f, 4.0f, 3.0f}; /* starts left of -0.5 */
    upscale(9, 13, 36, 52, full);
    upscale(9, 13, 20, 17, part);
    upscale(9, 13, 11, 7, left);
    upscale(1, 1, 4, 4, full);
    printf("ok\n");
    return 0;
}
