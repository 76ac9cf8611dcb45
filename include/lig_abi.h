/*
 * lig_abi.h -- C-ABI of the MI355X (gfx950) gsplat2d hot path.
 *
 * Every entry point takes raw device pointers (row-major, contiguous, FP32
 * unless stated), integer sizes and the caller's HIP stream (hipStream_t passed
 * as void*; NULL = the default stream).  The library never allocates device
 * memory: outputs and workspaces are the caller's.  Outputs need no
 * initialisation -- each call overwrites its outputs completely (the reference
 * zero-fills them with torch::zeros and then writes; the result is the same).
 * Calls are asynchronous on `stream` and reentrant; calls on one stream are
 * ordered, calls on different streams that share buffers are the caller's
 * business.  Return value: LIG_OK or a negative LIG_ERR_* code; lig_strerror()
 * names it.  Layouts: float2/float3 arrays are [N,2] / [N,3] row-major, images
 * are HWC [H,W,3], tile bins are [T,2] int32.
 *
 * Each function names the reference interface it replaces (paths relative to
 * the reference repository, gsplat2d/gsplat2d/cuda/csrc/ abbreviated csrc/).
 */
#ifndef LIG_ABI_H
#define LIG_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIG_OK 0
#define LIG_ERR_INVALID_ARGUMENT (-1)
#define LIG_ERR_WORKSPACE (-2)
#define LIG_ERR_LAUNCH (-3)
#define LIG_ERR_UNSUPPORTED (-4)
#define LIG_ERR_TOO_LARGE (-5)

/* Human-readable name of a status code. */
const char *lig_strerror(int status);

/* ABI version (major*10000 + minor*100 + patch). */
int lig_abi_version(void);

/* ---- projection ---------------------------------------------------------- */

/* Replaces compute_cov2d_bounds_tensor (csrc/bindings.cu:41-66, kernel
 * csrc/bindings.cu:19-39).  cov2d [n,3] (a,b,c upper triangle), opacities [n]
 * or NULL (= 1) -> conics [n,3], extents [n,2].  det == 0 gives zeros. */
int lig_compute_cov2d_bounds(int n, const float *cov2d, const float *opacities,
                             float *conics, float *extents, void *stream);

/* Replaces project_gaussians_forward_tensor (csrc/bindings.cu:68-117, kernel
 * csrc/forward.cu:13-53).  -> xys [n,2], extents [n,2], conics [n,3],
 * num_tiles_hit [n] int32.  block_width in [1, 16]. */
int lig_project_gaussians_forward(int n, const float *cov2d, const float *means2d,
                                  const float *opacities, int img_height, int img_width,
                                  int block_width, float *xys, float *extents, float *conics,
                                  int32_t *num_tiles_hit, void *stream);

/* Replaces project_gaussians_forward_cholesky_tensor (csrc/bindings.cu:153-202,
 * kernel csrc/forward.cu:55-105).  cholesky [n,3] = (l11, l21, l22). */
int lig_project_gaussians_forward_cholesky(int n, const float *cholesky, const float *means2d,
                                           const float *opacities, int img_height,
                                           int img_width, int block_width, float *xys,
                                           float *extents, float *conics,
                                           int32_t *num_tiles_hit, void *stream);

/* No reference counterpart: crop-native projection for rendering a window.
 * Each Gaussian is first mapped into the view with origin (origin_x,
 * origin_y) on the source canvas and (scale_x, scale_y) output pixels per
 * source pixel: mean' = (mean - origin) * scale and, for cholesky != 0,
 * (l11, l21, l22)' = (l11 sx, l21 sy, l22 sy) (the reference's scale_to,
 * minimal_trainer_upscale.py:139-159, with an origin), else
 * cov' = (a sx sx, b sx sy, c sy sy); FP32 without contraction.  Then exactly
 * lig_project_gaussians_forward(_cholesky) on an img_height x img_width
 * canvas.  scale_x, scale_y > 0. */
int lig_project_gaussians_view(int n, const float *par, const float *means2d,
                               const float *opacities, int cholesky, float origin_x,
                               float origin_y, float scale_x, float scale_y, int img_height,
                               int img_width, int block_width, float *xys, float *extents,
                               float *conics, int32_t *num_tiles_hit, void *stream);

/* Replaces project_gaussians_backward_tensor (csrc/bindings.cu:119-151, kernel
 * csrc/backward.cu:185-204).  -> v_cov2d [n,3], v_mean2d [n,2]. */
int lig_project_gaussians_backward(int n, const float *extents, const float *conics,
                                   const float *v_xy, const float *v_conic, float *v_cov2d,
                                   float *v_mean2d, void *stream);

/* Replaces project_gaussians_backward_cholesky_tensor (csrc/bindings.cu:204-237,
 * kernel csrc/backward.cu:206-247).  -> v_cholesky [n,3], v_mean2d [n,2]. */
int lig_project_gaussians_backward_cholesky(int n, const float *extents, const float *cholesky,
                                            const float *conics, const float *v_xy,
                                            const float *v_conic, float *v_cholesky,
                                            float *v_mean2d, void *stream);

/* ---- tile intersection ----------------------------------------------------- */

/* Replaces torch.cumsum(num_tiles_hit, dtype=int32) of compute_cumulative_intersects
 * (gsplat2d/gsplat2d/utils.py:61-67): inclusive int32 scan.  The workspace is
 * lig_scan_workspace_size(n) bytes. */
size_t lig_scan_workspace_size(int n);
int lig_inclusive_scan_i32(int n, const int32_t *in, int32_t *out, void *workspace,
                           size_t workspace_bytes, void *stream);

/* Replaces map_gaussian_to_intersects_tensor (csrc/bindings.cu:239-281, kernel
 * csrc/forward.cu:109-143).  -> isect_ids [num_intersects] int64
 * ((tile << 32) | sign-extended depth bits), gaussian_ids [num_intersects] int32.
 * Slots of Gaussians whose extent is 0 are written as (0, 0), as the
 * reference's zero fill leaves them. */
int lig_map_gaussian_to_intersects(int n, int num_intersects, const float *xys,
                                   const float *depths, const float *extents,
                                   const int32_t *cum_tiles_hit, int tiles_x, int tiles_y,
                                   int block_width, int64_t *isect_ids, int32_t *gaussian_ids,
                                   void *stream);

/* Replaces torch.sort(isect_ids) + torch.gather(gaussian_ids, sorted_indices)
 * (gsplat2d/gsplat2d/utils.py:97-98): stable LSD radix sort by signed 64-bit
 * key carrying int32 values.  Workspace: lig_sort_pairs_workspace_size(m, 64). */
size_t lig_sort_pairs_workspace_size(int m, int key_bits);
int lig_sort_isect_pairs(int m, const int64_t *keys_in, const int32_t *vals_in,
                         int64_t *keys_out, int32_t *vals_out, void *workspace,
                         size_t workspace_bytes, void *stream);

/* Replaces get_tile_bin_edges_tensor (csrc/bindings.cu:283-301, kernel
 * csrc/forward.cu:148-170).  -> tile_bins [tiles_x*tiles_y, 2], empty tiles (0,0). */
int lig_get_tile_bin_edges(int num_intersects, const int64_t *isect_ids_sorted, int tiles_x,
                           int tiles_y, int32_t *tile_bins, void *stream);

/* Fused replacement of bin_and_sort_gaussians (gsplat2d/gsplat2d/utils.py:70-100)
 * for the rasterisation path, where every depth is 0 (gsplat2d/gsplat2d/
 * rasterize.py:158): emits 32-bit tile keys, radix-sorts only the
 * ceil(log2(tiles)) key bits that vary, and builds the bins.  cum_tiles_hit is
 * the inclusive scan of num_tiles_hit and num_intersects its last element.
 * -> gaussian_ids_sorted [num_intersects] int32 in (tile, emission) order,
 * tile_bins [tiles,2], and, when emission_ids_sorted is not NULL, the emission
 * slot of each sorted intersection (Gaussian g's slots are [cum[g-1], cum[g]);
 * the deterministic backward needs them).
 * Workspace: lig_bin_workspace_size(num_intersects, tiles). */
size_t lig_bin_workspace_size(int num_intersects, int num_tiles);
/* No reference counterpart: the workspace of lig_bin_and_sort_gaussians's row
 * path (0 when it does not apply: more than 1024 tiles per row or more than
 * 4096 tile rows).  The row path writes one entry per (tile row, Gaussian) in
 * (row, gid) order straight from each Gaussian's tile rectangle and expands
 * each row's entries into its tiles with per-tile counts and ballot ranks; it
 * gives the same outputs bit for bit and runs when the workspace passed holds
 * this many bytes, the radix path otherwise (which needs
 * lig_bin_workspace_size). */
size_t lig_bin_rows_workspace_size(int n, int num_intersects, int img_height, int img_width,
                                   int block_width);
int lig_bin_and_sort_gaussians(int n, int num_intersects, const float *xys,
                               const float *extents, const int32_t *cum_tiles_hit,
                               int img_height, int img_width, int block_width,
                               int32_t *gaussian_ids_sorted, int32_t *tile_bins,
                               int32_t *emission_ids_sorted, void *workspace,
                               size_t workspace_bytes, void *stream);

/* No counterpart in the reference.  lig_bin_and_sort_gaussians without the
 * workspace argument, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the workspace is a fixed scratch array of
 * LIG_BIN_NOALLOC_SCRATCH_BYTES in the library's device memory, one per GPU,
 * held from the library's first use on that GPU until the process ends.  The
 * call allocates nothing.  Same arguments (minus the workspace) and the same
 * results bit for bit.  Sizes whose lig_bin_workspace_size(num_intersects,
 * tiles) exceeds the scratch (about 2.7 M intersections at 16 x 16 tiles) are
 * rejected with LIG_ERR_TOO_LARGE before anything is launched.  The scratch
 * used is the current device's; a non-NULL stream must belong to it
 * (LIG_ERR_INVALID_ARGUMENT otherwise).  Nothing in the scratch needs to start
 * at zero; every call on a GPU shares that GPU's scratch, so calls that overlap
 * on one GPU (different streams, or different host threads) race: the caller
 * serialises them. */
#define LIG_BIN_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int lig_bin_and_sort_gaussians_noalloc(int n, int num_intersects, const float *xys,
                                       const float *extents, const int32_t *cum_tiles_hit,
                                       int img_height, int img_width, int block_width,
                                       int32_t *gaussian_ids_sorted, int32_t *tile_bins,
                                       int32_t *emission_ids_sorted, void *stream);

/* ---- rasterisation ------------------------------------------------------- */

/* No reference counterpart: builds the per-Gaussian records (64 bytes each,
 * lig_packed_gaussians_bytes(n)) that both rasteriser kernels read, from
 * xys [n,2], conics [n,3], colors [n,3] and opacities [n] or NULL.
 * lig_rasterize_forward calls it itself; a backward that is not handed the
 * forward's buffer calls it first. */
size_t lig_packed_gaussians_bytes(int n);
int lig_pack_gaussians(int n, const float *xys, const float *conics, const float *colors,
                       const float *opacities, void *packed, void *stream);

/* Replaces rasterize_forward_tensor (csrc/bindings.cu:303-401, kernels
 * csrc/forward.cu:173-279 and csrc/gradient_aware_rasterizer.cu:351-495).
 * colors [n,3]; opacities [n] or NULL.  -> out_img [H,W,3], out_wsum [H,W],
 * final_idx [H,W] int32; when out_dx != NULL (compute_upscale_gradients) also
 * out_dx/out_dy/out_dxy [H,W,3].  `packed` is a caller buffer of
 * lig_packed_gaussians_bytes(n) bytes that the call fills with the per-Gaussian
 * records the kernels read; pass the same buffer, unchanged, to the backward.
 * extents [n,2] and cum_tiles_hit [n] may be NULL; when given (block_width 16,
 * opacities NULL, bins from lig_bin_and_sort_gaussians without emission ids)
 * `packed` also carries each Gaussian's record addressing, and the backward
 * (lig_rasterize_backward_records) then needs no emission ids.
 * out_wsum and final_idx may both be NULL (block_width 16, no out_dx): an
 * image-only render, e.g. for serving; final_idx alone may be NULL (the same
 * conditions) for a backward that takes none (lig_rasterize_backward). */
int lig_rasterize_forward(int n, int img_height, int img_width, int block_width,
                          const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                          const float *xys, const float *conics, const float *colors,
                          const float *opacities, const float *extents,
                          const int32_t *cum_tiles_hit, void *packed, float *out_img,
                          float *out_wsum, int32_t *final_idx, float *out_dx, float *out_dy,
                          float *out_dxy, void *stream);

/* No reference counterpart: lig_rasterize_forward with a workspace of
 * lig_tile_order_workspace_size(num_tiles) bytes (or more), in which the
 * 16 x 16 plain forward (no out_dx) orders its tiles' launch: heaviest (most
 * pairs) first when the launch is short or the image's density is uneven
 * enough that contiguous per-XCD shares of the tiles would leave some XCDs
 * running long after the others.  The order changes speed only, never a
 * result.  The backward does the same when its workspace
 * (lig_rasterize_backward_records / lig_rasterize_backward with emission ids)
 * has lig_tile_order_workspace_size(num_tiles) bytes to spare after the
 * records: pass lig_rasterize_backward_workspace_size(num_intersects) + 256 +
 * lig_tile_order_workspace_size(num_tiles). */
size_t lig_tile_order_workspace_size(int num_tiles);
int lig_rasterize_forward_ws(int n, int img_height, int img_width, int block_width,
                             const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                             const float *xys, const float *conics, const float *colors,
                             const float *opacities, const float *extents,
                             const int32_t *cum_tiles_hit, void *packed, float *out_img,
                             float *out_wsum, int32_t *final_idx, float *out_dx, float *out_dy,
                             float *out_dxy, void *workspace, size_t workspace_bytes,
                             void *stream);

/* Replaces rasterize_backward_tensor (csrc/bindings.cu:403-506, kernels
 * csrc/backward.cu:8-183 and csrc/gradient_aware_rasterizer.cu:118-349).
 * v_output [H,W,3], v_wsum [H,W] or NULL (= zeros); v_output_dx/dy/dxy
 * [H,W,3] or all NULL (plain).  -> v_xy [n,2], v_xy_abs [n,2], v_conic [n,3],
 * v_colors [n,3], v_opacity [n] (written only when has_opacities).  `packed`
 * holds the records of lig_pack_gaussians for the same Gaussians.
 * Two reductions:
 *   emission_ids_sorted == NULL: each Gaussian's per-tile sum is added with
 *     float atomics (one set per Gaussian per tile); summation order varies.
 *   emission_ids_sorted != NULL (from lig_bin_and_sort_gaussians, with its
 *     cum_tiles_hit and num_intersects): per-tile sums are written to the
 *     workspace (lig_rasterize_backward_workspace_size(num_intersects) bytes)
 *     and reduced per Gaussian in a fixed order that depends only on its own
 *     records (chunks of 16 slots from its first, each summed by a fixed
 *     tree, chunk sums in order): bit-reproducible, no atomics.
 * final_idx may be NULL (block_width 16, no out_dx gradients) when it would
 * be the forward's: in sum mode the final index only marks the last pair the
 * forward accepted, and the backward evaluates the same accept test on the
 * same FP32 values, so no pixel needs gating; each tile's Gaussians are then
 * grouped by the tile rows they reach (as in the fused fit step).  The
 * results equal those given the forward's final index (bit for bit with
 * records).  The same holds for lig_rasterize_backward_records. */
size_t lig_rasterize_backward_workspace_size(int num_intersects);
int lig_rasterize_backward(int n, int img_height, int img_width, int block_width,
                           const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                           const void *packed, const int32_t *final_idx,
                           const float *v_output, const float *v_wsum,
                           const float *v_output_dx, const float *v_output_dy,
                           const float *v_output_dxy, int has_opacities,
                           const int32_t *emission_ids_sorted, const int32_t *cum_tiles_hit,
                           int num_intersects, void *workspace, size_t workspace_bytes,
                           float *v_xy, float *v_xy_abs, float *v_conic, float *v_colors,
                           float *v_opacity, void *stream);

/* The two halves of the deterministic lig_rasterize_backward, for callers that
 * schedule (or time) them separately.  No reference counterpart.
 * lig_rasterize_backward_records writes each (tile, Gaussian) partial
 * gradient to its emission slot in the workspace;
 * lig_reduce_gradient_records sums each Gaussian's slots in order and writes
 * the gradients.  Same workspace, lig_rasterize_backward_workspace_size.
 * emission_ids_sorted may be NULL when `packed` came from a forward given
 * extents and cum_tiles_hit (16 x 16 tiles, no opacities): each record's slot
 * is then computed from the pack. */
int lig_rasterize_backward_records(int n, int img_height, int img_width, int block_width,
                                   const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                                   const void *packed, const int32_t *final_idx,
                                   const float *v_output, const float *v_wsum,
                                   const float *v_output_dx, const float *v_output_dy,
                                   const float *v_output_dxy, int has_opacities,
                                   const int32_t *emission_ids_sorted, int num_intersects,
                                   void *workspace, size_t workspace_bytes, void *stream);
int lig_reduce_gradient_records(int n, int num_intersects, const int32_t *cum_tiles_hit,
                                int has_opacities, const void *workspace,
                                size_t workspace_bytes, float *v_xy, float *v_xy_abs,
                                float *v_conic, float *v_colors, float *v_opacity, void *stream);

/* ---- fused fit step (16 x 16 tiles) --------------------------------------- */

/* No reference counterpart: the training step of the gsplat2d caller
 * (gaussianlig.py:89-103: rasterize, clamp(0, 1), mse; backward) with the
 * loss fused into the rasteriser.  Sum mode, block width 16, no opacities
 * (so no zero-extent phantom intersections), num_intersects >= 1 (the caller
 * takes the reference's I < 1 path otherwise).  Bins come from
 * lig_bin_and_sort_gaussians without emission ids.
 *
 * lig_fit_render_l2: packs the Gaussians into `packed` (adding each one's
 * record addressing from extents / cum_tiles_hit), rasterises out_img [H,W,3]
 * exactly as lig_rasterize_forward (same pair test, same order, so the same
 * bits) but without out_wsum / final_idx, and writes
 * loss[0] = mean((clamp(out_img,0,1) - target)^2) over H*W*3 values, summed
 * per tile in FP32 and over tiles in FP64 (deterministic).  batch_depth
 * (64, 128, 256; 0 = 64) is the number of Gaussians staged in LDS per batch;
 * it changes speed only, not results.  Workspace:
 * lig_fit_render_workspace_size(H, W), which also holds the tiles' launch
 * order: heaviest (most pairs) first, so that long tiles start early and the
 * XCDs get like loads; the order changes speed only.  The backward below and
 * the batched calls order their tiles the same way (the backward when its
 * workspace has room after the 32-byte records, as it has whenever
 * num_intersects >= num_tiles / 4 + 96).
 *
 * lig_fit_backward_records_l2: the records half of the deterministic backward
 * for upstream gradient dL/dI = (0 <= img <= 1) ? (2/(3HW))(img - target) *
 * grad_loss[0] : 0 (the expression of lig_l2_clamped_loss_backward), formed
 * per tile from img and target; no v_wsum, no final index (in sum mode it only
 * skips pairs that the accept test rejects, so the gradients equal those of
 * lig_rasterize_backward_records fed the forward's final_idx) and no v_xy_abs
 * (no parameter gradient depends on it).  Each 32-byte record lands at its
 * emission slot, computed from `packed`.
 *
 * lig_fit_reduce_records: sums each Gaussian's records (in the order of the
 * API reduction, so both paths give the same bits) into v_xy [n,2],
 * v_conic [n,3], v_colors [n,3].  Same workspace as the records
 * call, lig_rasterize_backward_workspace_size(num_intersects). */
size_t lig_fit_render_workspace_size(int img_height, int img_width);
int lig_fit_render_l2(int n, int img_height, int img_width, const int32_t *gaussian_ids_sorted,
                      const int32_t *tile_bins, const float *xys, const float *extents,
                      const int32_t *cum_tiles_hit, const float *conics, const float *colors,
                      void *packed, const float *target, float *out_img, float *loss,
                      int batch_depth, void *workspace, size_t workspace_bytes, void *stream);
int lig_fit_backward_records_l2(int n, int img_height, int img_width,
                                const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                                const void *packed, const float *img, const float *target,
                                const float *grad_loss, int num_intersects, void *workspace,
                                size_t workspace_bytes, void *stream);
int lig_fit_reduce_records(int n, int num_intersects, const int32_t *cum_tiles_hit,
                           const void *workspace, size_t workspace_bytes, float *v_xy,
                           float *v_conic, float *v_colors, void *stream);

/* No reference counterpart (a tuning override).  The 16x16 backward kernels
 * (API records or atomics, fit records, batched grid launch) may split each
 * tile's 64-Gaussian batches over several blocks, "shares"; every record has
 * its own slot, so the results do not depend on the split.  By default
 * (shares = 0) the library picks the count per launch from the tile count,
 * the mean pairs per tile and the resident blocks; 1..8 forces that count
 * for every later launch in the process.  Returns the previous setting, or
 * LIG_ERR_INVALID_ARGUMENT for a value outside 0..8. */
int lig_set_backward_shares(int shares);

/* ---- batched fit (many images, one launch per stage) ----------------------- */

/* No reference counterpart.  B images fitted together: `batch_table` is a
 * device array of B + 1 records {int tile_off, tiles_x, tiles_y, w, h,
 * gauss_off; int64 pix_off} (32 bytes each, 8-byte aligned; the last record a
 * sentinel whose tile_off / gauss_off are the totals).  Image b owns global
 * tiles [tile_off, +tiles_x*tiles_y) (16 x 16 tiles), pixels [pix_off,
 * +w*h) of the concatenated HWC buffers and Gaussians [gauss_off, next
 * gauss_off); coordinates inside an image are its own, so each image computes
 * exactly what lig_fit_render_l2 / lig_fit_backward_records_l2 compute for it
 * alone.  Cholesky parameters, no opacities.
 * lig_project_batched_cholesky: lig_project_gaussians_forward_cholesky per
 * image.  lig_bin_batched: keys are global tile ids; tile_bins [num_tiles,2].
 * lig_fit_render_l2_batched: loss_per_image[b] = image b's clamp + L2 mean,
 * loss[0] = their sum; workspace lig_fit_batched_workspace_size(num_tiles).
 * lig_fit_backward_records_l2_batched: records for upstream gradient
 * grad_loss[0] of loss[0]; follow with lig_fit_reduce_records. */
int lig_project_batched_cholesky(int n, const float *cholesky, const float *means2d,
                                 const void *batch_table, int num_images, float *xys,
                                 float *extents, float *conics, int32_t *num_tiles_hit,
                                 void *stream);
int lig_bin_batched(int n, int num_intersects, const float *xys, const float *extents,
                    const int32_t *cum_tiles_hit, const void *batch_table, int num_images,
                    int num_tiles, int32_t *gaussian_ids_sorted, int32_t *tile_bins,
                    void *workspace, size_t workspace_bytes, void *stream);
size_t lig_fit_batched_workspace_size(int num_tiles);
int lig_fit_render_l2_batched(int n, const int32_t *gaussian_ids_sorted,
                              const int32_t *tile_bins, const float *xys, const float *extents,
                              const int32_t *cum_tiles_hit, const float *conics,
                              const float *colors, void *packed, const float *target,
                              const void *batch_table, int num_images, int num_tiles,
                              float *out_img, float *loss_per_image, float *loss,
                              int batch_depth, void *workspace, size_t workspace_bytes,
                              void *stream);
int lig_fit_backward_records_l2_batched(int n, const int32_t *gaussian_ids_sorted,
                                        const int32_t *tile_bins, const void *packed,
                                        const float *img, const float *target,
                                        const float *grad_loss, const void *batch_table,
                                        int num_images, int num_tiles, int num_intersects,
                                        void *workspace, size_t workspace_bytes, void *stream);

/* Persistent forms of the two batched raster stages: the same arguments,
 * workspaces and results, bit for bit.  Instead of one workgroup per tile the
 * grid is the waves that fit on the device at once, and each wave takes tiles
 * from a work queue: one counter per XCD over that XCD's contiguous share of
 * the batch's tile list (images in table order), then tiles left in the other
 * XCDs' shares, so uneven images and uneven tiles balance across the chip.
 * The counters live in the workspace (2 KB, reset by the call): after the tile
 * partials in the render's, after the fit records in the backward's (when
 * num_intersects >= 144 leaves room; else the backward runs one workgroup
 * per tile). */
int lig_fit_render_l2_batched_persistent(int n, const int32_t *gaussian_ids_sorted,
                                         const int32_t *tile_bins, const float *xys,
                                         const float *extents, const int32_t *cum_tiles_hit,
                                         const float *conics, const float *colors, void *packed,
                                         const float *target, const void *batch_table,
                                         int num_images, int num_tiles, float *out_img,
                                         float *loss_per_image, float *loss, int batch_depth,
                                         void *workspace, size_t workspace_bytes, void *stream);
int lig_fit_backward_records_l2_batched_persistent(
    int n, const int32_t *gaussian_ids_sorted, const int32_t *tile_bins, const void *packed,
    const float *img, const float *target, const float *grad_loss, const void *batch_table,
    int num_images, int num_tiles, int num_intersects, void *workspace, size_t workspace_bytes,
    void *stream);

/* ---- batched crop rendering (serving) -------------------------------------- */

/* One view of a batched crop render (64 bytes): n Gaussians (par [n,3]
 * Cholesky factors when cholesky != 0, else covariances; means2d [n,2]; rgbs
 * [n,3]) seen through lig_project_gaussians_view's transform with this origin
 * and scale on an img_height x img_width canvas, written to slots
 * [gauss_off, gauss_off + n) of the batch's concatenated arrays. */
typedef struct LigView {
    const float *par;
    const float *means2d;
    const float *rgbs;
    int32_t n, cholesky;
    float origin_x, origin_y, scale_x, scale_y;
    int32_t img_height, img_width;
    int32_t gauss_off, reserved;
} LigView;

/* One crop of the batched composition (128 bytes), train.py:150-154 /
 * 187-196 for a window: out[h,w,3] = clamp(clamp(L2) * span + up - 0.5 + lo,
 * 0, 1) evaluated left to right in FP32 without contraction, where L2 is the
 * window at (off_x, off_y) of the level-2 view render `level2` (row pitch
 * view_w pixels; NULL: a view with no intersections, whose render is ones,
 * rasterize.py:80-90), span = (float)(hi - lo), lo = (float)lo, and `up` the
 * level-1 estimate: mode 0, lig_gradient_aware_upscale_forward of (f, fx,
 * fy, fxy) [src_h,src_w,3] over roi to h x w, clamped to [0, 1]; mode 1,
 * lig_bilinear_resize of f to full_h x full_w, window at (win_x, win_y).
 * block_off: the crop's first 256-pixel block in the launch (crops sorted by
 * it, each owning ceil(w*h/256) blocks). */
typedef struct LigCrop {
    const float *level2;
    const float *f, *fx, *fy, *fxy;
    float *out;
    int32_t view_w, off_x, off_y, w, h, mode, src_h, src_w;
    float roi_x1, roi_y1, roi_x2, roi_y2;
    int32_t full_h, full_w, win_x, win_y;
    float span, lo;
    int32_t block_off, reserved;
} LigCrop;

/* No reference counterpart: the crops a server has queued rendered as one
 * batch, one launch per stage.  Views may come from different
 * representations.  lig_project_views_batched: lig_project_gaussians_view for
 * every view of the device array `views` (sorted by gauss_off, ranges back to
 * back from 0, n_total = their sum; no opacities), bit for bit, and, when
 * colors is not NULL, the rgbs of each slot whose Gaussian hits a tile copied
 * to colors [n_total,3] (other slots are not written: nothing reads them); a
 * view with a non-positive size or scale emits nothing (validate on the host).
 * Binning: lig_bin_batched with the views as the batch's images.
 * lig_render_batched: the image-only sum render of every image of the batch
 * table (lig_fit_render_l2_batched's image, without target or loss), the
 * same bits as lig_rasterize_forward's image of each view alone; it builds
 * the records of Gaussian 0 and of the Gaussians with tile intersections
 * only; batch_depth as for lig_fit_render_l2_batched (a speed setting);
 * workspace lig_fit_batched_workspace_size(num_tiles).  lig_compose_crops: the crops
 * of the device array `crops` in one launch of total_blocks 256-pixel
 * blocks; the same bits as the per-crop chain (upscale or resize, clamps and
 * train.py's composition as FP32 tensor operations). */
int lig_project_views_batched(int num_views, const LigView *views, int n_total,
                              int block_width, float *xys, float *extents, float *conics,
                              int32_t *num_tiles_hit, float *colors, void *stream);
int lig_render_batched(int n, const int32_t *gaussian_ids_sorted, const int32_t *tile_bins,
                       const float *xys, const float *extents, const int32_t *cum_tiles_hit,
                       const float *conics, const float *colors, void *packed,
                       const void *batch_table, int num_images, int num_tiles, float *out_img,
                       int batch_depth, void *workspace, size_t workspace_bytes, void *stream);
int lig_compose_crops(int num_crops, const LigCrop *crops, int total_blocks, void *stream);

/* ---- captured fit step (no host read of the intersection count) ------------ */

/* No reference counterpart (the reference reads the count on the host every
 * step, utils.py:65).  For a fit step recorded once as a hipGraph and
 * replayed: buffers are sized for `capacity` intersections and the count is
 * read on the device.
 * lig_bin_batched_capacity: lig_bin_batched with the count taken from
 * cum_tiles_hit[n-1] on the device; workspace
 * lig_bin_workspace_size(capacity, num_tiles), gaussian_ids_sorted holds
 * capacity entries.  When the count exceeds the capacity nothing is emitted
 * and every bin is empty.  lig_fit_backward_records_l2_batched and
 * lig_fit_reduce_records take the capacity as num_intersects.
 * `control` is 4 device int32 {step, skipped, max_count, ok}.
 * lig_fit_step_begin: ok = (*count <= capacity && step < table_len); then
 * step += 1 when ok, else skipped += 1; max_count = max(max_count, *count).
 * Launch it before the optimiser steps of the same fit step.
 * lig_adam_step_device: lig_adam_step with the step number read from
 * control (after lig_fit_step_begin) and the bias corrections from
 * bias_table[2*(step-1) .. +1] (device floats written by
 * lig_adam_bias_corrections, so they equal lig_adam_step's bit for bit);
 * does nothing when ok == 0.  A step that was not admitted changes no
 * parameter or optimiser state.
 * lig_adam_bias_corrections: host function; out[2k], out[2k+1] = the
 * (1 - beta1^t, sqrt(1 - beta2^t)) lig_adam_step uses for t = first_step + k. */
int lig_bin_batched_capacity(int n, int capacity, const float *xys, const float *extents,
                             const int32_t *cum_tiles_hit, const void *batch_table,
                             int num_images, int num_tiles, int32_t *gaussian_ids_sorted,
                             int32_t *tile_bins, void *workspace, size_t workspace_bytes,
                             void *stream);
int lig_fit_step_begin(const int32_t *count, int capacity, int table_len, void *control,
                       void *stream);
int lig_adam_step_device(long long n, float *param, const float *grad, float *exp_avg,
                         float *exp_avg_sq, float lr, float beta1, float beta2, float eps,
                         float weight_decay, const void *control, const float *bias_table,
                         void *stream);
int lig_adam_bias_corrections(float beta1, float beta2, int first_step, int count, float *out);

/* No reference counterpart: workload counters for roofline accounting.
 * counts[0] = E, pixel/Gaussian pairs whose pixel centre lies inside the
 * Gaussian's clipped SnugBox box |p - mean| <= extent (Gaussians with
 * num_tiles_hit > 0); counts[1] = C, the pairs of E with sigma >= 0 and
 * alpha >= 1/255 (evaluated in FP64).  Independent of the tile size. */
int lig_count_pairs(int n, int img_height, int img_width, const float *xys,
                    const float *extents, const float *conics, const float *opacities,
                    const int32_t *num_tiles_hit, unsigned long long *counts, void *stream);

/* ---- fused loss ---------------------------------------------------------- */

/* Fused clamp + mean-squared-error of the fit step, replacing the caller's
 * torch.clamp(img, 0, 1) followed by F.mse_loss (gsplat2d caller
 * gaussianlig.py:95-103, utils.py:21-22).  n = number of floats.
 * forward: loss[0] = mean((clamp(img,0,1) - target)^2), deterministic;
 * workspace lig_l2_loss_workspace_size(n).  backward: grad_img =
 * (0 <= img <= 1) ? (2/n)(img - target) * grad_loss[0] : 0 (grad_loss is a
 * device scalar). */
size_t lig_l2_loss_workspace_size(long long n);
int lig_l2_clamped_loss_forward(long long n, const float *img, const float *target, float *loss,
                                void *workspace, size_t workspace_bytes, void *stream);
int lig_l2_clamped_loss_backward(long long n, const float *img, const float *target,
                                 const float *grad_loss, float *grad_img, void *stream);

/* ---- fused optimiser steps (one FP32 tensor per call) ----------------------- */

/* No reference counterpart in the CUDA extension: the optimiser step that
 * follows the backward (gaussianlig.py:64-65 uses torch Adam;
 * minimal_trainer_upscale.py:205-210 uses the reference's Adan).
 * lig_adam_step: torch.optim.Adam's update (decoupled from nothing: weight
 * decay is added to the gradient), step = 1-based count after this update.
 * lig_adan_step: optimizer.py:238-293 (_single_tensor_adan) per element, with
 * neg_pre_grad the state that holds -grad of the previous step (initialised
 * by the call itself when step == 1) and clip_global_grad_norm the factor the
 * caller computed from max_grad_norm (1 when unused).  grad is not modified. */
int lig_adam_step(long long n, float *param, const float *grad, float *exp_avg,
                  float *exp_avg_sq, float lr, float beta1, float beta2, float eps,
                  float weight_decay, int step, void *stream);
int lig_adan_step(long long n, float *param, const float *grad, float *exp_avg,
                  float *exp_avg_sq, float *exp_avg_diff, float *neg_pre_grad, float lr,
                  float beta1, float beta2, float beta3, float eps, float weight_decay,
                  int no_prox, float clip_global_grad_norm, int step, void *stream);

/* ---- gradient-aware upscaler ---------------------------------------------- */

/* Replaces torch.nn.functional.interpolate(mode="bilinear", align_corners=False)
 * of train.py's level composition (train.py:103-105, 193): src [src_h,src_w,3]
 * resized to a virtual full_h x full_w image, of which the out_h x out_w
 * window at (off_y, off_x) is written to output [out_h,out_w,3] (the whole
 * image when the window is).  Destination index d samples source coordinate
 * max((src/full) * (d + 0.5) - 0.5, 0), the upper neighbour clamped to the
 * last index.  The window must lie inside the full image. */
int lig_bilinear_resize(int src_h, int src_w, int full_h, int full_w, int off_y, int off_x,
                        int out_h, int out_w, const float *src, float *output, void *stream);

/* Replaces gradient_aware_upscale_forward_tensor (csrc/bindings.cu:508-559,
 * kernel csrc/gradient_aware_upscaler.cu:181-246).  render/dx/dy/dxy
 * [src_h,src_w,3] -> output [dst_h,dst_w,3]; roi = (x1,y1,x2,y2), x2>=x1, y2>=y1. */
int lig_gradient_aware_upscale_forward(int src_h, int src_w, int dst_h, int dst_w,
                                       float roi_x1, float roi_y1, float roi_x2,
                                       float roi_y2, const float *render, const float *dx,
                                       const float *dy, const float *dxy, float *output,
                                       void *stream);

/* Replaces gradient_aware_upscale_backward_tensor (csrc/bindings.cu:561-605,
 * kernel csrc/gradient_aware_upscaler.cu:248-371) with a deterministic
 * two-pass separable gather (no atomics).  -> grad_render/dx/dy/dxy
 * [src_h,src_w,3].  Workspace: lig_upscale_backward_workspace_size(...). */
size_t lig_upscale_backward_workspace_size(int src_h, int src_w, int dst_h, int dst_w);
int lig_gradient_aware_upscale_backward(int src_h, int src_w, int dst_h, int dst_w,
                                        float roi_x1, float roi_y1, float roi_x2,
                                        float roi_y2, const float *grad_output,
                                        float *grad_render, float *grad_dx, float *grad_dy,
                                        float *grad_dxy, void *workspace,
                                        size_t workspace_bytes, void *stream);

/* ---- image-quality metrics --------------------------------------------------- */

/* ---- bounds-checked build ---------------------------------------------------- */

/* No reference counterpart: a debugging aid.  A library built with LIG_CHECKED
 * (lig_amd/liblig_hip_checked.so) evaluates bounds checks at its index-heavy
 * device code (the sort's and the row expansion's LDS ranks and outputs, the
 * emission's owner lookups, the backward's row-group sort and record slots,
 * the launch order, the batched crop kernels); a failed check does not trap,
 * it is counted.  lig_check_status waits for the device, then returns the
 * failures since the last call, the first failing line and its source file,
 * and clears them; a normal build returns LIG_ERR_UNSUPPORTED. */
int lig_check_status(unsigned long long *failures, int *first_line, const char **first_file);

/* SSIM / MS-SSIM of two image batches, replacing the third-party
 * pytorch_msssim.ms_ssim / ssim the reference's evaluation and loss menu call
 * (train.py:200, utils.py:26-34).  x, y: [n, c, h, w] FP32 addressed through
 * element strides {n, c, h, w} (so NCHW and NHWC views both work, and x and y
 * may differ in layout).  win: win_size taps (odd, 3..11) applied separably
 * with no padding; C1 = (k1 data_range)^2, C2 = (k2 data_range)^2.  levels
 * scales (1..8) with 2 x 2 zero-padded average pools between them; every
 * scale must be at least win_size on both sides.  Outputs (any may be NULL,
 * not all): out_value[n*c] = prod_l relu(cs_l)^weights[l] with the last scale's
 * relu(ssim) in place of its cs; out_ssim / out_cs [levels][n*c] the per-scale
 * plane means.  win and weights are host arrays.  Deterministic.  Workspace:
 * lig_ssim_workspace_size(...), 0 when the sizes are invalid. */
size_t lig_ssim_workspace_size(int n, int c, int h, int w, int win_size, int levels);
int lig_ssim_multiscale(int n, int c, int h, int w, const float *x, const long long *x_strides,
                        const float *y, const long long *y_strides, float data_range,
                        int win_size, const float *win, float k1, float k2, int levels,
                        const float *weights, float *out_value, float *out_ssim, float *out_cs,
                        void *workspace, size_t workspace_bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* LIG_ABI_H */
