// Shared device/host helpers for the gfx950 gsplat2d kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lig_abi.h"

#define LIG_WAVE 64

namespace lig {

// Hand-off of LDS data between the lanes of one wave: every LDS write the
// wave issued before the call is visible to every lane's LDS reads after it.
// The wave barrier alone only keeps the compiler from moving code across it;
// the wavefront-scope release/acquire pair makes the ordering part of the
// memory model (also for accesses through generic pointers), not a
// consequence of LDS operations completing in issue order.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Bounds checks.  In a library built with LIG_CHECKED, LIG_CHECK(cond)
// evaluates cond in device code and, when it is false, counts the failure
// and records the first failing line in this translation unit's device state
// (no trap: the kernel carries on); lig_check_status() collects and clears the
// state of every unit.  In a normal build LIG_CHECK compiles to nothing.
typedef int (*CheckReader)(unsigned *count, unsigned *line, int clear);
int register_check_reader(const char *file, CheckReader reader);

}  // namespace lig

#ifdef LIG_CHECKED
static __device__ unsigned int lig_check_state[2];  // failures, first failing line
#define LIG_CHECK(cond)                                                   \
    do {                                                                  \
        if (!(cond)) {                                                    \
            atomicAdd(&lig_check_state[0], 1u);                           \
            atomicCAS(&lig_check_state[1], 0u, (unsigned)__LINE__);       \
        }                                                                 \
    } while (0)
static int lig_check_read(unsigned *count, unsigned *line, int clear) {
    unsigned st[2] = {0u, 0u};
    if (hipMemcpyFromSymbol(st, HIP_SYMBOL(lig_check_state), sizeof(st), 0,
                            hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    *count = st[0];
    *line = st[1];
    if (clear && st[0]) {
        const unsigned zero[2] = {0u, 0u};
        if (hipMemcpyToSymbol(HIP_SYMBOL(lig_check_state), zero, sizeof(zero), 0,
                              hipMemcpyHostToDevice) != hipSuccess)
            return -1;
    }
    return 0;
}
static const int lig_check_registered = lig::register_check_reader(__BASE_FILE__,
                                                                   lig_check_read);
#else
#define LIG_CHECK(cond) ((void)0)
#endif

namespace lig {

// Status of the last launch on this thread, mapped to the ABI codes.
inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LIG_OK : LIG_ERR_LAUNCH;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// IEEE FP32 helpers with a fixed meaning regardless of compiler flags:
// division and square root are computed in FP64 and rounded once (correctly
// rounded FP32 results, since FP64 carries more than 2*24+2 bits), the cast
// to int truncates and saturates with NaN -> 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float div_rn(float a, float b) {
    return (float)((double)a / (double)b);
}
__device__ __forceinline__ float sqrt_rn(float a) { return (float)__builtin_sqrt((double)a); }
__device__ __forceinline__ float log_rd(float a) { return (float)log((double)a); }
__device__ __forceinline__ int f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int)x;
}

// Tile-space bounding box of a Gaussian: inclusive min, exclusive max, clamped
// to the tile grid (helpers.cuh:12-45 semantics).
__device__ __forceinline__ void tile_bbox(float cx, float cy, float ex, float ey, int tiles_x,
                                          int tiles_y, int block_width, int &x0, int &y0,
                                          int &x1, int &y1) {
#pragma clang fp contract(off)
    const float fb = (float)block_width;
    const float tcx = div_rn(cx, fb), tcy = div_rn(cy, fb);
    const float tex = div_rn(ex, fb), tey = div_rn(ey, fb);
    x0 = min(max(0, f2i_sat(tcx - tex)), tiles_x);
    x1 = min(max(0, f2i_sat(tcx + tex + 1.0f)), tiles_x);
    y0 = min(max(0, f2i_sat(tcy - tey)), tiles_y);
    y1 = min(max(0, f2i_sat(tcy + tey + 1.0f)), tiles_y);
}

// One image of a batched fit: its tiles are [tile_off, tile_off + tiles_x *
// tiles_y) of the batch's global tile space, its pixels [pix_off, pix_off +
// w * h) of the concatenated HWC buffers, its Gaussians [gauss_off, next
// image's gauss_off).  Coordinates inside an image are the image's own, so a
// batched image computes exactly what the same image fitted alone does.  The
// table has one entry per image plus a sentinel whose tile_off / gauss_off are
// the totals.  Tiles are 16 x 16.
struct BatchImage {
    int tile_off, tiles_x, tiles_y, w, h, gauss_off;
    long long pix_off;
};

// Index of the entry of `tab` (nimg real entries, sorted offsets) whose
// [off, next off) range holds `key`, for the field selected by OFF.
template <int OFF>
__device__ __forceinline__ int batch_find(const BatchImage *__restrict__ tab, int nimg, int key) {
    int lo = 0, hi = nimg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int v = OFF == 0 ? tab[mid].tile_off : tab[mid].gauss_off;
        if (v <= key)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Map a linear workgroup id to a work item so that each XCD (workgroups are
// dispatched round-robin over the 8 XCDs) receives one contiguous range of
// items: neighbouring tiles then share the XCD's L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nb) {
    const unsigned q = nb >> 3, r = nb & 7u, x = b & 7u, i = b >> 3;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Tile work queue of a persistent launch: one counter per XCD (zeroed before
// the launch, kQueueStride ints apart so that each sits in its own 256-byte
// line) over that XCD's share of the tiles.  Without a launch order (order
// null, or order[0] = -1) the shares are the contiguous split of xcd_remap;
// with one, the tiles are dealt round-robin in that order (heaviest first):
// item i of XCD x's share is order[8 i + x], so every XCD starts on its part
// of the heaviest tiles.  A wave (blocks are dispatched to XCDs round-robin)
// takes kQueueGrab items at a time from its own XCD's share while any is
// left, then from the other shares in turn, so an XCD whose share runs long
// is helped by the others.  The next grab is issued before the current tiles
// are processed.
constexpr int kQueueStride = 64;
constexpr int kQueueGrab = 2;
constexpr int kQueueInts = 8 * kQueueStride;

struct TileQueue {
    int *ctr;  // kQueueInts ints
    int ntiles;
    const int32_t *order;  // launch order or null
};

__device__ __forceinline__ void xcd_share(int x, int nt, int &start, int &size) {
    const int q = nt >> 3, r = nt & 7;
    size = q + (x < r ? 1 : 0);
    start = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
}

// Lane 0 takes the next grab of its own XCD's share (the raw counter value;
// -1 on the other lanes).
__device__ __forceinline__ int queue_issue(const TileQueue &q) {
    int i = -1;
    if ((threadIdx.x & (LIG_WAVE - 1)) == 0)
        i = atomicAdd(&q.ctr[(blockIdx.x & 7) * kQueueStride], kQueueGrab);
    return i;
}

// First item and item count of the grab `own` (from queue_issue), or of a
// grab taken from another XCD's share when the own share is done, and the
// share's XCD; -1 when every share is taken.  Wave-uniform results.
__device__ __forceinline__ int queue_resolve(const TileQueue &q, int own, int &count, int &xcd) {
    int t = -1, c = 0, xs = 0;
    if ((threadIdx.x & (LIG_WAVE - 1)) == 0) {
        const int x0 = blockIdx.x & 7;
        int start, size;
        xcd_share(x0, q.ntiles, start, size);
        if (own < size) {
            t = own;
            c = min(kQueueGrab, size - own);
            xs = x0;
        } else {
            for (int k = 1; k < 8; ++k) {
                const int x = (x0 + k) & 7;
                xcd_share(x, q.ntiles, start, size);
                int *ctr = &q.ctr[x * kQueueStride];
                if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= size)
                    continue;  // share already taken
                const int i = atomicAdd(ctr, kQueueGrab);
                if (i < size) {
                    t = i;
                    c = min(kQueueGrab, size - i);
                    xs = x;
                    break;
                }
            }
        }
    }
    count = __builtin_amdgcn_readfirstlane(c);
    xcd = __builtin_amdgcn_readfirstlane(xs);
    return __builtin_amdgcn_readfirstlane(t);
}

// Tile of item i of XCD x's share.
__device__ __forceinline__ int queue_tile(const TileQueue &q, int x, int i) {
    if (q.order && q.order[0] >= 0) return q.order[8 * i + x];
    int start, size;
    xcd_share(x, q.ntiles, start, size);
    return start + i;
}

// Per-Gaussian record the rasteriser kernels read (64 bytes, 16-byte aligned):
//   p0 = (mean.x, mean.y, a', 2b')  p1 = (c', opacity, red, green)
//   p2 = (blue, a, b, c)            p3 = (cull_x, cull_y, slot base, slot width)
// with (a, b, c) the conic and (a', b', c') = K * (a, b, c), K = log2(e)/2, so
// that with d = mean - pixel, u' = a'dx + b'dy, v' = b'dx + c'dy:
//   sigma * log2(e) = dx*u' + dy*v'   and   alpha = o * 2^-(dx*u' + dy*v').
// (cull_x, cull_y) bound |pixel - mean| beyond which alpha < 1/255 is certain
// (infinite when the conic is not positive definite, negative when the
// Gaussian can never reach 1/255); kernels may skip pairs outside that box.
// The slot fields (int bits; 0 and 1 unless packed for the fit step) give the
// emission slot of the Gaussian's intersection with tile (tx, ty) as
// base + ty * width + tx.
struct __align__(16) GaussPack {
    float4 p0, p1, p2, p3;
};

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kHalfLog2e = 0.5f * kLog2e;
constexpr float kAlphaMin = 1.0f / 255.0f;
// log2(255): alpha >= 1/255 without opacity needs exp2(-s) >= 1/255
constexpr float kAcceptLog2 = 7.994353436858858f;

// Accept test of a pair without per-Gaussian opacity.  The reference keeps a
// pair unless sigma < 0 or alpha = min(exp(-sigma), clamp) < 1/255
// (forward.cu:254-263, backward.cu:105-111); with s = sigma * log2(e) and a
// clamp above 1/255 that is 0 <= s <= log2(255), decided here on s itself: s
// in [+0, kAcceptS] with kAcceptS = 0x40ffd1be = 7.99435329..., the largest
// FP32 value not above log2(255) (exp2(-kAcceptS) * 255 = 1 + 9.9e-8, the
// next value gives 1 - 2.3e-7), so the cut sits within 2.3e-7 of alpha =
// 1/255 instead of within the device exp2's rounding of it.  Comparing the bit
// patterns as unsigned integers takes one instruction and also rejects -0 and
// NaN (the reference, through fminf(NaN, clamp) = clamp, would keep a NaN
// pair at full alpha).  Every kernel without opacities uses this test, so the
// forward and every backward accept the same pairs.
constexpr uint32_t kAcceptSBits = 0x40ffd1beu;
__device__ __forceinline__ bool accept_s(float s) {
    return __float_as_uint(s) <= kAcceptSBits;
}

// sigma*log2(e) of a pair from the scaled conic (a', b', c'), evaluated with
// explicit FMAs so that every kernel computes bit-identical values for the
// same pair (the forward's accept test and the backward's must agree):
//   s = fma(dx, fma(a', dx, b2dy), c2dy),  b2dy = (2b') * dy,  c2dy = (c' * dy) * dy.
// The row terms depend on dy only and are hoisted out of loops over a row
// (row_terms takes 2b' as packed); bdy = b' dy and cdy = c' dy serve the
// backward's u' = a'dx + b'dy and v' = b'dx + c'dy.
struct RowTerms {
    float bdy, cdy;    // b' dy, c' dy
    float b2dy, c2dy;  // 2 b' dy, c' dy^2
};

__device__ __forceinline__ RowTerms row_terms(float bp2, float cp, float dy) {
    RowTerms r;
    r.b2dy = bp2 * dy;
    r.bdy = r.b2dy * 0.5f;  // exact
    r.cdy = cp * dy;
    r.c2dy = r.cdy * dy;
    return r;
}

__device__ __forceinline__ float pair_s(float ap, const RowTerms &r, float dx) {
    return __builtin_fmaf(dx, __builtin_fmaf(ap, dx, r.b2dy), r.c2dy);
}

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Wave-uniform value moved into a VGPR: an operand held in a VGPR issues at
// the full VALU rate where an SGPR operand does not, and LDS addresses built
// from it stay in one register with immediate offsets.
__device__ __forceinline__ int in_vgpr(int x) {
    int r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}
__device__ __forceinline__ float in_vgpr(float x) {
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

}  // namespace lig
