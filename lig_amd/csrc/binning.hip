// Tile intersection: int32 device scan, intersection emission, stable LSD radix
// sort, and tile bin edges.
//
// Ordering contract: the sorted order is a STABLE sort of the emission order,
// which is gid-major (Gaussian g's records are written at [cum[g-1], cum[g])
// in row-major tile order).  With the all-zero depths of the rasterisation
// path this is the (tile id, Gaussian id) ascending order.
//
// Reference behaviour followed (paths relative to the reference repository):
//   gsplat2d/gsplat2d/utils.py:61-100 (cumsum, .item, torch.sort, gather),
//   gsplat2d/gsplat2d/cuda/csrc/forward.cu:109-170, bindings.cu:239-301.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "lig_device.h"

namespace lig {
namespace {

// ---------------------------------------------------------------------------
// Device-wide int32 scan: reduce per block, scan the block sums (recursively),
// then scan each block with its offset.  2048 items per 256-thread block.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Items a kernel processes: `cap` when `dev` is NULL, else the count stored
// at *dev, or -1 when that count exceeds `cap` (the buffers were sized for
// `cap` items: the kernel then does nothing).
__device__ __forceinline__ int live_count(int cap, const int32_t *dev) {
    if (!dev) return cap;
    const int v = *dev;
    return v <= cap ? v : -1;
}

// Exclusive scan of one int per thread across a 256-thread block.  `wsum` is
// 4 ints of LDS.  Returns the exclusive prefix; *total gets the block sum.
template <int THREADS>
__device__ __forceinline__ int block_exclusive_scan_n(int v, int *wsum, int *total) {
    const int lane = threadIdx.x & (LIG_WAVE - 1);
    const int wave = threadIdx.x / LIG_WAVE;
    int x = v;
#pragma unroll
    for (int off = 1; off < LIG_WAVE; off <<= 1) {
        const int y = __shfl_up(x, off, LIG_WAVE);
        if (lane >= off) x += y;
    }
    if (lane == LIG_WAVE - 1) wsum[wave] = x;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < THREADS / LIG_WAVE; ++w) {
        const int s = wsum[w];
        before += w < wave ? s : 0;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__device__ __forceinline__ int block_exclusive_scan(int v, int *wsum, int *total) {
    const int lane = threadIdx.x & (LIG_WAVE - 1);
    const int wave = threadIdx.x / LIG_WAVE;
    int x = v;
#pragma unroll
    for (int off = 1; off < LIG_WAVE; off <<= 1) {
        const int y = __shfl_up(x, off, LIG_WAVE);
        if (lane >= off) x += y;
    }
    if (lane == LIG_WAVE - 1) wsum[wave] = x;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / LIG_WAVE; ++w) {
        const int s = wsum[w];
        before += w < wave ? s : 0;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const int32_t *__restrict__ in,
                                                                   int n,
                                                                   int32_t *__restrict__ sums) {
    __shared__ int wsum[kScanThreads / LIG_WAVE];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const size_t i = base + (size_t)k * kScanThreads + threadIdx.x;
        s += i < (size_t)n ? in[i] : 0;
    }
    int total;
    block_exclusive_scan(s, wsum, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Scans one tile of 2048 items (in place allowed) and adds offsets[block].
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(
    const int32_t *in, int n, int inclusive, const int32_t *__restrict__ offsets,
    int32_t *out) {
    __shared__ int tile[kScanTile];
    __shared__ int wsum[kScanThreads / LIG_WAVE];
    const size_t base = (size_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int j = k * kScanThreads + threadIdx.x;
        const size_t i = base + j;
        tile[j] = i < (size_t)n ? in[i] : 0;
    }
    __syncthreads();
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = tile[threadIdx.x * kScanItems + k];
        s += v[k];
    }
    int total;
    int run = block_exclusive_scan(s, wsum, &total);
    run += offsets ? offsets[blockIdx.x] : 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int incl = run + v[k];
        tile[threadIdx.x * kScanItems + k] = inclusive ? incl : run;
        run = incl;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int j = k * kScanThreads + threadIdx.x;
        const size_t i = base + j;
        if (i < (size_t)n) out[i] = tile[j];
    }
}

size_t scan_ws_elems(size_t n) {
    if (n <= (size_t)kScanTile) return 0;
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    return nb + scan_ws_elems(nb);
}

// Scans n int32 (exclusive unless `inclusive`); `ws` holds scan_ws_elems(n).
void scan_i32(const int32_t *in, int32_t *out, size_t n, bool inclusive, int32_t *ws,
              hipStream_t st) {
    if (n == 0) return;
    if (n <= (size_t)kScanTile) {
        hipLaunchKernelGGL(scan_apply_kernel, dim3(1), dim3(kScanThreads), 0, st, in, (int)n,
                           (int)inclusive, (const int32_t *)nullptr, out);
        return;
    }
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    int32_t *sums = ws;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in,
                       (int)n, sums);
    scan_i32(sums, sums, nb, false, ws + nb, st);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in,
                       (int)n, (int)inclusive, (const int32_t *)sums, out);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, uint32 value) pairs, RB-bit digits (RB <= 10).
// Each 512-thread block owns 4096 consecutive pairs: a histogram kernel
// counts digits per block (digit-major table), an exclusive scan turns the
// table into output offsets, and a scatter kernel ranks each key stably
// (wave ballots for lane order, LDS counters for wave and round order).
// A NULL value array stands for the identity (value = input index).
// ---------------------------------------------------------------------------
// 8 waves per block, 1024 pairs per wave.  With digits of at most 7 bits
// (digit_plan) a block's run of one digit averages 64 or more pairs, so the
// scatter writes whole cache lines; at 79 M (18-bit keys) and 39 M (14-bit)
// intersections this beat 4096-pair blocks with 9- and 10-bit digits by 19%
// and 11%, and 10240-pair blocks (one block per CU) by 20% and 35%.
constexpr int kSortThreads = 512;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;
constexpr int kSortWaves = kSortThreads / LIG_WAVE;

template <typename K, int RB>
__device__ __forceinline__ unsigned digit_of(K key, int shift, K flip) {
    return (unsigned)(((key ^ flip) >> shift) & (K)((1u << RB) - 1));
}

template <typename K, int RB>
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const K *__restrict__ keys,
                                                                  int cap, const int32_t *n_dev,
                                                                  int shift, K flip, int nblocks,
                                                                  int32_t *__restrict__ hist) {
    constexpr int NB = 1 << RB;
    const int n = live_count(cap, n_dev);
    if (n < 0) return;
    // one histogram per wave (fewer same-address LDS atomics); 32-bit keys
    // of full blocks are read 4 at a time
    __shared__ int h[kSortWaves][NB];
    for (int d = threadIdx.x; d < kSortWaves * NB; d += kSortThreads) (&h[0][0])[d] = 0;
    __syncthreads();
    int *hw = h[threadIdx.x / LIG_WAVE];
    const size_t base = (size_t)blockIdx.x * kSortTile;
    if (sizeof(K) == 4 && base + kSortTile <= (size_t)n) {
        const uint4 *k4 = reinterpret_cast<const uint4 *>(keys + base);
        uint4 v[kSortRounds / 4];
#pragma unroll
        for (int r = 0; r < kSortRounds / 4; ++r) v[r] = k4[r * kSortThreads + threadIdx.x];
#pragma unroll
        for (int r = 0; r < kSortRounds / 4; ++r) {
            atomicAdd(&hw[digit_of<K, RB>((K)v[r].x, shift, flip)], 1);
            atomicAdd(&hw[digit_of<K, RB>((K)v[r].y, shift, flip)], 1);
            atomicAdd(&hw[digit_of<K, RB>((K)v[r].z, shift, flip)], 1);
            atomicAdd(&hw[digit_of<K, RB>((K)v[r].w, shift, flip)], 1);
        }
    } else {
#pragma unroll 4
        for (int r = 0; r < kSortRounds; ++r) {
            const size_t i = base + (size_t)r * kSortThreads + threadIdx.x;
            if (i < (size_t)n) atomicAdd(&hw[digit_of<K, RB>(keys[i], shift, flip)], 1);
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < NB; d += kSortThreads) {
        int s = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) s += h[w][d];
        hist[(size_t)d * nblocks + blockIdx.x] = s;
    }
}

// Scatter of one pass.  Each wave owns 512 consecutive pairs of the block's
// 4096 and ranks them round by round (64 at a time) against wave-private
// digit counters in LDS: lanes with equal digits are matched by RB ballots,
// the round's leader per digit advances the counter.  No block barrier is
// needed until all rounds are ranked; then the per-wave counts are turned into
// per-(wave, digit) starts, the pairs are placed in LDS in (digit, input
// order) order, and written out so that each digit's run goes to consecutive
// global addresses.  Order within a digit is wave, round, lane: input order,
// so the pass is stable.
template <typename K, int RB>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const K *__restrict__ keys_in, const uint32_t *__restrict__ vals_in,
    K *__restrict__ keys_out, uint32_t *__restrict__ vals_out,
    const int32_t *__restrict__ offsets, int cap, const int32_t *n_dev, int shift, K flip,
    int nblocks) {
    constexpr int NB = 1 << RB;
    const int n = live_count(cap, n_dev);
    if (n < 0) return;
    constexpr int PER = NB / kSortThreads > 0 ? NB / kSortThreads : 1;  // digits per thread
    constexpr int WSPAN = kSortRounds * LIG_WAVE;                        // pairs per wave
    __shared__ K skey[kSortTile];
    __shared__ uint32_t sval[kSortTile];
    __shared__ int wcnt[kSortWaves][NB];  // per-wave digit counts, then per-wave digit starts
    __shared__ int lstart[NB];            // block-local exclusive start of each digit
    __shared__ int wsum[kSortWaves];
    const int tid = threadIdx.x;
    const int lane = tid & (LIG_WAVE - 1);
    const int wave = tid / LIG_WAVE;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (LIG_WAVE - lane));
    const size_t base = (size_t)blockIdx.x * kSortTile;
    // blocks past a device-side count get no pairs
    const int count = (size_t)n > base ? (int)min((size_t)kSortTile, (size_t)n - base) : 0;
    const int wbase = wave * WSPAN;

    for (int d = tid; d < NB; d += kSortThreads) {
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) wcnt[w][d] = 0;
    }
    K key[kSortRounds];
    uint32_t val[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const int j = wbase + r * LIG_WAVE + lane;
        key[r] = 0;
        val[r] = 0;
        if (j < count) {
            key[r] = keys_in[base + j];
            val[r] = vals_in ? vals_in[base + j] : (uint32_t)(base + j);
        }
    }
    __syncthreads();
    // wave-local stable ranking: lpos = position among the wave's pairs of the same digit
    int lpos[kSortRounds];
    int *cnt_w = wcnt[wave];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        lpos[r] = 0;
        if (wbase + r * LIG_WAVE >= count) continue;  // wave-uniform
        const bool valid = wbase + r * LIG_WAVE + lane < count;
        const unsigned d = valid ? digit_of<K, RB>(key[r], shift, flip) : 0u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const int bit = (d >> b) & 1;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const int rank = __popcll(m & lt_mask);
        const int before = valid ? cnt_w[d] : 0;
        lpos[r] = before + rank;
        // one lane per digit advances the wave's counter; the hand-off below
        // makes the new value visible to the next round
        if (valid && rank == 0) cnt_w[d] = before + __popcll(m);
        wave_lds_sync();
    }
    __syncthreads();
    // block exclusive scan of the digit totals; per-(wave, digit) starts
    int c[PER], sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int d = tid * PER + q;
        int t = 0;
        if (d < NB) {
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) t += wcnt[w][d];
        }
        c[q] = t;
        sum += t;
    }
    int total;
    int pre = block_exclusive_scan_n<kSortThreads>(sum, wsum, &total);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int d = tid * PER + q;
        if (d < NB) {
            lstart[d] = pre;
            int run = pre;
#pragma unroll
            for (int w = 0; w < kSortWaves; ++w) {
                const int t = wcnt[w][d];
                wcnt[w][d] = run;
                run += t;
            }
        }
        pre += c[q];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        if (wbase + r * LIG_WAVE + lane < count) {
            const int pos = cnt_w[digit_of<K, RB>(key[r], shift, flip)] + lpos[r];
            LIG_CHECK(pos >= 0 && pos < count);
            skey[pos] = key[r];
            sval[pos] = val[r];
        }
    }
    // global position of local slot j of digit d: offsets[d] + (j - lstart[d])
    for (int d = tid; d < NB; d += kSortThreads)
        lstart[d] = offsets[(size_t)d * nblocks + blockIdx.x] - lstart[d];
    __syncthreads();
    for (int j = tid; j < count; j += kSortThreads) {
        const K k = skey[j];
        const int pos = lstart[digit_of<K, RB>(k, shift, flip)] + j;
        LIG_CHECK(pos >= 0 && pos < n);
        keys_out[pos] = k;
        vals_out[pos] = sval[j];
    }
}

template <typename K>
__global__ void copy_pairs_kernel(const K *__restrict__ ki, const uint32_t *__restrict__ vi,
                                  K *__restrict__ ko, uint32_t *__restrict__ vo, int cap,
                                  const int32_t *n_dev) {
    const int n = live_count(cap, n_dev);
    if (n < 0) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (size_t)n) {
        ko[i] = ki[i];
        vo[i] = vi ? vi[i] : (uint32_t)i;
    }
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Digit plan for `key_bits` bits: up to 10 bits in one pass (small tile
// grids, where launches dominate), else the fewest passes of at most 7 bits
// with the bits split evenly (18 -> 3 x 6, 14 -> 2 x 7, 11 -> 2 x 6);
// 64-bit keys use 8 x 8.
void digit_plan(int key_bits, int *passes, int *rb) {
    if (key_bits > 32) {
        *passes = (key_bits + 7) / 8;
        *rb = 8;
        return;
    }
    *passes = key_bits <= 10 ? (key_bits > 0 ? 1 : 0) : (key_bits + 6) / 7;
    *rb = *passes ? (key_bits + *passes - 1) / *passes : 0;
    if (*rb < 4 && *passes) *rb = 4;
}

template <typename K>
size_t sort_ws_bytes(size_t m, int key_bits) {
    int passes, rb;
    digit_plan(key_bits, &passes, &rb);
    const size_t nblocks = (m + kSortTile - 1) / kSortTile;
    const size_t nh = ((size_t)1 << (rb > 0 ? rb : 1)) * (nblocks ? nblocks : 1);
    return align_up(m * sizeof(K)) + align_up(m * 4) + align_up(nh * 4) +
           align_up(scan_ws_elems(nh) * 4) + 256;
}

template <typename K, int RB>
void sort_pass(const K *ksrc, const uint32_t *vsrc, K *kdst, uint32_t *vdst, size_t m,
               const int32_t *m_dev, int shift, K flip, int nblocks, int32_t *hist,
               int32_t *scan_ws, hipStream_t st) {
    hipLaunchKernelGGL((radix_hist_kernel<K, RB>), dim3(nblocks), dim3(kSortThreads), 0, st, ksrc,
                       (int)m, m_dev, shift, flip, nblocks, hist);
    scan_i32(hist, hist, ((size_t)1 << RB) * nblocks, false, scan_ws, st);
    hipLaunchKernelGGL((radix_scatter_kernel<K, RB>), dim3(nblocks), dim3(kSortThreads), 0, st,
                       ksrc, vsrc, kdst, vdst, (const int32_t *)hist, (int)m, m_dev, shift, flip,
                       nblocks);
}

// Sorts m pairs by the low `key_bits` bits of the key (XOR `flip` first: the
// sign bit for signed 64-bit keys).  Input is read only; vals_in may be NULL
// (identity values); the result lands in keys_out / vals_out.  With m_dev,
// m is the capacity and the pairs sorted are the first *m_dev (none when
// *m_dev > m).
template <typename K>
int radix_sort(const K *keys_in, const uint32_t *vals_in, K *keys_out, uint32_t *vals_out,
               size_t m, int key_bits, K flip, char *ws, hipStream_t st,
               const int32_t *m_dev = nullptr) {
    if (m == 0) return LIG_OK;
    int passes, rb;
    digit_plan(key_bits, &passes, &rb);
    if (passes == 0) {
        hipLaunchKernelGGL(copy_pairs_kernel<K>, dim3(ceil_div(m, 256)), dim3(256), 0, st,
                           keys_in, vals_in, keys_out, vals_out, (int)m, m_dev);
        return launch_status();
    }
    const int nblocks = (int)((m + kSortTile - 1) / kSortTile);
    char *p = ws;
    K *ktmp = reinterpret_cast<K *>(p);
    p += align_up(m * sizeof(K));
    uint32_t *vtmp = reinterpret_cast<uint32_t *>(p);
    p += align_up(m * 4);
    int32_t *hist = reinterpret_cast<int32_t *>(p);
    p += align_up(((size_t)1 << rb) * nblocks * 4);
    int32_t *scan_ws = reinterpret_cast<int32_t *>(p);
    const K *ksrc = keys_in;
    const uint32_t *vsrc = vals_in;
    for (int pass = 0; pass < passes; ++pass) {
        const bool to_out = ((passes - 1 - pass) % 2) == 0;
        K *kdst = to_out ? keys_out : ktmp;
        uint32_t *vdst = to_out ? vals_out : vtmp;
        const int shift = rb * pass;
        switch (rb) {
            case 4: sort_pass<K, 4>(ksrc, vsrc, kdst, vdst, m, m_dev, shift, flip, nblocks, hist, scan_ws, st); break;
            case 5: sort_pass<K, 5>(ksrc, vsrc, kdst, vdst, m, m_dev, shift, flip, nblocks, hist, scan_ws, st); break;
            case 6: sort_pass<K, 6>(ksrc, vsrc, kdst, vdst, m, m_dev, shift, flip, nblocks, hist, scan_ws, st); break;
            case 7: sort_pass<K, 7>(ksrc, vsrc, kdst, vdst, m, m_dev, shift, flip, nblocks, hist, scan_ws, st); break;
            case 8: sort_pass<K, 8>(ksrc, vsrc, kdst, vdst, m, m_dev, shift, flip, nblocks, hist, scan_ws, st); break;
            case 9: sort_pass<K, 9>(ksrc, vsrc, kdst, vdst, m, m_dev, shift, flip, nblocks, hist, scan_ws, st); break;
            default: sort_pass<K, 10>(ksrc, vsrc, kdst, vdst, m, m_dev, shift, flip, nblocks, hist, scan_ws, st); break;
        }
        ksrc = kdst;
        vsrc = vdst;
    }
    return launch_status();
}

// ---------------------------------------------------------------------------
// Emission and bin edges.
// ---------------------------------------------------------------------------
constexpr int kEmitThreads = 256;

// Reference layout: int64 key (tile << 32) | sign-extended depth bits.
__global__ __launch_bounds__(kEmitThreads) void emit_isect_i64_kernel(
    int n, int m, const float *__restrict__ xys, const float *__restrict__ depths,
    const float *__restrict__ extents, const int32_t *__restrict__ cum, int tiles_x,
    int tiles_y, int block_width, int64_t *__restrict__ isect_ids,
    int32_t *__restrict__ gids) {
    const int g = blockIdx.x * kEmitThreads + threadIdx.x;
    if (g >= n) return;
    int cur = g == 0 ? 0 : cum[g - 1];
    const int end = min(cum[g], m);
    const float ex = extents[2 * g], ey = extents[2 * g + 1];
    if (ex <= 0.0f || ey <= 0.0f) {
        for (; cur < end; ++cur) {
            isect_ids[cur] = 0;
            gids[cur] = 0;
        }
        return;
    }
    int x0, y0, x1, y1;
    tile_bbox(xys[2 * g], xys[2 * g + 1], ex, ey, tiles_x, tiles_y, block_width, x0, y0, x1, y1);
    const int64_t depth_id = (int64_t)__float_as_int(depths[g]);
    for (int y = y0; y < y1; ++y)
        for (int x = x0; x < x1; ++x) {
            if (cur >= end) return;
            const int64_t tile = (int64_t)y * tiles_x + x;
            isect_ids[cur] = (int64_t)((uint64_t)tile << 32) | depth_id;
            gids[cur] = g;
            ++cur;
        }
}

// Rasterisation layout: 32-bit tile key per emission slot, plus slot_gid, the
// Gaussian id of each slot (0 for the slots of zero-extent Gaussians, as the
// reference's zero fill leaves them).  The sort's values are the slots
// themselves (implicit identity), so the sorted order maps back to each
// Gaussian's contiguous slot range.  Each wave computes the tile rectangles
// of its 64 Gaussians, parks them in LDS, then covers their joint slot range
// 64 slots at a time: a lane finds the Gaussian that owns its slot by
// counting the range starts (non-decreasing, as cum is) at or before it, in
// two levels (every eighth start from scalar registers, then the seven inside
// the chosen group of eight from LDS), decodes the tile from the slot's offset
// in that Gaussian's row-major rectangle (row by a reciprocal estimate with an
// exact correction), and every store is coalesced.
//
// Batched (tab != NULL, 16 x 16 tiles): each Gaussian's rectangle is taken in
// its own image's tile grid and its keys are offset by the image's first
// global tile.
__global__ __launch_bounds__(kEmitThreads) void emit_tile_keys_kernel(
    int n, int cap, const int32_t *m_dev, const float *__restrict__ xys,
    const float *__restrict__ extents, const int32_t *__restrict__ cum, int tiles_x, int tiles_y,
    int block_width, const BatchImage *__restrict__ tab, int nimg, uint32_t *__restrict__ keys,
    int32_t *__restrict__ slot_gid) {
    const int m = live_count(cap, m_dev);
    if (m < 0) return;
    const int lane = threadIdx.x & (LIG_WAVE - 1);
    const int g0 = blockIdx.x * kEmitThreads + (threadIdx.x - lane);
    if (g0 >= n) return;  // wave-uniform
    const int g = g0 + lane;
    // slots of the wave's Gaussians: [wstart, wend)
    const int wend = min(cum[min(g0 + LIG_WAVE, n) - 1], m);
    int start = wend, x0 = 0, y0 = 0, w = 0, row_len = tiles_x, toff = 0;
    if (g < n) {
        start = min(g == 0 ? 0 : cum[g - 1], m);
        int txs = tiles_x, tys = tiles_y;
        if (tab) {
            const BatchImage &im = tab[batch_find<1>(tab, nimg, g)];
            txs = im.tiles_x;
            tys = im.tiles_y;
            toff = im.tile_off;
            row_len = txs;
        }
        const float ex = extents[2 * g], ey = extents[2 * g + 1];
        if (ex > 0.0f && ey > 0.0f) {  // else w = 0: slots get key 0 and gid 0
            int x1, y1;
            tile_bbox(xys[2 * g], xys[2 * g + 1], ex, ey, txs, tys, block_width, x0, y0, x1, y1);
            w = max(x1 - x0, 1);
        }
    }
    // the wave's rectangles in LDS; only this wave reads its rows (handed
    // over by wave_lds_sync, no block barrier)
    constexpr int kWaves = kEmitThreads / LIG_WAVE;
    __shared__ __attribute__((aligned(16))) int s_start[kWaves][LIG_WAVE];  // read as int4
    __shared__ int4 s_rect[kWaves][LIG_WAVE];  // (start, x0, y0, w)
    __shared__ int4 s_grid[kWaves][LIG_WAVE];  // (row length, tile offset, 1/w bits, -)
    const int wv = threadIdx.x / LIG_WAVE;
    const float rcp = w > 0 ? __builtin_amdgcn_rcpf((float)w) : 0.0f;
    s_start[wv][lane] = start;
    s_rect[wv][lane] = make_int4(start, x0, y0, w);
    s_grid[wv][lane] = make_int4(row_len, toff, __float_as_int(rcp), 0);
    wave_lds_sync();
    // every eighth start, wave-uniform
    int b8[LIG_WAVE / 8];
#pragma unroll
    for (int i = 1; i < LIG_WAVE / 8; ++i) b8[i] = __builtin_amdgcn_readlane(start, 8 * i);
    const int wstart = __builtin_amdgcn_readfirstlane(start);
    for (int base = wstart; base < wend; base += LIG_WAVE) {
        const int pos = base + lane;
        // owner: the last lane whose range starts at or before pos (its range
        // ends where the next lane's starts, past pos); starts are
        // non-decreasing, so that is the count of starts 1..63 at or before pos
        int j8 = 0;
#pragma unroll
        for (int i = 1; i < LIG_WAVE / 8; ++i) j8 += b8[i] <= pos ? 8 : 0;
        const int4 lo4 = *reinterpret_cast<const int4 *>(&s_start[wv][j8]);
        const int4 hi4 = *reinterpret_cast<const int4 *>(&s_start[wv][j8 + 4]);
        const int j = j8 + (lo4.y <= pos) + (lo4.z <= pos) + (lo4.w <= pos) + (hi4.x <= pos) +
                      (hi4.y <= pos) + (hi4.z <= pos) + (hi4.w <= pos);
        LIG_CHECK(j >= 0 && j < LIG_WAVE);
        const int4 rc = s_rect[wv][j];
        const int4 gr = s_grid[wv][j];
        if (pos < wend) {
            LIG_CHECK(pos >= rc.x);
            uint32_t key = 0;
            int gid = 0;
            const int rw = rc.w;
            if (rw > 0) {
                const int k = pos - rc.x;
                // row = k / rw: the estimate is within one of it while k < 2^22,
                // and the exact division covers anything further off
                int row = (int)((float)k * __int_as_float(gr.z));
                int col = k - row * rw;
                if (col >= rw) {
                    ++row;
                    col -= rw;
                }
                if (col < 0) {
                    --row;
                    col += rw;
                }
                if ((unsigned)col >= (unsigned)rw) {
                    row = k / rw;
                    col = k - row * rw;
                }
                key = (uint32_t)gr.y + (uint32_t)(rc.z + row) * (uint32_t)gr.x +
                      (uint32_t)(rc.y + col);
                LIG_CHECK(col >= 0 && col < rw && row >= 0);
                gid = g0 + j;
            }
            keys[pos] = key;
            slot_gid[pos] = gid;
        }
    }
}

// Bin edges from the sorted tile keys plus the sorted Gaussian ids.
__global__ void finish_bins_kernel(int m, const uint32_t *__restrict__ keys_sorted,
                                   const uint32_t *__restrict__ slots_sorted,
                                   const int32_t *__restrict__ slot_gid, int num_tiles,
                                   int32_t *__restrict__ bins, int32_t *__restrict__ gids_sorted) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    gids_sorted[i] = slot_gid[slots_sorted[i]];
    const int cur = (int)keys_sorted[i];
    if (cur >= num_tiles) return;
    if (i == 0) bins[2 * cur] = 0;
    if (i == m - 1) bins[2 * cur + 1] = m;
    if (i == 0) return;
    const int prev = (int)keys_sorted[i - 1];
    if (prev != cur) {
        if (prev < num_tiles) bins[2 * prev + 1] = i;
        bins[2 * cur] = i;
    }
}

__global__ void fill_i32_kernel(int32_t *__restrict__ p, size_t count, int32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// Bin edges from tile ids sorted ascending (forward.cu:148-170 semantics on
// a zero-filled bins array).  Each thread handles 4 consecutive entries.
__device__ __forceinline__ void bin_edge_at(int i, int m, int prev, int cur, int num_tiles,
                                            int32_t *__restrict__ bins) {
    if (cur < 0 || cur >= num_tiles) return;
    if (i == 0) bins[2 * cur] = 0;
    if (i == m - 1) bins[2 * cur + 1] = m;
    if (i == 0) return;
    if (prev != cur) {
        if (prev >= 0 && prev < num_tiles) bins[2 * prev + 1] = i;
        bins[2 * cur] = i;
    }
}

// Bin of each key value (key & kmask) >> SHIFT in a sorted run.
template <typename K, int SHIFT>
__global__ void bin_edges_kernel(int cap, const int32_t *m_dev, const K *__restrict__ sorted,
                                 int num_tiles, int32_t *__restrict__ bins, K kmask) {
    const int m = live_count(cap, m_dev);
    if (m < 0) return;
    const size_t i0 = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= (size_t)m) return;
    int t[5];
    t[0] = i0 > 0 ? (int)((sorted[i0 - 1] & kmask) >> SHIFT) : -1;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        t[k + 1] = i0 + k < (size_t)m ? (int)((sorted[i0 + k] & kmask) >> SHIFT) : -1;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (i0 + k < (size_t)m) bin_edge_at((int)(i0 + k), m, t[k], t[k + 1], num_tiles, bins);
}

int bits_for(int num_tiles) {
    int b = 0;
    while (b < 31 && (1 << b) < num_tiles) ++b;
    return b;
}

// ---------------------------------------------------------------------------
// Row-expansion binning (one image, at most kRowMaxTiles tiles per row and
// kMaxRowsPacked rows).
//
// A Gaussian's tiles form a rectangle, so its intersections are laid out by
// tile rows: one row entry (row, gid) per tile row a Gaussian spans, about
// sqrt(tiles per Gaussian) times fewer than intersections (12.6 M against
// 79 M at 8192^2 x 2 M).  The entries are written straight from the
// rectangles in (row, gid) order (the direct row pass below, a counting
// sort by row); each entry's key carries its Gaussian's column range above
// the row bits, so no later pass gathers rectangles.  Each row's entries are
// split into P segments (P chosen by the host from the grid size); a
// counting pass writes, for every (row, segment, tile), how many of the
// segment's entries cover the tile.  Within a row, a prefix over segments per
// tile and a scan of the tile totals, plus a scan of the row totals, give
// every segment's first output position in every tile and the bins.  The
// scatter pass gives every (row, segment, window of kRowWin tiles) one wave:
// it filters the segment's entries down to those that reach the window, in
// gid order, and takes them 64 at a time, one per lane, through the window's
// tiles; per tile a ballot of the lanes covering it ranks them, and they
// store their gids at consecutive positions of the tile's bin.  The output is
// the stable (tile, emission) order of the radix path, bit for bit.
// ---------------------------------------------------------------------------
constexpr int kRowMaxTiles = 1024;     // tiles per row (10-bit column fields)
constexpr uint32_t kPhantomBit = 0x80000000u;

// Tile rectangle of a Gaussian: x = x0 | x1 << 16, y = y0 | y1 << 16, bit 31
// of y set for a zero-extent Gaussian with a tile (its one slot is binned as
// tile 0, gid 0: the reference's zero fill).
__device__ __forceinline__ void unpack_rect(uint2 rc, int &x0, int &y0, int &x1, int &y1,
                                            bool &ph) {
    x0 = (int)(rc.x & 0xFFFFu);
    x1 = (int)(rc.x >> 16);
    y0 = (int)(rc.y & 0xFFFFu);
    y1 = (int)((rc.y >> 16) & 0x7FFFu);
    ph = (rc.y & kPhantomBit) != 0;
}

// Row entry key: the row in the low `kshift` bits (enough for the image's
// rows), then the first column x0 (10 bits) and the width - 1 (10 bits).
__device__ __forceinline__ uint32_t row_key(int row, int x0, int x1, int kshift) {
    return (uint32_t)row | ((uint32_t)x0 << kshift) | ((uint32_t)(x1 - x0 - 1) << (kshift + 10));
}
__device__ __forceinline__ void rowkey_cols(uint32_t k, int kshift, int &x0, int &x1) {
    x0 = (int)((k >> kshift) & 1023u);
    x1 = x0 + (int)((k >> (kshift + 10)) & 1023u) + 1;
}

// Per Gaussian: its tile rectangle ((0, 0) without a slot; row 0, columns
// [0, 1) for a phantom slot).
__global__ __launch_bounds__(256) void rect_rows_kernel(int n, const float *__restrict__ xys,
                                                        const float *__restrict__ extents,
                                                        const int32_t *__restrict__ cum,
                                                        int tiles_x, int tiles_y,
                                                        int block_width,
                                                        uint2 *__restrict__ rect) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int cnt = cum[g] - (g == 0 ? 0 : cum[g - 1]);
    if (cnt <= 0) {
        rect[g] = make_uint2(0u, 0u);
        return;
    }
    const float ex = extents[2 * g], ey = extents[2 * g + 1];
    if (ex > 0.0f && ey > 0.0f) {
        int x0, y0, x1, y1;
        tile_bbox(xys[2 * g], xys[2 * g + 1], ex, ey, tiles_x, tiles_y, block_width, x0, y0, x1,
                  y1);
        rect[g] = make_uint2((uint32_t)x0 | ((uint32_t)x1 << 16),
                             (uint32_t)y0 | ((uint32_t)y1 << 16));
    } else {
        rect[g] = make_uint2(1u << 16, (1u << 16) | kPhantomBit);
    }
}

// Row entries straight from the rectangles, in (row, gid) order: the
// Gaussians are cut into segments of `gseg` consecutive gids; a counting pass
// gives every (segment, row) the number of the segment's Gaussians spanning
// the row, a prefix over segments per row and a scan of the row totals give
// each segment's first position in each row (and the row bins), and a
// scatter pass gives every (segment, window of kRowWin rows) one wave, which
// filters the segment's Gaussians down to those spanning a row of the window
// and, row by row, ranks them by ballot and writes their entries at
// consecutive positions.  An entry is key = row_key(row, x0, x1), value = gid
// (phantom bit set for a phantom slot, whose key is 0: row 0, columns
// [0, 1)); the order is that of a stable sort by row of the gid-major entries.
constexpr int kRowWin = 32;           // rows (tiles) per window of the scatter passes
constexpr int kRowQueue = 512;        // LDS ring of queued entries
constexpr int kScatterWaves = 16384;  // scatter waves aimed for: segment counts follow
constexpr int kMaxRowsPacked = 4096;  // rows the key's row field holds (12 bits)
constexpr int kRowCountThreads = 256;

// Gaussians per segment of the direct row pass.
int gauss_segment(int n, int tiles_y) {
    const long long nrw = (tiles_y + kRowWin - 1) / kRowWin;
    const long long target = std::max(1LL, (kScatterWaves + nrw - 1) / nrw);
    return (int)std::max(256LL, ((long long)n + target - 1) / target);
}

__global__ __launch_bounds__(kRowCountThreads) void gauss_row_count_kernel(
    int n, int gseg, int tiles_y, const uint2 *__restrict__ rect, int32_t *__restrict__ counts) {
    __shared__ int diff[kMaxRowsPacked + 1];
    __shared__ int wsum[kRowCountThreads / LIG_WAVE];
    const int s = blockIdx.x;
    const int g0 = s * gseg, g1 = min(n, g0 + gseg);
    for (int r = threadIdx.x; r <= tiles_y; r += kRowCountThreads) diff[r] = 0;
    __syncthreads();
    for (int g = g0 + threadIdx.x; g < g1; g += kRowCountThreads) {
        int x0, y0, x1, y1;
        bool ph;
        unpack_rect(rect[g], x0, y0, x1, y1, ph);
        if (y1 > y0) {
            LIG_CHECK(y1 <= tiles_y);
            atomicAdd(&diff[y0], 1);
            atomicAdd(&diff[y1], -1);
        }
    }
    __syncthreads();
    // prefix over rows: thread i takes rows i kPer .. i kPer + kPer - 1
    constexpr int kPer = kMaxRowsPacked / kRowCountThreads;
    int d[kPer];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int r = threadIdx.x * kPer + k;
        d[k] = r < tiles_y ? diff[r] : 0;
        mine += d[k];
    }
    int total;
    int run = block_exclusive_scan_n<kRowCountThreads>(mine, wsum, &total);
    int32_t *out = counts + (size_t)s * tiles_y;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int r = threadIdx.x * kPer + k;
        run += d[k];
        if (r < tiles_y) out[r] = run;
    }
}

// Per row (one block): the exclusive prefix of the segments' counts (pre,
// same layout as counts) and the row total (rowtot[r]; rowtot[rows] = 0 for
// the scan that follows).
__global__ __launch_bounds__(kRowCountThreads) void gauss_row_prefix_kernel(
    int nseg, int tiles_y, const int32_t *__restrict__ counts, int32_t *__restrict__ pre,
    int32_t *__restrict__ rowtot) {
    __shared__ int wsum[kRowCountThreads / LIG_WAVE];
    const int r = blockIdx.x;
    const int per = (nseg + kRowCountThreads - 1) / kRowCountThreads;
    const int s0 = min(nseg, (int)threadIdx.x * per), s1 = min(nseg, s0 + per);
    int mine = 0;
    for (int q = s0; q < s1; ++q) mine += counts[(size_t)q * tiles_y + r];
    int total;
    int run = block_exclusive_scan_n<kRowCountThreads>(mine, wsum, &total);
    for (int q = s0; q < s1; ++q) {
        pre[(size_t)q * tiles_y + r] = run;
        run += counts[(size_t)q * tiles_y + r];
    }
    if (threadIdx.x == 0) {
        rowtot[r] = total;
        if (r == 0) rowtot[tiles_y] = 0;
    }
}

// Row bins from the scanned row totals, clamped to the entry buffer's m.
__global__ void entry_bins_kernel(int rows, int m, const int32_t *__restrict__ off,
                                  int2 *__restrict__ rowbins) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) rowbins[r] = make_int2(min(off[r], m), min(off[r + 1], m));
}

// The ballot rounds of a window: lane l holds queue entry head + l (cnt of
// them); for each of the window's wn lines k (rows or tiles) the lanes with
// (unsigned)(d + k) < wd cover it and write at consecutive positions from
// the line's cursor.  `cur` holds the cursors (fast: all 32 lines present and
// every position below m, below 2^30) or, in lane k of `curv`, the cursor of
// line k (checked against m).  put(pos, k) stores one lane's output.
template <typename Put>
__device__ __forceinline__ void ballot_lines(bool fast, int wn, int m, int d, unsigned wd,
                                             int (&cur)[kRowWin], int &curv, Put put) {
    const int lane = threadIdx.x & (LIG_WAVE - 1);
    if (fast) {
#pragma unroll
        for (int k = 0; k < kRowWin; ++k) {
            const bool c = (unsigned)(d + k) < wd;
            const uint64_t mk = __ballot(c);
            // the cursor joins the rank as mbcnt's addend
            const uint32_t pos = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(mk >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mk, (uint32_t)cur[k]));
            if (c) {
                LIG_CHECK((int)pos >= 0 && (int)pos < m);
                put(pos, k);
            }
            cur[k] += __popcll(mk);
        }
    } else {
        for (int k = 0; k < wn; ++k) {
            const bool c = (unsigned)(d + k) < wd;
            const uint64_t mk = __ballot(c);
            const int ck = __builtin_amdgcn_readlane(curv, k);
            const int pos = ck + (int)__builtin_amdgcn_mbcnt_hi(
                                     (uint32_t)(mk >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u));
            if (c && pos < m) put((uint32_t)pos, k);
            curv = lane == k ? ck + __popcll(mk) : curv;
        }
    }
}

// The window's cursors: lane k < wn holds line k's first position (start)
// and its end; fast when the window is full and every end is at most m <= 2^30.
__device__ __forceinline__ bool window_cursors(int start, int end, int wn, int m,
                                               int (&cur)[kRowWin]) {
    bool fast = wn == kRowWin && m <= (1 << 30);
#pragma unroll
    for (int off = LIG_WAVE / 2; off > 0; off >>= 1) end = max(end, __shfl_xor(end, off, LIG_WAVE));
    fast = fast && end <= m;
#pragma unroll
    for (int k = 0; k < kRowWin; ++k) cur[k] = __builtin_amdgcn_readlane(start, k);
    return fast;
}

__device__ __forceinline__ void put32(int32_t *__restrict__ base, uint32_t pos, int32_t v) {
    *reinterpret_cast<int32_t *>(reinterpret_cast<char *>(base) + (size_t)(pos * 4u)) = v;
}

// Scatter of the direct row pass: one wave per (segment, window of kRowWin
// rows), a segment's windows on one XCD.
template <bool SLOTS>
__global__ __launch_bounds__(LIG_WAVE) void gauss_row_scatter_kernel(
    int n, int gseg, int nrw, int tiles_y, int kshift, int m, const uint2 *__restrict__ rect,
    const int32_t *__restrict__ cum, const int32_t *__restrict__ counts,
    const int32_t *__restrict__ pre, const int32_t *__restrict__ off,
    uint32_t *__restrict__ rowkey, uint32_t *__restrict__ rowval,
    int32_t *__restrict__ rowslot) {
    __shared__ uint32_t qv[kRowQueue];  // gid | phantom bit
    __shared__ uint32_t qk[kRowQueue];  // key of the Gaussian's row 0
    __shared__ uint32_t qy[kRowQueue];  // y0 | y1 << 16
    const int lane = threadIdx.x;
    const int u = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int s = u / nrw;
    const int w = u - s * nrw;
    const int g0 = s * gseg, g1 = min(n, g0 + gseg);
    if (g0 >= g1) return;  // wave-uniform
    const int r0 = w * kRowWin;
    const int rn = min(kRowWin, tiles_y - r0);
    const int r1 = r0 + rn;
    int start = 0, end = 0;
    if (lane < rn) {
        const size_t c = (size_t)s * tiles_y + r0 + lane;
        start = off[r0 + lane] + pre[c];
        end = start + counts[c];
    }
    int cur[kRowWin];
    const bool fast = window_cursors(start, end, rn, m, cur);

    auto round = [&](int head, int cnt) {
        wave_lds_sync();
        const bool valid = lane < cnt;
        const int q = (head + lane) & (kRowQueue - 1);
        const uint32_t v = valid ? qv[q] : 0u;
        const uint32_t kb = valid ? qk[q] : 0u;
        const uint32_t yy = valid ? qy[q] : 0u;
        wave_lds_sync();  // the slots just read are refilled after the round
        const int y0 = (int)(yy & 0xFFFFu), y1 = (int)(yy >> 16);
        // emission slot of the entry's row at column 0: cum[g - 1] + (row - y0)
        // (x1 - x0) - x0 (a phantom's single slot: cum[g - 1])
        int sb = 0, sw = 0;
        if (SLOTS && valid) {
            const int g = (int)(v & ~kPhantomBit);
            int x0, x1;
            rowkey_cols(kb, kshift, x0, x1);
            const bool ph = (v & kPhantomBit) != 0;
            sb = (g == 0 ? 0 : cum[g - 1]) + (ph ? 0 : -y0 * (x1 - x0) - x0);
            sw = ph ? 0 : x1 - x0;
        }
        ballot_lines(fast, rn, m, r0 - y0, (unsigned)(y1 - y0), cur, start,
                     [&](uint32_t pos, int k) {
                         put32(reinterpret_cast<int32_t *>(rowkey), pos,
                               (int32_t)(kb + (uint32_t)(r0 + k)));
                         put32(reinterpret_cast<int32_t *>(rowval), pos, (int32_t)v);
                         if (SLOTS) put32(rowslot, pos, sb + (r0 + k) * sw);
                     });
    };

    int head = 0, tail = 0;
    auto rect_at = [&](int g) { return rect[min(g, g1 - 1)]; };
    auto filter = [&](uint2 rc, int b) {
        const int g = b + lane;
        int x0, y0, x1, y1;
        bool ph;
        unpack_rect(rc, x0, y0, x1, y1, ph);
        const bool ov = g < g1 && y1 > y0 && y0 < r1 && y1 > r0;
        const uint64_t mo = __ballot(ov);
        if (ov) {
            const int q = (tail + (int)__builtin_amdgcn_mbcnt_hi(
                                      (uint32_t)(mo >> 32),
                                      __builtin_amdgcn_mbcnt_lo((uint32_t)mo, 0u))) &
                          (kRowQueue - 1);
            qv[q] = (uint32_t)g | (ph ? kPhantomBit : 0u);
            qk[q] = row_key(0, x0, x1, kshift);  // a phantom's: row_key(0, 0, 1) = 0
            qy[q] = (uint32_t)y0 | ((uint32_t)y1 << 16);
        }
        tail += __popcll(mo);
    };
    uint2 c0 = rect_at(g0 + lane), c1 = rect_at(g0 + LIG_WAVE + lane);
    uint2 c2 = rect_at(g0 + 2 * LIG_WAVE + lane), c3 = rect_at(g0 + 3 * LIG_WAVE + lane);
    for (int base = g0; base < g1; base += 4 * LIG_WAVE) {
        const int nb = base + 4 * LIG_WAVE + lane;
        const uint2 n0 = rect_at(nb), n1 = rect_at(nb + LIG_WAVE);
        const uint2 n2 = rect_at(nb + 2 * LIG_WAVE), n3 = rect_at(nb + 3 * LIG_WAVE);
        filter(c0, base);
        filter(c1, base + LIG_WAVE);
        filter(c2, base + 2 * LIG_WAVE);
        filter(c3, base + 3 * LIG_WAVE);
        const bool last = base + 4 * LIG_WAVE >= g1;
        while (tail - head >= LIG_WAVE || (last && tail > head)) {
            const int cnt = min(LIG_WAVE, tail - head);
            round(head, cnt);
            head += cnt;
        }
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
}

// Row r's entries [lo, hi) of segment s (of P).
__device__ __forceinline__ void row_segment(const int2 *__restrict__ rowbins, int r, int s, int P,
                                            int &lo, int &hi) {
    const int2 rb = rowbins[r];
    const int cnt = rb.y - rb.x;
    const int len = (cnt + P - 1) / P;
    lo = rb.x + min(cnt, s * len);
    hi = rb.x + min(cnt, (s + 1) * len);
}

__device__ __forceinline__ int wave_exclusive_sum(int v) {
    const int lane = threadIdx.x & (LIG_WAVE - 1);
    int x = v;
#pragma unroll
    for (int off = 1; off < LIG_WAVE; off <<= 1) {
        const int y = __shfl_up(x, off, LIG_WAVE);
        if (lane >= off) x += y;
    }
    return x - v;
}

// Counting pass: one block per (row, segment), a row's segments on one XCD.
// counts[(r P + s) tiles_x + t] = the segment's entries whose column range
// covers tile t (a difference array over the row's columns in LDS, then its
// prefix sum).
constexpr int kCountThreads = 256;
__global__ __launch_bounds__(kCountThreads) void row_count_kernel(
    int P, int tiles_x, int kshift, const int2 *__restrict__ rowbins,
    const uint32_t *__restrict__ rowkey, int32_t *__restrict__ counts) {
    __shared__ int diff[kRowMaxTiles + 1];
    __shared__ int wsum[kCountThreads / LIG_WAVE];
    const int u = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int r = u / P;
    const int s = u - r * P;
    int lo, hi;
    row_segment(rowbins, r, s, P, lo, hi);
    for (int t = threadIdx.x; t <= tiles_x; t += kCountThreads) diff[t] = 0;
    __syncthreads();
    for (int j = lo + threadIdx.x; j < hi; j += kCountThreads) {
        int x0, x1;
        rowkey_cols(rowkey[j], kshift, x0, x1);
        LIG_CHECK(x0 >= 0 && x0 < x1 && x1 <= tiles_x);
        atomicAdd(&diff[x0], 1);
        atomicAdd(&diff[min(x1, tiles_x)], -1);
    }
    __syncthreads();
    // prefix: thread i takes columns i kPer .. i kPer + kPer - 1
    constexpr int kPer = kRowMaxTiles / kCountThreads;
    int d[kPer];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int t = threadIdx.x * kPer + k;
        d[k] = t < tiles_x ? diff[t] : 0;
        mine += d[k];
    }
    int total;
    int run = block_exclusive_scan_n<kCountThreads>(mine, wsum, &total);
    int32_t *out = counts + ((size_t)r * P + s) * tiles_x;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int t = threadIdx.x * kPer + k;
        run += d[k];
        if (t < tiles_x) out[t] = run;
    }
}

// Per row (one block): each tile's column of segment counts becomes its
// exclusive prefix over segments (scol, same layout); the tile totals go to
// tt, their exclusive prefix within the row to te, the row total to
// rowtot[r] (rowtot[rows] = 0 for the scan that follows).  Threads take
// consecutive columns (coalesced) and load four segments ahead; the totals
// meet in LDS for the scan in column order.
constexpr int kPrefixThreads = 256;
__global__ __launch_bounds__(kPrefixThreads) void row_prefix_kernel(
    int P, int tiles_x, int rows, const int32_t *__restrict__ counts,
    int32_t *__restrict__ scol, int32_t *__restrict__ tt, int32_t *__restrict__ te,
    int32_t *__restrict__ rowtot) {
    __shared__ int wsum[kPrefixThreads / LIG_WAVE];
    __shared__ int s_tot[kRowMaxTiles];
    const int r = blockIdx.x;
    const size_t rbase = (size_t)r * P * tiles_x;
    for (int t = threadIdx.x; t < tiles_x; t += kPrefixThreads) {
        const int32_t *c = counts + rbase + t;
        int32_t *o = scol + rbase + t;
        int run = 0, s = 0;
        for (; s + 4 <= P; s += 4) {
            const int v0 = c[(size_t)s * tiles_x], v1 = c[(size_t)(s + 1) * tiles_x];
            const int v2 = c[(size_t)(s + 2) * tiles_x], v3 = c[(size_t)(s + 3) * tiles_x];
            o[(size_t)s * tiles_x] = run;
            o[(size_t)(s + 1) * tiles_x] = run + v0;
            o[(size_t)(s + 2) * tiles_x] = run + v0 + v1;
            o[(size_t)(s + 3) * tiles_x] = run + v0 + v1 + v2;
            run += v0 + v1 + v2 + v3;
        }
        for (; s < P; ++s) {
            const int v = c[(size_t)s * tiles_x];
            o[(size_t)s * tiles_x] = run;
            run += v;
        }
        s_tot[t] = run;
    }
    __syncthreads();
    constexpr int kPer = kRowMaxTiles / kPrefixThreads;
    int tot[kPer];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        // thread i takes columns i * kPer .. i * kPer + kPer - 1 (row order for the scan)
        const int t = threadIdx.x * kPer + k;
        tot[k] = t < tiles_x ? s_tot[t] : 0;
        mine += tot[k];
    }
    int total;
    int excl = block_exclusive_scan_n<kPrefixThreads>(mine, wsum, &total);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int t = threadIdx.x * kPer + k;
        if (t < tiles_x) {
            tt[(size_t)r * tiles_x + t] = tot[k];
            te[(size_t)r * tiles_x + t] = excl;
        }
        excl += tot[k];
    }
    if (threadIdx.x == 0) {
        rowtot[r] = total;
        if (r == 0) rowtot[rows] = 0;
    }
}

// Bins: tile (r, t) holds [rowoff[r] + te, + tt), (0, 0) when empty.
__global__ void row_bins_kernel(int num_tiles, int tiles_x, const int32_t *__restrict__ rowoff,
                                const int32_t *__restrict__ tt, const int32_t *__restrict__ te,
                                int32_t *__restrict__ bins) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= num_tiles) return;
    const int c = tt[t];
    const int st = rowoff[t / tiles_x] + te[t];
    bins[2 * t] = c > 0 ? st : 0;
    bins[2 * t + 1] = c > 0 ? st + c : 0;
}

// Scatter pass: one wave per (row, segment, window of kRowWin tiles), a row's
// waves on one XCD (its entries stay in that L2).  Lane k < kRowWin loads the
// segment's first output position in the window's tile k, the tile's cursor.
// The wave reads the segment's row entries four rounds of 64 at a time (the
// next four requested before these are used) and queues those whose column
// range meets the window, in order, in an LDS ring of 512; every 64 queued
// entries (and the rest at the end) make a round: one entry per lane, and
// for each tile k of the window a ballot of the lanes covering it gives each
// such lane its rank (mbcnt) and the cursor's advance (popcount), so the
// covering lanes store their gids -- and emission slots when asked, from
// each entry's slot at column 0 -- at consecutive positions of the tile's
// bin.  Positions past the output's capacity m are not written.
template <bool SLOTS>
__global__ __launch_bounds__(LIG_WAVE) void row_scatter_kernel(
    int P, int nwin, int tiles_x, int kshift, int m, const int2 *__restrict__ rowbins,
    const uint32_t *__restrict__ rowkey, const uint32_t *__restrict__ rowval,
    const int32_t *__restrict__ rowslot, const int32_t *__restrict__ counts,
    const int32_t *__restrict__ scol, const int32_t *__restrict__ te,
    const int32_t *__restrict__ rowoff, int32_t *__restrict__ gids_out,
    int32_t *__restrict__ slots_out) {
    __shared__ uint32_t qv[kRowQueue];  // queued row entries: value (gid | phantom bit)
    __shared__ uint32_t qx[kRowQueue];  // and column range x0 | x1 << 16
    __shared__ int qs[SLOTS ? kRowQueue : 1];  // and slot at column 0 (emission slots)
    const int lane = threadIdx.x;
    const int u = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int rs = u / nwin;  // r P + s
    const int w = u - rs * nwin;
    const int r = rs / P;
    const int s = rs - r * P;
    int lo, hi;
    row_segment(rowbins, r, s, P, lo, hi);
    if (lo >= hi) return;  // wave-uniform
    const int wlo = w * kRowWin;
    const int wn = min(kRowWin, tiles_x - wlo);
    const int whi = wlo + wn;
    int start = 0, end = 0;
    if (lane < wn) {
        const int t = wlo + lane;
        start = rowoff[r] + te[(size_t)r * tiles_x + t] + scol[(size_t)rs * tiles_x + t];
        end = start + counts[(size_t)rs * tiles_x + t];
    }
    int cur[kRowWin];
    const bool fast = window_cursors(start, end, wn, m, cur);

    // one round: queue slots head .. head + cnt - 1, one per lane
    auto round = [&](int head, int cnt) {
        wave_lds_sync();
        const bool valid = lane < cnt;
        const int q = (head + lane) & (kRowQueue - 1);
        const uint32_t v = valid ? qv[q] : 0u;
        const uint32_t xx = valid ? qx[q] : 0u;
        const int slot0 = SLOTS && valid ? qs[q] : 0;
        wave_lds_sync();  // the slots just read are refilled after the round
        const int x0 = (int)(xx & 0xFFFFu), x1 = (int)(xx >> 16);
        const bool ph = (v & kPhantomBit) != 0;
        const int gid = ph ? 0 : (int)(v & ~kPhantomBit);
        ballot_lines(fast, wn, m, wlo - x0, (unsigned)(x1 - x0), cur, start,
                     [&](uint32_t pos, int k) {
                         put32(gids_out, pos, gid);
                         // a phantom's only tile is tile 0 (wlo + k = 0)
                         if (SLOTS) put32(slots_out, pos, slot0 + wlo + k);
                     });
    };

    int head = 0, tail = 0;  // queue positions (wave-uniform)
    // unconditional loads (clamped to the segment's last entry; lanes past
    // hi are masked by the overlap test): the next four rounds' keys and
    // values are requested before this iteration's are filtered
    auto key_at = [&](int jj) { return rowkey[min(jj, hi - 1)]; };
    auto val_at = [&](int jj) { return rowval[min(jj, hi - 1)]; };
    auto slot_at = [&](int jj) { return SLOTS ? rowslot[min(jj, hi - 1)] : 0; };
    auto filter = [&](uint32_t kc, uint32_t vc, int sc, int b) {
        const int j = b + lane;
        int x0 = 0, x1 = 0;
        rowkey_cols(kc, kshift, x0, x1);
        const bool ov = j < hi && x0 < whi && x1 > wlo;
        const uint64_t mo = __ballot(ov);
        if (ov) {
            const int q = (tail + (int)__builtin_amdgcn_mbcnt_hi(
                                      (uint32_t)(mo >> 32),
                                      __builtin_amdgcn_mbcnt_lo((uint32_t)mo, 0u))) &
                          (kRowQueue - 1);
            qv[q] = vc;
            qx[q] = (uint32_t)x0 | ((uint32_t)x1 << 16);
            if (SLOTS) qs[q] = sc;
        }
        tail += __popcll(mo);
    };
    uint32_t k0 = key_at(lo + lane), k1 = key_at(lo + LIG_WAVE + lane);
    uint32_t k2 = key_at(lo + 2 * LIG_WAVE + lane), k3 = key_at(lo + 3 * LIG_WAVE + lane);
    uint32_t v0 = val_at(lo + lane), v1 = val_at(lo + LIG_WAVE + lane);
    uint32_t v2 = val_at(lo + 2 * LIG_WAVE + lane), v3 = val_at(lo + 3 * LIG_WAVE + lane);
    int s0 = slot_at(lo + lane), s1 = slot_at(lo + LIG_WAVE + lane);
    int s2 = slot_at(lo + 2 * LIG_WAVE + lane), s3 = slot_at(lo + 3 * LIG_WAVE + lane);
    for (int base = lo; base < hi; base += 4 * LIG_WAVE) {
        const int nb = base + 4 * LIG_WAVE + lane;
        const uint32_t n0 = key_at(nb), n1 = key_at(nb + LIG_WAVE);
        const uint32_t n2 = key_at(nb + 2 * LIG_WAVE), n3 = key_at(nb + 3 * LIG_WAVE);
        const uint32_t w0 = val_at(nb), w1 = val_at(nb + LIG_WAVE);
        const uint32_t w2 = val_at(nb + 2 * LIG_WAVE), w3 = val_at(nb + 3 * LIG_WAVE);
        const int t0 = slot_at(nb), t1 = slot_at(nb + LIG_WAVE);
        const int t2 = slot_at(nb + 2 * LIG_WAVE), t3 = slot_at(nb + 3 * LIG_WAVE);
        // up to 4 x 64 entries join the (at most 63) queued: 319 < kRowQueue
        filter(k0, v0, s0, base);
        filter(k1, v1, s1, base + LIG_WAVE);
        filter(k2, v2, s2, base + 2 * LIG_WAVE);
        filter(k3, v3, s3, base + 3 * LIG_WAVE);
        const bool last = base + 4 * LIG_WAVE >= hi;
        while (tail - head >= LIG_WAVE || (last && tail > head)) {
            const int cnt = min(LIG_WAVE, tail - head);
            round(head, cnt);
            head += cnt;
        }
        k0 = n0;
        k1 = n1;
        k2 = n2;
        k3 = n3;
        v0 = w0;
        v1 = w1;
        v2 = w2;
        v3 = w3;
        s0 = t0;
        s1 = t1;
        s2 = t2;
        s3 = t3;
    }
}

// Bits of a row entry key's row field (the column fields sit above them).
int row_key_shift(int rows) { return bits_for(rows); }

int row_windows(int tiles_x) { return (tiles_x + kRowWin - 1) / kRowWin; }

// Segments per row: enough for about kScatterWaves scatter waves, but a
// segment keeps about 1024 intersections on an average row.
int row_segments(int m, int tiles_x, int tiles_y) {
    const long long rw = (long long)tiles_y * row_windows(tiles_x);
    const long long by_waves = (kScatterWaves + rw - 1) / rw;
    const long long by_len = std::max(1LL, (long long)m / ((long long)tiles_y * 1024));
    return (int)std::max(1LL, std::min(std::min(by_waves, by_len), 4096LL));
}

// Workspace of the row path (0 when it does not apply): rectangles, the
// direct row pass's (segment, row) counts and their prefixes, row totals and
// their scan and its workspace, the row entries (keys, values, slots at
// column 0 for the emission slots), the row bins,
// the (row, segment, tile) counts and their segment prefixes, tile totals and
// their row prefixes, row totals and their scan.
size_t rows_ws_bytes(int n, int m, int tiles_x, int tiles_y, int *P_out) {
    if (n <= 0 || m <= 0 || tiles_x < 1 || tiles_y < 1 || tiles_x > kRowMaxTiles ||
        tiles_y > kMaxRowsPacked || row_key_shift(tiles_y) + 20 > 32)
        return 0;
    const int P = row_segments(m, tiles_x, tiles_y);
    const size_t T = (size_t)tiles_x * tiles_y;
    const size_t nctr = (size_t)P * T;
    const int gseg = gauss_segment(n, tiles_y);
    const size_t nseg = ((size_t)n + gseg - 1) / gseg;
    const size_t gctr = nseg * (size_t)tiles_y;
    if (nctr > (size_t)INT32_MAX || gctr > (size_t)INT32_MAX) return 0;
    if (P_out) *P_out = P;
    const size_t nn = (size_t)n, mm = (size_t)m, ty = (size_t)tiles_y;
    return align_up(8 * nn) + 2 * align_up(4 * gctr) + align_up(4 * (ty + 1)) +
           align_up(scan_ws_elems(ty + 1) * 4) + 3 * align_up(4 * mm) + align_up(8 * ty) +
           2 * align_up(4 * nctr) + 2 * align_up(4 * T) + align_up(4 * (ty + 1)) +
           align_up(scan_ws_elems(ty + 1) * 4) + 256;
}

// The row path of bin_and_sort (workspace checked by the caller).
int bin_rows(int n, int m, const float *xys, const float *extents, const int32_t *cum,
             int tiles_x, int tiles_y, int block_width, int32_t *gaussian_ids_sorted,
             int32_t *tile_bins, int32_t *emission_ids_sorted, char *ws, hipStream_t st) {
    int P = 1;
    rows_ws_bytes(n, m, tiles_x, tiles_y, &P);
    const size_t nn = (size_t)n, mm = (size_t)m, T = (size_t)tiles_x * tiles_y;
    const size_t nctr = (size_t)P * T, ty = (size_t)tiles_y;
    const int gseg = gauss_segment(n, tiles_y);
    const int nseg = (int)((nn + gseg - 1) / gseg);
    const size_t gctr = (size_t)nseg * ty;
    char *p = ws;
    auto take = [&p](size_t bytes) {
        char *q = p;
        p += align_up(bytes);
        return q;
    };
    uint2 *rect = reinterpret_cast<uint2 *>(take(8 * nn));
    int32_t *gcounts = reinterpret_cast<int32_t *>(take(4 * gctr));
    int32_t *gpre = reinterpret_cast<int32_t *>(take(4 * gctr));
    int32_t *eoff = reinterpret_cast<int32_t *>(take(4 * (ty + 1)));
    int32_t *eoff_ws = reinterpret_cast<int32_t *>(take(scan_ws_elems(ty + 1) * 4));
    uint32_t *rowkey = reinterpret_cast<uint32_t *>(take(4 * mm));
    uint32_t *rowval = reinterpret_cast<uint32_t *>(take(4 * mm));
    int32_t *rowslot = reinterpret_cast<int32_t *>(take(4 * mm));
    int2 *rowbins = reinterpret_cast<int2 *>(take(8 * ty));
    int32_t *counts = reinterpret_cast<int32_t *>(take(4 * nctr));
    int32_t *scol = reinterpret_cast<int32_t *>(take(4 * nctr));
    int32_t *tt = reinterpret_cast<int32_t *>(take(4 * T));
    int32_t *te = reinterpret_cast<int32_t *>(take(4 * T));
    int32_t *rowoff = reinterpret_cast<int32_t *>(take(4 * (ty + 1)));
    int32_t *row_ws = reinterpret_cast<int32_t *>(take(scan_ws_elems(ty + 1) * 4));

    const int ks = row_key_shift(tiles_y);
    hipLaunchKernelGGL(rect_rows_kernel, dim3(ceil_div(nn, 256)), dim3(256), 0, st, n, xys,
                       extents, cum, tiles_x, tiles_y, block_width, rect);
    // row entries in (row, gid) order, straight from the rectangles
    hipLaunchKernelGGL(gauss_row_count_kernel, dim3(nseg), dim3(kRowCountThreads), 0, st, n,
                       gseg, tiles_y, (const uint2 *)rect, gcounts);
    hipLaunchKernelGGL(gauss_row_prefix_kernel, dim3(tiles_y), dim3(kRowCountThreads), 0, st,
                       nseg, tiles_y, (const int32_t *)gcounts, gpre, eoff);
    scan_i32(eoff, eoff, ty + 1, false, eoff_ws, st);
    hipLaunchKernelGGL(entry_bins_kernel, dim3(ceil_div(ty, 256)), dim3(256), 0, st, tiles_y, m,
                       (const int32_t *)eoff, rowbins);
    const int nrw = (tiles_y + kRowWin - 1) / kRowWin;
    const bool slots = emission_ids_sorted != nullptr;
    if (slots)
        hipLaunchKernelGGL(gauss_row_scatter_kernel<true>, dim3((unsigned)nseg * (unsigned)nrw),
                           dim3(LIG_WAVE), 0, st, n, gseg, nrw, tiles_y, ks, m,
                           (const uint2 *)rect, cum, (const int32_t *)gcounts,
                           (const int32_t *)gpre, (const int32_t *)eoff, rowkey, rowval, rowslot);
    else
        hipLaunchKernelGGL(gauss_row_scatter_kernel<false>, dim3((unsigned)nseg * (unsigned)nrw),
                           dim3(LIG_WAVE), 0, st, n, gseg, nrw, tiles_y, ks, m,
                           (const uint2 *)rect, cum, (const int32_t *)gcounts,
                           (const int32_t *)gpre, (const int32_t *)eoff, rowkey, rowval, rowslot);
    // tiles from the row entries
    const unsigned segs = (unsigned)tiles_y * (unsigned)P;
    hipLaunchKernelGGL(row_count_kernel, dim3(segs), dim3(kCountThreads), 0, st, P, tiles_x, ks,
                       (const int2 *)rowbins, (const uint32_t *)rowkey, counts);
    hipLaunchKernelGGL(row_prefix_kernel, dim3(tiles_y), dim3(kPrefixThreads), 0, st, P, tiles_x,
                       tiles_y, (const int32_t *)counts, scol, tt, te, rowoff);
    scan_i32(rowoff, rowoff, ty + 1, false, row_ws, st);
    hipLaunchKernelGGL(row_bins_kernel, dim3(ceil_div(T, 256)), dim3(256), 0, st, (int)T,
                       tiles_x, (const int32_t *)rowoff, (const int32_t *)tt,
                       (const int32_t *)te, tile_bins);
    const int nwin = row_windows(tiles_x);
    const unsigned waves = segs * (unsigned)nwin;
    if (slots)
        hipLaunchKernelGGL(row_scatter_kernel<true>, dim3(waves), dim3(LIG_WAVE), 0, st, P, nwin,
                           tiles_x, ks, m, (const int2 *)rowbins, (const uint32_t *)rowkey,
                           (const uint32_t *)rowval, (const int32_t *)rowslot,
                           (const int32_t *)counts, (const int32_t *)scol, (const int32_t *)te,
                           (const int32_t *)rowoff, gaussian_ids_sorted, emission_ids_sorted);
    else
        hipLaunchKernelGGL(row_scatter_kernel<false>, dim3(waves), dim3(LIG_WAVE), 0, st, P, nwin,
                           tiles_x, ks, m, (const int2 *)rowbins, (const uint32_t *)rowkey,
                           (const uint32_t *)rowval, (const int32_t *)rowslot,
                           (const int32_t *)counts, (const int32_t *)scol, (const int32_t *)te,
                           (const int32_t *)rowoff, gaussian_ids_sorted, emission_ids_sorted);
    return launch_status();
}

}  // namespace
}  // namespace lig

using namespace lig;

extern "C" size_t lig_scan_workspace_size(int n) {
    return n > 0 ? scan_ws_elems((size_t)n) * 4 + 256 : 0;
}

extern "C" int lig_inclusive_scan_i32(int n, const int32_t *in, int32_t *out, void *workspace,
                                      size_t workspace_bytes, void *stream) {
    if (n < 0 || (n > 0 && (!in || !out))) return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0) return LIG_OK;
    if (workspace_bytes < scan_ws_elems((size_t)n) * 4 ||
        (scan_ws_elems((size_t)n) > 0 && !workspace))
        return LIG_ERR_WORKSPACE;
    scan_i32(in, out, (size_t)n, true, reinterpret_cast<int32_t *>(workspace),
             as_stream(stream));
    return launch_status();
}

extern "C" int lig_map_gaussian_to_intersects(int n, int num_intersects, const float *xys,
                                              const float *depths, const float *extents,
                                              const int32_t *cum_tiles_hit, int tiles_x,
                                              int tiles_y, int block_width, int64_t *isect_ids,
                                              int32_t *gaussian_ids, void *stream) {
    if (n < 0 || num_intersects < 0 || tiles_x <= 0 || tiles_y <= 0 || block_width < 1)
        return LIG_ERR_INVALID_ARGUMENT;
    if (n == 0 || num_intersects == 0) return LIG_OK;
    if (!xys || !depths || !extents || !cum_tiles_hit || !isect_ids || !gaussian_ids)
        return LIG_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(emit_isect_i64_kernel, dim3(ceil_div(n, kEmitThreads)),
                       dim3(kEmitThreads), 0, as_stream(stream), n, num_intersects, xys, depths,
                       extents, cum_tiles_hit, tiles_x, tiles_y, block_width, isect_ids,
                       gaussian_ids);
    return launch_status();
}

extern "C" size_t lig_sort_pairs_workspace_size(int m, int key_bits) {
    if (m <= 0) return 0;
    return key_bits > 32 ? sort_ws_bytes<uint64_t>((size_t)m, key_bits)
                         : sort_ws_bytes<uint32_t>((size_t)m, key_bits);
}

extern "C" int lig_sort_isect_pairs(int m, const int64_t *keys_in, const int32_t *vals_in,
                                    int64_t *keys_out, int32_t *vals_out, void *workspace,
                                    size_t workspace_bytes, void *stream) {
    if (m < 0 || (m > 0 && (!keys_in || !vals_in || !keys_out || !vals_out)))
        return LIG_ERR_INVALID_ARGUMENT;
    if (m == 0) return LIG_OK;
    if (!workspace || workspace_bytes < sort_ws_bytes<uint64_t>((size_t)m, 64))
        return LIG_ERR_WORKSPACE;
    return radix_sort<uint64_t>(reinterpret_cast<const uint64_t *>(keys_in),
                                reinterpret_cast<const uint32_t *>(vals_in),
                                reinterpret_cast<uint64_t *>(keys_out),
                                reinterpret_cast<uint32_t *>(vals_out), (size_t)m, 64,
                                (uint64_t)1 << 63, reinterpret_cast<char *>(workspace),
                                as_stream(stream));
}

extern "C" int lig_get_tile_bin_edges(int num_intersects, const int64_t *isect_ids_sorted,
                                      int tiles_x, int tiles_y, int32_t *tile_bins,
                                      void *stream) {
    if (num_intersects < 0 || tiles_x <= 0 || tiles_y <= 0 || !tile_bins)
        return LIG_ERR_INVALID_ARGUMENT;
    const size_t nt = (size_t)tiles_x * tiles_y;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(2 * nt, 256)), dim3(256), 0, st, tile_bins,
                       2 * nt, 0);
    if (num_intersects > 0) {
        if (!isect_ids_sorted) return LIG_ERR_INVALID_ARGUMENT;
        hipLaunchKernelGGL((bin_edges_kernel<int64_t, 32>), dim3(ceil_div(num_intersects, 1024)),
                           dim3(256), 0, st, num_intersects, (const int32_t *)nullptr,
                           isect_ids_sorted, (int)nt, tile_bins, ~(int64_t)0);
    }
    return launch_status();
}

// Workspace layout of the fused path: tile keys and slot gids (emission
// order), sorted keys, sorted slots (unless the caller keeps them), then the
// sort's own workspace.
extern "C" size_t lig_bin_workspace_size(int num_intersects, int num_tiles) {
    const size_t m = num_intersects > 0 ? (size_t)num_intersects : 0;
    return 4 * align_up(m * 4) + sort_ws_bytes<uint32_t>(m, bits_for(num_tiles)) + 256;
}

extern "C" size_t lig_bin_rows_workspace_size(int n, int num_intersects, int img_height,
                                              int img_width, int block_width) {
    if (img_height <= 0 || img_width <= 0 || block_width < 1 || block_width > 16) return 0;
    const int tx = (img_width + block_width - 1) / block_width;
    const int ty = (img_height + block_width - 1) / block_width;
    return rows_ws_bytes(n, num_intersects, tx, ty, nullptr);
}

namespace lig {
namespace {

// Body of lig_bin_and_sort_gaussians and of its no-allocation form: the
// arguments are checked, the workspace is `workspace_bytes` at `workspace`.
int bin_and_sort(int n, int num_intersects, const float *xys, const float *extents,
                 const int32_t *cum_tiles_hit, int img_height, int img_width, int block_width,
                 int32_t *gaussian_ids_sorted, int32_t *tile_bins, int32_t *emission_ids_sorted,
                 void *workspace, size_t workspace_bytes, hipStream_t st) {
    const int tx = (img_width + block_width - 1) / block_width;
    const int ty = (img_height + block_width - 1) / block_width;
    const int nt = tx * ty;
    const size_t rows_bytes = rows_ws_bytes(n, num_intersects, tx, ty, nullptr);
    const bool rows_path = rows_bytes > 0 && workspace_bytes >= rows_bytes;
    if (!rows_path && workspace_bytes < lig_bin_workspace_size(num_intersects, nt))
        return LIG_ERR_WORKSPACE;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(2 * (size_t)nt, 256)), dim3(256), 0, st,
                       tile_bins, 2 * (size_t)nt, 0);
    if (num_intersects == 0 || n == 0) return launch_status();
    if (!xys || !extents || !cum_tiles_hit || !gaussian_ids_sorted || !workspace)
        return LIG_ERR_INVALID_ARGUMENT;
    if (rows_path)
        return bin_rows(n, num_intersects, xys, extents, cum_tiles_hit, tx, ty, block_width,
                        gaussian_ids_sorted, tile_bins, emission_ids_sorted,
                        reinterpret_cast<char *>(workspace), st);
    const size_t m = (size_t)num_intersects;
    char *p = reinterpret_cast<char *>(workspace);
    auto take = [&p, m]() {
        char *q = p;
        p += align_up(m * 4);
        return q;
    };
    uint32_t *keys = reinterpret_cast<uint32_t *>(take());
    int32_t *slot_gid = reinterpret_cast<int32_t *>(take());
    uint32_t *keys_sorted = reinterpret_cast<uint32_t *>(take());
    uint32_t *slots_sorted = reinterpret_cast<uint32_t *>(take());
    hipLaunchKernelGGL(emit_tile_keys_kernel, dim3(ceil_div(n, kEmitThreads)), dim3(kEmitThreads),
                       0, st, n, num_intersects, (const int32_t *)nullptr, xys, extents,
                       cum_tiles_hit, tx, ty, block_width, (const BatchImage *)nullptr, 0, keys,
                       slot_gid);
    if (!emission_ids_sorted) {
        // sort (tile, gid) pairs directly: the stable sort of the emission
        // order gives the same gids, and no gather is needed afterwards
        int rc = radix_sort<uint32_t>(keys, reinterpret_cast<const uint32_t *>(slot_gid),
                                      keys_sorted,
                                      reinterpret_cast<uint32_t *>(gaussian_ids_sorted), m,
                                      bits_for(nt), 0u, p, st);
        if (rc != LIG_OK) return rc;
        hipLaunchKernelGGL((bin_edges_kernel<uint32_t, 0>), dim3(ceil_div(m, 1024)), dim3(256), 0,
                           st, num_intersects, (const int32_t *)nullptr,
                           (const uint32_t *)keys_sorted, nt, tile_bins, ~0u);
        return launch_status();
    }
    slots_sorted = reinterpret_cast<uint32_t *>(emission_ids_sorted);
    int rc = radix_sort<uint32_t>(keys, nullptr, keys_sorted, slots_sorted, m, bits_for(nt), 0u,
                                  p, st);
    if (rc != LIG_OK) return rc;
    hipLaunchKernelGGL(finish_bins_kernel, dim3(ceil_div(m, 256)), dim3(256), 0, st,
                       num_intersects, (const uint32_t *)keys_sorted,
                       (const uint32_t *)slots_sorted, (const int32_t *)slot_gid, nt, tile_bins,
                       gaussian_ids_sorted);
    return launch_status();
}

bool bin_args_ok(int n, int num_intersects, int img_height, int img_width, int block_width,
                 const int32_t *tile_bins) {
    return n >= 0 && num_intersects >= 0 && img_height > 0 && img_width > 0 &&
           block_width >= 1 && block_width <= 16 && tile_bins;
}

}  // namespace
}  // namespace lig

extern "C" int lig_bin_and_sort_gaussians(int n, int num_intersects, const float *xys,
                                          const float *extents, const int32_t *cum_tiles_hit,
                                          int img_height, int img_width, int block_width,
                                          int32_t *gaussian_ids_sorted, int32_t *tile_bins,
                                          int32_t *emission_ids_sorted, void *workspace,
                                          size_t workspace_bytes, void *stream) {
    if (!bin_args_ok(n, num_intersects, img_height, img_width, block_width, tile_bins))
        return LIG_ERR_INVALID_ARGUMENT;
    return bin_and_sort(n, num_intersects, xys, extents, cum_tiles_hit, img_height, img_width,
                        block_width, gaussian_ids_sorted, tile_bins, emission_ids_sorted,
                        workspace, workspace_bytes, as_stream(stream));
}

// Scratch of lig_bin_and_sort_gaussians_noalloc: each device that runs the
// library holds its own copy for the life of the process.  Host code never
// passes the array's name to a kernel (on the host it is a stand-in, not a
// device address); it uses the device address hipGetSymbolAddress returns,
// looked up once per device and kept in g_scratch_addr.
static __device__ __align__(256) unsigned char scratch[LIG_BIN_NOALLOC_SCRATCH_BYTES];

namespace lig {
namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxScratchDevices];

// Device address of the current device's scratch (looked up on the first
// call for that device).
int scratch_address(void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return LIG_ERR_LAUNCH;
    if (dev < 0 || dev >= kMaxScratchDevices) return LIG_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return LIG_ERR_LAUNCH;
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return LIG_OK;
}

}  // namespace
}  // namespace lig

extern "C" int lig_bin_and_sort_gaussians_noalloc(int n, int num_intersects, const float *xys,
                                                  const float *extents,
                                                  const int32_t *cum_tiles_hit, int img_height,
                                                  int img_width, int block_width,
                                                  int32_t *gaussian_ids_sorted,
                                                  int32_t *tile_bins,
                                                  int32_t *emission_ids_sorted, void *stream) {
    if (!bin_args_ok(n, num_intersects, img_height, img_width, block_width, tile_bins))
        return LIG_ERR_INVALID_ARGUMENT;
    const int nt = ((img_width + block_width - 1) / block_width) *
                   ((img_height + block_width - 1) / block_width);
    if (lig_bin_workspace_size(num_intersects, nt) > LIG_BIN_NOALLOC_SCRATCH_BYTES)
        return LIG_ERR_TOO_LARGE;
    hipStream_t st = as_stream(stream);
    if (st) {
        // the scratch is the current device's: the stream must belong to it
        int cur = -1, sdev = -1;
        if (hipGetDevice(&cur) != hipSuccess || hipStreamGetDevice(st, &sdev) != hipSuccess)
            return LIG_ERR_LAUNCH;
        if (cur != sdev) return LIG_ERR_INVALID_ARGUMENT;
    }
    void *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != LIG_OK) return rc;
    return bin_and_sort(n, num_intersects, xys, extents, cum_tiles_hit, img_height, img_width,
                        block_width, gaussian_ids_sorted, tile_bins, emission_ids_sorted, ws,
                        LIG_BIN_NOALLOC_SCRATCH_BYTES, st);
}

namespace lig {
namespace {

// Batched binning into the global tile space.  m_dev NULL: m is the exact
// intersection count.  m_dev set: m is the capacity the buffers hold and the
// count is read on the device from *m_dev; a count above the capacity leaves
// every bin empty.
int bin_batched(int n, int m, const int32_t *m_dev, const float *xys, const float *extents,
                const int32_t *cum_tiles_hit, const void *batch_table, int num_images,
                int num_tiles, int32_t *gaussian_ids_sorted, int32_t *tile_bins, void *workspace,
                size_t workspace_bytes, hipStream_t st) {
    if (n < 0 || m < 0 || num_images < 1 || num_tiles < 1 || !batch_table || !tile_bins)
        return LIG_ERR_INVALID_ARGUMENT;
    if (workspace_bytes < lig_bin_workspace_size(m, num_tiles)) return LIG_ERR_WORKSPACE;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(2 * (size_t)num_tiles, 256)), dim3(256), 0,
                       st, tile_bins, 2 * (size_t)num_tiles, 0);
    if (m == 0 || n == 0) return launch_status();
    if (!xys || !extents || !cum_tiles_hit || !gaussian_ids_sorted || !workspace)
        return LIG_ERR_INVALID_ARGUMENT;
    const size_t mm = (size_t)m;
    char *p = reinterpret_cast<char *>(workspace);
    auto take = [&p, mm]() {
        char *q = p;
        p += align_up(mm * 4);
        return q;
    };
    uint32_t *keys = reinterpret_cast<uint32_t *>(take());
    int32_t *slot_gid = reinterpret_cast<int32_t *>(take());
    uint32_t *keys_sorted = reinterpret_cast<uint32_t *>(take());
    take();
    hipLaunchKernelGGL(emit_tile_keys_kernel, dim3(ceil_div(n, kEmitThreads)), dim3(kEmitThreads),
                       0, st, n, m, m_dev, xys, extents, cum_tiles_hit, 0, 0, 16,
                       reinterpret_cast<const BatchImage *>(batch_table), num_images, keys,
                       slot_gid);
    int rc = radix_sort<uint32_t>(keys, reinterpret_cast<const uint32_t *>(slot_gid), keys_sorted,
                                  reinterpret_cast<uint32_t *>(gaussian_ids_sorted), mm,
                                  bits_for(num_tiles), 0u, p, st, m_dev);
    if (rc != LIG_OK) return rc;
    hipLaunchKernelGGL((bin_edges_kernel<uint32_t, 0>), dim3(ceil_div(mm, 1024)), dim3(256), 0,
                       st, m, m_dev, (const uint32_t *)keys_sorted, num_tiles, tile_bins, ~0u);
    return launch_status();
}

}  // namespace
}  // namespace lig

extern "C" int lig_bin_batched(int n, int num_intersects, const float *xys, const float *extents,
                               const int32_t *cum_tiles_hit, const void *batch_table,
                               int num_images, int num_tiles, int32_t *gaussian_ids_sorted,
                               int32_t *tile_bins, void *workspace, size_t workspace_bytes,
                               void *stream) {
    return bin_batched(n, num_intersects, nullptr, xys, extents, cum_tiles_hit, batch_table,
                       num_images, num_tiles, gaussian_ids_sorted, tile_bins, workspace,
                       workspace_bytes, as_stream(stream));
}

extern "C" int lig_bin_batched_capacity(int n, int capacity, const float *xys,
                                        const float *extents, const int32_t *cum_tiles_hit,
                                        const void *batch_table, int num_images, int num_tiles,
                                        int32_t *gaussian_ids_sorted, int32_t *tile_bins,
                                        void *workspace, size_t workspace_bytes, void *stream) {
    if (n < 1 || capacity < 1 || !cum_tiles_hit) return LIG_ERR_INVALID_ARGUMENT;
    return bin_batched(n, capacity, cum_tiles_hit + (n - 1), xys, extents, cum_tiles_hit,
                       batch_table, num_images, num_tiles, gaussian_ids_sorted, tile_bins,
                       workspace, workspace_bytes, as_stream(stream));
}
