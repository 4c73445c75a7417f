// kscr_lbvh.hip -- LBVH build / refit / self query for gfx950.
//
// Reference: KE/KittenGpuLBVH/lbvh.cu.  The Morton codes, the sorted order
// and the tree are bit-exact with the reference's (the sort is stable, the
// Karras build is deterministic given the sorted keys); refit results do not
// depend on thread order (min/max and the stack-depth recursion commute).
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "kscr_lbvh.h"
#include "kscr_sync.h"

#include <type_traits>

namespace kscr {
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kTile = kSortThreads * kSortItems;  // keys per radix tile
constexpr int kScanChunk = 1024;
// Sorts of at most kFusedTiles tiles run one launch per digit pass
// (radix_scatter_kernel<true, kFusedBits>): four passes of 8-bit digits
// cover the 30-bit Morton codes (three 10-bit passes measured 21 us each
// against 14: every workgroup reads four times the digit counts).  Their
// per-tile digit counts are kept digit-major with kFusedTiles entries per
// digit, one such block per pass.  Larger sorts also take 8-bit passes.
constexpr int kFusedTiles = 64;
constexpr int kFusedBits = 8;
constexpr int kFusedPasses = 4;
constexpr int kFusedHist = (1 << kFusedBits) * kFusedTiles;
static_assert(kFusedBits * kFusedPasses >= 30, "the passes cover the Morton code");
constexpr int kQueryBlock = 128;
constexpr int kQueryStack = 64;
// entries per query of the two-query first pass above 2^20 objects (deeper
// queries are redone with kQueryStack; 12 measured slower: fewer waves per SIMD)
constexpr int kQueryStackFirst2 = 8;
constexpr unsigned kRedoBlocks = 512;   // grid of the redo pass (grid-stride)

__device__ inline uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// float -> int toward zero, saturating, NaN -> 0 (the GPU conversion the
// reference relies on for empty boxes, whose centres are NaN)
__device__ inline int cvt_rzi(float f) {
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)0x80000000;
    return (int)f;
}

__device__ inline bool intersects(const float* a, const float* b) {
    return !(a[3] <= b[0] || a[4] <= b[1] || a[5] <= b[2] || a[0] >= b[3] || a[1] >= b[4] || a[2] >= b[5]);
}

// ---------------------------------------------------------------- root box
__global__ void root_partial_kernel(const float* __restrict__ aabb, int n, float* __restrict__ partial) {
    __shared__ float s[6][256];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* a = aabb + 6 * (size_t)i;
        for (int k = 0; k < 3; k++) {
            mn[k] = fminf(mn[k], a[k]);
            mx[k] = fmaxf(mx[k], a[3 + k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = mn[k];
        s[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; k++) {
                s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + w]);
                s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline uint32_t morton_code(const float* __restrict__ a, const float* root) {
    int c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float centre = (a[k] + a[3 + k]) * 0.5f;
        const float coord = (centre - root[k]) / (root[3 + k] - root[k]) * 1024.f;
        const int v = cvt_rzi(coord);
        c[k] = v < 0 ? 0 : (v > 1023 ? 1023 : v);
    }
    return expand_bits((uint32_t)c[0]) * 4 + expand_bits((uint32_t)c[1]) * 2 + expand_bits((uint32_t)c[2]);
}

// ---------------------------------------------------------------- radix sort
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys, int n, int shift,
                                                                  int tiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
    for (int r = 0; r < kSortItems; r++) {
        const int i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// mortonKernel (lbvh.cu:53-60) over one radix tile (kTile objects), fused
// with the final reduction of the root box (every workgroup reduces the
// partials of root_partial_kernel; min and max give the same box in any
// order) and with the first sort pass's digit histogram of the tile
// (hist[digit * stride + tile], BITS-bit digits).  For a one-launch-per-pass
// sort, each workgroup also clears its tile's counts of the later passes
// (hist_next, kFusedHist words per pass) and workgroup 0 the word at zero_word.
template <int BITS>
__global__ __launch_bounds__(kSortThreads) void morton_hist_kernel(const float* __restrict__ aabb,
                                                                   const float* __restrict__ partial, int pblocks,
                                                                   int n, int stride, float* __restrict__ root_out,
                                                                   uint32_t* __restrict__ codes,
                                                                   uint32_t* __restrict__ ids,
                                                                   uint32_t* __restrict__ hist,
                                                                   uint32_t* __restrict__ hist_next, int next_passes,
                                                                   uint32_t* __restrict__ zero_word) {
    constexpr int D = 1 << BITS;
    __shared__ float s[6][kSortThreads];
    __shared__ uint32_t cnt[D];
    const int t = threadIdx.x;
    const int base = blockIdx.x * kTile;
    // the tile's boxes are loaded before the root box is reduced, so the two
    // loads' latencies overlap
    float box[kSortItems][6];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int i = base + r * kSortThreads + t;
        for (int k = 0; k < 6; k++) box[r][k] = i < n ? aabb[6 * (size_t)i + k] : 0.0f;
    }
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = t; b < pblocks; b += kSortThreads)
        for (int k = 0; k < 6; k++) v[k] = k < 3 ? fminf(v[k], partial[6 * b + k]) : fmaxf(v[k], partial[6 * b + k]);
    for (int k = 0; k < 6; k++) s[k][t] = v[k];
    for (int d = t; d < D; d += kSortThreads) cnt[d] = 0;
    __syncthreads();
    for (int w = kSortThreads / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 6; k++)
                s[k][t] = k < 3 ? fminf(s[k][t], s[k][t + w]) : fmaxf(s[k][t], s[k][t + w]);
        __syncthreads();
    }
    float root[6];
    for (int k = 0; k < 6; k++) root[k] = s[k][0];
    if (blockIdx.x == 0 && t < 6) root_out[t] = root[t];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int i = base + r * kSortThreads + t;
        if (i < n) {
            const uint32_t code = morton_code(box[r], root);
            codes[i] = code;
            ids[i] = (uint32_t)i;
            atomicAdd(&cnt[code & (uint32_t)(D - 1)], 1u);
        }
    }
    __syncthreads();
    for (int d = t; d < D; d += kSortThreads) hist[(size_t)d * stride + blockIdx.x] = cnt[d];
    if (hist_next)
        for (int k = 0; k < next_passes; k++)
            for (int d = t; d < D; d += kSortThreads)
                hist_next[(size_t)k * kFusedHist + d * kFusedTiles + blockIdx.x] = 0u;
    if (zero_word && blockIdx.x == 0 && t == 0) *zero_word = 0u;
}

// Exclusive scan of at most kSmallScanMax values in one workgroup: the values
// are staged in LDS with coalesced loads, each lane sums a run of consecutive
// values there, the lanes' sums are scanned, and each lane writes its run's
// prefixes back to LDS, which is stored coalesced.  (Integer sums: the same
// values as the three-kernel scan below.)
constexpr int kSmallScanThreads = 1024;
constexpr int kSmallScanMax = 1 << 15;
__global__ __launch_bounds__(kSmallScanThreads) void scan_small_kernel(uint32_t* __restrict__ v, int total) {
    __shared__ uint32_t vals[kSmallScanMax];
    __shared__ uint32_t s[kSmallScanThreads];
    const int t = threadIdx.x;
    for (int i = t; i < total; i += kSmallScanThreads) vals[i] = v[i];
    __syncthreads();
    const int per = (total + kSmallScanThreads - 1) / kSmallScanThreads;
    const int b = t * per;
    const int e = min(b + per, total);
    uint32_t sum = 0;
    for (int i = b; i < e; i++) sum += vals[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < kSmallScanThreads; off <<= 1) {
        const uint32_t x = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    uint32_t run = t > 0 ? s[t - 1] : 0u;
    for (int i = b; i < e; i++) {
        const uint32_t x = vals[i];
        vals[i] = run;
        run += x;
    }
    __syncthreads();
    for (int i = t; i < total; i += kSmallScanThreads) v[i] = vals[i];
}

// exclusive scan of chunks of kScanChunk values; chunk totals to sums
__global__ __launch_bounds__(256) void scan_chunk_kernel(uint32_t* __restrict__ v, size_t total,
                                                          uint32_t* __restrict__ sums) {
    __shared__ uint32_t s[kScanChunk];
    const size_t base = (size_t)blockIdx.x * kScanChunk;
    for (int k = threadIdx.x; k < kScanChunk; k += 256) s[k] = (base + k < total) ? v[base + k] : 0u;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 values with 256 threads x 4
    for (int off = 1; off < kScanChunk; off <<= 1) {
        uint32_t tmp[4];
        for (int j = 0; j < 4; j++) {
            const int k = threadIdx.x + 256 * j;
            tmp[j] = (k >= off) ? s[k - off] : 0u;
        }
        __syncthreads();
        for (int j = 0; j < 4; j++) s[threadIdx.x + 256 * j] += tmp[j];
        __syncthreads();
    }
    for (int k = threadIdx.x; k < kScanChunk; k += 256)
        if (base + k < total) v[base + k] = (k == 0) ? 0u : s[k - 1];
    if (threadIdx.x == 0) sums[blockIdx.x] = s[kScanChunk - 1];
}

// exclusive scan of the chunk totals in one workgroup (sequential over
// blocks of 256 values with a running carry)
__global__ __launch_bounds__(256) void scan_sums_kernel(uint32_t* __restrict__ sums, int m) {
    __shared__ uint32_t s[256];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < m; base += 256) {
        const int k = base + threadIdx.x;
        s[threadIdx.x] = k < m ? sums[k] : 0u;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const uint32_t t = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        const uint32_t incl = s[threadIdx.x];
        const uint32_t excl = carry + incl - (k < m ? sums[k] : 0u);
        __syncthreads();
        if (k < m) sums[k] = excl;
        if (threadIdx.x == 255) carry += incl;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void scan_add_kernel(uint32_t* __restrict__ v, size_t total,
                                                        const uint32_t* __restrict__ sums) {
    const size_t base = (size_t)blockIdx.x * kScanChunk;
    const uint32_t add = sums[blockIdx.x];
    for (int k = threadIdx.x; k < kScanChunk; k += 256)
        if (base + k < total) v[base + k] += add;
}

// Stable scatter: within a tile, items keep their index order; tiles keep
// their order through the digit-major, tile-minor offsets.  The kernel reads
// none of its global writes, so its barriers order LDS only.  BITS-bit
// digits at `shift`.
// FUSED = false (8-bit digits): offsets is the exclusive scan of the per-tile
// digit counts (digit-major, `tiles` per digit).
// FUSED = true (at most kFusedTiles tiles): offsets holds the per-tile digit
// counts themselves (kFusedTiles per digit) and every workgroup forms its own
// offsets from them: the digit totals below its digit plus its digit's counts
// in the earlier tiles, the same values as the scan.  With hist_next (cleared
// beforehand), each key also counts its next digit in the tile of its
// destination, one global atomic per key: the next pass's per-tile counts,
// so the pass needs no launch of its own for them.
template <bool FUSED, int BITS>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(const uint32_t* __restrict__ keys,
                                                                     const uint32_t* __restrict__ vals, int n, int shift,
                                                                     int tiles, const uint32_t* __restrict__ offsets,
                                                                     uint32_t* __restrict__ keys_out,
                                                                     uint32_t* __restrict__ vals_out,
                                                                     uint32_t* __restrict__ hist_next) {
    constexpr int D = 1 << BITS;
    constexpr int DPT = D / kSortThreads;  // digits per thread, consecutive
    static_assert(D % kSortThreads == 0, "whole digits per thread");
    static_assert(FUSED || DPT == 1, "the scanned offsets are one digit per thread");
    constexpr uint32_t kMask = D - 1;
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t running[D];
    __shared__ uint32_t wave_cnt[kWaves][D];
    __shared__ uint32_t wave_sum[kWaves];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    // Wave w owns the tile's keys w * kWaveKeys .. (w + 1) * kWaveKeys - 1,
    // loaded before any is used (one latency, not one per round).  It ranks
    // them round by round (64 keys a round) against its own digit counts in
    // LDS, wave-synchronously; the waves' counts then give each wave's offset
    // per digit.  A key's place is (digits below, earlier tiles, earlier
    // waves, earlier rounds, earlier lanes) with the same digit: its index
    // order, so the pass is stable.
    constexpr int kWaveKeys = kSortItems * 64;
    const int base = blockIdx.x * kTile + wave * kWaveKeys;
    uint32_t kreg[kSortItems], vreg[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int i = base + r * 64 + lane;
        kreg[r] = i < n ? keys[i] : 0u;
        vreg[r] = i < n ? vals[i] : 0u;
    }
    for (int w = 0; w < kWaves; w++)
        for (int k = 0; k < DPT; k++) wave_cnt[w][t * DPT + k] = 0;
    if (FUSED) {
        // this thread's digits: their totals and their counts in the earlier
        // tiles (entries past `tiles` are never written: masked, not summed)
        uint32_t before[DPT], total[DPT];
#pragma unroll
        for (int k = 0; k < DPT; k++) {
            const uint32_t* row = offsets + (size_t)(t * DPT + k) * kFusedTiles;
            before[k] = total[k] = 0;
#pragma unroll
            for (int b = 0; b < kFusedTiles; b += 4) {
                const uint4 v = *reinterpret_cast<const uint4*>(row + b);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t x = b + j < tiles ? w[j] : 0u;
                    total[k] += x;
                    before[k] += b + j < (int)blockIdx.x ? x : 0u;
                }
            }
        }
        // exclusive sum of the digit totals over the digits below each
        uint32_t tsum = 0;
#pragma unroll
        for (int k = 0; k < DPT; k++) tsum += total[k];
        uint32_t incl = tsum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == 63) wave_sum[wave] = incl;
        lds_barrier();
        uint32_t below = incl - tsum;
        for (int w = 0; w < wave; w++) below += wave_sum[w];
#pragma unroll
        for (int k = 0; k < DPT; k++) {
            running[t * DPT + k] = below + before[k];
            below += total[k];
        }
    } else {
        running[t] = offsets[(size_t)t * tiles + blockIdx.x];
    }
    lds_barrier();  // the wave counts are zero
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const bool valid = base + r * 64 + lane < n;
        const uint32_t digit = valid ? (kreg[r] >> shift) & kMask : 0u;
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; b++) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t ones = __ballot(valid && bit);
            match &= bit ? ones : ~ones;
        }
        const uint32_t earlier = wave_cnt[wave][digit];  // this digit in the wave's earlier rounds
        rank[r] = earlier + (uint32_t)__popcll(match & lt_mask);
        __builtin_amdgcn_wave_barrier();  // every lane has read before the count moves on
        if (valid && (match & lt_mask) == 0) wave_cnt[wave][digit] = earlier + (uint32_t)__popcll(match);
        __builtin_amdgcn_wave_barrier();
    }
    lds_barrier();
    // each wave's offset per digit: the counts of the waves before it
#pragma unroll
    for (int k = 0; k < DPT; k++) {
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const uint32_t c = wave_cnt[w][t * DPT + k];
            wave_cnt[w][t * DPT + k] = acc;
            acc += c;
        }
    }
    lds_barrier();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        if (base + r * 64 + lane >= n) continue;
        const uint32_t key = kreg[r];
        const uint32_t digit = (key >> shift) & kMask;
        const uint32_t off = running[digit] + wave_cnt[wave][digit] + rank[r];
        keys_out[off] = key;
        vals_out[off] = vreg[r];
        if (FUSED && hist_next)
            atomicAdd(hist_next + ((key >> (shift + BITS)) & kMask) * kFusedTiles + off / kTile, 1u);
    }
}

// ---------------------------------------------------------------- build
__device__ inline int clz64(uint64_t x) { return x ? __clzll(x) : 64; }
__device__ inline uint64_t sort_key(const uint32_t* codes, int i) { return ((uint64_t)codes[i] << 32) | (uint32_t)i; }

// Karras's internal node construction (lbvh.cu:68-150) for node i: its leaf
// range [lo, hi] and the split (left child ends at split).
__device__ inline void karras_node(const uint32_t* __restrict__ codes, int n, int i, int& lo, int& hi, int& split) {
    if (i == 0) {
        lo = 0;
        hi = n - 1;
    } else {
        const uint64_t self = sort_key(codes, i);
        const int dl = clz64(self ^ sort_key(codes, i - 1));
        const int dr = clz64(self ^ sort_key(codes, i + 1));
        const int d = dr > dl ? 1 : -1;
        const int dmin = min(dl, dr);
        int lmax = 2;
        for (;;) {
            const int j = i + d * lmax;
            if (j < 0 || j >= n) break;
            if (clz64(self ^ sort_key(codes, j)) <= dmin) break;
            lmax <<= 1;
        }
        int l = 0;
        for (int t = lmax >> 1; t > 0; t >>= 1) {
            const int j = i + (l + t) * d;
            if (j >= 0 && j < n && clz64(self ^ sort_key(codes, j)) > dmin) l += t;
        }
        const int j = i + l * d;
        lo = min(i, j);
        hi = max(i, j);
    }
    const uint64_t fk = sort_key(codes, lo);
    const int dnode = clz64(fk ^ sort_key(codes, hi));
    int s = lo, stride = hi - lo;
    do {
        stride = (stride + 1) >> 1;
        const int mid = s + stride;
        if (mid < hi && clz64(fk ^ sort_key(codes, mid)) > dnode) s = mid;
    } while (stride > 1);
    split = s;
}

// Writes node i's links and its children's parent links; returns the child
// words (leaf bit in the MSB).
__device__ inline uint2 write_links(LbvhNode* __restrict__ nodes, uint32_t* __restrict__ leaf_parent, int i, int lo,
                                    int hi, int split) {
    if (i == 0) nodes[0].parent = 0;  // the root has no parent (value-initialised in the reference)
    nodes[i].fence = (i == lo) ? (uint32_t)hi : (uint32_t)lo;
    uint32_t left, right;
    if (lo == split) {
        leaf_parent[split] = (uint32_t)i;
        left = (uint32_t)split | 0x80000000u;
    } else {
        left = (uint32_t)split;
        nodes[split].parent = (uint32_t)i;
    }
    if (hi == split + 1) {
        leaf_parent[split + 1] = (uint32_t)i | 0x80000000u;
        right = (uint32_t)(split + 1) | 0x80000000u;
    } else {
        right = (uint32_t)(split + 1);
        nodes[split + 1].parent = (uint32_t)i | 0x80000000u;
    }
    nodes[i].left = left;
    nodes[i].right = right;
    return make_uint2(left, right);
}

// All internal nodes; also clears what the height pass accumulates into: each
// internal node's arrival word and the per-height counters (2 kMaxLevels words).
__global__ void build_internal_kernel(const uint32_t* __restrict__ codes, int n, LbvhNode* __restrict__ nodes,
                                      uint32_t* __restrict__ leaf_parent, int* __restrict__ word,
                                      uint32_t* __restrict__ level_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * kMaxLevels) level_count[i] = 0;
    if (i >= n - 1) return;
    word[i] = 0;
    int lo, hi, split;
    karras_node(codes, n, i, lo, hi, split);
    write_links(nodes, leaf_parent, i, lo, hi, split);
}

// Bottom-up refit (lbvh.cu:153-188): the second thread to reach a node
// merges both children and continues; depth = max(dL, dR + 1).
// Tree levels, once per build.  Every leaf walks up; at each internal node the
// first arriving child leaves its height, DFS stack depth and side in the
// node's word with one atomicAdd and stops; the second reads them back from
// the atomic's return value and continues.  No other memory passes between the
// two, so no fence is needed (fences would write back and invalidate the L2 of
// every XCD).  height(leaf) = 0, height(node) = 1 + max(children); stack depth
// (the DFS bound of mergeUpKernel, lbvh.cu:153-188): leaf 1, node max(dL, dR + 1).
__global__ void height_kernel(const uint32_t* __restrict__ leaf_parent, const LbvhNode* __restrict__ nodes,
                              uint32_t* __restrict__ word, int n, uint32_t* __restrict__ node_height,
                              int* __restrict__ max_depth) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint32_t h = 0, d = 1;
    uint32_t parent = leaf_parent[t];
    for (;;) {
        const uint32_t right = parent >> 31;
        const uint32_t p = parent & 0x7FFFFFFFu;
        const uint32_t old = atomicAdd(word + p, 1u | (h << 2) | (d << 10) | (right << 18));
        if ((old & 3u) == 0) return;  // first child here
        const uint32_t oh = (old >> 2) & 255u, od = (old >> 10) & 255u;
        const uint32_t dl = right ? od : d, dr = right ? d : od;
        h = 1 + (h > oh ? h : oh);
        d = dl > dr + 1 ? dl : dr + 1;
        node_height[p] = h;
        if (p == 0) {
            *max_depth = (int)d;
            return;
        }
        parent = nodes[p].parent;
    }
}

// Internal nodes per height that cross a refit tile boundary, and those nodes
// grouped by height.  A workgroup counts its kLevelChunk nodes in LDS and touches each
// global counter once: per-node global atomics on a few dozen counters
// serialise (most nodes sit at the lowest heights).  A height past the bound
// can only come from a corrupt tree; the host then finds the counts short and
// reports an error.
constexpr int kLevelBlock = 256, kLevelPerThread = 16, kLevelChunk = kLevelBlock * kLevelPerThread;

// Refit tiles: kRefitTile consecutive leaves (sorted positions; 256, 512 and
// 1024 measure within 2 % of each other on 10M segments).  An internal
// node whose leaf range lies inside one tile has its whole subtree there;
// tile_fit_kernel fits those in one workgroup per tile (bottom-up in LDS, for
// trees of every size), and for larger trees only the nodes whose ranges
// cross a tile boundary go into the per-height lists.
constexpr int kRefitTile = 512;

__device__ inline bool crosses_tiles(const LbvhNode* __restrict__ nodes, int p) {
    const uint32_t f = nodes[p].fence;
    const uint32_t lo = min((uint32_t)p, f), hi = max((uint32_t)p, f);
    return lo / kRefitTile != hi / kRefitTile;
}

__global__ __launch_bounds__(kLevelBlock) void level_count_kernel(const uint32_t* __restrict__ node_height,
                                                                  const LbvhNode* __restrict__ nodes, int m,
                                                                  uint32_t* __restrict__ level_count) {
    __shared__ uint32_t bins[kMaxLevels];
    for (int i = threadIdx.x; i < kMaxLevels; i += kLevelBlock) bins[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * kLevelChunk;
    for (int j = 0; j < kLevelPerThread; j++) {
        const int p = base + j * kLevelBlock + threadIdx.x;
        if (p < m && node_height[p] < (uint32_t)kMaxLevels && crosses_tiles(nodes, p))
            atomicAdd(&bins[node_height[p]], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMaxLevels; i += kLevelBlock)
        if (bins[i]) atomicAdd(level_count + i, bins[i]);
}

// The offset of each height's nodes is the exclusive sum of the counts below
// it; every workgroup forms it in LDS, and workgroup 0 publishes (offset,
// count) per height for the refit kernels.
__global__ __launch_bounds__(kLevelBlock) void level_scatter_kernel(const uint32_t* __restrict__ node_height,
                                                                    const LbvhNode* __restrict__ nodes, int m,
                                                                    uint32_t* __restrict__ level_count,
                                                                    int2* __restrict__ level_range,
                                                                    uint32_t* __restrict__ level_nodes) {
    __shared__ uint32_t bins[kMaxLevels], start[kMaxLevels], offset[kMaxLevels];
    uint32_t* cursor = level_count + kMaxLevels;
    for (int i = threadIdx.x; i < kMaxLevels; i += kLevelBlock) bins[i] = 0;
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int h = 0; h < kMaxLevels; h++) {
            offset[h] = acc;
            acc += level_count[h];
        }
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int h = threadIdx.x; h < kMaxLevels; h += kLevelBlock)
            level_range[h] = make_int2((int)offset[h], (int)level_count[h]);
    const int base = blockIdx.x * kLevelChunk;
    uint32_t h[kLevelPerThread], rank[kLevelPerThread];
    for (int j = 0; j < kLevelPerThread; j++) {
        const int p = base + j * kLevelBlock + threadIdx.x;
        h[j] = (p < m && crosses_tiles(nodes, p)) ? node_height[p] : (uint32_t)kMaxLevels;
        if (h[j] < (uint32_t)kMaxLevels) rank[j] = atomicAdd(&bins[h[j]], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMaxLevels; i += kLevelBlock)
        if (bins[i]) start[i] = atomicAdd(cursor + i, bins[i]);
    __syncthreads();
    for (int j = 0; j < kLevelPerThread; j++) {
        const int p = base + j * kLevelBlock + threadIdx.x;
        if (h[j] < (uint32_t)kMaxLevels) level_nodes[offset[h[j]] + start[h[j]] + rank[j]] = (uint32_t)p;
    }
}

// Child boxes of the internal nodes of one height; their children are leaves
// or nodes of smaller height, finished by earlier launches.  Unions take glm's
// min/max (Bound::absorb), so a NaN coordinate propagates as in the reference.
__device__ inline void refit_node(const float* __restrict__ aabb, const uint32_t* __restrict__ ids,
                                  LbvhNode* __restrict__ nodes, uint32_t p) {
    const uint32_t child[2] = {nodes[p].left, nodes[p].right};
    for (int side = 0; side < 2; side++) {
        const uint32_t c = child[side];
        float b[6];
        if (c & 0x80000000u) {
            const float* a = aabb + 6 * (size_t)ids[c & 0x7FFFFFFFu];
            for (int j = 0; j < 6; j++) b[j] = a[j];
        } else {
            const LbvhNode& cn = nodes[c];
            for (int j = 0; j < 3; j++) {
                b[j] = cn.b[1][j] < cn.b[0][j] ? cn.b[1][j] : cn.b[0][j];
                b[3 + j] = cn.b[0][3 + j] < cn.b[1][3 + j] ? cn.b[1][3 + j] : cn.b[0][3 + j];
            }
        }
        for (int j = 0; j < 6; j++) nodes[p].b[side][j] = b[j];
    }
}

// All crossing nodes of height h (grid-stride; the level's range from the device).
__global__ void refit_level_kernel(const float* __restrict__ aabb, const uint32_t* __restrict__ ids,
                                   LbvhNode* __restrict__ nodes, const uint32_t* __restrict__ level_nodes,
                                   const int2* __restrict__ level_range, int h) {
    const int2 lv = level_range[h];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < lv.y; k += gridDim.x * blockDim.x)
        refit_node(aabb, ids, nodes, level_nodes[lv.x + k]);
}

// The crossing nodes of heights h0 and up, in one workgroup: a barrier
// between heights orders them (the workgroup's writes are visible to it after
// the barrier), which saves one launch per height.  Heights without crossing
// nodes are skipped (the test is uniform, so no barrier is needed there).
constexpr int kTopBlock = 1024;
__global__ __launch_bounds__(kTopBlock) void refit_top_kernel(const float* __restrict__ aabb,
                                                              const uint32_t* __restrict__ ids,
                                                              LbvhNode* __restrict__ nodes,
                                                              const uint32_t* __restrict__ level_nodes,
                                                              const int2* __restrict__ level_range, int h0) {
    __shared__ int2 s_range[kMaxLevels];
    for (int h = threadIdx.x; h < kMaxLevels; h += kTopBlock) s_range[h] = level_range[h];
    __syncthreads();
    for (int h = h0; h < kMaxLevels; h++) {
        const int2 lv = s_range[h];
        if (lv.y == 0) continue;
        for (int k = threadIdx.x; k < lv.y; k += kTopBlock) refit_node(aabb, ids, nodes, level_nodes[lv.x + k]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------- small trees
// Trees over at most kSmallBuild objects are fit without heights or level
// lists, bottom-up in LDS as mergeUpKernel (lbvh.cu:153-188) fits in global
// memory: tile_fit_kernel fits the nodes inside each tile of kRefitTile
// leaves, and cross_fit_kernel the nodes that cross a tile boundary, in one
// workgroup.  Every finished node's lane takes an LDS atomic arrival at its
// parent; the second child to arrive forms the parent's union and goes on up.
// Each union is written before the arrival that publishes it (release) and
// read after the arrival that finds it (acquire), LDS-scoped fences.  The
// child boxes go to the nodes after one barrier.  The unions are formed as
// refit_node forms a child's union from the child's two boxes, so every node
// box equals the height-ordered refit's bit for bit.
constexpr int kSmallBuild = kFusedTiles * kTile;
constexpr uint16_t kNoParent = 0xFFFFu;

__device__ inline void child_union(const float* b0, const float* b1, float* u) {
    for (int j = 0; j < 3; j++) {
        u[j] = b1[j] < b0[j] ? b1[j] : b0[j];
        u[3 + j] = b0[3 + j] < b1[3 + j] ? b1[3 + j] : b0[3 + j];
    }
}

__device__ inline void store_child_boxes(LbvhNode* __restrict__ node, const float (&b)[2][6]) {
    float4* q = reinterpret_cast<float4*>(node->b[0]);
    q[0] = make_float4(b[0][0], b[0][1], b[0][2], b[0][3]);
    q[1] = make_float4(b[0][4], b[0][5], b[1][0], b[1][1]);
    q[2] = make_float4(b[1][2], b[1][3], b[1][4], b[1][5]);
}

__device__ inline bool range_crosses(uint32_t p, uint32_t fence) {
    return min(p, fence) / kRefitTile != max(p, fence) / kRefitTile;
}

// The arrival of a child at its parent's LDS counter: everything the lane
// wrote to LDS before is visible to the lane that finds the count it left,
// and that lane sees what the first arrival wrote.  Returns the count before.
__device__ inline uint32_t lds_arrive(uint32_t* counter) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    const uint32_t old = atomicAdd(counter, 1u);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    return old;
}

// In the crossing list's child words: the child is an internal node whose
// range crosses a tile boundary too (indices of small trees stay below 2^30).
constexpr uint32_t kCrossBit = 0x40000000u;

// One workgroup per tile of kRefitTile sorted leaves and the internal nodes
// with the same indices.  BUILD: the nodes are built here (karras_node,
// write_links, as build_internal_kernel), each in-tile node's DFS stack depth
// (leaf 1, node max(dL, dR + 1): the bound mergeUpKernel computes,
// lbvh.cu:153-188) goes to node_dep and the root's to *max_depth, and the
// nodes whose ranges cross a tile boundary are listed for cross_fit_kernel
// (cross_list[slot] = node, cross_child[slot] = its child words with
// kCrossBit on crossing children, slot_of[node] = slot; the slot order varies
// from run to run, the boxes do not).  Otherwise the links are read.  A node
// whose range lies in the tile has its whole subtree there: its children
// register with it in LDS, and each leaf's lane walks up until it arrives
// first or reaches a node whose parent crosses.
template <bool BUILD>
__global__ __launch_bounds__(kRefitTile) void tile_fit_kernel(const uint32_t* __restrict__ codes,
                                                              const float* __restrict__ aabb,
                                                              const uint32_t* __restrict__ ids,
                                                              LbvhNode* __restrict__ nodes,
                                                              uint32_t* __restrict__ leaf_parent, int n,
                                                              uint32_t* __restrict__ node_dep,
                                                              uint32_t* __restrict__ cross_list,
                                                              uint2* __restrict__ cross_child,
                                                              uint32_t* __restrict__ slot_of,
                                                              uint32_t* __restrict__ cross_count,
                                                              int* __restrict__ max_depth) {
    __shared__ float s_leaf[6][kRefitTile];
    __shared__ float s_node[6][kRefitTile];
    __shared__ uint32_t s_child[2][kRefitTile];
    __shared__ uint32_t s_arrive[kRefitTile];
    __shared__ uint16_t s_par_leaf[kRefitTile], s_par_node[kRefitTile];  // local parent | side << 15
    __shared__ uint8_t s_dep[kRefitTile];
    __shared__ uint32_t s_ncross, s_base;
    const int t = threadIdx.x;
    const int L = blockIdx.x * kRefitTile;
    const int p = L + t, m = n - 1;
    if (t == 0) s_ncross = 0;
    s_arrive[t] = 0;
    s_par_leaf[t] = s_par_node[t] = kNoParent;
    if (p < n) {
        const float* a = aabb + 6 * (size_t)ids[p];
        for (int j = 0; j < 6; j++) s_leaf[j][t] = a[j];
    }
    bool in = false, cross = false;
    uint32_t child[2] = {0, 0}, cross_mark = 0;
    if (p < m) {
        int lo, hi;
        if (BUILD) {
            int split;
            karras_node(codes, n, p, lo, hi, split);
            const uint2 c = write_links(nodes, leaf_parent, p, lo, hi, split);
            child[0] = c.x;
            child[1] = c.y;
            // internal children whose own ranges cross a tile boundary, for
            // the crossing list's child words
            if (!(c.x & 0x80000000u) && lo / kRefitTile != split / kRefitTile) cross_mark |= 1u;
            if (!(c.y & 0x80000000u) && (split + 1) / kRefitTile != hi / kRefitTile) cross_mark |= 2u;
        } else {
            const uint4 w = *reinterpret_cast<const uint4*>(nodes + p);  // parent, left, right, fence
            lo = min(p, (int)w.w);
            hi = max(p, (int)w.w);
            child[0] = w.y;
            child[1] = w.z;
        }
        in = lo / kRefitTile == hi / kRefitTile;
        cross = !in;
    }
    s_child[0][t] = child[0];
    s_child[1][t] = child[1];
    // the kernel reads none of its global writes: its barriers order LDS only
    lds_barrier();
    if (in)
        for (int side = 0; side < 2; side++) {
            const uint32_t c = child[side];
            const int k = (int)(c & 0x7FFFFFFFu) - L;  // in this tile, as the node is
            const uint16_t up = (uint16_t)(t | (side << 15));
            if (c & 0x80000000u)
                s_par_leaf[k] = up;
            else
                s_par_node[k] = up;
        }
    uint32_t rank = 0;
    if (BUILD && cross) rank = atomicAdd(&s_ncross, 1u);
    lds_barrier();
    if (BUILD && t == 0) s_base = s_ncross ? atomicAdd(cross_count, s_ncross) : 0u;
    // Walk from leaf t.  The walker holds the union it brings; at the parent
    // it reads the sibling's (a node's union and depth are in LDS before its
    // arrival).  The parent's links are read before the arrival: they do not
    // change during the walks.
    if (p < n) {
        uint32_t up = s_par_leaf[t];
        uint32_t dep = 1;
        float own[6];
        for (int j = 0; j < 6; j++) own[j] = s_leaf[j][t];
        // a subtree is at most kMaxLevels tall: the bound only ends a walk on a corrupt tree
        for (int step = 0; step <= kMaxLevels && up != kNoParent; step++) {
            const int q = (int)(up & 0x7FFFu);
            const int side = (int)(up >> 15);
            const uint32_t sib = s_child[side ^ 1][q];
            const uint32_t next = s_par_node[q];
            if (lds_arrive(&s_arrive[q]) == 0) break;  // the sibling goes on
            const int k = (int)(sib & 0x7FFFFFFFu) - L;
            const bool leaf = sib & 0x80000000u;
            // left and right by select (an index by `side` would select
            // over every element)
            float bl[6], br[6];
            for (int j = 0; j < 6; j++) {
                const float o = leaf ? s_leaf[j][k] : s_node[j][k];
                bl[j] = side ? o : own[j];
                br[j] = side ? own[j] : o;
            }
            const uint32_t ds = leaf ? 1u : s_dep[k];
            const uint32_t d0 = side ? ds : dep, d1 = side ? dep : ds;
            child_union(bl, br, own);
            for (int j = 0; j < 6; j++) s_node[j][q] = own[j];
            dep = d0 > d1 + 1 ? d0 : d1 + 1;
            s_dep[q] = (uint8_t)dep;
            up = next;
        }
    }
    lds_barrier();
    // each in-tile node's child boxes, from the finished unions
    if (in) {
        float b[2][6];
        for (int side = 0; side < 2; side++) {
            const uint32_t c = child[side];
            const int k = (int)(c & 0x7FFFFFFFu) - L;
            const bool leaf = c & 0x80000000u;
            for (int j = 0; j < 6; j++) b[side][j] = leaf ? s_leaf[j][k] : s_node[j][k];
        }
        store_child_boxes(nodes + p, b);
    }
    if (BUILD) {
        if (in) {
            const uint32_t dep = s_dep[t];
            node_dep[p] = dep;
            if (p == 0) *max_depth = (int)dep;
        }
        if (cross) {
            const uint32_t slot = s_base + rank;
            cross_list[slot] = (uint32_t)p;
            cross_child[slot] = make_uint2(child[0] | (cross_mark & 1u ? kCrossBit : 0u),
                                           child[1] | (cross_mark & 2u ? kCrossBit : 0u));
            slot_of[p] = slot;
        }
    }
}

// The nodes that cross a tile boundary (*cross_count of them, listed by
// tile_fit_kernel<true>), in one workgroup.  Up to `slots` of them (at most
// kCrossCap) are fit in LDS: each lane owns kCrossPer slots and gathers once
// the boxes of their children that are leaves or in-tile nodes (fit by
// tile_fit_kernel), keeping them in registers for the final writes.  A slot
// with one such child starts with that child's box and depth in its LDS entry
// and one arrival counted; a slot with two is fit at once and its lane walks
// up the crossing nodes.  Beyond `slots`, rounds over the slots with the
// boxes in global memory (refit_node) and a done word per slot (`done`, the
// round in which the slot was fit).  BUILD: DFS stack depths as in
// tile_fit_kernel, the root's to *max_depth.
constexpr int kCrossBlock = 1024;
constexpr int kCrossPer = 4;
constexpr int kCrossCap = kCrossBlock * kCrossPer;
static_assert(kCrossCap == kLbvhFitSlots, "LDS slots of the crossing-node fit");
static_assert(kSmallBuild == 131072, "small-tree bound stated in kscr_lbvh.h");
static_assert(kCrossCap < (int)kNoParent, "slot indices fit 15 bits with the side bit");

__device__ inline bool internal_crosses(const LbvhNode* __restrict__ nodes, uint32_t c) {
    return !(c & 0x80000000u) && range_crosses(c, nodes[c].fence);
}

template <bool BUILD>
__global__ __launch_bounds__(kCrossBlock) void cross_fit_kernel(const float* __restrict__ aabb,
                                                                const uint32_t* __restrict__ ids,
                                                                LbvhNode* __restrict__ nodes,
                                                                uint32_t* __restrict__ node_dep,
                                                                const uint32_t* __restrict__ cross_list,
                                                                const uint2* __restrict__ cross_child,
                                                                const uint32_t* __restrict__ slot_of,
                                                                const uint32_t* __restrict__ cross_count, int slots,
                                                                uint32_t* __restrict__ done,
                                                                int* __restrict__ max_depth) {
    __shared__ float s_box[6][kCrossCap];         // a slot's union (first: its external child's box)
    __shared__ uint32_t s_arrive[kCrossCap];
    __shared__ uint32_t s_ref[kCrossCap];         // crossing children's slots, kNoParent for external (16 bits each)
    __shared__ uint16_t s_par[kCrossCap];         // parent slot | side << 15
    __shared__ uint8_t s_dep[kCrossCap];
    const int t = threadIdx.x;
    const int count = (int)*cross_count;
    if (count > slots) {
        for (int s = t; s < count; s += kCrossBlock) done[s] = 0;
        __syncthreads();
        for (uint32_t r = 1; r <= (uint32_t)kMaxLevels + 1; r++) {
            int pending = 0;
            for (int s = t; s < count; s += kCrossBlock) {
                if (done[s]) continue;
                const uint32_t p = cross_list[s];
                const uint32_t child[2] = {nodes[p].left, nodes[p].right};
                bool ready = true;
                for (int side = 0; side < 2; side++)
                    if (internal_crosses(nodes, child[side])) {
                        const uint32_t dc = done[slot_of[child[side]]];
                        ready = ready && dc != 0 && dc < r;  // fit in an earlier round
                    }
                if (!ready) {
                    pending = 1;
                    continue;
                }
                refit_node(aabb, ids, nodes, p);
                if (BUILD) {
                    uint32_t d[2];
                    for (int side = 0; side < 2; side++)
                        d[side] = (child[side] & 0x80000000u) ? 1u : node_dep[child[side]];
                    const uint32_t dep = d[0] > d[1] + 1 ? d[0] : d[1] + 1;
                    node_dep[p] = dep;
                    if (p == 0) *max_depth = (int)dep;
                }
                done[s] = r;
            }
            if (!__syncthreads_or(pending)) break;
        }
        return;
    }
    uint32_t node[kCrossPer], dep_ext[kCrossPer];
    int ref[kCrossPer][2];
    float ext[kCrossPer][6];
    bool start[kCrossPer];
#pragma unroll
    for (int j = 0; j < kCrossPer; j++) {
        const int s = t + j * kCrossBlock;
        start[j] = false;
        node[j] = 0;
        dep_ext[j] = 0;
        ref[j][0] = ref[j][1] = -1;
        for (int k = 0; k < 6; k++) ext[j][k] = 0.0f;
        if (s < count) s_par[s] = kNoParent;
    }
    lds_barrier();
    // Two slots per lane at a time, in three steps whose loads are all in
    // flight together: the slot's node and child words; per child its slot
    // (crossing), object id (leaf) or boxes and depth (in-tile node); the
    // leaves' boxes.
#pragma unroll
    for (int jb = 0; jb < kCrossPer; jb += 2) {
        if (jb * kCrossBlock >= count) break;
        bool valid[2];
        uint32_t cw[2][2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int s = t + (jb + q) * kCrossBlock;
            valid[q] = s < count;
            const uint2 c = valid[q] ? cross_child[s] : make_uint2(0u, 0u);
            node[jb + q] = valid[q] ? cross_list[s] : 0u;
            cw[q][0] = c.x;
            cw[q][1] = c.y;
        }
        float b[2][2][12];
        uint32_t aux[2][2];
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int side = 0; side < 2; side++) {
                const uint32_t c = cw[q][side];
                aux[q][side] = 0;
                for (int k = 0; k < 12; k++) b[q][side][k] = 0.0f;
                if (!valid[q]) continue;
                if (c & 0x80000000u) {
                    aux[q][side] = ids[c & 0x7FFFFFFFu];
                } else if (c & kCrossBit) {
                    aux[q][side] = slot_of[c & ~kCrossBit];
                } else {
                    const float4* cb = reinterpret_cast<const float4*>(nodes[c].b[0]);
                    const float4 x0 = cb[0], x1 = cb[1], x2 = cb[2];
                    const float v[12] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w, x2.x, x2.y, x2.z, x2.w};
                    for (int k = 0; k < 12; k++) b[q][side][k] = v[k];
                    aux[q][side] = node_dep[c];
                }
            }
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int side = 0; side < 2; side++)
                if (valid[q] && (cw[q][side] & 0x80000000u)) {
                    const float* a = aabb + 6 * (size_t)aux[q][side];
                    for (int k = 0; k < 6; k++) b[q][side][k] = a[k];
                }
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int j = jb + q;
            if (!valid[q]) continue;
            const int s = t + j * kCrossBlock;
            float box[2][6];
            uint32_t d[2];
            for (int side = 0; side < 2; side++) {
                const uint32_t c = cw[q][side];
                d[side] = 1;
                if (c & 0x80000000u) {
                    for (int k = 0; k < 6; k++) box[side][k] = b[q][side][k];
                } else if (c & kCrossBit) {
                    ref[j][side] = (int)aux[q][side];
                    s_par[aux[q][side]] = (uint16_t)(s | (side << 15));
                    for (int k = 0; k < 6; k++) box[side][k] = 0.0f;
                } else {
                    child_union(&b[q][side][0], &b[q][side][6], box[side]);
                    d[side] = aux[q][side];
                }
            }
            s_ref[s] = (uint32_t)(ref[j][0] < 0 ? kNoParent : ref[j][0]) |
                       ((uint32_t)(ref[j][1] < 0 ? kNoParent : ref[j][1]) << 16);
            if (ref[j][0] < 0 && ref[j][1] < 0) {
                store_child_boxes(nodes + node[j], box);
                float u[6];
                child_union(box[0], box[1], u);
                for (int k = 0; k < 6; k++) s_box[k][s] = u[k];
                const uint32_t dep = d[0] > d[1] + 1 ? d[0] : d[1] + 1;
                s_dep[s] = (uint8_t)dep;
                s_arrive[s] = 0;
                start[j] = true;
                if (BUILD && node[j] == 0) *max_depth = (int)dep;
            } else {
                const int e = ref[j][0] < 0 ? 0 : 1;  // the external side, if any
                for (int k = 0; k < 6; k++) ext[j][k] = box[e][k];
                dep_ext[j] = d[e];
                const bool one = ref[j][0] < 0 || ref[j][1] < 0;
                if (one) {
                    for (int k = 0; k < 6; k++) s_box[k][s] = box[e][k];
                    s_dep[s] = (uint8_t)d[e];
                }
                s_arrive[s] = one ? 1u : 0u;  // the external child counts as arrived
            }
        }
    }
    lds_barrier();
    // Walks from the slots fit above, one after another on each lane (the
    // next starts as soon as one ends).  The walker holds the union it
    // brings; at the parent it reads the other side's box (a crossing child's
    // union, or the external child's box staged in the parent's own entry).
    // The parent's links are read before the arrival: they do not change
    // during the walks.
    {
        uint32_t todo = 0;
#pragma unroll
        for (int j = 0; j < kCrossPer; j++) todo |= start[j] ? 1u << j : 0u;
        uint32_t up = kNoParent, dep = 0;
        float own[6] = {};
        // each walk is at most kMaxLevels steps: the bound only ends the loop on a corrupt tree
        for (int step = 0; step < kCrossPer * (kMaxLevels + 2); step++) {
            if (up == kNoParent) {
                if (!todo) break;
                const int s = t + __builtin_ctz(todo) * kCrossBlock;
                todo &= todo - 1;
                for (int i = 0; i < 6; i++) own[i] = s_box[i][s];
                dep = s_dep[s];
                up = s_par[s];
                continue;
            }
            const int q = (int)(up & 0x7FFFu);
            const int side = (int)(up >> 15);
            const uint32_t rw = s_ref[q];
            const uint32_t next = s_par[q];
            if (lds_arrive(&s_arrive[q]) == 0) {  // the other child goes on
                up = kNoParent;
                continue;
            }
            const uint32_t other = side ? (rw & 0xFFFFu) : (rw >> 16);
            const int k = other == kNoParent ? q : (int)other;
            // left and right by select (an index by `side` would select
            // over every element)
            float bl[6], br[6];
            for (int i = 0; i < 6; i++) {
                const float o = s_box[i][k];
                bl[i] = side ? o : own[i];
                br[i] = side ? own[i] : o;
            }
            const uint32_t ds = s_dep[k];
            const uint32_t d0 = side ? ds : dep, d1 = side ? dep : ds;
            child_union(bl, br, own);
            for (int i = 0; i < 6; i++) s_box[i][q] = own[i];
            dep = d0 > d1 + 1 ? d0 : d1 + 1;
            s_dep[q] = (uint8_t)dep;
            up = next;
            if (BUILD && up == kNoParent) *max_depth = (int)dep;  // the root is the only slot without one
        }
    }
    lds_barrier();
    // the child boxes of the slots with a crossing child, from the finished unions
#pragma unroll
    for (int j = 0; j < kCrossPer; j++) {
        if (ref[j][0] < 0 && ref[j][1] < 0) continue;  // fit above, or no slot
        float b[2][6];
        for (int side = 0; side < 2; side++) {
            const int r = ref[j][side];
            for (int k = 0; k < 6; k++) b[side][k] = r >= 0 ? s_box[k][r] : ext[j][k];
        }
        store_child_boxes(nodes + node[j], b);
    }
}

// ---------------------------------------------------------------- query
// Self query with IGNORE_SELF (lbvh.cu:196-289).  Each lane walks the tree
// depth first: of the two children of a node it continues with the left one
// and keeps the right one on its LDS stack (a column per lane) only when both
// boxes overlap; leaf children are tested where they are found, so a pass
// visits one internal node and finds up to two pairs.  Subtrees whose leaf
// ranges end at or before the query's own sorted index are skipped (each
// pair is reported once, by its lower index).
//
// Pairs collect in an LDS buffer per wave and go to the global list kWaveBuf
// at a time, one atomic per flush (as the reference's shared pair buffer,
// lbvh.cu:211-280): an atomic on the one global counter per found pair
// serialises the whole query.
constexpr int kWaveBuf = 128;
constexpr int kQueryWaves = kQueryBlock / 64;

// REDO = false: one query per lane over the sorted objects, with a STACK-entry
// LDS stack.  A query that needs more stack stops and appends (its sorted
// index, the pairs it has reported) to the redo list.
// REDO = true: the listed queries again with a deeper stack (grid-stride over
// the list).  The traversal order does not depend on the stack size, so a
// redone query finds the same pairs in the same order and reports only those
// after the ones its first pass reported.
template <int STACK, bool REDO>
__global__ __launch_bounds__(kQueryBlock) void query_kernel(const float* __restrict__ aabb,
                                                            const uint32_t* __restrict__ ids,
                                                            const LbvhNode* __restrict__ nodes, int n,
                                                            int2* __restrict__ pairs, int cap, int min_gap,
                                                            int* __restrict__ count, int* __restrict__ error,
                                                            int2* __restrict__ redo, int* __restrict__ redo_count,
                                                            int redo_cap) {
    __shared__ uint32_t stack[STACK][kQueryBlock];
    __shared__ int2 wbuf[kQueryWaves][kWaveBuf];
    __shared__ int wfill[kQueryWaves];
    const int w = threadIdx.x >> 6;
    const uint64_t lane_below = (1ull << (threadIdx.x & 63)) - 1;
    if ((threadIdx.x & 63) == 0) wfill[w] = 0;
    __builtin_amdgcn_wave_barrier();
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < 2) return;
    if (!REDO && gid >= n) return;

    // copies the wave's buffered pairs to the global list with the wave's
    // active lanes; every active lane sees the same fill
    auto flush = [&](int fill) {
        const uint64_t act = __ballot(1);
        const int leader = __ffsll((unsigned long long)act) - 1;
        int base = 0;
        if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(count, fill);
        base = __shfl(base, leader);
        const int nact = __popcll(act);
        for (int k = __popcll(act & lane_below); k < fill; k += nact)
            if (base + k < cap) pairs[base + k] = wbuf[w][k];
    };
    constexpr uint32_t kLeaf = 0x80000000u, kNone = 0xFFFFFFFFu;
    // the query of sorted object t; the first `skip` pairs it finds are not
    // reported (a redo's first pass reported them)
    auto run = [&](const int t, const int skip) {
    const int qid = (int)ids[t];
    float q[6];
    for (int k = 0; k < 6; k++) q[k] = aabb[6 * (size_t)qid + k];
    if (!(q[0] <= q[3])) return;  // empty box (strand tip): intersects nothing
    // object id of a leaf paired with the query, or -1
    auto leaf_pair = [&](uint32_t leaf) -> int {
        if ((int)leaf <= t) return -1;
        const int fid = (int)ids[leaf];
        if (min_gap >= 0 && abs(fid - qid) <= min_gap) return -1;
        return fid;
    };
    int sp = 0, seen = 0;
    uint32_t node = 0;  // the root (internal: n >= 2)
    while (node != kNone) {
        int f0 = -1, f1 = -1;
        uint32_t next = kNone;
        // the whole 64-byte node in one round trip (four 16-byte loads)
        const float4* np = reinterpret_cast<const float4*>(nodes + node);
        const float4 w0 = np[0], w1 = np[1], w2 = np[2], w3 = np[3];
        const uint32_t fence = __float_as_uint(w0.w);
        bool overflow = false;
        if (max(node, fence) > (uint32_t)t) {
            const uint32_t lc = __float_as_uint(w0.y), rc = __float_as_uint(w0.z);
            const float b0[6] = {w1.x, w1.y, w1.z, w1.w, w2.x, w2.y};
            const float b1[6] = {w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
            bool h0 = intersects(b0, q), h1 = intersects(b1, q);
            if (h0 && (lc & kLeaf)) {
                f0 = leaf_pair(lc & ~kLeaf);
                h0 = false;
            }
            if (h1 && (rc & kLeaf)) {
                f1 = leaf_pair(rc & ~kLeaf);
                h1 = false;
            }
            if (h0 && h1) {
                if (sp + 1 > STACK) {  // deeper than the LDS stack: stop this query
                    if (REDO && error) atomicOr(error, 4);
                    overflow = true;
                } else {
                    stack[sp++][threadIdx.x] = rc;
                    next = lc;
                }
            } else if (h0) {
                next = lc;
            } else if (h1) {
                next = rc;
            }
        }
        if (!overflow && next == kNone && sp > 0) next = stack[--sp][threadIdx.x];
        if (REDO) {
            if (f0 >= 0 && seen++ < skip) f0 = -1;
            if (f1 >= 0 && seen++ < skip) f1 = -1;
        } else {
            seen += (f0 >= 0) + (f1 >= 0);
        }
        // the lanes of the wave that found pairs in this pass append them to
        // the wave's buffer in lane order
        const uint64_t m0 = __ballot(f0 >= 0), m1 = __ballot(f1 >= 0);
        const int found = __popcll(m0) + __popcll(m1);
        if (found) {
            int fill = wfill[w];
            if (fill + found > kWaveBuf) {
                flush(fill);
                fill = 0;
            }
            if (f0 >= 0) wbuf[w][fill + __popcll(m0 & lane_below)] = make_int2(f0, qid);
            if (f1 >= 0) wbuf[w][fill + __popcll(m0) + __popcll(m1 & lane_below)] = make_int2(f1, qid);
            __builtin_amdgcn_wave_barrier();
            if ((int)(threadIdx.x & 63) == __ffsll((unsigned long long)(m0 | m1)) - 1) wfill[w] = fill + found;
            __builtin_amdgcn_wave_barrier();
        }
        if (!REDO && overflow) {  // the redo pass takes it from here
            const int slot = atomicAdd(redo_count, 1);
            if (slot < redo_cap)
                redo[slot] = make_int2(t, seen);
            else if (error)
                atomicOr(error, 4);
        }
        node = overflow ? kNone : next;
    }
    };
    if (REDO) {
        const int m = min(*redo_count, redo_cap);
        for (int r = gid; r < m; r += gridDim.x * blockDim.x) {
            const int2 e = redo[r];
            run(e.x, e.y);
        }
    } else {
        run(gid, 0);
    }
    // the lanes still here are the wave's lanes that ran queries
    __builtin_amdgcn_wave_barrier();
    const int fill = wfill[w];
    if (fill > 0) flush(fill);
}

// The first pass over large trees: two queries per lane, the sorted objects
// base + lane and base + 64 + lane of the wave's 128.  A query's walk is a
// chain of dependent node loads; the two walks advance in lockstep with both
// node loads issued before either is used, so a lane keeps two loads in
// flight.  Each query walks, reports and stops exactly as query_kernel's
// first pass does (same order, same pairs, same redo entry), so the redo
// pass is unchanged.
template <int STACK>
__global__ __launch_bounds__(kQueryBlock) void query2_kernel(const float* __restrict__ aabb,
                                                             const uint32_t* __restrict__ ids,
                                                             const LbvhNode* __restrict__ nodes, int n,
                                                             int2* __restrict__ pairs, int cap, int min_gap,
                                                             int* __restrict__ count, int* __restrict__ error,
                                                             int2* __restrict__ redo, int* __restrict__ redo_count,
                                                             int redo_cap) {
    __shared__ uint32_t stack[STACK][2 * kQueryBlock];  // a column per query
    __shared__ int2 wbuf[kQueryWaves][kWaveBuf];
    __shared__ int wfill[kQueryWaves];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const uint64_t lane_below = (1ull << lane) - 1;
    if (lane == 0) wfill[w] = 0;
    __builtin_amdgcn_wave_barrier();
    const int base = blockIdx.x * (2 * kQueryBlock) + w * 128;
    if (n < 2 || base + lane >= n) return;  // neither of the lane's queries exists

    auto flush = [&](int fill) {
        const uint64_t act = __ballot(1);
        const int leader = __ffsll((unsigned long long)act) - 1;
        int b = 0;
        if (lane == leader) b = atomicAdd(count, fill);
        b = __shfl(b, leader);
        const int nact = __popcll(act);
        for (int k = __popcll(act & lane_below); k < fill; k += nact)
            if (b + k < cap) pairs[b + k] = wbuf[w][k];
    };
    // appends the pass's pairs (f0, qid), (f1, qid) of the active lanes to the
    // wave's buffer in lane order
    auto append = [&](int f0, int f1, int qid) {
        const uint64_t m0 = __ballot(f0 >= 0), m1 = __ballot(f1 >= 0);
        const int found = __popcll(m0) + __popcll(m1);
        if (!found) return;
        int fill = wfill[w];
        if (fill + found > kWaveBuf) {
            flush(fill);
            fill = 0;
        }
        if (f0 >= 0) wbuf[w][fill + __popcll(m0 & lane_below)] = make_int2(f0, qid);
        if (f1 >= 0) wbuf[w][fill + __popcll(m0) + __popcll(m1 & lane_below)] = make_int2(f1, qid);
        __builtin_amdgcn_wave_barrier();
        if (lane == __ffsll((unsigned long long)(m0 | m1)) - 1) wfill[w] = fill + found;
        __builtin_amdgcn_wave_barrier();
    };
    constexpr uint32_t kLeaf = 0x80000000u, kNone = 0xFFFFFFFFu;
    struct Walk {
        int t, qid, sp, seen;
        uint32_t node;  // kNone: finished (or no query)
        float q[6];
    };
    auto start = [&](Walk& s, int t) {
        s.t = t;
        s.sp = 0;
        s.seen = 0;
        s.qid = 0;
        s.node = kNone;
        if (t >= n) return;
        s.qid = (int)ids[t];
        for (int k = 0; k < 6; k++) s.q[k] = aabb[6 * (size_t)s.qid + k];
        if (s.q[0] <= s.q[3]) s.node = 0;  // an empty box (strand tip) intersects nothing
    };
    // one pass of a walk over the node (w0..w3): the pairs it reports and its
    // next node
    auto pass = [&](Walk& s, int col, float4 w0, float4 w1, float4 w2, float4 w3, int& f0, int& f1) {
        f0 = f1 = -1;
        if (s.node == kNone) return;
        auto leaf_pair = [&](uint32_t leaf) -> int {
            if ((int)leaf <= s.t) return -1;
            const int fid = (int)ids[leaf];
            if (min_gap >= 0 && abs(fid - s.qid) <= min_gap) return -1;
            return fid;
        };
        uint32_t next = kNone;
        const uint32_t fence = __float_as_uint(w0.w);
        bool overflow = false;
        if (max(s.node, fence) > (uint32_t)s.t) {
            const uint32_t lc = __float_as_uint(w0.y), rc = __float_as_uint(w0.z);
            const float b0[6] = {w1.x, w1.y, w1.z, w1.w, w2.x, w2.y};
            const float b1[6] = {w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
            bool h0 = intersects(b0, s.q), h1 = intersects(b1, s.q);
            if (h0 && (lc & kLeaf)) {
                f0 = leaf_pair(lc & ~kLeaf);
                h0 = false;
            }
            if (h1 && (rc & kLeaf)) {
                f1 = leaf_pair(rc & ~kLeaf);
                h1 = false;
            }
            if (h0 && h1) {
                if (s.sp + 1 > STACK) {  // deeper than the LDS stack: the redo pass takes it
                    overflow = true;
                } else {
                    stack[s.sp++][col] = rc;
                    next = lc;
                }
            } else if (h0) {
                next = lc;
            } else if (h1) {
                next = rc;
            }
        }
        if (!overflow && next == kNone && s.sp > 0) next = stack[--s.sp][col];
        s.seen += (f0 >= 0) + (f1 >= 0);
        if (overflow) {
            const int slot = atomicAdd(redo_count, 1);
            if (slot < redo_cap)
                redo[slot] = make_int2(s.t, s.seen);
            else if (error)
                atomicOr(error, 4);
        }
        s.node = overflow ? kNone : next;
    };
    Walk A, B;
    start(A, base + lane);
    start(B, base + 64 + lane);
    while (A.node != kNone || B.node != kNone) {
        // both node loads first (a finished walk loads the root, which stays cached)
        const float4* pa = reinterpret_cast<const float4*>(nodes + (A.node != kNone ? A.node : 0u));
        const float4* pb = reinterpret_cast<const float4*>(nodes + (B.node != kNone ? B.node : 0u));
        const float4 a0 = pa[0], a1 = pa[1], a2 = pa[2], a3 = pa[3];
        const float4 c0 = pb[0], c1 = pb[1], c2 = pb[2], c3 = pb[3];
        int fa0, fa1, fb0, fb1;
        pass(A, threadIdx.x, a0, a1, a2, a3, fa0, fa1);
        pass(B, threadIdx.x + kQueryBlock, c0, c1, c2, c3, fb0, fb1);
        append(fa0, fa1, A.qid);
        append(fb0, fb1, B.qid);
    }
    __builtin_amdgcn_wave_barrier();
    const int fill = wfill[w];
    if (fill > 0) flush(fill);
}

__global__ void fill_i32_kernel(int* p, int value, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

template <class T>
hipError_t dmalloc(T*& p, size_t count) {
    return hipMalloc((void**)&p, count * sizeof(T) + 64);
}

}  // namespace

hipError_t fill_i32(int* p, int value, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, value, n);
    return hipGetLastError();
}

void lbvh_release(LbvhDev& d) {
    void* ps[] = {d.aabb,  d.codes,       d.ids,         d.codes_tmp, d.ids_tmp,   d.hist,
                  d.hist_sums, d.partial,   d.root,        d.nodes,     d.leaf_parent, d.flags,
                  d.max_depth, d.count,     d.node_height, d.level_nodes, d.level_count, d.level_range,
                  d.redo};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    const int slots = d.fit_slots;
    d = LbvhDev{};
    d.fit_slots = slots;
}

// Digit-count words of the sort: the one-launch-per-pass layout of small
// sorts (kFusedPasses blocks of kFusedHist) or 256 per tile, whichever is larger, so
// a reservation for n also fits every smaller n.
static size_t hist_words(int tiles) { return std::max((size_t)kFusedPasses * kFusedHist, (size_t)256 * tiles); }

hipError_t lbvh_reserve(LbvhDev& d, int n) {
    if (n <= d.cap) return hipSuccess;
    lbvh_release(d);
    const int tiles = (n + kTile - 1) / kTile;
    const size_t hist_total = hist_words(tiles);
    const size_t chunks = (hist_total + kScanChunk - 1) / kScanChunk;
    hipError_t e;
    if ((e = dmalloc(d.aabb, (size_t)6 * n)) != hipSuccess || (e = dmalloc(d.codes, n)) != hipSuccess ||
        (e = dmalloc(d.ids, n)) != hipSuccess || (e = dmalloc(d.codes_tmp, n)) != hipSuccess ||
        (e = dmalloc(d.ids_tmp, n)) != hipSuccess || (e = dmalloc(d.hist, hist_total)) != hipSuccess ||
        (e = dmalloc(d.hist_sums, chunks)) != hipSuccess || (e = dmalloc(d.partial, 6 * 1024)) != hipSuccess ||
        (e = dmalloc(d.root, 6)) != hipSuccess || (e = dmalloc(d.nodes, (size_t)(n > 1 ? n - 1 : 1))) != hipSuccess ||
        (e = dmalloc(d.leaf_parent, n)) != hipSuccess || (e = dmalloc(d.flags, n)) != hipSuccess ||
        (e = dmalloc(d.max_depth, 1)) != hipSuccess || (e = dmalloc(d.count, 4)) != hipSuccess ||
        (e = dmalloc(d.redo, (size_t)n)) != hipSuccess ||
        (e = dmalloc(d.node_height, (size_t)n)) != hipSuccess || (e = dmalloc(d.level_nodes, (size_t)n)) != hipSuccess ||
        (e = dmalloc(d.level_count, (size_t)2 * kMaxLevels)) != hipSuccess ||
        (e = dmalloc(d.level_range, (size_t)kMaxLevels)) != hipSuccess) {
        lbvh_release(d);
        return e;
    }
    d.redo_count = d.count + 1;
    d.cap = n;
    return hipSuccess;
}

size_t lbvh_carve(LbvhDev& d, int n, char* base, size_t bytes) {
    const int tiles = (n + kTile - 1) / kTile;
    const size_t hist_total = hist_words(tiles);
    const size_t chunks = (hist_total + kScanChunk - 1) / kScanChunk;
    size_t off = 0;
    LbvhDev v;
    auto take = [&](auto*& ptr, size_t count) {
        using T = std::remove_reference_t<decltype(*ptr)>;
        off = (off + 255) & ~(size_t)255;
        ptr = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += count * sizeof(T);
    };
    take(v.aabb, (size_t)6 * n);
    take(v.codes, n);
    take(v.ids, n);
    take(v.codes_tmp, n);
    take(v.ids_tmp, n);
    take(v.hist, hist_total);
    take(v.hist_sums, chunks);
    take(v.partial, 6 * 1024);
    take(v.root, 6);
    take(v.nodes, (size_t)(n > 1 ? n - 1 : 1));
    take(v.leaf_parent, n);
    take(v.flags, n);
    take(v.max_depth, 1);
    take(v.count, 4);
    take(v.redo, (size_t)n);
    take(v.node_height, (size_t)n);
    take(v.level_nodes, (size_t)n);
    take(v.level_count, (size_t)2 * kMaxLevels);
    take(v.level_range, (size_t)kMaxLevels);
    v.redo_count = v.count ? v.count + 1 : nullptr;
    if (base && off <= bytes) {
        v.fit_slots = d.fit_slots;
        d = v;
        d.cap = n;
        d.n = 0;
        d.built = false;
    }
    return off;
}

// Heights at or below this are refit one launch per height; the rest by
// refit_top_kernel.  In Karras trees over spread-out boxes the node count of
// height h falls by about 0.66 per height from ~0.4 n (100k boxes: 36k, 23k,
// 15k, 10k, 6.4k, 4.4k, 2.6k, ...); the launches stop near 2 kTopBlock nodes.
// Any value gives the same boxes.
// Estimated crossing nodes of height h: all nodes of that height (about
// 0.4 n 0.66^(h-1) in Karras trees over spread-out boxes), but at most a
// couple per tile boundary.
static double crossing_estimate(int n, int h) {
    double c = 0.4 * n;
    for (int k = 1; k < h; k++) c *= 0.66;
    return std::min(c, 2.0 * n / kRefitTile + 64.0);
}

// Heights that get launches of their own: up to the last one estimated to
// hold two workgroups' worth of crossing nodes.
static int bottom_levels(int n) {
    int last = 0;
    for (int h = 1; h < kMaxLevels - 1; h++)
        if (crossing_estimate(n, h) >= 2.0 * kTopBlock) last = h;
    return last;
}

// Grid of the refit launch for height h (grid-stride over the actual count).
static unsigned level_blocks(int n, int h) {
    const double b = std::ceil(crossing_estimate(n, h) / 256.0);
    return (unsigned)std::max(8.0, std::min(b, 65536.0));
}

// The crossing list's child words of a small tree live in the query's redo
// buffer (n int2, unused outside a query), and the global-memory fit's done
// words in the sort's second key buffer (free after the sort).
static uint2* cross_child(const LbvhDev& d) { return reinterpret_cast<uint2*>(d.redo); }

hipError_t lbvh_build(LbvhDev& d, int n, hipStream_t stream) {
    if (n > d.cap) return hipErrorInvalidValue;
    d.n = n;
    d.built = false;
    d.max_depth_host = -1;
    if (n <= 0) return hipSuccess;
    const float* src = d.boxes();
    const int nb = (n + 255) / 256;
    const int pblocks = nb < 1024 ? nb : 1024;
    hipLaunchKernelGGL(root_partial_kernel, dim3(pblocks), dim3(256), 0, stream, src, n, d.partial);
    const int tiles = (n + kTile - 1) / kTile;
    d.tiles = tiles;
    const bool small = tiles <= kFusedTiles;  // n <= kSmallBuild
    d.small = small;
    uint32_t *kin = d.codes, *vin = d.ids, *kout = d.codes_tmp, *vout = d.ids_tmp;
    if (small) {
        // one launch per pass; each pass counts the next pass's digits
        hipLaunchKernelGGL(morton_hist_kernel<kFusedBits>, dim3(tiles), dim3(kSortThreads), 0, stream, src, d.partial,
                           pblocks, n, kFusedTiles, d.root, d.codes, d.ids, d.hist, d.hist + kFusedHist,
                           kFusedPasses - 1, d.level_count);
        for (int pass = 0; pass < kFusedPasses; pass++) {
            uint32_t* hist = d.hist + (size_t)pass * kFusedHist;
            hipLaunchKernelGGL((radix_scatter_kernel<true, kFusedBits>), dim3(tiles), dim3(kSortThreads), 0, stream,
                               kin, vin, n, pass * kFusedBits, tiles, hist, kout, vout,
                               pass + 1 < kFusedPasses ? hist + kFusedHist : nullptr);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        // an odd number of passes leaves the sorted data in the second buffers
        if (kin != d.codes) {
            std::swap(d.codes, d.codes_tmp);
            std::swap(d.ids, d.ids_tmp);
        }
    } else {
        hipLaunchKernelGGL(morton_hist_kernel<8>, dim3(tiles), dim3(kSortThreads), 0, stream, src, d.partial, pblocks,
                           n, tiles, d.root, d.codes, d.ids, d.hist, nullptr, 0, nullptr);
        const size_t total = (size_t)256 * tiles;
        const int chunks = (int)((total + kScanChunk - 1) / kScanChunk);
        for (int shift = 0; shift < 32; shift += 8) {
            if (shift > 0)  // the first pass's histogram comes from morton_hist_kernel
                hipLaunchKernelGGL(radix_hist_kernel, dim3(tiles), dim3(kSortThreads), 0, stream, kin, n, shift, tiles,
                                   d.hist);
            if (total <= (size_t)kSmallScanMax) {
                hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(kSmallScanThreads), 0, stream, d.hist, (int)total);
            } else {
                hipLaunchKernelGGL(scan_chunk_kernel, dim3(chunks), dim3(256), 0, stream, d.hist, total, d.hist_sums);
                hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, stream, d.hist_sums, chunks);
                hipLaunchKernelGGL(scan_add_kernel, dim3(chunks), dim3(256), 0, stream, d.hist, total, d.hist_sums);
            }
            hipLaunchKernelGGL((radix_scatter_kernel<false, 8>), dim3(tiles), dim3(kSortThreads), 0, stream, kin, vin,
                               n, shift, tiles, d.hist, kout, vout, nullptr);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
    }
    // the sorted data is in d.codes / d.ids
    hipError_t e;
    if (n <= 1) {
        if ((e = hipGetLastError()) != hipSuccess) return e;
        d.built = true;
        return fill_i32(d.max_depth, 1, 1, stream);
    }
    if (small) {
        hipLaunchKernelGGL(tile_fit_kernel<true>, dim3((n + kRefitTile - 1) / kRefitTile), dim3(kRefitTile), 0, stream,
                           d.codes, src, d.ids, d.nodes, d.leaf_parent, n, d.node_height, d.level_nodes,
                           cross_child(d), (uint32_t*)d.flags, d.level_count, d.max_depth);
        if (n > kRefitTile)
            hipLaunchKernelGGL(cross_fit_kernel<true>, dim3(1), dim3(kCrossBlock), 0, stream, src, d.ids, d.nodes,
                               d.node_height, d.level_nodes, cross_child(d), (const uint32_t*)d.flags, d.level_count,
                               std::clamp(d.fit_slots, 0, kCrossCap), d.codes_tmp,
                               d.max_depth);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        d.built = true;
        return hipSuccess;
    }
    const int m = n - 1;
    const int bthreads = std::max(m, 2 * kMaxLevels);
    hipLaunchKernelGGL(build_internal_kernel, dim3((bthreads + 255) / 256), dim3(256), 0, stream, d.codes, n, d.nodes,
                       d.leaf_parent, d.flags, d.level_count);
    // heights and stack depths (topology only), then the nodes grouped by height
    hipLaunchKernelGGL(height_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d.leaf_parent, d.nodes,
                       (uint32_t*)d.flags, n, d.node_height, d.max_depth);
    const dim3 level_grid((m + kLevelChunk - 1) / kLevelChunk);
    hipLaunchKernelGGL(level_count_kernel, level_grid, dim3(kLevelBlock), 0, stream, d.node_height, d.nodes, m,
                       d.level_count);
    hipLaunchKernelGGL(level_scatter_kernel, level_grid, dim3(kLevelBlock), 0, stream, d.node_height, d.nodes, m,
                       d.level_count, d.level_range, d.level_nodes);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    d.built = true;
    return lbvh_refit(d, stream);
}

// The nodes inside tiles in one launch, then the crossing nodes: one launch
// per low height, then one workgroup for the rest.  Launch boundaries and
// workgroup barriers order the levels (no atomics, no fences); the level
// ranges live on the device, so nothing is read back.
hipError_t lbvh_refit(LbvhDev& d, hipStream_t stream) {
    if (!d.built) return hipErrorInvalidValue;
    if (d.n < 2) return hipSuccess;
    const float* src = d.boxes();
    if (d.small) {
        hipLaunchKernelGGL(tile_fit_kernel<false>, dim3((d.n + kRefitTile - 1) / kRefitTile), dim3(kRefitTile), 0,
                           stream, nullptr, src, d.ids, d.nodes, nullptr, d.n, nullptr, nullptr, nullptr, nullptr,
                           nullptr, nullptr);
        if (d.n > kRefitTile)
            hipLaunchKernelGGL(cross_fit_kernel<false>, dim3(1), dim3(kCrossBlock), 0, stream, src, d.ids, d.nodes,
                               d.node_height, d.level_nodes, cross_child(d), (const uint32_t*)d.flags, d.level_count,
                               std::clamp(d.fit_slots, 0, kCrossCap), d.codes_tmp, nullptr);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(tile_fit_kernel<false>, dim3((unsigned)((d.n + kRefitTile - 1) / kRefitTile)), dim3(kRefitTile),
                       0, stream, nullptr, src, d.ids, d.nodes, nullptr, d.n, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr);
    const int K = bottom_levels(d.n);
    for (int h = 1; h <= K; h++)
        hipLaunchKernelGGL(refit_level_kernel, dim3(level_blocks(d.n, h)), dim3(256), 0, stream, src, d.ids,
                           d.nodes, d.level_nodes, d.level_range, h);
    hipLaunchKernelGGL(refit_top_kernel, dim3(1), dim3(kTopBlock), 0, stream, src, d.ids, d.nodes, d.level_nodes,
                       d.level_range, K + 1);
    return hipGetLastError();
}

hipError_t lbvh_read_depth(LbvhDev& d, hipStream_t stream) {
    if (!d.built) return hipErrorInvalidValue;
    int depth = 0;
    hipError_t e;
    if ((e = hipMemcpyAsync(&depth, d.max_depth, sizeof depth, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
        return e;
    d.max_depth_host = depth;
    return hipSuccess;
}

// Up to this many objects the query runs in one pass with the full stack:
// its LDS (32 KB per 128-lane workgroup) then still admits enough lanes per
// CU for all queries.
constexpr int kOnePassMax = 1 << 20;

hipError_t lbvh_query(LbvhDev& d, int2* pairs, int cap, int min_gap, int* error, hipStream_t stream) {
    if (!d.built) return hipErrorInvalidValue;
    // the pair and redo counters (and the library handle's error word, when
    // that is the one passed) in one fill
    hipError_t e = fill_i32(d.count, 0, error == d.count + 2 ? 3 : 2, stream);
    if (e != hipSuccess) return e;
    if (d.n < 2) return hipSuccess;
    const float* src = d.boxes();
    const dim3 grid((d.n + kQueryBlock - 1) / kQueryBlock);
    if (d.n <= kOnePassMax) {
        // a Karras tree over 62-bit keys is at most 63 levels tall: a
        // kQueryStack-entry stack never overflows
        hipLaunchKernelGGL((query_kernel<kQueryStack, false>), grid, dim3(kQueryBlock), 0, stream, src, d.ids,
                           d.nodes, d.n, pairs, cap, min_gap, d.count, error, d.redo, d.redo_count, d.n);
        return hipGetLastError();
    }
    // The first pass walks two queries per lane with a kQueryStackFirst2-entry
    // LDS stack each; the few queries that need more are redone with the full
    // stack (the redo pass returns at once when none was listed).
    const dim3 grid2((d.n + 2 * kQueryBlock - 1) / (2 * kQueryBlock));
    hipLaunchKernelGGL((query2_kernel<kQueryStackFirst2>), grid2, dim3(kQueryBlock), 0, stream, src, d.ids, d.nodes,
                       d.n, pairs, cap, min_gap, d.count, error, d.redo, d.redo_count, d.n);
    const dim3 rgrid(std::min<unsigned>(grid.x, kRedoBlocks));
    hipLaunchKernelGGL((query_kernel<kQueryStack, true>), rgrid, dim3(kQueryBlock), 0, stream, src, d.ids, d.nodes,
                       d.n, pairs, cap, min_gap, d.count, error, d.redo, d.redo_count, d.n);
    return hipGetLastError();
}

int lbvh_query_stack_limit() { return kQueryStack; }

}  // namespace kscr
