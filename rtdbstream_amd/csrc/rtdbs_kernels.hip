// HIP kernels of the MI355X streaming DBSCAN slide step (gfx950, wave64).
//
// One slide of every stream in a batch is eight launches:
//   sort_expire (stride sort in runs + survivors per window tile) ->
//   rank (stride rank merge + survivor compaction) -> merge_final -> cells ->
//   rows -> count -> union -> border (outputs)
// (large batches: count_points -> fcore -> rep -> union -> border_points).
// Each kernel covers all streams of the batch in one launch: a block finds its
// stream from prefix tables of per-stream work units (tiles of positions,
// 64-position chunks) or, for the cell kernels whose sizes are only known on
// the device, walks a device-side work list in a persistent loop.
//
// Semantics follow the reference's hot path (RTDBStream level 9,
// src/dbscan/dbscan.cpp:2276-2362) and its checker (util.cpp:383-442):
// FP32 distances ((dx*dx + dy*dy) + dz*dz) without contraction, strict
// d2 < eps2, self-inclusive counts, core iff count >= min_pts, union-find
// hooking the larger root under the smaller (dbscan.cu:22-43) with slots as
// the order, so every cluster id is its minimum core slot.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include <hip/hip_ext.h>

#include "rtdbs_dev.h"
#include "rtdbs_devfn.h"

namespace rtdbs {

const char *const kPhaseNames[kNumPhases] = {"sort_expire", "rank_compact", "merge_final",
                                             "cells", "rows", "count", "union", "border"};

// ---------------------------------------------------------------------------
// small helpers


// Largest s in [0, n) with pre[s] <= g (pre[0] = 0, g < pre[n]); pre in LDS.
template <class T>
__device__ __forceinline__ int find_stream(const T *pre, int n, long long g) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if ((long long)pre[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The stream of cell g for a wave that walks its cells in increasing order:
// the largest s >= s0 with pre[s] <= g (pre[s0] <= g), by galloping from the
// previous cell's stream - one probe while the stream stays the same -
// instead of a full binary search per cell. Scalar loads (constant memory).
__device__ __forceinline__ int find_stream_from(const int32_t *pre, int n, int g, int s0) {
    int lo = s0, hi = s0 + 1, step = 1;
    while (hi < n && ldc(pre[hi]) <= g) {
        lo = hi;
        step <<= 1;
        hi = min(lo + step, n);
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ldc(pre[mid]) <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The same over a per-batch prefix table in global memory, which no kernel
// writes during a launch: scalar loads (the table stays in the scalar cache),
// not a dependent vector load per step.
__device__ __forceinline__ int find_stream_c(const int32_t *pre, int n, long long g) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if ((long long)ldc(pre[mid]) <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Logical index of workgroup b of n. The dispatcher hands workgroups to the
// 8 XCDs round-robin (b runs on XCD b mod 8), and each XCD has its own L2.
// With `on`, the workgroups of XCD x take the x-th contiguous eighth of the
// logical range, so work items that are neighbours in that range (cells or
// positions whose stencils share candidates) run on one XCD and share its
// L2. A bijection on [0, n) for any n: XCD x holds n/8 workgroups, one more
// when x < n mod 8. Measured: in a 3.2 M-point batch it shortens the merge
// (385 -> 373 us) and the point-parallel output kernel (200 -> 184 us) and
// lengthens the count and rows kernels (190 -> 210 us, 127 -> 161 us); on
// single streams of 0.1-1 M points it lengthens every kernel. So the engine
// enables it per kernel (BatchArgs::xcd_groups).
__device__ __forceinline__ int xcd_block(int b, int n, bool on) {
    if (!on) return b;
    const int q = n >> 3, r = n & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

__device__ __forceinline__ uint32_t slot_mask(const StreamDev &S) {
    return (uint32_t)((1ull << S.slot_bits) - 1ull);
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += xor_lane(v, o);
    return v;
}

// Value of lane `l` (uniform) of a per-lane int.
__device__ __forceinline__ int lane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Sum over a 256-thread block.
__device__ __forceinline__ int block_sum(int v, int *lds4) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) lds4[w] = v;
    __syncthreads();
    int tot = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return tot;
}

// Exclusive rank of this thread's flag within a 256-thread block, plus the
// block total (ordered stream compaction by wave ballot/popcount).
__device__ __forceinline__ int block_rank_flags(bool f, int *total, int *lds4) {
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t m = __ballot(f);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) lds4[w] = __popcll(m);
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int c = lds4[i];
        base += (i < w) ? c : 0;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return base + below;
}

// Inclusive prefix (into pre[1..n], pre[0] = 0) of per-stream values over a
// block of 256 threads; n <= kMaxBatch. Values are 64-bit.
template <class F>
__device__ void block_stream_prefix(int n, long long *pre, long long *tmp4, F value) {
    const int tid = threadIdx.x;
    // each thread owns 4 consecutive streams
    long long v[4];
    long long local = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int i = tid * 4 + q;
        v[q] = (i < n) ? value(i) : 0;
        local += v[q];
    }
    // wave inclusive scan of `local`
    const int lane = lane_id(), w = tid >> 6;
    long long x = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        long long y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp4[w] = x;
    __syncthreads();
    long long wbase = 0;
    for (int i = 0; i < w; i++) wbase += tmp4[i];
    long long run = wbase + x - local;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int i = tid * 4 + q;
        run += v[q];
        if (i < n) pre[i + 1] = run;
    }
    if (tid == 0) pre[0] = 0;
    __syncthreads();
}

__device__ __forceinline__ int slot_of(const StreamDev &S, int pos) {
    return (int)((uint32_t)S.skey[pos] & (uint32_t)((1ull << S.slot_bits) - 1ull));
}

// ---------------------------------------------------------------------------
// grid

// The grid of a stream read once as constant memory (uniform per block), for
// kernels that compute many keys.
struct KeyGrid {
    double lo0, lo1, lo2, inv_side;
    int64_t n0, n1, n2;
    int dim, slot_bits;
};

__device__ __forceinline__ KeyGrid key_grid(const StreamDev &S) {
    KeyGrid g;
    g.lo0 = ldc(S.lo[0]);
    g.lo1 = ldc(S.lo[1]);
    g.lo2 = ldc(S.lo[2]);
    g.inv_side = ldc(S.inv_side);
    g.n0 = ldc(S.n[0]);
    g.n1 = ldc(S.n[1]);
    g.n2 = ldc(S.n[2]);
    g.dim = ldc(S.dim);
    g.slot_bits = ldc(S.slot_bits);
    return g;
}

// Cell index per axis: floor((x - lo) / side) evaluated in double and clamped
// into the grid. Clamping keeps the search exact: |clamp(i) - clamp(j)| <=
// |i - j|, so points within reach stay within reach.
__device__ __forceinline__ int64_t bin_axis(double x, double lo, double inv, int64_t n) {
    double v = floor((x - lo) * inv);
    if (!(v >= 0.0)) v = 0.0;  // negative or NaN
    const double mx = (double)(n - 1);
    if (v > mx) v = mx;
    return (int64_t)v;
}

// Packed cell key (i0 * n1 + i1) * n2 + i2; unused axes have index 0.
__device__ __forceinline__ uint64_t cell_key_g(const KeyGrid &G, float x, float y, float z) {
    const int64_t i0 = bin_axis((double)x, G.lo0, G.inv_side, G.n0);
    const int64_t i1 = G.dim > 1 ? bin_axis((double)y, G.lo1, G.inv_side, G.n1) : 0;
    const int64_t i2 = G.dim > 2 ? bin_axis((double)z, G.lo2, G.inv_side, G.n2) : 0;
    return (uint64_t)((i0 * G.n1 + i1) * G.n2 + i2);
}

// Quotient and remainder of a cell key (< 2^52, so exact in a double) by a
// grid extent (<= 2^30), via a double-precision estimate, off by at most one,
// and an exact integer fix-up (64-bit integer division is a long software
// sequence on the GPU).
__device__ __forceinline__ void divmod_key(uint64_t x, int64_t d, int64_t *q, int64_t *r) {
    int64_t qq = (int64_t)((double)x / (double)d);
    int64_t rr = (int64_t)x - qq * d;
    while (rr < 0) {
        qq--;
        rr += d;
    }
    while (rr >= d) {
        qq++;
        rr -= d;
    }
    *q = qq;
    *r = rr;
}

// ---------------------------------------------------------------------------
// sorting: register bitonic tile sort, rank merge of the stride's tiles,
// merge-path merges of the window

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    const int lo = xor_lane((int)(uint32_t)v, m);
    const int hi = xor_lane((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Compare-exchange of lanes l and l ^ m (m a power of two below 64): the
// smaller key where keep_min, else the larger. For m = 16 and 32 the row-pair
// and half-wave swaps of both copies of a word leave the lower partner's
// value in one register and the upper partner's in the other, in every lane,
// so no select picks the partner out first.
__device__ __forceinline__ uint64_t lane_minmax(uint64_t v, int m, bool keep_min) {
    uint64_t x, y;
    if (m == 16 || m == 32) {
        const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
        const auto rl = m == 32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                                : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto rh = m == 32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                                : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        x = ((uint64_t)(uint32_t)rh[0] << 32) | (uint32_t)rl[0];  // lane l & ~m
        y = ((uint64_t)(uint32_t)rh[1] << 32) | (uint32_t)rl[1];  // lane l | m
    } else {
        x = v;
        y = shfl_xor_u64(v, m);
    }
    return ((y < x) == keep_min) ? y : x;
}

// Sorts N keys ascending, E per thread: thread t (of N/E) holds positions
// Et..Et+E-1 in v. Bitonic network: exchanges between a thread's own elements
// stay in registers, exchanges within a wave use lane exchanges, and only the
// stages whose partners sit in another wave (j >= 64E) go through LDS (`lds`,
// N keys) with block barriers: none for N = 64E (one wave), three stages for
// a 512-key run over four waves (E = 2) and for N = 2048 over four waves
// (E = 8). Every thread of the block takes part in the same stages.
template <int N, int E>
__device__ __forceinline__ void bitonic_sort_regs(uint64_t (&v)[E], uint64_t *lds, int tid) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= E) {
                const int m = j / E;  // partner thread = tid ^ m, same element index
                const bool keep_min = ((tid & m) == 0) == (((tid * E) & k) == 0);
                if (j >= 64 * E) {
                    __syncthreads();
#pragma unroll
                    for (int e = 0; e < E; e++) lds[tid * E + e] = v[e];
                    __syncthreads();
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        const uint64_t o = lds[(tid ^ m) * E + e];
                        v[e] = ((o < v[e]) == keep_min) ? o : v[e];
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < E; e++) v[e] = lane_minmax(v[e], m, keep_min);
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; e++) {
                    if (e & j) continue;
                    const bool up = (((tid * E) + e) & k) == 0;
                    const uint64_t x = v[e], y = v[e | j];
                    const bool sw = (x > y) == up;
                    v[e] = sw ? y : x;
                    v[e | j] = sw ? x : y;
                }
            }
        }
    }
}

// Loads the packed keys (cell << slot_bits | slot) of points [tile0, tile0 +
// N) of `in` (len points, slots from slot0) into v, E per thread (tid < N/E),
// and sorts them; padding keys are ~0 and sort last. Thread tid takes points
// tid, tid + N/E, ... so each load instruction reads consecutive points (the
// network sorts any assignment; a key carries its own slot).
template <int N, int E>
__device__ __forceinline__ void sort_tile_keys(const KeyGrid &G, gptr<const float> in, int len, int tile0, int slot0,
                                               uint64_t (&v)[E], uint64_t *lds, int tid) {
    const int dim = G.dim;
    float c[E][3];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int gi = min(tile0 + e * (N / E) + tid, len - 1);
#pragma unroll
        for (int a = 0; a < 3; a++) c[e][a] = a < dim ? in[(size_t)gi * dim + a] : 0.f;
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int gi = tile0 + e * (N / E) + tid;
        v[e] = gi < len ? (cell_key_g(G, c[e][0], c[e][1], c[e][2]) << G.slot_bits) | (uint64_t)(slot0 + gi) : ~0ull;
    }
    bitonic_sort_regs<N, E>(v, lds, tid);
}

// First window: keys and coordinates of each sorted tile (init ping-pong A).
__global__ __launch_bounds__(kTileThreads) void k_init_tiles(BatchArgs a) {
    __shared__ uint64_t lds[kTile];
    const int g = blockIdx.x;
    const int s = find_stream_c(a.pre_itile, a.n, g);
    const StreamDev &S = a.s[s];
    const IO io = a.use_io1 ? a.io1 : a.io[s];
    const int tile0 = (g - ldc(a.pre_itile[s])) * kTile;
    uint64_t v[8];
    sort_tile_keys<kTile, 8>(key_grid(S), gl(io.in), ldc(S.W), tile0, 0, v, lds, threadIdx.x);
    const int nvalid = min(kTile, S.W - tile0);
    const uint32_t m = slot_mask(S);
    const int dim = S.dim;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int i = threadIdx.x * 8 + e;
        if (i < nvalid) {
            const int gi = (int)((uint32_t)v[e] & m);
            const float *p = io.in + (size_t)gi * dim;
            S.akey[tile0 + i] = v[e];
            S.ax[tile0 + i] = p[0];
            if (dim > 1) S.ay[tile0 + i] = p[1];
            if (dim > 2) S.az[tile0 + i] = p[2];
        }
    }
}


// Survivors per window tile (the reference's stable_partition,
// thrust_helper.cu:37-39, split into count + compact).
__device__ __forceinline__ void expire_count(const BatchArgs &a, int g, int *lds4) {
    const int s = find_stream_c(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - ldc(a.pre_wtile[s]);
    const int W = ldc(S.W), SS = ldc(S.S);
    const gptr<const uint64_t> skey = gl((const uint64_t *)ldc(S.skey));
    const uint32_t sm = (uint32_t)((1ull << ldc(S.slot_bits)) - 1ull);
    const int p0 = tile * kWTile, p1 = min(p0 + kWTile, W);
    const int lo_slot = ldc(ldc(S.mut)->u) * SS;
    uint64_t k[kWTile / kTileThreads];
#pragma unroll
    for (int r = 0; r < kWTile / kTileThreads; r++) {
        const int p = p0 + threadIdx.x + r * kTileThreads;
        k[r] = p < p1 ? skey[p] : 0ull;
    }
    int cnt = 0;
#pragma unroll
    for (int r = 0; r < kWTile / kTileThreads; r++) {
        const int slot = (int)((uint32_t)k[r] & sm);
        cnt += (p0 + (int)threadIdx.x + r * kTileThreads < p1) && (slot < lo_slot || slot >= lo_slot + SS);
    }
    cnt = block_sum(cnt, lds4);
    if (threadIdx.x == 0) gl(ldc(S.tile_a))[tile] = cnt;
}

// Launch 1 of a slide: the first kSortParts blocks of each stride tile sort
// the stride by key (the reference's compute_cell_id + sortByCellIdAndOrder,
// aabb.cu:589-611, thrust_helper.cu:21-29) in runs of kRun keys into bkey0,
// four waves per run (kSortE keys per thread: a quarter of the network's
// instructions per wave of a one-wave run, for three LDS stages); the
// remaining blocks count the survivors of each window tile.
constexpr int kRun = 512;                                // keys per sorted run of the stride
constexpr int kSortE = 2;                                // keys per thread
constexpr int kSortRunThreads = kRun / kSortE;           // 256: four waves per run
constexpr int kSortBlockKeys = kTileThreads * kSortE;    // 512: one run per block
constexpr int kSortParts = kTile / kSortBlockKeys;       // blocks per stride tile

__global__ __launch_bounds__(kTileThreads) void k_sort_expire(BatchArgs a) {
    __shared__ uint64_t lds[kTile];
    const int g = blockIdx.x;
    const int nsb = ldc(a.pre_stile[a.n]) * kSortParts;
    if (g >= nsb) {
        expire_count(a, g - nsb, (int *)lds);
        return;
    }
    const int gt = g / kSortParts;
    const int s = find_stream_c(a.pre_stile, a.n, gt);
    const StreamDev &S = a.s[s];
    const IO io = a.use_io1 ? a.io1 : a.io[s];
    const int SS = ldc(S.S);
    const int tile0 = (gt - ldc(a.pre_stile[s])) * kTile + (g % kSortParts) * kSortBlockKeys;
    if (tile0 >= SS) return;  // the whole block: past the stride's end
    const int r = threadIdx.x / kSortRunThreads, t = threadIdx.x % kSortRunThreads;
    uint64_t v[kSortE];
    sort_tile_keys<kRun, kSortE>(key_grid(S), gl(io.in), SS, tile0 + r * kRun, ldc(ldc(S.mut)->u) * SS, v,
                                 lds + r * kRun, t);
    // coalesced store through LDS (thread i holds positions kSortE * i ..)
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kSortE; e++) lds[threadIdx.x * kSortE + e] = v[e];
    __syncthreads();
    const int nvalid = min(kSortBlockKeys, SS - tile0);
    const gptr<uint64_t> bkey0 = gl(ldc(S.bkey0));
    for (int i = threadIdx.x; i < nvalid; i += kTileThreads) bkey0[tile0 + i] = lds[i];
}

// Ordered compaction of the survivors into the A arrays. Each tile's offset
// is the sum of the preceding tiles' counts (no spin-waiting look-back).
__device__ __forceinline__ void expire_compact(const BatchArgs &a, int g, int *lds4) {
    const int s = find_stream_c(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - ldc(a.pre_wtile[s]);
    const int p0 = tile * kWTile;
    const int W = ldc(S.W), SS = ldc(S.S), dim = ldc(S.dim);
    const uint32_t sm = (uint32_t)((1ull << ldc(S.slot_bits)) - 1ull);
    const int lo_slot = ldc(ldc(S.mut)->u) * SS;
    const gptr<const uint64_t> skey = gl((const uint64_t *)ldc(S.skey));
    const gptr<const float> sx = gl((const float *)ldc(S.sx)), sy = gl((const float *)ldc(S.sy)),
                            sz = gl((const float *)ldc(S.sz));
    const gptr<const int32_t> tile_a = gl((const int32_t *)ldc(S.tile_a));
    constexpr int per = kWTile / kTileThreads;
    // every load of the tile issued before the scan and the stores
    uint64_t key[per];
    float cx[per], cy[per], cz[per];
#pragma unroll
    for (int r = 0; r < per; r++) {
        const int p = p0 + r * kTileThreads + threadIdx.x;
        const int pp = p < W ? p : 0;
        key[r] = skey[pp];
        cx[r] = sx[pp];
        cy[r] = dim > 1 ? sy[pp] : 0.f;
        cz[r] = dim > 2 ? sz[pp] : 0.f;
    }
    int part = 0;
    for (int t = threadIdx.x; t < tile; t += kTileThreads) part += tile_a[t];
    int run = block_sum(part, lds4);
    const gptr<uint64_t> akey = gl(ldc(S.akey));
    const gptr<float> ax = gl(ldc(S.ax)), ay = gl(ldc(S.ay)), az = gl(ldc(S.az));
#pragma unroll
    for (int r = 0; r < per; r++) {
        const int p = p0 + r * kTileThreads + threadIdx.x;
        const int slot = (int)((uint32_t)key[r] & sm);
        const bool f = p < W && (slot < lo_slot || slot >= lo_slot + SS);
        int tot;
        const int rank = block_rank_flags(f, &tot, lds4);
        if (f) {
            const int d = run + rank;
            akey[d] = key[r];
            ax[d] = cx[r];
            if (dim > 1) ay[d] = cy[r];
            if (dim > 2) az[d] = cz[r];
        }
        run += tot;
    }
}

// Launch 2 of a slide: blocks [0, rank_blocks) place each stride key at its rank
// in the whole stride - its index in its sorted run plus, for every other
// run, the number of that run's keys below it - and write keys and
// coordinates to the B1 arrays: the stride sorted by (cell, slot). The other
// runs are staged in LDS kStageRuns at a time (the next ones load while the
// current ones are searched). With SPLIT == 1 a block ranks 256 keys, one per
// thread, and every thread searches every staged run. With SPLIT == 4 a
// block ranks 64 keys and its four waves split each stage's runs between
// them, summing their partial ranks in LDS at the end: four times the blocks
// and a quarter of the searches per thread, for strides too small to fill the
// device with 256-key blocks. The blocks after them compact the window's
// survivors (expire_compact), which depends on the same launch before.
constexpr int kStageRuns = 8;  // 32 KB of LDS

template <int SPLIT>
__global__ __launch_bounds__(kTileThreads) void k_rank(BatchArgs a, int rank_blocks) {
    constexpr int kWaves = kTileThreads / 64;
    constexpr int kKeys = kTileThreads / SPLIT;  // keys per block, all in one run (kKeys divides kRun)
    constexpr int kParts = kTile / kKeys;        // blocks per stride tile
    constexpr int kMine = kStageRuns / SPLIT;    // staged runs each thread searches
    __shared__ uint64_t lds[kStageRuns * kRun];
    __shared__ int part[SPLIT > 1 ? kWaves : 1][64];
    if ((int)blockIdx.x >= rank_blocks) {  // the survivor compaction shares the launch
        expire_compact(a, blockIdx.x - rank_blocks, (int *)part);
        return;
    }
    const int g = blockIdx.x;
    const int gb = g / kParts;
    const int s = find_stream_c(a.pre_stile, a.n, gb);
    const StreamDev &S = a.s[s];
    // the stream's fields as constant memory and its arrays as global memory:
    // global loads wait on their own counter, so the next stage's keys load
    // while the LDS searches of this stage run
    const int SS = ldc(S.S), dim = ldc(S.dim);
    const int i0 = (gb - ldc(a.pre_stile[s])) * kTile + (g % kParts) * kKeys;
    if (i0 >= SS) return;  // the whole block: past the stride's end
    const gptr<const uint64_t> bkey0 = gl((const uint64_t *)ldc(S.bkey0));
    const int wv = threadIdx.x >> 6, lane = lane_id();
    const int sigma = wv % SPLIT;  // which of each stage's runs this wave searches
    const int i = i0 + (SPLIT > 1 ? lane : (int)threadIdx.x);
    const bool valid = i < SS;
    const uint64_t key = valid ? bkey0[i] : ~0ull;
    const int myrun = i0 / kRun, nr = (SS + kRun - 1) / kRun;
    constexpr int per = kStageRuns * kRun / kTileThreads;
    auto run_of = [&](int o) { return o < myrun ? o : o + 1; };  // the o-th run other than the block's
    // keys of the staged runs o0 .. o0 + kStageRuns - 1, ~0 past their ends
    auto load_stage = [&](int o0, uint64_t(&st)[per]) {
#pragma unroll
        for (int e = 0; e < per; e++) {
            const int slot = threadIdx.x + e * kTileThreads;  // [0, kStageRuns * kRun)
            const int q = slot / kRun, j = slot % kRun;
            const int o = o0 + q;
            const int r = run_of(o);
            st[e] = (o < nr - 1 && j < min(kRun, SS - r * kRun)) ? bkey0[r * kRun + j] : ~0ull;
        }
    };
    uint64_t cur[per], nxt[per];
    if (nr > 1) load_stage(0, cur);
    // the point's coordinates are needed only at the end: loaded now
    const IO io = a.use_io1 ? a.io1 : a.io[s];
    float px = 0.f, py = 0.f, pz = 0.f;
    if (sigma == 0) {
        const int gi = valid ? (int)((uint32_t)key & slot_mask(S)) - ldc(ldc(S.mut)->u) * SS : 0;
        const gptr<const float> p = gl(io.in) + (size_t)gi * dim;
        px = p[0];
        py = dim > 1 ? p[1] : 0.f;
        pz = dim > 2 ? p[2] : 0.f;
    }
    int rank = 0;
    for (int o0 = 0; o0 < nr - 1; o0 += kStageRuns) {
        const bool more = o0 + kStageRuns < nr - 1;
        if (more) load_stage(o0 + kStageRuns, nxt);
        int len[kMine];
#pragma unroll
        for (int q = 0; q < kMine; q++) {
            const int o = o0 + q * SPLIT + sigma;
            len[q] = o < nr - 1 ? min(kRun, SS - run_of(o) * kRun) : 0;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < per; e++) lds[threadIdx.x + e * kTileThreads] = cur[e];
        __syncthreads();
        int pos[kMine];
#pragma unroll
        for (int q = 0; q < kMine; q++) pos[q] = 0;
#pragma unroll
        for (int step = kRun; step > 0; step >>= 1) {
#pragma unroll
            for (int q = 0; q < kMine; q++) {
                const int probe = pos[q] + step;
                const uint64_t v = lds[(q * SPLIT + sigma) * kRun + min(probe, kRun) - 1];
                if (probe <= len[q] && v < key) pos[q] = probe;
            }
        }
#pragma unroll
        for (int q = 0; q < kMine; q++) rank += pos[q];
        if (more) {
#pragma unroll
            for (int e = 0; e < per; e++) cur[e] = nxt[e];
        }
    }
    if (SPLIT > 1) {
        part[wv][lane] = rank;
        __syncthreads();
        if (sigma != 0) return;
#pragma unroll
        for (int w = 1; w < SPLIT; w++) rank += part[w][lane];
    }
    if (!valid) return;
    rank += i - myrun * kRun;
    gl(S.bkey1)[rank] = key;
    gl(S.bx1)[rank] = px;
    if (dim > 1) gl(S.by1)[rank] = py;
    if (dim > 2) gl(S.bz1)[rank] = pz;
}


// Number of elements of A among the first d of merge(A, B); keys are unique.
// P(i) = A[i] < B[d-1-i] holds exactly for i below the answer. One wave
// samples P at 64 evenly spaced points per round, so the search costs
// ceil(log64(range)) dependent round trips instead of log2(range).
__device__ __forceinline__ int merge_path_wave(const uint64_t *A, int la, const uint64_t *B, int lb, int d) {
    const int lane = lane_id();
    int lo = max(0, d - lb), hi = min(d, la);
    while (hi > lo) {
        const int n = hi - lo;
        const int step = (n + 63) >> 6;
        const int i = lo + lane * step;
        bool p = false;
        if (i < hi) p = A[i] < B[d - 1 - i];
        const int cnt = __popcll(__ballot(p));
        if (step == 1) return lo + cnt;
        if (cnt == 0) return lo;
        const int nlo = lo + (cnt - 1) * step + 1;
        const int nhi = min(hi, lo + cnt * step);
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

template <int T>
struct MergeLds {
    uint64_t in[T];
    uint64_t out[T];
    int32_t src[T];  // >= 0: index into A; < 0: ~index into B
    int32_t sp[2];
    uint64_t pred;
};

// Merges outputs [d0, d1) (d1 - d0 <= T) of merge(A, B) into m.out / m.src
// (block-wide).
template <int T>
__device__ __forceinline__ void merge_tile(const uint64_t *A, int la, const uint64_t *B, int lb, int d0,
                                           int d1, MergeLds<T> &m, int *a0_out, int *b0_out) {
    const int tid = threadIdx.x;
    if (tid < 128) {
        const int sp = merge_path_wave(A, la, B, lb, tid < 64 ? d0 : d1);
        if (tid == 0) m.sp[0] = sp;
        if (tid == 64) m.sp[1] = sp;
    }
    __syncthreads();
    const int a0 = m.sp[0], a1 = m.sp[1];
    const int b0 = d0 - a0, b1 = d1 - a1;
    const int na = a1 - a0, nb = b1 - b0, n = na + nb;
    constexpr int per = T / kTileThreads;
    {
        uint64_t v[per];  // every load issued before the LDS stores
#pragma unroll
        for (int r = 0; r < per; r++) {
            const int j = tid + r * kTileThreads;
            v[r] = j < na ? A[a0 + j] : (j < n ? B[b0 + j - na] : 0ull);
        }
#pragma unroll
        for (int r = 0; r < per; r++) m.in[tid + r * kTileThreads] = v[r];
    }
    __syncthreads();
    const int diag = tid * per;
    if (diag < n) {
        int lo = max(0, diag - nb), hi = min(diag, na);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (m.in[mid] < m.in[na + diag - 1 - mid]) lo = mid + 1;
            else hi = mid;
        }
        int ai = lo, bi = diag - lo;
        for (int q = 0; q < per && diag + q < n; q++) {
            const bool takeA = (bi >= nb) || (ai < na && m.in[ai] < m.in[na + bi]);
            if (takeA) {
                m.out[diag + q] = m.in[ai];
                m.src[diag + q] = a0 + ai;
                ai++;
            } else {
                m.out[diag + q] = m.in[na + bi];
                m.src[diag + q] = ~(b0 + bi);
                bi++;
            }
        }
    }
    __syncthreads();
    *a0_out = a0;
    *b0_out = b0;
}

// One pass of the bottom-up merge sort of the first window: merges runs of
// length L pairwise, ping-pong between akey (sel 0: a->s) and skey (sel 1).
__global__ __launch_bounds__(kTileThreads) void k_init_merge(BatchArgs a, int L, int sel) {
    __shared__ MergeLds<kTile> m;
    const int g = blockIdx.x;
    const int s = find_stream_c(a.pre_itile, a.n, g);
    const StreamDev &S = a.s[s];
    const int len = S.W;
    const int o0 = (g - ldc(a.pre_itile[s])) * kTile;
    const int o1 = min(o0 + kTile, len);
    const uint64_t *sk = sel == 0 ? S.akey : S.skey;
    const float *sx = sel == 0 ? S.ax : S.sx, *sy = sel == 0 ? S.ay : S.sy, *sz = sel == 0 ? S.az : S.sz;
    uint64_t *dk = sel == 0 ? S.skey : S.akey;
    float *dx = sel == 0 ? S.sx : S.ax, *dy = sel == 0 ? S.sy : S.ay, *dz = sel == 0 ? S.sz : S.az;
    const long long L2 = 2ll * L;
    const int base = (int)((o0 / L2) * L2);
    const int la = min(L, len - base);
    const int lb = min(L, len - base - la);
    int a0, b0;
    merge_tile(sk + base, la, sk + base + la, lb, o0 - base, o1 - base, m, &a0, &b0);
    const int dim = S.dim;
    for (int i = threadIdx.x; i < o1 - o0; i += kTileThreads) {
        const int sidx = m.src[i];
        const int gi = sidx >= 0 ? base + sidx : base + la + ~sidx;
        dk[o0 + i] = m.out[i];
        dx[o0 + i] = sx[gi];
        if (dim > 1) dy[o0 + i] = sy[gi];
        if (dim > 2) dz[o0 + i] = sz[gi];
    }
}

// ---------------------------------------------------------------------------
// window update: merge the survivors with the sorted stride

// merge(survivors, sorted stride) -> window, plus cell starts per tile
// (thrust::merge at thrust_helper.cu:47-50 and the run starts of
// compute_offsets_of_cells, aabb.cu:377-395).
__global__ __launch_bounds__(kTileThreads) void k_merge_final(BatchArgs a) {
    __shared__ MergeLds<kWTile> m;
    __shared__ int lds4[4];
    const int g = blockIdx.x;
    const int s = find_stream_c(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - ldc(a.pre_wtile[s]);
    const int o0 = tile * kWTile, o1 = min(o0 + kWTile, S.W);
    const int la = S.W - S.S, lb = S.S;
    const uint64_t *B = S.bkey1;
    int a0, b0;
    merge_tile(S.akey, la, B, lb, o0, o1, m, &a0, &b0);
    if (threadIdx.x == 0) {
        uint64_t pk = 0;
        if (a0 > 0) pk = S.akey[a0 - 1];
        if (b0 > 0 && B[b0 - 1] > pk) pk = B[b0 - 1];
        m.pred = pk;
    }
    __syncthreads();
    const int dim = S.dim, sb = S.slot_bits;
    constexpr int per = kWTile / kTileThreads;
    float cx[per], cy[per], cz[per];  // every gather issued before the stores
#pragma unroll
    for (int r = 0; r < per; r++) {
        const int i = threadIdx.x + r * kTileThreads;
        const int sidx = i < o1 - o0 ? m.src[i] : 0;
        const bool fromA = sidx >= 0;
        const int j = fromA ? sidx : ~sidx;
        const float *px = fromA ? S.ax : S.bx1, *py = fromA ? S.ay : S.by1, *pz = fromA ? S.az : S.bz1;
        cx[r] = px[j];
        cy[r] = dim > 1 ? py[j] : 0.f;
        cz[r] = dim > 2 ? pz[j] : 0.f;
    }
    int starts = 0;
#pragma unroll
    for (int r = 0; r < per; r++) {
        const int i = threadIdx.x + r * kTileThreads;
        if (i < o1 - o0) {
            const uint64_t key = m.out[i];
            S.skey[o0 + i] = key;
            S.sx[o0 + i] = cx[r];
            if (dim > 1) S.sy[o0 + i] = cy[r];
            if (dim > 2) S.sz[o0 + i] = cz[r];
            const uint64_t prev = i > 0 ? m.out[i - 1] : m.pred;
            starts += (o0 + i == 0) || ((key >> sb) != (prev >> sb));
        }
    }
    starts = block_sum(starts, lds4);
    if (threadIdx.x == 0) S.tile_b[tile] = starts;
}

// Cell table: ordered compaction of run starts -> cstart/ckey and position ->
// cell. Replaces compute_offsets_of_cells + thrust_sort
// (aabb.cu:377-410, thrust_helper.cu:16-19): the order comes from the scan,
// so no sort and no host round trip (dbscan.cpp:549) are needed. Also resets
// the union-find (every slot its own root) in index order, coalesced, rather
// than slot by slot from the count kernels.
__global__ __launch_bounds__(kTileThreads) void k_cells(BatchArgs a) {
    __shared__ int lds4[4];
    const int g = blockIdx.x;
    const int s = find_stream_c(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - ldc(a.pre_wtile[s]);
    const int ntile = ldc(a.pre_wtile[s + 1]) - ldc(a.pre_wtile[s]);
    const int p0 = tile * kWTile;
    const int sb = S.slot_bits;
    constexpr int per = kWTile / kTileThreads;
    uint64_t kc[per], kp[per];  // keys of each position and its predecessor, loaded up front
#pragma unroll
    for (int r = 0; r < per; r++) {
        const int p = p0 + r * kTileThreads + threadIdx.x;
        kc[r] = p < S.W ? S.skey[p] : 0ull;
        kp[r] = (p < S.W && p > 0) ? S.skey[p - 1] : 0ull;
    }
    int part = 0;
    for (int t = threadIdx.x; t < tile; t += kTileThreads) part += S.tile_b[t];
    const int base = block_sum(part, lds4);
    int run = 0;
#pragma unroll
    for (int r = 0; r < per; r++) {
        const int p = p0 + r * kTileThreads + threadIdx.x;
        const bool valid = p < S.W;
        uint64_t ck = 0;
        bool f = false;
        if (valid) {
            ck = kc[r] >> sb;
            f = (p == 0) || (ck != (kp[r] >> sb));
        }
        int tot;
        const int rank = block_rank_flags(f, &tot, lds4);
        if (valid) {
            S.pcell[p] = base + run + rank + (f ? 1 : 0) - 1;
            S.parent[p] = p;  // union-find over slots [0, W), each its own root; written in order here
        }
        if (f) {
            const int c = base + run + rank;
            S.cstart[c] = p;
            S.ckey[c] = ck;
        }
        run += tot;
    }
    if (tile == ntile - 1 && threadIdx.x == 0) {
        const int total = base + run;
        S.mut->ncell = total;
        S.cstart[total] = S.W;
    }
}

// ---------------------------------------------------------------------------
// stencil rows: for every cell and every row of its (2k+1)^dim stencil, the
// contiguous range of window positions holding that row's cells. A row varies
// the last used axis, whose key stride is 1 (unused axes have extent 1 and
// index 0), so its cells are the cells with keys in [lo, hi] = the target
// column's keys at zlo..zhi, consecutive in the sorted cell table. For one
// row offset, lo and hi rise with the cell's own key (the target column is
// the cell's column shifted by a constant), so one wave serves 64 consecutive
// cells: a cooperative search brackets the tile's bounds in the cell table,
// the bracket is staged in LDS and each lane binary-searches its own bounds
// there. No hashing, and a few memory round trips per 64 (cell, row) items.

constexpr int kRowTile = 64;    // cells per row-kernel item (one wave)
constexpr int kRowCap = 192;    // bracket cells staged in LDS per wave
constexpr int kRowWaves = kTileThreads / 64;

// Wave-uniform view of one stream for the row kernel (scalar descriptor
// loads, global addressing).
struct RowsView {
    gptr<const uint64_t> ckey;
    gptr<const int32_t> cstart;
    gptr<uint64_t> rows;  // int2 (first, end) per stencil row, as one word
    gptr<uint8_t> cflag;
    gptr<int32_t> fcore;
    int64_t n0, n1, n2;
    int k, R, dim, min_pts, dense_ok, ncell;
};

__device__ __forceinline__ RowsView rows_view(const StreamDev &S) {
    RowsView v;
    v.ckey = gl(ldc(S.ckey));
    v.cstart = gl(ldc(S.cstart));
    v.rows = gl((uint64_t *)ldc(S.rows));
    v.cflag = gl(ldc(S.cflag));
    v.fcore = gl(ldc(S.fcore));
    v.n0 = ldc(S.n[0]);
    v.n1 = ldc(S.n[1]);
    v.n2 = ldc(S.n[2]);
    v.k = ldc(S.k);
    v.R = ldc(S.R);
    v.dim = ldc(S.dim);
    v.min_pts = ldc(S.min_pts);
    v.dense_ok = ldc(S.dense_ok);
    v.ncell = ldc(ldc(S.mut)->ncell);  // written by the cell-table kernel
    return v;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Ranks (number of keys below) of two wave-uniform keys in the sorted
// ckey[0, n), n >= 1: two 64-ary searches in lockstep, one probe per lane per
// key and round, then a last round over at most 64 keys.
__device__ __forceinline__ void wave_rank2(gptr<const uint64_t> ckey, int n, uint64_t k1, uint64_t k2, int *r1,
                                           int *r2) {
    const int lane = lane_id();
    int lo1 = 0, hi1 = n, lo2 = 0, hi2 = n;  // rank in [lo, hi]
    while (hi1 - lo1 > 64 || hi2 - lo2 > 64) {
        const int s1 = hi1 - lo1, s2 = hi2 - lo2;
        const int p1 = lo1 + (int)((long long)(lane + 1) * s1 / 65);  // in [lo, hi) when s > 64
        const int p2 = lo2 + (int)((long long)(lane + 1) * s2 / 65);
        const uint64_t v1 = ckey[min(p1, n - 1)];
        const uint64_t v2 = ckey[min(p2, n - 1)];
        if (s1 > 64) {
            const int c = __popcll(__ballot(v1 < k1));  // the probes below k1 are a prefix of the lanes
            const int nlo = c > 0 ? __builtin_amdgcn_readlane(p1, c - 1) + 1 : lo1;
            hi1 = c < 64 ? __builtin_amdgcn_readlane(p1, c) : hi1;
            lo1 = nlo;
        }
        if (s2 > 64) {
            const int c = __popcll(__ballot(v2 < k2));
            const int nlo = c > 0 ? __builtin_amdgcn_readlane(p2, c - 1) + 1 : lo2;
            hi2 = c < 64 ? __builtin_amdgcn_readlane(p2, c) : hi2;
            lo2 = nlo;
        }
    }
    const bool b1 = lo1 + lane < hi1 && ckey[lo1 + lane] < k1;
    const bool b2 = lo2 + lane < hi2 && ckey[lo2 + lane] < k2;
    *r1 = lo1 + __popcll(__ballot(b1));
    *r2 = lo2 + __popcll(__ballot(b2));
}

// First index in sorted a[0, n) (n >= 1) whose key is >= key.
template <class P>
__device__ __forceinline__ int lower_index(P a, int n, uint64_t key) {
    int base = 0;
    while (n > 1) {
        const int half = n >> 1;
        if (a[base + half] < key) base += half;
        n -= half;
    }
    return base + (a[base] < key ? 1 : 0);
}

// Cell classes (written once per slide by k_rows):
//   kCellSafe  - the grid is dense-safe (k*k > dim: any two points of one
//                cell are closer than eps) and the cell is interior; border
//                cells may hold points clamped in from outside the grid
//                bounds, which the side-length argument does not cover.
//   kCellDense - safe and holding >= min_pts points: all of them are core.
__device__ __forceinline__ uint8_t cell_class(const RowsView &V, int64_t i0, int64_t i1, int64_t i2, int c) {
    if (!V.dense_ok) return 0;
    if (i0 <= 0 || i0 >= V.n0 - 1) return 0;
    if (V.dim > 1 && (i1 <= 0 || i1 >= V.n1 - 1)) return 0;
    if (V.dim > 2 && (i2 <= 0 || i2 >= V.n2 - 1)) return 0;
    const int cnt = V.cstart[c + 1] - V.cstart[c];
    return cnt >= V.min_pts ? (kCellSafe | kCellDense) : kCellSafe;
}

// Storage order of a cell's stencil rows: nearest first, group by group. An
// offset d in [-k, k] ranks as 0, -1, +1, -2, +2, ... (|d| first, the
// negative side first). In 3-D row (d1, d2) - offsets on the first two axes -
// is stored at rank(d1) * (2k+1) + rank(d2): the rows of one first-axis
// offset (one rows-kernel item) stay adjacent, and the groups come nearest
// first. In 2-D row d1 is stored at rank(d1). The centre row is row 0. The
// count kernels walk rows in storage order, so a query's count meets the
// rows holding most of its neighbours first and reaches min_pts sooner; the
// other consumers need no particular order.
__device__ __forceinline__ int offset_rank(int d) { return d == 0 ? 0 : 2 * (d < 0 ? -d : d) - (d < 0 ? 1 : 0); }

// Rows per item: in 3-D one item holds the 2k+1 rows of one offset o0 on the
// first axis (o1 = -k..k on the second), whose target columns are adjacent,
// so one bracket serves all of them; in 2-D and 1-D an item is one row.
__device__ __forceinline__ int rows_groups(int dim, int k, int R) { return dim == 3 ? 2 * k + 1 : R; }

// Item g of tile t: the rows with offset o0 = g - k (3-D), or row g; the
// 64 cells of the tile one per lane. The item holding the centre row also
// records each cell's class. lk/lc: this wave's LDS. NS >= the item's row
// count (2k + 1 <= 7 in 3-D, 1 otherwise): the lower bounds of all its rows
// in the staged bracket are searched in lockstep, one LDS round per step for
// all of them.
template <int NS>
__device__ __forceinline__ void rows_tile(const RowsView &V, int t, int g, uint64_t *lk, int32_t *lc) {
    const int lane = lane_id();
    const int c = t * kRowTile + lane;
    const bool cv = c < V.ncell;
    const uint64_t ck = cv ? V.ckey[c] : 0ull;
    int64_t q, i0, i1, i2;
    divmod_key(ck, V.n2, &q, &i2);
    divmod_key((uint64_t)q, V.n1, &i0, &i1);
    const int k = V.k, w = 2 * k + 1;
    const int nsub = V.dim == 3 ? w : 1;
    // offsets of the item's rows (uniform): o0 fixed, o1 over [o1lo, o1lo + nsub)
    int o0 = 0, o1lo = 0;
    if (V.dim == 3) {
        o0 = g - k;
        o1lo = -k;
    } else if (V.dim == 2) {
        o0 = g - k;
    }
    int64_t zc, zn;
    if (V.dim == 3) {
        zc = i2;
        zn = V.n2;
    } else if (V.dim == 2) {
        zc = i1;
        zn = V.n1;
    } else {
        zc = i0;
        zn = V.n0;
    }
    // a bracket of the cell table holding every row bound of the item: the
    // bounds rise with the lane and with o1, and lo >= key + shift - k,
    // hi <= key + shift + k
    const uint64_t cvm = __ballot(cv);  // non-empty: t < ceil(ncell / 64)
    const int64_t klo = (int64_t)readlane64(ck, __ffsll((unsigned long long)cvm) - 1);
    const int64_t khi = (int64_t)readlane64(ck, 63 - __clzll((long long)cvm));
    const int64_t sh_lo = ((int64_t)o0 * V.n1 + o1lo) * V.n2, sh_hi = ((int64_t)o0 * V.n1 + o1lo + nsub - 1) * V.n2;
    const int64_t bmin = klo + sh_lo - k, bmax = khi + sh_hi + k;
    int b0 = 0, b1 = 0;
    if (bmax >= 0) wave_rank2(V.ckey, V.ncell, (uint64_t)max(bmin, (int64_t)0), (uint64_t)bmax + 1, &b0, &b1);
    const int B = b1 - b0;
    const bool staged = B > 0 && B <= kRowCap;
    if (staged) {
        for (int i = lane; i < B; i += 64) lk[i] = V.ckey[b0 + i];
        for (int i = lane; i <= B; i += 64) lc[i] = V.cstart[b0 + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // key bounds of each row: the target column's keys at zlo..zhi
    uint64_t lo[NS], hi1[NS];
    bool ok[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const int o1 = o1lo + j;
        bool okj = cv && j < nsub;
        if (V.dim == 3) okj = okj && i0 + o0 >= 0 && i0 + o0 < V.n0 && i1 + o1 >= 0 && i1 + o1 < V.n1;
        else if (V.dim == 2) okj = okj && i0 + o0 >= 0 && i0 + o0 < V.n0;
        const int64_t colbase = (int64_t)ck - zc + ((int64_t)o0 * V.n1 + (V.dim == 3 ? o1 : 0)) * V.n2;
        lo[j] = (uint64_t)(colbase + (zc - k < 0 ? 0 : zc - k));
        hi1[j] = (uint64_t)(colbase + (zc + k >= zn ? zn - 1 : zc + k)) + 1ull;
        ok[j] = okj;
    }
    int f[NS], e[NS];  // first and end positions of each row's cells in the bracket
    if (staged) {
#pragma unroll
        for (int j = 0; j < NS; j++) f[j] = e[j] = 0;
        for (int n = B; n > 1;) {
            const int half = n >> 1;
#pragma unroll
            for (int j = 0; j < NS; j++) {
                if (lk[f[j] + half] < lo[j]) f[j] += half;
                if (lk[e[j] + half] < hi1[j]) e[j] += half;
            }
            n -= half;
        }
#pragma unroll
        for (int j = 0; j < NS; j++) {
            f[j] += lk[f[j]] < lo[j] ? 1 : 0;
            e[j] += lk[e[j]] < hi1[j] ? 1 : 0;
        }
    }
#pragma unroll
    for (int j = 0; j < NS; j++) {
        if (j >= nsub) break;
        uint64_t row = 0ull;
        if (ok[j] && B > 0) {
            if (staged) {
                if (f[j] < e[j]) row = (uint64_t)(uint32_t)lc[f[j]] | ((uint64_t)(uint32_t)lc[e[j]] << 32);
            } else {
                const int fg = b0 + lower_index(V.ckey + b0, B, lo[j]);
                const int eg = b0 + lower_index(V.ckey + b0, B, hi1[j]);
                if (fg < eg) row = (uint64_t)(uint32_t)V.cstart[fg] | ((uint64_t)(uint32_t)V.cstart[eg] << 32);
            }
        }
        // stencil row number: nearest groups and rows first, the centre row 0
        const int r = V.dim == 3 ? offset_rank(o0) * w + offset_rank(o1lo + j) : (V.dim == 2 ? offset_rank(o0) : 0);
        if (cv) {
            V.rows[(size_t)c * V.R + r] = row;
            if (r == 0) {
                V.cflag[c] = cell_class(V, i0, i1, i2, c);
                V.fcore[c] = kNoCore;
            }
        }
    }
    if (staged) __builtin_amdgcn_wave_barrier();  // reads done before the next item's stores
}

// Persistent: walks the device-side list of (stream, 64-cell tile, row group)
// items, one per wave. Block 0 also publishes the per-stream prefix of cell
// counts for the cell kernels.
__global__ __launch_bounds__(kTileThreads) void k_rows(BatchArgs a) {
    __shared__ long long pre[kMaxBatch + 1];
    __shared__ long long tmp4[4];
    __shared__ uint64_t lkey[kRowWaves][kRowCap];
    __shared__ int32_t lcs[kRowWaves][kRowCap + 1];
    block_stream_prefix(a.n, pre, tmp4, [&](int i) {
        const StreamDev &S = a.s[i];
        return (long long)((S.mut->ncell + kRowTile - 1) / kRowTile) * rows_groups(S.dim, S.k, S.R);
    });
    if (blockIdx.x == 0) {
        long long run = 0;
        if (threadIdx.x == 0) {
            a.cell_pre[0] = 0;
            for (int i = 0; i < a.n; i++) {
                run += a.s[i].mut->ncell;
                a.cell_pre[i + 1] = (int32_t)run;
            }
        }
    }
    const long long total = pre[a.n];
    const int wv = uni(threadIdx.x >> 6);
    for (long long wi = (long long)blockIdx.x * kRowWaves + wv; wi < total; wi += (long long)gridDim.x * kRowWaves) {
        const int s = uni(find_stream(pre, a.n, wi));
        const RowsView V = rows_view(a.s[s]);
        const int local = uni((int)(wi - pre[s]));
        const int G = rows_groups(V.dim, V.k, V.R);
        const int t = local / G;
        if (V.dim == 3) rows_tile<7>(V, t, local - t * G, lkey[wv], lcs[wv]);
        else rows_tile<1>(V, t, local - t * G, lkey[wv], lcs[wv]);
    }
}

// ---------------------------------------------------------------------------
// cell-parallel neighbourhood kernels: one wave per (stream, cell), walking
// the device-side cell list in a persistent loop. The candidates of a cell
// are the concatenation of its stencil rows; lanes 0..R-1 hold one row each.

// Index r of the row holding candidate j (largest r < R with pre(r) <= j).
// Every lane takes part; j may differ per lane.
__device__ __forceinline__ int row_of(int pre, int R, int j) {
    int lo = 0, hi = R;
#pragma unroll
    for (int it = 0; it < 6; it++) {
        const int mid = (lo + hi) >> 1;
        const int pm = __shfl(pre, mid);
        if (hi - lo > 1) {
            if (pm <= j) lo = mid;
            else hi = mid;
        }
    }
    return lo;
}

struct CellCtx {
    int s, c;
};

__device__ __forceinline__ CellCtx cell_of(const BatchArgs &a, int gc) {
    CellCtx x;
    x.s = find_stream_c(a.cell_pre, a.n, gc);
    x.c = gc - ldc(a.cell_pre[x.s]);
    return x;
}


struct ChunkCtx {
    int s, p0, p1;
};

__device__ __forceinline__ ChunkCtx chunk_of(const BatchArgs &a, int g) {
    ChunkCtx x;
    x.s = find_stream_c(a.pre_chunk, a.n, g);
    const int W = a.s[x.s].W;
    x.p0 = (g - ldc(a.pre_chunk[x.s])) * kChunk;
    x.p1 = min(x.p0 + kChunk, W);
    return x;
}

// The union-find element (a slot) representing a core: the slot of the first
// core of its cell when the cell is safe (all its points are within eps of
// each other, so all its cores belong to one cluster), else its own slot.
__device__ __forceinline__ int rep_slot(const StreamDev &S, int j) {
    if (!S.dense_ok) return slot_of(S, j);
    const int cj = S.pcell[j];
    return slot_of(S, (S.cflag[cj] & kCellSafe) ? S.fcore[cj] : j);
}

// Wave-uniform view of one stream for the cell-parallel kernels. The
// descriptor is read with scalar loads (no kernel writes it during a launch)
// and the arrays are addressed as global memory, so the per-lane accesses
// are global loads and stores off a scalar base.
struct CellView {
    gptr<const float> sx, sy, sz;
    gptr<const uint64_t> skey;
    gptr<const int32_t> cstart;
    gptr<const uint64_t> rows;  // int2 (first, end) per stencil row, as one word
    gptr<const uint8_t> cflag;
    gptr<const int32_t> pcell;
    gptr<uint8_t> flags;
    gptr<int32_t> fcore, parent, rep;
    int R, dim, min_pts;
    uint32_t smask;
    float eps2;
};

__device__ __forceinline__ CellView cell_view(const StreamDev &S) {
    CellView v;
    v.sx = gl(ldc(S.sx));
    v.sy = gl(ldc(S.sy));
    v.sz = gl(ldc(S.sz));
    v.skey = gl(ldc(S.skey));
    v.cstart = gl(ldc(S.cstart));
    v.rows = gl((const uint64_t *)ldc(S.rows));
    v.cflag = gl(ldc(S.cflag));
    v.pcell = gl(ldc(S.pcell));
    v.flags = gl(ldc(S.flags));
    v.fcore = gl(ldc(S.fcore));
    v.parent = gl(ldc(S.parent));
    v.rep = gl(ldc(S.rep));
    v.R = ldc(S.R);
    v.dim = ldc(S.dim);
    v.min_pts = ldc(S.min_pts);
    v.smask = (uint32_t)((1ull << ldc(S.slot_bits)) - 1ull);
    v.eps2 = ldc(S.eps2);
    return v;
}

__device__ __forceinline__ int vslot(const CellView &V, int pos) { return (int)((uint32_t)V.skey[pos] & V.smask); }

// Element at byte offset `off` of a wave-uniform array: the load takes the
// array's scalar base and a 32-bit lane offset, with no 64-bit address
// arithmetic per lane (window positions are below 2^24, so every offset into
// a per-position array fits).
template <class T>
__device__ __forceinline__ T ldo(gptr<const T> base, uint32_t off) {
    return *(gptr<const T>)((gptr<const char>)base + off);
}
template <class T>
__device__ __forceinline__ T ldo(gptr<T> base, uint32_t off) {
    return *(gptr<const T>)((gptr<const char>)base + off);
}

constexpr int kRowGroup = 5;  // stencil rows fetched together
constexpr int kUnroll = 4;    // candidates fetched together

// Walks the stencil candidates of cell c for one query point, kRowGroup rows
// and kUnroll candidates per memory round trip (a thread's walk is a chain of
// dependent loads; batching shortens it). visit(j, x, y, z) returns true to
// stop the walk.
// With `step` > 1 the thread takes rows first, first+step, ... only (the
// other rows go to sibling threads of the same query).
// Candidates at positions >= limit are skipped (rows are position ranges).
template <class Visit>
__device__ __forceinline__ void walk_candidates(const CellView &V, int c, Visit visit, int first = 0,
                                                int step = 1, int limit = 0x7fffffff) {
    const int dim = V.dim, R = V.R;
    for (int r0 = first; r0 < R; r0 += kRowGroup * step) {
        int rx[kRowGroup], ry[kRowGroup];
#pragma unroll
        for (int i = 0; i < kRowGroup; i++) {
            const uint64_t rw = (r0 + i * step < R) ? V.rows[(size_t)c * R + r0 + i * step] : 0ull;
            rx[i] = (int)(uint32_t)rw;
            ry[i] = min((int)(uint32_t)(rw >> 32), limit);
        }
#pragma unroll
        for (int i = 0; i < kRowGroup; i++) {
            for (int j = rx[i]; j < ry[i]; j += kUnroll) {
                float cx[kUnroll], cy[kUnroll], cz[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; u++) {
                    const int jj = min(j + u, ry[i] - 1);
                    cx[u] = V.sx[jj];
                    cy[u] = dim > 1 ? V.sy[jj] : 0.f;
                    cz[u] = dim > 2 ? V.sz[jj] : 0.f;
                }
#pragma unroll
                for (int u = 0; u < kUnroll; u++)
                    if (j + u < ry[i] && visit(j + u, cx[u], cy[u], cz[u])) return;
            }
        }
    }
}

// Query ownership of a block of a point-parallel kernel with LPQ lanes per
// query: block g covers chunk g / LPQ; each query is served by LPQ adjacent
// lanes that split its stencil rows. More lanes per query shorten each
// thread's chain of dependent loads when a launch has few queries.
template <int LPQ>
struct QueryLane {
    ChunkCtx x;
    int p;    // window position of the query (may be >= x.p1: idle)
    int sub;  // 0..LPQ-1
};

template <int LPQ>
__device__ __forceinline__ QueryLane<LPQ> query_lane(const BatchArgs &a, int vb) {
    QueryLane<LPQ> q;
    const int chunk = vb / LPQ, part = vb % LPQ;
    q.x = chunk_of(a, chunk);
    q.sub = lane_id() % LPQ;
    q.p = q.x.p0 + part * (64 / LPQ) + lane_id() / LPQ;
    return q;
}

template <int LPQ>
__device__ __forceinline__ int group_sum(int v) {
#pragma unroll
    for (int o = 1; o < LPQ; o <<= 1) v += xor_lane(v, o);
    return v;
}

template <int LPQ>
__device__ __forceinline__ int group_min(int v) {
#pragma unroll
    for (int o = 1; o < LPQ; o <<= 1) v = min(v, xor_lane(v, o));
    return v;
}

// ---------------------------------------------------------------------------
// cell-parallel neighbourhood kernels: one wave per cell. The lanes hold 64
// of the cell's stencil candidates at a time - consecutive positions of its
// rows, so the loads coalesce and each candidate is read once per cell rather
// than once per query - and every query of the cell is tested against all of
// them with one ballot.

constexpr int kCellWaves = 4;  // waves (cells) per block

__device__ __forceinline__ float lane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Work counters of one wave (uniform values), flushed to the per-stream
// totals with one atomic per counter when the wave moves to another stream
// or ends. The cell kernels are instantiated with WaveStats when counting is
// on (a.stats non-null) and with NoStats otherwise, so that normal launches
// carry no counting instructions.
struct WaveStats {
    int s = -1;
    unsigned long long v[kNumStats] = {};
    __device__ __forceinline__ void add(int k, unsigned long long x) { v[k] += x; }
};

struct NoStats {
    __device__ __forceinline__ void add(int, unsigned long long) {}
};

__device__ __forceinline__ void stats_flush(const BatchArgs &, NoStats &) {}
__device__ __forceinline__ void stats_stream(const BatchArgs &, NoStats &, int) {}

__device__ __forceinline__ void stats_flush(const BatchArgs &a, WaveStats &w) {
    if (a.stats && w.s >= 0 && lane_id() == 0)
        for (int k = 0; k < kNumStats; k++)
            if (w.v[k]) atomicAdd(a.stats + (size_t)w.s * kNumStats + k, w.v[k]);
    for (int k = 0; k < kNumStats; k++) w.v[k] = 0;
}

__device__ __forceinline__ void stats_stream(const BatchArgs &a, WaveStats &w, int s) {
    if (s != w.s) {
        stats_flush(a, w);
        w.s = s;
    }
}

// A cell's stencil candidates, 64 at a time. Lane r < R holds stencil row r:
// its first window position and its length. A walk in scalar registers opens
// the non-empty rows in order (a ballot mask; each row read with readlane)
// and gives every lane the position of its candidate, which belongs to the
// last row opened at or before it - no cross-lane search per chunk.
struct RowWalk {
    int rx = 0, len = 0;  // lane r: row r
    uint64_t rest = 0;    // non-empty rows not opened yet (uniform)
    int acc = 0;          // first candidate of the next row to open (uniform)
    int base = 0;         // position minus candidate index in the last row opened (uniform)
};

// Row r of cell A in lane r < R (int2 as one word), 0 in the other lanes.
__device__ __forceinline__ uint64_t row_word(const CellView &V, int A) {
    const int lane = lane_id();
    return lane < V.R ? V.rows[(size_t)A * V.R + lane] : 0ull;
}

// The walk over rows (row_word), each cut at window position `limit`.
__device__ __forceinline__ RowWalk row_walk(uint64_t rw, int limit) {
    RowWalk w;
    w.rx = (int)(uint32_t)rw;
    w.len = max(0, min((int)(uint32_t)(rw >> 32), limit) - w.rx);
    w.rest = __ballot(w.len > 0);
    return w;
}

// Waits here for v, loaded earlier. The cell kernels issue every load that
// depends on the same thing together and then wait for all of them at once,
// so that none is moved past a branch into a round trip of its own.
template <class T>
__device__ __forceinline__ void arrived(T v) {
    asm volatile("" ::"v"(v));
}

// Opens the rows that start before jb + 64 and returns this lane's candidate
// position in chunk [jb, jb + 64); lanes past the last candidate get the last
// one. *end is the candidate count once every row is open, else INT_MAX.
// Callers stop at jb >= *end before using the position.
__device__ __forceinline__ int walk_chunk(RowWalk &w, int jb, int *end) {
    const int j = jb + lane_id();
    int base = w.base;
    while (w.rest != 0ull && w.acc < jb + 64) {
        const int r = __ffsll((unsigned long long)w.rest) - 1;
        w.rest &= w.rest - 1ull;
        const int b = __builtin_amdgcn_readlane(w.rx, r) - w.acc;
        if (j >= w.acc) base = b;
        w.base = b;
        w.acc += __builtin_amdgcn_readlane(w.len, r);
    }
    *end = w.rest != 0ull ? INT_MAX : w.acc;
    return base + min(j, *end - 1);
}

// A cell's header as each lane loaded it: run start and end, class, and
// stencil row `lane` (lanes < R). One round trip, issued ahead of its use.
struct CellHdr {
    int csv, cev, cfv;
    uint64_t rw;
};

__device__ __forceinline__ CellHdr load_hdr(const StreamDev &S, int A) {
    const gptr<const int32_t> cstart = gl((const int32_t *)ldc(S.cstart));
    const gptr<const uint8_t> cflag = gl((const uint8_t *)ldc(S.cflag));
    const gptr<const uint64_t> rows = gl((const uint64_t *)ldc(S.rows));
    const int R = ldc(S.R);
    CellHdr h;
    h.csv = cstart[A];
    h.cev = cstart[A + 1];
    h.cfv = cflag[A];
    // every lane loads (no branch around the load); lanes >= R are cleared
    // where the row is used
    h.rw = rows[(size_t)A * R + min(lane_id(), R - 1)];
    return h;
}

// Range count = core identification (the reference's __raygen__grid +
// __intersection__grid, dbscan.cu:155-194,250-270, and set_label,
// aabb.cu:575-587), for cell A whose header h was loaded ahead. Dense cells
// are core without counting (all pairs within eps). A cell whose whole
// stencil holds fewer than min_pts points cannot hold a core: its points are
// non-core without a distance test. Otherwise each query counts its
// neighbours, itself included, and stops once the count reaches min_pts.
// The queries and the first candidate chunk load in one round, after which
// `prefetch` issues the next cell's header, so that it arrives while this
// cell computes; each further chunk is a round of its own. `prefetch` runs
// exactly once on every path.
template <class ST, class Pre>
__device__ __forceinline__ void count_cell(const CellView &V, int A, const CellHdr &h, ST &st, Pre &&prefetch) {
    const int lane = lane_id();
    const int dim = V.dim, minpts = V.min_pts;
    const float e2 = V.eps2;
    const int cs = uni(h.csv), nA = uni(h.cev) - cs;
    const int cf = uni(h.cfv);
    if (cf & kCellDense) {
        const int rsv = vslot(V, cs);
        prefetch();
        st.add(kStatDensePoints, nA);
        const int rs = uni(rsv);
        for (int i = lane; i < nA; i += 64) {
            V.flags[cs + i] = kFlagCore | kFlagDense;
            V.rep[cs + i] = rs;
        }
        if (lane == 0) V.fcore[A] = cs;
        return;
    }
    const bool safe = cf & kCellSafe;
    const RowWalk w0 = row_walk(lane < V.R ? h.rw : 0ull, INT_MAX);
    st.add(kStatQueries, nA);
    if (uni(wave_sum(w0.len)) < minpts) {
        // fewer candidates than min_pts: no core (rep is read for cores only)
        prefetch();
        for (int i = lane; i < nA; i += 64) V.flags[cs + i] = 0;
        if (lane == 0) V.fcore[A] = kNoCore;
        return;
    }
    int fc = kNoCore, fslot = 0;
    // one block of up to 64 queries; the first block issues the prefetch
    auto block = [&](const int q0, const bool first) __attribute__((always_inline)) {
        const int nq = min(64, nA - q0);
        const int pq = cs + q0 + min(lane, nq - 1);
        const float qx = V.sx[pq];
        const float qy = dim > 1 ? V.sy[pq] : 0.f;
        const float qz = dim > 2 ? V.sz[pq] : 0.f;
        const int qslot = vslot(V, pq);
        RowWalk w = w0;
        int end;
        int pos = walk_chunk(w, 0, &end);
        const float cx = V.sx[pos];
        const float cy = dim > 1 ? V.sy[pos] : 0.f;
        const float cz = dim > 2 ? V.sz[pos] : 0.f;
        if (first) prefetch();
        arrived(qx);
        arrived(qy);
        arrived(qz);
        arrived(cx);
        arrived(cy);
        arrived(cz);
        int mine = 0;
        uint64_t open = nq == 64 ? ~0ull : ((1ull << nq) - 1ull);
        // the open queries against candidate chunk [jb, jb + 64)
        auto test = [&](const int jb, const float x, const float y, const float z) __attribute__((always_inline)) {
            const bool v = jb + lane < end;
            uint64_t m = open;
            st.add(kStatCountTests, (unsigned long long)__popcll(m) * (unsigned)min(64, end - jb));
            while (m) {
                const int qi = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const int c =
                    __popcll(__ballot(v & (sqd(dim, lane_f(qx, qi), lane_f(qy, qi), lane_f(qz, qi), x, y, z) < e2)));
                if (lane == qi) mine += c;
            }
            open &= ~__ballot(mine >= minpts);
        };
        // the first chunk outside the loop: its wait leaves the prefetch in flight
        test(0, cx, cy, cz);
        for (int jb = 64; open; jb += 64) {
            pos = walk_chunk(w, jb, &end);
            if (jb >= end) break;
            const float x = V.sx[pos];
            const float y = dim > 1 ? V.sy[pos] : 0.f;
            const float z = dim > 2 ? V.sz[pos] : 0.f;
            test(jb, x, y, z);
        }
        if (lane < nq) {
            V.flags[cs + q0 + lane] = mine >= minpts ? kFlagCore : 0;
            if (!safe) V.rep[cs + q0 + lane] = qslot;  // unsafe cell: every core its own element
        }
        const uint64_t cm = __ballot(lane < nq && mine >= minpts);
        if (fc == kNoCore && cm != 0ull) {
            const int l = __ffsll((unsigned long long)cm) - 1;
            fc = cs + q0 + l;
            fslot = lane_i(qslot, l);
        }
    };
    block(0, true);
    for (int q0 = 64; q0 < nA; q0 += 64) block(q0, false);
    if (lane == 0) V.fcore[A] = fc;  // first (minimum-slot) core: the cell's representative
    // safe cell: one union-find element, the representative's slot
    if (safe)
        for (int i = lane; i < nA; i += 64) V.rep[cs + i] = fc != kNoCore ? fslot : 0;
}

// One wave per cell in a grid-stride walk over all cells of the launch; the
// next cell's header is in flight while the current cell computes.
template <class ST>
__global__ __launch_bounds__(64 * kCellWaves) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_count(BatchArgs a) {
    const int total = uni(ldc(a.cell_pre[a.n]));
    const int nw = gridDim.x * kCellWaves;
    ST st;
    int gc = uni(blockIdx.x * kCellWaves + (threadIdx.x >> 6));
    if (gc < total) {
        int s = uni(find_stream_from(a.cell_pre, a.n, gc, 0));  // cells rise, so streams do
        int c = gc - uni(ldc(a.cell_pre[s]));
        CellHdr h = load_hdr(a.s[s], c);
        while (true) {
            const int gn = gc + nw;
            int sn = s, cn = 0;
            CellHdr hn = h;
            // (past the last cell it reloads the last one: no branch around the loads)
            auto prefetch = [&]() {
                const int gq = min(gn, total - 1);
                sn = uni(find_stream_from(a.cell_pre, a.n, gq, s));
                cn = gq - uni(ldc(a.cell_pre[sn]));
                hn = load_hdr(a.s[sn], cn);
            };
            stats_stream(a, st, s);
            count_cell(cell_view(a.s[s]), c, h, st, prefetch);
            if (gn >= total) break;
            gc = gn;
            s = sn;
            c = cn;
            h = hn;
        }
    }
    stats_flush(a, st);
}

// Root of x's union-find tree, read only (no path halving): for kernels that
// run after the merge has finished, when the trees no longer change.
__device__ __forceinline__ int uf_root(const int32_t *parent, int x) {
    int p = uf_load(parent + x);
    while (p != x) {
        x = p;
        p = uf_load(parent + x);
    }
    return x;
}

// Outputs of one cell, after the merge (the reference's post_cluster_t,
// aabb.cu:173-196, and border claim, dbscan.cu:370-377): a core takes the
// root of its union-find element - the minimum core slot of its cluster; a
// non-core point takes the minimum root among the cores within eps (border,
// the deterministic form of the reference's claim) or stays noise. The roots
// are found here, so no separate flatten pass runs. The wave of each
// stream's first cell also advances the ring cursor. Loads in rounds, as in
// count_cell.
template <class ST>
__device__ void output_cell(const BatchArgs &a, int s, const CellView &V, const IO &io, int A, ST &st) {
    const int lane = lane_id();
    const int dim = V.dim;
    const float e2 = V.eps2;
    const int32_t *parent = ldc(a.s[s].parent);
    const int cfv = V.cflag[A], csv = V.cstart[A], cev = V.cstart[A + 1];
    const uint64_t rw = row_word(V, A);
    arrived(rw);
    if (A == 0 && lane == 0) {
        StreamMut *m = ldc(a.s[s].mut);
        if (a.stats) a.stats[(size_t)s * kNumStats + kStatCells] += (unsigned long long)m->ncell;
        m->t = m->t + 1;
        m->u = m->u + 1 == ldc(a.s[s].units) ? 0 : m->u + 1;
    }
    const int cs = uni(csv), nA = uni(cev) - cs;
    const bool dense = uni(cfv) & kCellDense;
    gptr<int32_t> label = gl(io.label), cidv = gl(io.cid);
    const RowWalk w0 = row_walk(rw, INT_MAX);
    for (int q0 = 0; q0 < nA; q0 += 64) {
        const int nq = min(64, nA - q0);
        const int pq = cs + q0 + min(lane, nq - 1);
        const int qf = V.flags[pq];
        const uint64_t qk = V.skey[pq];
        const int qr = V.rep[pq];
        const float qx = V.sx[pq];
        const float qy = dim > 1 ? V.sy[pq] : 0.f;
        const float qz = dim > 2 ? V.sz[pq] : 0.f;
        arrived(qk);
        arrived(qr);
        arrived(qx);
        arrived(qy);
        arrived(qz);
        const int qslot = (int)((uint32_t)qk & V.smask);
        if (lane < nq && (qf & kFlagCore)) {
            label[qslot] = 0;
            cidv[qslot] = uf_root(parent, qr);
        }
        if (dense) continue;
        const bool isq = lane < nq && !(qf & kFlagCore);
        const uint64_t open = __ballot(isq);
        if (open == 0ull) continue;
        int best = INT_MAX;
        RowWalk w = w0;
        for (int jb = 0;; jb += 64) {
            int end;
            const int pos = walk_chunk(w, jb, &end);
            if (jb >= end) break;
            const int f = V.flags[pos];
            const float cx = V.sx[pos];
            const float cy = dim > 1 ? V.sy[pos] : 0.f;
            const float cz = dim > 2 ? V.sz[pos] : 0.f;
            const int rp = V.rep[pos];  // meaningful for cores only
            arrived(cx);
            arrived(cy);
            arrived(cz);
            arrived(rp);
            const bool core = jb + lane < end && (f & kFlagCore);
            uint64_t m = open;
            st.add(kStatBorderTests, (unsigned long long)__popcll(m) * (unsigned)min(64, end - jb));
            // roots of the chunk's cores, found once a query of the cell hits one
            int cid = INT_MAX;
            bool have = false;
            while (m) {
                const int qi = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const bool hit = core & (sqd(dim, lane_f(qx, qi), lane_f(qy, qi), lane_f(qz, qi), cx, cy, cz) < e2);
                if (__ballot(hit) == 0ull) continue;
                if (!have) {
                    if (core) cid = uf_root(parent, rp);
                    have = true;
                }
                const int mn = wave_min(hit ? cid : INT_MAX);
                if (lane == qi) best = min(best, mn);
            }
        }
        if (isq) {
            label[qslot] = best != INT_MAX ? 1 : 2;
            cidv[qslot] = best != INT_MAX ? best : qslot;
        }
    }
}

template <class ST>
__global__ __launch_bounds__(64 * kCellWaves) void k_border(BatchArgs a) {
    const int total = uni(ldc(a.cell_pre[a.n]));
    const int nw = gridDim.x * kCellWaves;
    ST st;
    int s = 0;  // stream of the wave's current cell (cells rise, so streams do)
    for (int gc = uni(blockIdx.x * kCellWaves + (threadIdx.x >> 6)); gc < total; gc += nw) {
        s = uni(find_stream_from(a.cell_pre, a.n, gc, s));
        const int c = gc - uni(ldc(a.cell_pre[s]));
        stats_stream(a, st, s);
        output_cell(a, s, cell_view(a.s[s]), a.use_io1 ? a.io1 : a.io[s], c, st);
    }
    stats_flush(a, st);
}

// Touching (element, element) pairs of one wave, united in parallel rounds:
// the pairs of several candidate chunks and cells - of any streams of the
// launch - are collected in LDS and each lane then walks the union-find of
// its pair's stream, so the dependent loads of the walks overlap instead of
// adding up cell after cell. 320 pairs (3.75 KB per wave) still let 8 blocks
// of 4 waves share a CU's LDS; a fuller queue flushes less often where unites
// are many (skewed3d 429 -> 419 us, a 62-stream served batch 1084 -> 1059 us,
// 64 streams unchanged; 384 gained more on served batches but cost 1-2% at 64
// streams and on walks3d).
constexpr int kEdgeCap = 320;

// One cell's header as the merge needs it: stream, cell within the stream,
// first core position (kNoCore: none, or no cell), flags and position range.
struct CellHead {
    int s, c, fc, cf, cs, ce;
};

// Lane l's cell g = g0 + l * nw of a wave's grid-stride sequence: its stream
// by a per-lane binary search over the batch's cell prefix (a small table
// that stays in cache), then the header, all lanes' loads in one round. When
// every lane's cell is in one stream (always for a single stream) the array
// bases come from scalar loads of that stream's descriptor, otherwise from
// per-lane loads of each lane's descriptor. A cell past `total` reads as a
// cell without cores. Called by the whole wave.
__device__ __forceinline__ CellHead cell_head(const BatchArgs &a, long long g, int total) {
    CellHead h{0, 0, kNoCore, 0, 0, 0};
    const bool valid = g < total;
    const gptr<const int32_t> pre = gl((const int32_t *)a.cell_pre);
    int lo = 0, hi = a.n, plo = 0;  // pre[0] = 0
    while (valid && hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        const int v = pre[mid];
        if ((long long)v <= g) {
            lo = mid;
            plo = v;
        } else {
            hi = mid;
        }
    }
    h.s = lo;
    h.c = valid ? (int)(g - plo) : 0;
    const int s0 = uni(h.s);  // lane 0's cell is valid
    gptr<const int32_t> fcore, cstart;
    gptr<const uint8_t> cflag;
    if (__ballot(valid && h.s != s0) == 0ull) {
        const StreamDev &S = a.s[s0];
        fcore = gl((const int32_t *)ldc(S.fcore));
        cflag = gl((const uint8_t *)ldc(S.cflag));
        cstart = gl((const int32_t *)ldc(S.cstart));
    } else {
        const gptr<const StreamDev> sd = gl(a.s) + h.s;
        fcore = gl((const int32_t *)sd->fcore);
        cflag = gl((const uint8_t *)sd->cflag);
        cstart = gl((const int32_t *)sd->cstart);
    }
    if (valid) {
        h.fc = fcore[h.c];
        h.cf = cflag[h.c];
        h.cs = cstart[h.c];
        h.ce = cstart[h.c + 1];
    }
    return h;
}

// Lane j's header, wave-uniform.
__device__ __forceinline__ CellHead lane_head(const CellHead &m, int j) {
    return CellHead{lane_i(m.s, j), lane_i(m.c, j), lane_i(m.fc, j), lane_i(m.cf, j), lane_i(m.cs, j),
                    lane_i(m.ce, j)};
}

struct Edge {
    int x, y;  // union-find elements (slots)
    int s;     // stream
};

struct EdgeBuf {
    Edge *e;  // this wave's LDS slice
    int n;    // pairs held (uniform)
};

__device__ __forceinline__ void edge_flush(const BatchArgs &a, EdgeBuf &eb) {
    for (int i = lane_id(); i < eb.n; i += 64) {
        const Edge e = eb.e[i];
        uf_unite(ldc(a.s[e.s].parent), e.x, e.y);
    }
    eb.n = 0;
}

// Queues (target, rep) for every lane that wants it, once per run of lanes
// holding the same representative (a safe cell's points are adjacent in a
// candidate chunk and share one representative).
template <class ST>
__device__ __forceinline__ void unite_once_per_rep(const BatchArgs &a, EdgeBuf &eb, int s, bool want, int rep,
                                                   int target, ST &st) {
    const int lane = lane_id();
    const uint64_t m = __ballot(want);
    if (m == 0ull) return;
    const uint64_t below = m & ((1ull << lane) - 1ull);
    const int prev = below ? 63 - __clzll((long long)below) : lane;
    const int prep = __shfl(rep, prev);
    const bool first = want && (below == 0ull || prep != rep);
    const uint64_t fm = __ballot(first);
    const int nf = __popcll(fm);
    st.add(kStatMergePairs, __popcll(m));
    st.add(kStatUnites, nf);
    if (eb.n + nf > kEdgeCap) edge_flush(a, eb);
    if (first) eb.e[eb.n + __popcll(fm & ((1ull << lane) - 1ull))] = Edge{target, rep, s};
    eb.n += nf;
}

// Candidate positions before which a safe cell's merge goes over candidate
// cells first (below it, scanning the positions costs fewer round trips).
constexpr int kCellLevelMin = 192;

// Cluster merge (the reference's __raygen__cluster + unite,
// dbscan.cu:279-321,354-394,9-43), cell-parallel: one wave per cell A with
// cores. Its candidates are the stencil positions before A's own first
// position - exactly the points of the stencil cells ordered before A, so each
// unordered pair of cells is tested once, from its later cell - loaded 64 at
// a time, coalesced. A safe cell A (all its points within eps of each other)
// is one union-find element, its representative: a candidate core within eps
// of any core of A joins it. The cores of an unsafe cell are separate
// elements, and the cell's own earlier points are candidates too. A
// candidate's element is its safe cell's representative or its own slot.
//
// A safe cell with many candidates (dense regions) goes over its candidate
// cells first, 64 per round: one safe candidate cell joins A as soon as its
// first core is within eps of a core of A - one test per cell pair instead of
// a scan of the cell's points, which is all the merge needs (the analogue of
// the reference's dense-cell spheres, dbscan.cu:378-393). Only the candidate
// cells not settled that way (unsafe cells, whose cores are separate
// elements, and safe cells whose first core is out of reach) have their
// points scanned.
template <class ST>
__device__ void union_cell(const BatchArgs &a, int s, const CellView &V, const CellHead &h, ST &st, EdgeBuf &eb,
                           uint64_t rw) {
    const int lane = lane_id();
    const int dim = V.dim;
    const float e2 = V.eps2;
    // the header comes screened (cell_head) and the stencil rows (rw, lane r:
    // row r, every lane loaded) came in with the previous cell; the first
    // queries and the first candidate chunk load in one round, then each
    // further chunk is a round of its own
    const bool safeA = h.cf & kCellSafe, denseA = h.cf & kCellDense;
    const int cs = h.cs, nA = h.ce - h.cs;
    const RowWalk w0 = row_walk(lane < V.R ? rw : 0ull, safeA ? cs : cs + nA);
    if (w0.rest == 0ull) return;
    const bool by_cells = safeA && uni(wave_sum(w0.len)) >= kCellLevelMin;
    for (int q0 = 0; q0 < nA; q0 += 64) {
        const int nq = min(64, nA - q0);
        const int pq = cs + q0 + min(lane, nq - 1);
        const int qf = V.flags[pq];
        const float qx = V.sx[pq];
        const float qy = dim > 1 ? V.sy[pq] : 0.f;
        const float qz = dim > 2 ? V.sz[pq] : 0.f;
        const int qslot = V.rep[pq];  // read for the cores of unsafe cells: the point's own slot
        // the point scan's first chunk, in the same round (not for the
        // cell-level pass, which scans other walks)
        RowWalk w1 = w0;
        int end1 = 0;
        const int pos1 = walk_chunk(w1, 0, &end1);
        const int f1 = by_cells ? 0 : V.flags[pos1];
        const float x1 = by_cells ? 0.f : V.sx[pos1];
        const float y1 = by_cells || dim < 2 ? 0.f : V.sy[pos1];
        const float z1 = by_cells || dim < 3 ? 0.f : V.sz[pos1];
        const int r1 = by_cells ? 0 : V.rep[pos1];
        arrived(qx);
        arrived(qy);
        arrived(qz);
        arrived(qslot);
        arrived(x1);
        arrived(y1);
        arrived(z1);
        arrived(r1);
        // a safe cell's element: its representative's slot, which every
        // position of the cell holds in rep - lane 0's query slot, so it comes
        // with the queries instead of a round trip of its own (unsafe cells:
        // each core its own slot, repA unused)
        const int repA = lane_i(qslot, 0);
        const bool qcore = lane < nq && (denseA || (qf & kFlagCore));
        const uint64_t qm = __ballot(qcore);
        if (qm == 0ull) continue;
        // The cores qm of this block against candidate chunk [jb, jb + 64) of
        // a walk whose candidates end at `end`.
        auto test = [&](const int jb, const int end, const int pos, const int f, const float cx, const float cy,
                        const float cz, const int rp) __attribute__((always_inline)) {
                const bool bcore = jb + lane < end && (f & kFlagCore);
                const uint64_t bm = __ballot(bcore);
                if (bm == 0ull) return;
                const int rep = bcore ? rp : 0;
                const unsigned chunk = (unsigned)min(64, end - jb);
                uint64_t m = qm;
                if (safeA) {
                    // lanes holding one element form a run (a safe cell's points are
                    // adjacent); one touching lane settles its whole run
                    const int prev_rep = prev_lane(rep);  // every lane takes part in the moves
                    const int prev_core = prev_lane((int)bcore);
                    const uint64_t starts = __ballot(lane == 0 || prev_rep != rep || !prev_core);
                    const uint64_t after = starts & ~((2ull << lane) - 1ull);  // run starts above this lane
                    const uint64_t upto = after ? (after & (~after + 1ull)) - 1ull : ~0ull;
                    const uint64_t below = starts & ((2ull << lane) - 1ull);
                    const uint64_t from = ~((1ull << (63 - __clzll((long long)below))) - 1ull);
                    const uint64_t run = upto & from;  // this lane's run
                    // touched candidates as one wave mask: the loop carries a scalar,
                    // not a lane predicate
                    uint64_t t = 0;
                    while (m) {
                        const int qi = __ffsll((unsigned long long)m) - 1;
                        m &= m - 1;
                        st.add(kStatMergeTests, chunk);
                        // every lane evaluates the test, no branch per query
                        t |= __ballot(sqd(dim, lane_f(qx, qi), lane_f(qy, qi), lane_f(qz, qi), cx, cy, cz) < e2) & bm;
                        if ((__ballot((t & run) != 0ull) | ~bm) == ~0ull) break;  // every candidate run settled
                    }
                    unite_once_per_rep(a, eb, s, (t >> lane) & 1ull, rep, repA, st);
                } else {
                    while (m) {
                        const int qi = __ffsll((unsigned long long)m) - 1;
                        m &= m - 1;
                        st.add(kStatMergeTests, chunk);
                        const bool hit = bcore & (pos < cs + q0 + qi) &
                                         (sqd(dim, lane_f(qx, qi), lane_f(qy, qi), lane_f(qz, qi), cx, cy, cz) < e2);
                        unite_once_per_rep(a, eb, s, hit, rep, lane_i(qslot, qi), st);
                    }
                }
        };
        // the cores qm of this block against the candidate positions of walk w
        auto scan = [&](RowWalk w) __attribute__((always_inline)) {
            for (int jb = 0;; jb += 64) {
                int end;
                const int pos = walk_chunk(w, jb, &end);
                if (jb >= end) break;
                const int f = V.flags[pos];
                const float cx = V.sx[pos];
                const float cy = dim > 1 ? V.sy[pos] : 0.f;
                const float cz = dim > 2 ? V.sz[pos] : 0.f;
                const int rp = V.rep[pos];  // meaningful for cores only
                arrived(cx);
                arrived(cy);
                arrived(cz);
                arrived(rp);
                test(jb, end, pos, f, cx, cy, cz, rp);
            }
        };
        if (!by_cells) {
            // the first chunk came with the queries
            if (end1 > 0) test(0, end1, pos1, f1, x1, y1, z1, r1);
            for (int jb = 64;; jb += 64) {
                int end;
                const int pos = walk_chunk(w1, jb, &end);
                if (jb >= end) break;
                const int f = V.flags[pos];
                const float cx = V.sx[pos];
                const float cy = dim > 1 ? V.sy[pos] : 0.f;
                const float cz = dim > 2 ? V.sz[pos] : 0.f;
                const int rp = V.rep[pos];
                arrived(cx);
                arrived(cy);
                arrived(cz);
                arrived(rp);
                test(jb, end, pos, f, cx, cy, cz, rp);
            }
            continue;
        }
        // candidate cells: row r (lane r) covers positions [rx, rx + len), all
        // before cs, i.e. the cells pcell[rx] .. pcell[rx + len - 1] (< A)
        const int clo = w0.len > 0 ? V.pcell[w0.rx] : 0;
        const int chi = w0.len > 0 ? V.pcell[w0.rx + w0.len - 1] + 1 : 0;
        RowWalk cw;
        cw.rx = clo;
        cw.len = chi - clo;
        cw.rest = __ballot(cw.len > 0);
        for (int jb = 0;; jb += 64) {
            int endc;
            const int B = walk_chunk(cw, jb, &endc);
            if (jb >= endc) break;
            const bool valid = jb + lane < endc;
            const int fb = V.fcore[B];
            const int cfb = V.cflag[B];
            const int bs = V.cstart[B];
            const int be = V.cstart[B + 1];
            arrived(fb);
            arrived(be);
            const bool has = valid && fb != kNoCore;
            const int fp = has ? fb : cs;  // a valid position either way
            const float bx = V.sx[fp];
            const float by = dim > 1 ? V.sy[fp] : 0.f;
            const float bz = dim > 2 ? V.sz[fp] : 0.f;
            const int bk = V.rep[fp];  // a safe cell's element: its first core's slot
            arrived(bx);
            arrived(by);
            arrived(bz);
            arrived(bk);
            // safe candidate cells: A's cores against each cell's first core
            const uint64_t cand = __ballot(has && (cfb & kCellSafe));
            uint64_t hit = 0;
            for (uint64_t m = qm; m && hit != cand;) {
                const int qi = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                st.add(kStatMergeTests, (unsigned)__popcll(cand & ~hit));
                hit |= __ballot(sqd(dim, lane_f(qx, qi), lane_f(qy, qi), lane_f(qz, qi), bx, by, bz) < e2) & cand;
            }
            // a safe cell's element is its first core's slot; one lane per cell
            unite_once_per_rep(a, eb, s, (hit >> lane) & 1ull, bk, repA, st);
            // the cells with cores not settled: scan their points
            RowWalk pw;
            pw.rx = bs;
            pw.len = (has && !((hit >> lane) & 1ull)) ? be - bs : 0;
            pw.rest = __ballot(pw.len > 0);
            if (pw.rest) scan(pw);
        }
    }
}

template <class ST>
__global__ __launch_bounds__(64 * kCellWaves) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_union(BatchArgs a) {
    __shared__ Edge edges[kCellWaves][kEdgeCap];
    const int total = uni(ldc(a.cell_pre[a.n]));
    const int nw = gridDim.x * kCellWaves;
    ST st;
    EdgeBuf eb{edges[threadIdx.x >> 6], 0};
    const int w = uni(xcd_block(blockIdx.x, gridDim.x, a.xcd_groups & kXcdUnion) * kCellWaves + (threadIdx.x >> 6));
    // stencil rows of a cell, lane r: row r (every lane loads, clamped to the
    // last row; lanes >= R are cleared where the rows are used)
    auto rows_of = [&](const CellHead &x) __attribute__((always_inline)) {
        const StreamDev &S = a.s[x.s];
        const int R = ldc(S.R);
        return gl((const uint64_t *)ldc(S.rows))[(size_t)x.c * R + min(lane_id(), R - 1)];
    };
    for (long long g0 = w; g0 < total; g0 += 64ll * nw) {
        const CellHead mine = cell_head(a, g0 + (long long)lane_id() * nw, total);
        uint64_t m = __ballot(mine.fc != kNoCore);
        if (m == 0ull) continue;
        CellHead h = lane_head(mine, __ffsll((unsigned long long)m) - 1);
        uint64_t rw = rows_of(h);
        while (true) {
            m &= m - 1ull;
            // the next cell's rows are in flight while this cell merges (past
            // the last cell the current one's reload: no branch around the load)
            const CellHead hn = m ? lane_head(mine, __ffsll((unsigned long long)m) - 1) : h;
            const uint64_t rwn = rows_of(hn);
            stats_stream(a, st, h.s);
            union_cell(a, h.s, cell_view(a.s[h.s]), h, st, eb, rw);
            if (m == 0ull) break;
            h = hn;
            rw = rwn;
        }
    }
    edge_flush(a, eb);
    stats_flush(a, st);
}

// Range count, point-parallel form (large batches, where the device is
// saturated and per-query work matters more than coalescing): dense cells are
// core without counting (all pairs within eps); otherwise the LPQ lanes of a
// query count its neighbours (itself included) over its cell's stencil rows,
// each stopping once its own count reaches min_pts. Queries of one cell are
// adjacent and read the same candidates. D is the stream's dimension, so the
// distance test and the loads carry no per-candidate dimension selects.
//
// Flat walks: a lane's rows, a group at a time, walked as one sequence of
// candidates. The group's row words come in one round trip, then kFlat
// candidates per round trip whatever the rows' lengths, so that a row holding
// a cell or two does not cost a round trip of its own. Candidate t of a group
// lies in the last row i with pre[i] <= t (pre: the candidates of the group's
// earlier rows; an empty row shares its pre with the next row, which
// overrides it), at position t + base[i].
constexpr int kFlat = 8;        // candidates per round trip of a flat walk
constexpr int kBorderRows = 8;  // rows per group of the outputs' walk (no early exit)

// One lane's share of a query's stencil: rows first, first + step, ..., in
// flat groups of kRowGroup rows, nearest rows first. The count is checked
// against min_pts once per round trip.
template <int D>
__device__ __forceinline__ int count_walk(const CellView &V, int c, float qx, float qy, float qz, int first,
                                          int step, int &tests) {
    const int minpts = V.min_pts;
    const float e2 = V.eps2;
    int cnt = 0;
    for (int r0 = first; r0 < V.R; r0 += kRowGroup * step) {
        int base[kRowGroup], pre[kRowGroup];
        int T = 0;
#pragma unroll
        for (int i = 0; i < kRowGroup; i++) {
            const uint64_t rw = (r0 + i * step < V.R) ? V.rows[(size_t)c * V.R + r0 + i * step] : 0ull;
            const int rx = (int)(uint32_t)rw, ry = (int)(uint32_t)(rw >> 32);  // rx < ry, or both 0
            pre[i] = T;
            base[i] = rx - T;
            T += ry - rx;
        }
        for (int t0 = 0; t0 < T; t0 += kFlat) {
            float cx[kFlat], cy[kFlat], cz[kFlat];
#pragma unroll
            for (int u = 0; u < kFlat; u++) {
                // clamped to the last candidate: every load reads a real one
                const int t = min(t0 + u, T - 1);
                int b = base[0];
#pragma unroll
                for (int i = 1; i < kRowGroup; i++) b = t >= pre[i] ? base[i] : b;
                const uint32_t o = (uint32_t)(t + b) << 2;
                cx[u] = ldo(V.sx, o);
                cy[u] = D > 1 ? ldo(V.sy, o) : 0.f;
                cz[u] = D > 2 ? ldo(V.sz, o) : 0.f;
            }
            const int nv = T - t0;
            tests += min(nv, kFlat);
#pragma unroll
            for (int u = 0; u < kFlat; u++) cnt += (u < nv) & (sqd(D, qx, qy, qz, cx[u], cy[u], cz[u]) < e2);
            if (cnt >= minpts) return cnt;
        }
    }
    return cnt;
}

template <int LPQ, int D>
__device__ __forceinline__ void count_vblock(const BatchArgs &a, const QueryLane<LPQ> &ql, int s,
                                             const CellView &V) {
    const int p = ql.p;
    const bool valid = p < ql.x.p1;
    const int minpts = V.min_pts;
    int cnt = 0;
    bool dense = false;
    int c = 0;
    // the query's cell and coordinates, one round trip
    const uint32_t po = (uint32_t)(valid ? p : ql.x.p0) << 2;
    const int cv = ldo(V.pcell, po);
    const float qx = ldo(V.sx, po);
    const float qy = D > 1 ? ldo(V.sy, po) : 0.f;
    const float qz = D > 2 ? ldo(V.sz, po) : 0.f;
    arrived(qx);
    arrived(qy);
    arrived(qz);
    if (valid) {
        c = cv;
        dense = V.cflag[c] & kCellDense;
    }
    int tests = 0;
    if (valid && !dense) cnt = count_walk<D>(V, c, qx, qy, qz, ql.sub, LPQ, tests);
    if (a.stats) {
        const unsigned long long t = (unsigned long long)wave_sum(tests);
        const int nd = __popcll(__ballot(valid && dense && ql.sub == 0));
        const int nq = __popcll(__ballot(valid && !dense && ql.sub == 0));
        if (lane_id() == 0) {
            unsigned long long *st = a.stats + (size_t)s * kNumStats;
            atomicAdd(st + kStatCountTests, t);
            if (nd) atomicAdd(st + kStatDensePoints, (unsigned long long)nd);
            if (nq) atomicAdd(st + kStatQueries, (unsigned long long)nq);
        }
    }
    cnt = group_sum<LPQ>(cnt);
    if (valid && ql.sub == 0) V.flags[p] = dense ? (kFlagCore | kFlagDense) : (cnt >= minpts ? kFlagCore : 0);
}

template <int LPQ>
__global__ __launch_bounds__(64) void k_count_points(BatchArgs a) {
    const QueryLane<LPQ> ql = query_lane<LPQ>(a, blockIdx.x);
    const int s = uni(ql.x.s);
    const CellView V = cell_view(a.s[s]);
    if (V.dim == 3) count_vblock<LPQ, 3>(a, ql, s, V);
    else if (V.dim == 2) count_vblock<LPQ, 2>(a, ql, s, V);
    else count_vblock<LPQ, 1>(a, ql, s, V);
}

// Border / noise, point-parallel form (large batches): LPQ lanes per
// non-core point, which takes the minimum root among the cores within eps.
// The kernel also writes the cores' outputs (the root of each core's
// union-find element, found here: no separate flatten pass) and advances each
// stream's ring cursor, as the cell-parallel output kernel does. D is the
// stream's dimension.
//
// One lane's share of a non-core point's stencil in flat groups of
// kBorderRows rows (larger than the count's: no early exit here), without an
// early exit: the minimum root of the cores within eps. A candidate's flags
// and element load only when it is within eps.
template <int D>
__device__ __forceinline__ int border_walk(const CellView &V, const int32_t *parent, int c, float qx, float qy,
                                           float qz, int first, int step, int &tests) {
    const float e2 = V.eps2;
    int best = INT_MAX;
    for (int r0 = first; r0 < V.R; r0 += kBorderRows * step) {
        int base[kBorderRows], pre[kBorderRows];
        int T = 0;
#pragma unroll
        for (int i = 0; i < kBorderRows; i++) {
            const uint64_t rw = (r0 + i * step < V.R) ? V.rows[(size_t)c * V.R + r0 + i * step] : 0ull;
            const int rx = (int)(uint32_t)rw, ry = (int)(uint32_t)(rw >> 32);
            pre[i] = T;
            base[i] = rx - T;
            T += ry - rx;
        }
        for (int t0 = 0; t0 < T; t0 += kFlat) {
            float cx[kFlat], cy[kFlat], cz[kFlat];
            int pos[kFlat];
#pragma unroll
            for (int u = 0; u < kFlat; u++) {
                const int t = min(t0 + u, T - 1);
                int b = base[0];
#pragma unroll
                for (int i = 1; i < kBorderRows; i++) b = t >= pre[i] ? base[i] : b;
                pos[u] = t + b;
                const uint32_t o = (uint32_t)pos[u] << 2;
                cx[u] = ldo(V.sx, o);
                cy[u] = D > 1 ? ldo(V.sy, o) : 0.f;
                cz[u] = D > 2 ? ldo(V.sz, o) : 0.f;
            }
            const int nv = T - t0;
            tests += min(nv, kFlat);
#pragma unroll
            for (int u = 0; u < kFlat; u++)
                if ((u < nv) & (sqd(D, qx, qy, qz, cx[u], cy[u], cz[u]) < e2) &&
                    (ldo(V.flags, (uint32_t)pos[u]) & kFlagCore))
                    best = min(best, uf_root(parent, ldo(V.rep, (uint32_t)pos[u] << 2)));
        }
    }
    return best;
}

template <int LPQ, int D>
__device__ __forceinline__ void border_vblock(const BatchArgs &a, const QueryLane<LPQ> &ql, int vb, int s,
                                              const CellView &V) {
    const int p = ql.p;
    // the query's flags, cell, coordinates and slot, one round trip
    const int pp = p < ql.x.p1 ? p : ql.x.p0;
    const uint32_t po = (uint32_t)pp << 2;
    const int qf = ldo(V.flags, (uint32_t)pp);
    const int c = ldo(V.pcell, po);
    const float qx = ldo(V.sx, po);
    const float qy = D > 1 ? ldo(V.sy, po) : 0.f;
    const float qz = D > 2 ? ldo(V.sz, po) : 0.f;
    const uint64_t qk = ldo(V.skey, po << 1);
    const int qr = ldo(V.rep, po);
    arrived(c);
    arrived(qx);
    arrived(qy);
    arrived(qz);
    arrived(qk);
    arrived(qr);
    const int32_t *parent = ldc(a.s[s].parent);
    const IO io = a.use_io1 ? a.io1 : a.io[s];
    if (vb == ldc(a.pre_chunk[s]) * LPQ && lane_id() == 0) {  // the stream's first block
        StreamMut *m = ldc(a.s[s].mut);
        if (a.stats) a.stats[(size_t)s * kNumStats + kStatCells] += (unsigned long long)m->ncell;
        m->t = m->t + 1;
        m->u = m->u + 1 == ldc(a.s[s].units) ? 0 : m->u + 1;
    }
    if (p < ql.x.p1 && (qf & kFlagCore) && ql.sub == 0) {
        const int slot = (int)((uint32_t)qk & V.smask);
        gl(io.label)[slot] = 0;
        gl(io.cid)[slot] = uf_root(parent, qr);
    }
    const bool mine = p < ql.x.p1 && !(qf & kFlagCore);
    int best = INT_MAX;
    int tests = 0;
    if (mine) best = border_walk<D>(V, parent, c, qx, qy, qz, ql.sub, LPQ, tests);
    if (a.stats) {
        const unsigned long long t = (unsigned long long)wave_sum(tests);
        if (lane_id() == 0 && t) atomicAdd(a.stats + (size_t)s * kNumStats + kStatBorderTests, t);
    }
    best = group_min<LPQ>(best);
    if (mine && ql.sub == 0) {
        const int slot = (int)((uint32_t)qk & V.smask);
        gl(io.label)[slot] = best != INT_MAX ? 1 : 2;
        gl(io.cid)[slot] = best != INT_MAX ? best : slot;
    }
}

template <int LPQ>
__global__ __launch_bounds__(64) void k_border_points(BatchArgs a) {
    const int vb = xcd_block(blockIdx.x, gridDim.x, a.xcd_groups & kXcdBorderPoints);
    const QueryLane<LPQ> ql = query_lane<LPQ>(a, vb);
    const int s = uni(ql.x.s);
    const CellView V = cell_view(a.s[s]);
    if (V.dim == 3) border_vblock<LPQ, 3>(a, ql, vb, s, V);
    else if (V.dim == 2) border_vblock<LPQ, 2>(a, ql, vb, s, V);
    else border_vblock<LPQ, 1>(a, ql, vb, s, V);
}

// Border / noise, point-parallel with regrouping (large batches): one wave per
// 64-position chunk, one lane per position. A core writes the root of its
// element at once; the chunk's non-core points are then regrouped 16 at a time
// onto four lanes each (a point's rank among the chunk's non-cores picks its
// group, its query passed through LDS), which split its stencil rows as in
// border_vblock. A core needs one lane and a non-core four, so a wave covers
// four times the positions of the four-lanes-per-position form with about as
// many round trips.
struct BorderQ {
    int c;
    float x, y, z;
    int slot;
};

template <int D>
__device__ __forceinline__ void border_chunk(const BatchArgs &a, const ChunkCtx &x, int g, int s, const CellView &V,
                                             BorderQ *lq) {
    const int lane = lane_id();
    const int p = x.p0 + lane;
    const bool valid = p < x.p1;
    // the query's flags, cell, coordinates and slot, one round trip
    const int pp = valid ? p : x.p0;
    const uint32_t po = (uint32_t)pp << 2;
    const int qf = ldo(V.flags, (uint32_t)pp);
    const int c = ldo(V.pcell, po);
    const float qx = ldo(V.sx, po);
    const float qy = D > 1 ? ldo(V.sy, po) : 0.f;
    const float qz = D > 2 ? ldo(V.sz, po) : 0.f;
    const uint64_t qk = ldo(V.skey, po << 1);
    const int qr = ldo(V.rep, po);
    arrived(c);
    arrived(qx);
    arrived(qy);
    arrived(qz);
    arrived(qk);
    arrived(qr);
    const int32_t *parent = ldc(a.s[s].parent);
    const IO io = a.use_io1 ? a.io1 : a.io[s];
    if (g == ldc(a.pre_chunk[s]) && lane == 0) {  // the stream's first block
        StreamMut *m = ldc(a.s[s].mut);
        if (a.stats) a.stats[(size_t)s * kNumStats + kStatCells] += (unsigned long long)m->ncell;
        m->t = m->t + 1;
        m->u = m->u + 1 == ldc(a.s[s].units) ? 0 : m->u + 1;
    }
    const int slot = (int)((uint32_t)qk & V.smask);
    const bool core = valid && (qf & kFlagCore);
    if (core) {
        gl(io.label)[slot] = 0;
        gl(io.cid)[slot] = uf_root(parent, qr);
    }
    const uint64_t nm = __ballot(valid && !core);
    if (valid && !core) lq[__popcll(nm & ((1ull << lane) - 1ull))] = BorderQ{c, qx, qy, qz, slot};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nn = __popcll(nm);
    const int grp = lane >> 2, sub = lane & 3;
    int tests = 0;
    for (int b0 = 0; b0 < nn; b0 += 16) {
        const int i = b0 + grp;
        const bool have = i < nn;
        const BorderQ q = lq[have ? i : 0];
        int best = INT_MAX;
        if (have) best = border_walk<D>(V, parent, q.c, q.x, q.y, q.z, sub, 4, tests);
        best = group_min<4>(best);
        if (have && sub == 0) {
            gl(io.label)[q.slot] = best != INT_MAX ? 1 : 2;
            gl(io.cid)[q.slot] = best != INT_MAX ? best : q.slot;
        }
    }
    if (a.stats) {
        const unsigned long long t = (unsigned long long)wave_sum(tests);
        if (lane == 0 && t) atomicAdd(a.stats + (size_t)s * kNumStats + kStatBorderTests, t);
    }
}

__global__ __launch_bounds__(64) void k_border_regroup(BatchArgs a) {
    __shared__ BorderQ lq[64];
    const int g = xcd_block(blockIdx.x, gridDim.x, a.xcd_groups & kXcdBorderPoints);
    const ChunkCtx x = chunk_of(a, g);
    const int s = uni(x.s);
    const CellView V = cell_view(a.s[s]);
    if (V.dim == 3) border_chunk<3>(a, x, g, s, V, lq);
    else if (V.dim == 2) border_chunk<2>(a, x, g, s, V, lq);
    else border_chunk<1>(a, x, g, s, V, lq);
}

// First core of every cell and union-find element of every position, in one
// pass after the point-parallel count: one thread per position, a wave on 64
// consecutive positions. Positions within a cell are in slot order, so a
// cell's first core is its minimum-slot core: the lowest core flag of the
// wave's ballot within the cell's lanes when the cell starts in this wave,
// otherwise (it starts in an earlier wave, or holds no core in this one and
// runs past it) a scan of its flags. The thread at the cell's first position
// records fcore.
__global__ __launch_bounds__(kTileThreads) void k_fcore_rep(BatchArgs a) {
    const int g = blockIdx.x;
    const int s = find_stream_c(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int W = ldc(S.W);
    const int lane = lane_id();
    const int p0 = (g - ldc(a.pre_wtile[s])) * kWTile;
#pragma unroll
    for (int r = 0; r < kWTile / kTileThreads; r++) {
        const int p = p0 + r * kTileThreads + threadIdx.x;
        const int wb = p - lane;  // the wave's first position
        // a wave past the window's end (the last tile's tail) has no position;
        // in the others the lanes past the end read the wave's first position
        if (wb >= W) break;
        const bool v = p < W;
        const int pp = v ? p : wb;
        const int c = S.pcell[pp];
        const int f = S.flags[pp];
        const int cf = S.cflag[c];
        const int cs = S.cstart[c], ce = S.cstart[c + 1];
        const uint64_t corem = __ballot(v && (f & kFlagCore));
        int fc = kNoCore;
        if (cf & kCellDense) {
            fc = cs;
        } else {
            int from = cs;  // positions still to scan: [from, ce)
            if (cs >= wb) {
                const int lo = cs - wb, hi = min(ce - wb, 64);
                const uint64_t m = corem & (~0ull << lo) & (hi >= 64 ? ~0ull : ((1ull << hi) - 1ull));
                if (m) fc = wb + __ffsll((unsigned long long)m) - 1;
                from = m ? ce : wb + 64;
            }
            for (int i = from; i < ce; i++)
                if (S.flags[i] & kFlagCore) {
                    fc = i;
                    break;
                }
        }
        if (v && p == cs) S.fcore[c] = fc;
        if (v) S.rep[p] = slot_of(S, (cf & kCellSafe) && fc != kNoCore ? fc : p);
    }
}

// Cluster merge, cell level, for safe cells: all cores of a safe cell belong
// to one cluster (its points are within eps of each other), so two safe cells
// join as soon as one core pair is within eps - the analogue of the
// reference's dense-cell spheres, where a ray unites with the first point of
// the cell within eps and stops (dbscan.cu:378-393). Only stencil cells
// B < A are tested, so each unordered pair of cells is tested once.
//
// A wave serves kUGCells query cells per round. Each query cell's cores are
// staged in LDS; the stencil rows of all the round's cells are spread over the
// lanes, their candidate cells are concatenated into one list, and each lane
// tests one (query cell, candidate) pair at a time. Touching pairs go to an
// LDS edge list that the lanes unite together at the end of the round, so a
// round costs a few dependent memory round trips for all its cells. Pairs
// needing more than kUGWork point tests are queued and tested by the whole wave;
// query cells with more than kUGCap cores are served by the whole wave
// (union_cell_wave).
constexpr int kUG = 64 / kUGCells;            // lanes staging one query cell
constexpr int kUGCap = 64;                    // cores of a query cell staged in LDS
constexpr int kUGWork = 256;                  // larger pair tests (cores x points): whole wave
constexpr int kUGQueue = 128;                 // queued (query cell, candidate) pairs
constexpr int kUGRows = kUGCells * kMaxRows;  // stencil rows of one round
constexpr int kUGEdges = 256;                 // buffered touching pairs
constexpr int kUnionA = kUGCells * kUGCap;    // LDS capacity for a large query cell
constexpr int kUnionB = 8;                    // register capacity for a candidate cell

struct UnionLds {
    float x[kUGCells * kUGCap], y[kUGCells * kUGCap], z[kUGCells * kUGCap];
    uint8_t lc[kUnionA];       // whole-wave path: core flags of the query cell
    int ncore[kUGCells];       // cores staged for the group's cell
    int rep[kUGCells];         // representative slot of the group's cell
    int sidx[kUGCells];        // stream of the group's cell
    int cell[kUGCells];        // the group's cell (stream-local)
    int mode[kUGCells];        // 0 idle, 1 staged, 2 whole-wave path
    int rowbase[kUGCells + 1]; // first row of each group in the round's row list
    int rowpre[kUGRows + 1];   // exclusive prefix of candidate counts
    int rowclo[kUGRows];       // first candidate cell of the row
    int2 q[kUGQueue];          // (group, candidate cell)
    int2 edge[kUGEdges];       // (group, representative slot of the candidate)
    int nq;
};

__device__ __forceinline__ void flush_edges(const BatchArgs &a, UnionLds &L, int &ne) {
    for (int i = lane_id(); i < ne; i += 64) {
        const int2 e = L.edge[i];
        uf_unite(a.s[L.sidx[e.x]].parent, L.rep[e.x], e.y);
    }
    ne = 0;
}

// Appends the touching pairs of the wave (lanes with want set) to the edge
// list, in lane order; all lanes call it together.
__device__ __forceinline__ void push_edges(const BatchArgs &a, UnionLds &L, int &ne, bool want, int g, int v) {
    const uint64_t m = __ballot(want);
    if (m == 0ull) return;
    const int n = __popcll(m);
    if (ne + n > kUGEdges) {
        __syncthreads();
        flush_edges(a, L, ne);
        __syncthreads();
    }
    if (want) L.edge[ne + __popcll(m & ((1ull << lane_id()) - 1ull))] = make_int2(g, v);
    ne += n;
}

// Whole-wave test of one (query cell, large candidate cell B) pair: lanes
// over B's points, a loop over the query cell's staged cores, stop at the
// first pair within eps.
__device__ void union_pair_wave(const StreamDev &S, const float *qx, const float *qy, const float *qz, int ncore,
                                int rep, int B) {
    const int lane = lane_id(), dim = S.dim;
    const float e2 = S.eps2;
    const uint8_t cfB = S.cflag[B];
    const int fcB = S.fcore[B];
    const int csB = S.cstart[B], ceB = S.cstart[B + 1];
    bool touch = false;
    for (int b0 = csB; b0 < ceB && !touch; b0 += 64) {
        const int pb = b0 + lane;
        bool vb = pb < ceB;
        float bx = 0.f, by = 0.f, bz = 0.f;
        if (vb) {
            vb = (cfB & kCellDense) || (S.flags[pb] & kFlagCore);
            bx = S.sx[pb];
            if (dim > 1) by = S.sy[pb];
            if (dim > 2) bz = S.sz[pb];
        }
        for (int i = 0; i < ncore; i++) {
            if (__ballot(vb && sqd(dim, qx[i], qy[i], qz[i], bx, by, bz) < e2) != 0ull) {
                touch = true;
                break;
            }
        }
    }
    if (touch && lane == 0) uf_unite(S.parent, rep, slot_of(S, fcB));
}

// Whole-wave merge of one large safe query cell A with its stencil cells
// B < A: lanes take candidate cells; A's points sit in LDS when they fit,
// and a small B is held in the lane's registers.
__device__ void union_cell_wave(const StreamDev &S, int A, UnionLds &L) {
    const int lane = lane_id();
    const uint8_t cfA = S.cflag[A];
    const int fcA = S.fcore[A];
    const int dim = S.dim;
    const float e2 = S.eps2;
    float *lx = L.x, *ly = L.y, *lz = L.z;
    // stencil cells B < A, row by row: [pcell[rb], min(pcell[re-1], A-1)]
    int clo = 0, cnt = 0;
    if (lane < S.R) {
        const int2 r = S.rows[(size_t)A * S.R + lane];
        if (r.x < r.y) {
            clo = S.pcell[r.x];
            const int chi = min(S.pcell[r.y - 1], A - 1);
            cnt = max(0, chi - clo + 1);
        }
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const int pre = incl - cnt;
    const int TC = __shfl(incl, 63);
    if (TC == 0) return;
    const int csA = S.cstart[A], ceA = S.cstart[A + 1], nA = ceA - csA;
    const bool inl = nA <= kUnionA;
    const int sA = slot_of(S, fcA);
    __syncthreads();
    if (inl) {
        for (int i = lane; i < nA; i += 64) {
            lx[i] = S.sx[csA + i];
            ly[i] = dim > 1 ? S.sy[csA + i] : 0.f;
            lz[i] = dim > 2 ? S.sz[csA + i] : 0.f;
            L.lc[i] = (cfA & kCellDense) ? (uint8_t)1 : (uint8_t)(S.flags[csA + i] & kFlagCore);
        }
    }
    __syncthreads();
    for (int jb = 0; jb < TC; jb += 64) {
        const int j = jb + lane;
        const int jj = j < TC ? j : TC - 1;
        const int r = row_of(pre, S.R, jj);
        const int B = __shfl(clo, r) + (jj - __shfl(pre, r));
        bool big = false;
        bool touch = false;
        int fcB = kNoCore;
        if (j < TC) {
            const uint8_t cfB = S.cflag[B];
            fcB = S.fcore[B];
            if ((cfB & kCellSafe) && fcB != kNoCore) {
                const int csB = S.cstart[B], nB = S.cstart[B + 1] - csB;
                if (inl && nB <= kUnionB) {
                    float bx[kUnionB], by[kUnionB], bz[kUnionB];
                    bool bc[kUnionB];
#pragma unroll
                    for (int m = 0; m < kUnionB; m++) {
                        bc[m] = false;
                        bx[m] = by[m] = bz[m] = 0.f;
                        if (m < nB) {
                            bc[m] = (cfB & kCellDense) || (S.flags[csB + m] & kFlagCore);
                            bx[m] = S.sx[csB + m];
                            if (dim > 1) by[m] = S.sy[csB + m];
                            if (dim > 2) bz[m] = S.sz[csB + m];
                        }
                    }
                    for (int i = 0; i < nA && !touch; i++) {
                        if (!L.lc[i]) continue;
                        const float qx = lx[i], qy = ly[i], qz = lz[i];
#pragma unroll
                        for (int m = 0; m < kUnionB; m++)
                            touch = touch || (bc[m] && sqd(dim, qx, qy, qz, bx[m], by[m], bz[m]) < e2);
                    }
                } else {
                    big = true;
                }
            }
        }
        if (touch) uf_unite(S.parent, sA, slot_of(S, fcB));
        // large candidate cells of this batch: one at a time, lanes over B's
        // points, a loop over A's cores, stop at the first pair within eps
        uint64_t bigm = __ballot(big);
        while (bigm) {
            const int l = __ffsll((unsigned long long)bigm) - 1;
            bigm &= bigm - 1;
            const int Bl = lane_i(B, l);
            const uint8_t cfB = S.cflag[Bl];
            const int fcBl = S.fcore[Bl];
            const int csB = S.cstart[Bl], ceB = S.cstart[Bl + 1];
            bool t = false;
            for (int b0 = csB; b0 < ceB && !t; b0 += 64) {
                const int pb = b0 + lane;
                bool vb = pb < ceB;
                float bx = 0.f, by = 0.f, bz = 0.f;
                if (vb) {
                    vb = (cfB & kCellDense) || (S.flags[pb] & kFlagCore);
                    bx = S.sx[pb];
                    if (dim > 1) by = S.sy[pb];
                    if (dim > 2) bz = S.sz[pb];
                }
                for (int i = 0; i < nA; i++) {
                    float qx, qy, qz;
                    bool qc;
                    if (inl) {
                        qc = L.lc[i];
                        qx = lx[i];
                        qy = ly[i];
                        qz = lz[i];
                    } else {
                        qc = (cfA & kCellDense) || (S.flags[csA + i] & kFlagCore);
                        qx = S.sx[csA + i];
                        qy = dim > 1 ? S.sy[csA + i] : 0.f;
                        qz = dim > 2 ? S.sz[csA + i] : 0.f;
                    }
                    if (!qc) continue;
                    if (__ballot(vb && sqd(dim, qx, qy, qz, bx, by, bz) < e2) != 0ull) {
                        t = true;
                        break;
                    }
                }
            }
            if (t && lane == 0) uf_unite(S.parent, sA, slot_of(S, fcBl));
        }
    }
}

__global__ __launch_bounds__(64) void k_union_cells(BatchArgs a) {
    __shared__ UnionLds L;
    const int lane = lane_id(), grp = lane / kUG, sub = lane % kUG;
    const uint64_t gmask = ((1ull << kUG) - 1ull) << (grp * kUG);
    const int total = ldc(a.cell_pre[a.n]);
    // Round r of wave b serves cells b + (r * kUGCells + g) * gridDim.x: the
    // cells of a round are far apart, so dense regions (long pair tests) are
    // spread over many waves instead of falling to a few.
    const int G = gridDim.x;
    for (int base = 0; base + (int)blockIdx.x < total; base += G * kUGCells) {
        // 1. stage each group's query cell: its cores, compacted, in LDS
        const int gc = (int)blockIdx.x + base + grp * G;
        int mode = 0, ncore = 0, A = 0, sidx = 0, rep = 0, R = 0;
        if (gc < total) {
            const CellCtx x = cell_of(a, gc);
            sidx = x.s;
            A = x.c;
            const StreamDev &S = a.s[sidx];
            const uint8_t cfA = S.cflag[A];
            const int fcA = S.fcore[A];
            if ((cfA & kCellSafe) && fcA != kNoCore) {
                const int csA = S.cstart[A], nA = S.cstart[A + 1] - csA;
                const bool denseA = cfA & kCellDense;
                const int dim = S.dim;
                float *gx = L.x + grp * kUGCap, *gy = L.y + grp * kUGCap, *gz = L.z + grp * kUGCap;
                rep = slot_of(S, fcA);
                for (int i0 = 0; i0 < nA; i0 += kUG) {
                    const int i = i0 + sub;
                    bool c = false;
                    float px = 0.f, py = 0.f, pz = 0.f;
                    if (i < nA) {
                        c = denseA || (S.flags[csA + i] & kFlagCore);
                        px = S.sx[csA + i];
                        if (dim > 1) py = S.sy[csA + i];
                        if (dim > 2) pz = S.sz[csA + i];
                    }
                    const uint64_t m = __ballot(c) & gmask;
                    const int pos = ncore + __popcll(m & ((1ull << lane) - 1ull));
                    if (c && pos < kUGCap) {
                        gx[pos] = px;
                        gy[pos] = py;
                        gz[pos] = pz;
                    }
                    ncore += __popcll(m);
                }
                mode = ncore <= kUGCap ? 1 : 2;
                R = mode == 1 ? S.R : 0;
            }
        }
        if (sub == 0) {
            L.mode[grp] = mode;
            L.ncore[grp] = ncore;
            L.rep[grp] = rep;
            L.sidx[grp] = sidx;
            L.cell[grp] = A;
        }
        // rows of group g occupy [rbase[g], rbase[g+1]) of the round's row list
        int rbase[kUGCells + 1];
        rbase[0] = 0;
#pragma unroll
        for (int g = 0; g < kUGCells; g++) rbase[g + 1] = rbase[g] + __shfl(R, g * kUG);
        if (lane <= kUGCells) {
            int v = 0;
#pragma unroll
            for (int g = 0; g <= kUGCells; g++) v = lane == g ? rbase[g] : v;
            L.rowbase[lane] = v;
        }
        const int NR = rbase[kUGCells];
        __syncthreads();
        // 2. rows of all groups over the lanes: candidate cells B < A per row
        int run = 0;
        for (int t0 = 0; t0 < NR; t0 += 64) {
            const int t = t0 + lane;
            int cnt = 0, clo = 0;
            if (t < NR) {
                int g = 0;
#pragma unroll
                for (int h = 1; h < kUGCells; h++) g += t >= rbase[h];
                const StreamDev &S = a.s[L.sidx[g]];
                const int Ag = L.cell[g];
                const int2 rr = S.rows[(size_t)Ag * S.R + (t - L.rowbase[g])];
                if (rr.x < rr.y) {
                    clo = S.pcell[rr.x];
                    const int chi = min(S.pcell[rr.y - 1], Ag - 1);
                    cnt = max(0, chi - clo + 1);
                }
                L.rowclo[t] = clo;
            }
            int incl = cnt;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            if (t < NR) L.rowpre[t] = run + incl - cnt;
            run += __shfl(incl, 63);
        }
        const int TC = run;
        if (lane == 0) {
            L.rowpre[NR] = TC;
            L.nq = 0;
        }
        __syncthreads();
        // 3. (query cell, candidate) pairs over the lanes
        int ne = 0;
        for (int jb = 0; jb < TC; jb += 64) {
            const int j = jb + lane;
            bool touch = false;
            int g = 0, vrep = 0;
            if (j < TC) {
                int lo = 0, hi = NR;  // last row t with rowpre[t] <= j
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (L.rowpre[mid] <= j) lo = mid;
                    else hi = mid;
                }
                const int t = lo;
#pragma unroll
                for (int h = 1; h < kUGCells; h++) g += t >= rbase[h];
                const int B = L.rowclo[t] + (j - L.rowpre[t]);
                const StreamDev &S = a.s[L.sidx[g]];
                const uint8_t cfB = S.cflag[B];
                const int fcB = S.fcore[B];
                if ((cfB & kCellSafe) && fcB != kNoCore) {
                    const int csB = S.cstart[B], nB = S.cstart[B + 1] - csB;
                    bool queued = false;
                    if (L.ncore[g] * nB > kUGWork) {
                        const int e = atomicAdd(&L.nq, 1);
                        if (e < kUGQueue) {
                            L.q[e] = make_int2(g, B);
                            queued = true;
                        }
                    }
                    if (!queued) {
                        const int dim = S.dim, nc = L.ncore[g];
                        const float e2 = S.eps2;
                        const float *qx = L.x + g * kUGCap, *qy = L.y + g * kUGCap, *qz = L.z + g * kUGCap;
                        const bool denseB = cfB & kCellDense;
                        for (int b0 = 0; b0 < nB && !touch; b0 += 4) {
                            float bx[4], by[4], bz[4];
                            bool bc[4];
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                const int jj = csB + min(b0 + u, nB - 1);
                                bc[u] = (b0 + u < nB) && (denseB || (S.flags[jj] & kFlagCore));
                                bx[u] = S.sx[jj];
                                by[u] = dim > 1 ? S.sy[jj] : 0.f;
                                bz[u] = dim > 2 ? S.sz[jj] : 0.f;
                            }
                            for (int i = 0; i < nc && !touch; i++) {
                                const float ax = qx[i], ay = qy[i], az = qz[i];
#pragma unroll
                                for (int u = 0; u < 4; u++)
                                    touch = touch || (bc[u] && sqd(dim, ax, ay, az, bx[u], by[u], bz[u]) < e2);
                            }
                        }
                        if (touch) vrep = slot_of(S, fcB);
                    }
                }
            }
            push_edges(a, L, ne, touch, g, vrep);
        }
        __syncthreads();
        flush_edges(a, L, ne);
        __syncthreads();
        // 4. queued large candidate cells, then large query cells: whole wave
        const int nq = min(L.nq, kUGQueue);
        for (int e = 0; e < nq; e++) {
            const int2 it = L.q[e];
            const int g = it.x;
            union_pair_wave(a.s[L.sidx[g]], L.x + g * kUGCap, L.y + g * kUGCap, L.z + g * kUGCap, L.ncore[g],
                            L.rep[g], it.y);
        }
        for (int g = 0; g < kUGCells; g++) {
            if (L.mode[g] != 2) continue;
            union_cell_wave(a.s[L.sidx[g]], L.cell[g], L);
        }
        __syncthreads();  // the next round overwrites the LDS
    }
}

// Cluster merge, point level, for the cores of unsafe cells (grid border, or
// grids whose cells are not dense-safe): each unites with every core within
// eps (the reference's core branch of __intersection__hybrid_radius_sphere,
// dbscan.cu:364-369), through the representatives of safe cells.
__device__ __forceinline__ void union_points_chunk(const BatchArgs &a, int g) {
    const ChunkCtx x = chunk_of(a, g);
    const StreamDev &S = a.s[x.s];
    const int p = x.p0 + lane_id();
    if (p >= x.p1) return;
    if (!(S.flags[p] & kFlagCore)) return;
    const int c = S.pcell[p];
    if (S.cflag[c] & kCellSafe) return;  // k_union_cells
    const int dim = S.dim;
    const float e2 = S.eps2;
    const float qx = S.sx[p];
    const float qy = dim > 1 ? S.sy[p] : 0.f;
    const float qz = dim > 2 ? S.sz[p] : 0.f;
    const int sp = slot_of(S, p);
    walk_candidates(cell_view(S), c, [&](int j, float cx, float cy, float cz) {
        if (j != p && sqd(dim, qx, qy, qz, cx, cy, cz) < e2 && (S.flags[j] & kFlagCore))
            uf_unite(S.parent, sp, rep_slot(S, j));
        return false;
    });
}

__global__ __launch_bounds__(64) void k_union_points(BatchArgs a) { union_points_chunk(a, blockIdx.x); }

// Plain device copy of 16-byte words: a kernel of known HBM traffic (reads and
// writes `n16` * 16 bytes) used to scale the profiler's byte counters.
__global__ __launch_bounds__(kTileThreads) void k_copy16(const int4 *__restrict__ src, int4 *__restrict__ dst,
                                                         size_t n16) {
    for (size_t i = (size_t)blockIdx.x * kTileThreads + threadIdx.x; i < n16; i += (size_t)gridDim.x * kTileThreads)
        dst[i] = src[i];
}

hipError_t launch_copy16(const void *src, void *dst, size_t n16, hipStream_t st) {
    if (n16 == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n16 + kTileThreads - 1) / kTileThreads, 65536);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(kTileThreads), 0, st, (const int4 *)src, (int4 *)dst,
                       n16);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers

#define RTDBS_LAUNCH(kernel, grid, block, ...)                        \
    do {                                                              \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, __VA_ARGS__); \
        hipError_t e_ = hipGetLastError();                            \
        if (e_ != hipSuccess) return e_;                              \
    } while (0)

static inline hipError_t mark(hipEvent_t *ev, uint32_t mask, int idx, hipStream_t st) {
    if (!ev || !((mask >> (idx / 2)) & 1u)) return hipSuccess;
    return hipEventRecord(ev[idx], st);
}

// Launches `kernel`; with events, they take the dispatch's own timestamps.
template <class K, class... Args>
static hipError_t launch_timed(K kernel, int grid, int block, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                               Args... args) {
    if (e0) hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, args...);
    return hipGetLastError();
}

hipError_t launch_slide(const BatchArgs &a, const LaunchShape &sh, hipStream_t st, hipEvent_t *ev,
                        uint32_t phase_mask) {
    hipError_t e;
#define PH_BEGIN(ph) if ((e = mark(ev, phase_mask, 2 * (ph), st)) != hipSuccess) return e
#define PH_END(ph) if ((e = mark(ev, phase_mask, 2 * (ph) + 1, st)) != hipSuccess) return e
    PH_BEGIN(kPhStrideSort);
    RTDBS_LAUNCH(k_sort_expire, sh.stiles * kSortParts + sh.wtiles, kTileThreads, a);
    PH_END(kPhStrideSort);
    PH_BEGIN(kPhMerge);
    if (sh.rank_split) {
        const int rb = sh.stiles * (kTile / 64);
        RTDBS_LAUNCH(k_rank<4>, rb + sh.wtiles, kTileThreads, a, rb);
    } else {
        const int rb = sh.stiles * (kTile / kTileThreads);
        RTDBS_LAUNCH(k_rank<1>, rb + sh.wtiles, kTileThreads, a, rb);
    }
    PH_END(kPhMerge);
    PH_BEGIN(kPhMergeFinal);
    RTDBS_LAUNCH(k_merge_final, sh.wtiles, kTileThreads, a);
    PH_END(kPhMergeFinal);
    PH_BEGIN(kPhCells);
    RTDBS_LAUNCH(k_cells, sh.wtiles, kTileThreads, a);
    PH_END(kPhCells);
    PH_BEGIN(kPhRows);
    RTDBS_LAUNCH(k_rows, sh.rows_blocks, kTileThreads, a);
    PH_END(kPhRows);
    {
        // The range-count kernel is timed by its own dispatch: the events take
        // the kernel's begin and end timestamps, with no marker packets of
        // their own between it and its neighbours.
        const bool timed = ev && ((phase_mask >> kPhCount) & 1u);
        const hipEvent_t e0 = timed ? ev[2 * kPhCount] : nullptr, e1 = timed ? ev[2 * kPhCount + 1] : nullptr;
        if (sh.count_cell && a.stats)
            e = launch_timed(k_count<WaveStats>, sh.cellpar_blocks, 64 * kCellWaves, st, e0, e1, a);
        else if (sh.count_cell)
            e = launch_timed(k_count<NoStats>, sh.cellpar_blocks, 64 * kCellWaves, st, e0, e1, a);
        else if (sh.lpq >= 4) e = launch_timed(k_count_points<4>, sh.chunks * 4, 64, st, e0, e1, a);
        else if (sh.lpq == 2) e = launch_timed(k_count_points<2>, sh.chunks * 2, 64, st, e0, e1, a);
        else e = launch_timed(k_count_points<1>, sh.chunks, 64, st, e0, e1, a);
        if (e != hipSuccess) return e;
    }
    PH_BEGIN(kPhUnion);
    // the cell-parallel count wrote every cell's first core and every
    // position's union-find element; after the point-parallel count one pass
    // over the positions does
    if (!sh.count_cell) RTDBS_LAUNCH(k_fcore_rep, sh.wtiles, kTileThreads, a);
    if (sh.union_cells) {
        RTDBS_LAUNCH(k_union_cells, sh.cell_blocks, 64, a);
        RTDBS_LAUNCH(k_union_points, sh.chunks, 64, a);
    } else if (a.stats) {
        RTDBS_LAUNCH(k_union<WaveStats>, sh.cellpar_blocks, 64 * kCellWaves, a);
    } else {
        RTDBS_LAUNCH(k_union<NoStats>, sh.cellpar_blocks, 64 * kCellWaves, a);
    }
    PH_END(kPhUnion);
    PH_BEGIN(kPhBorder);
    if (sh.cellpar && a.stats) RTDBS_LAUNCH(k_border<WaveStats>, sh.cellpar_blocks, 64 * kCellWaves, a);
    else if (sh.cellpar) RTDBS_LAUNCH(k_border<NoStats>, sh.cellpar_blocks, 64 * kCellWaves, a);
    else if (sh.border_regroup) RTDBS_LAUNCH(k_border_regroup, sh.chunks, 64, a);
    else if (sh.lpq_border >= 4) RTDBS_LAUNCH(k_border_points<4>, sh.chunks * 4, 64, a);
    else if (sh.lpq_border == 2) RTDBS_LAUNCH(k_border_points<2>, sh.chunks * 2, 64, a);
    else RTDBS_LAUNCH(k_border_points<1>, sh.chunks, 64, a);
    PH_END(kPhBorder);
#undef PH_BEGIN
#undef PH_END
    return hipSuccess;
}

hipError_t launch_init(const BatchArgs &a, int32_t itiles, int32_t passes, hipStream_t st) {
    RTDBS_LAUNCH(k_init_tiles, itiles, kTileThreads, a);
    for (int p = 0; p < passes; p++) RTDBS_LAUNCH(k_init_merge, itiles, kTileThreads, a, kTile << p, p & 1);
    return hipSuccess;
}

}  // namespace rtdbs
