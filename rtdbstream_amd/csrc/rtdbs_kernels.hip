// HIP kernels of the MI355X streaming DBSCAN slide step (gfx950, wave64).
//
// One slide of every stream in a batch is the sequence
//   stride_tiles -> merge_pass* -> expire_count -> expire_compact ->
//   merge_final -> cells -> rows -> count -> union_sparse -> union_dense ->
//   flatten -> border
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

#include "rtdbs_dev.h"

namespace rtdbs {

const char *const kPhaseNames[kNumPhases] = {"stride_sort", "merge_passes", "expire", "merge_final",
                                             "cells", "rows", "count", "union",
                                             "flatten", "border"};

// ---------------------------------------------------------------------------
// small helpers

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Largest s in [0, n) with pre[s] <= g (pre[0] = 0, g < pre[n]).
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

__device__ __forceinline__ uint32_t slot_mask(const StreamDev &S) {
    return (uint32_t)((1ull << S.slot_bits) - 1ull);
}

// FP32 squared distance in the reference's order, never contracted:
// util.cpp:310-313 / dbscan.cu:48-49 compute O.x*O.x + O.y*O.y + O.z*O.z.
__device__ __forceinline__ float sqd(int dim, float ax, float ay, float az, float bx, float by,
                                     float bz) {
    float dx = __fsub_rn(ax, bx);
    float d = __fmul_rn(dx, dx);
    if (dim > 1) {
        float dy = __fsub_rn(ay, by);
        d = __fadd_rn(d, __fmul_rn(dy, dy));
    }
    if (dim > 2) {
        float dz = __fsub_rn(az, bz);
        d = __fadd_rn(d, __fmul_rn(dz, dz));
    }
    return d;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

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

// Union-find over window positions. Loads bypass the non-coherent vector L1
// (agent-scope relaxed atomics) so a walk sees the hooks other CUs made.
__device__ __forceinline__ int uf_load(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void uf_store(int32_t *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving. Every stored value is an ancestor of the old
// one and roots only change through the CAS in uf_unite, so concurrent
// halving is benign (the reference compresses the same way, dbscan.cu:9-20).
__device__ __forceinline__ int uf_find(int32_t *parent, int x) {
    int p = uf_load(parent + x);
    while (p != x) {
        int gp = uf_load(parent + p);
        if (gp != p) uf_store(parent + x, gp);
        x = p;
        p = gp;
    }
    return x;
}

// Hooks the root with the larger slot under the root with the smaller slot
// (dbscan.cu:22-43 hooks by index; here the order is the ring slot, so the
// final root of every component is its minimum slot).
__device__ __forceinline__ void uf_unite(int32_t *parent, const uint64_t *skey, uint32_t smask, int a,
                                         int b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        uint32_t sa = (uint32_t)skey[a] & smask, sb = (uint32_t)skey[b] & smask;
        if (sa < sb) {
            int t = a;
            a = b;
            b = t;
        }
        int old = atomicCAS(parent + a, a, b);
        if (old == a) return;
    }
}

// ---------------------------------------------------------------------------
// grid

// Cell index per axis: floor((x - lo) / side) evaluated in double and clamped
// into the grid. Clamping keeps the search exact: |clamp(i) - clamp(j)| <=
// |i - j|, so points within reach stay within reach.
__device__ __forceinline__ uint64_t cell_key_of(const StreamDev &S, const float *p) {
    int64_t ix[3] = {0, 0, 0};
    for (int a = 0; a < S.dim; a++) {
        double v = ((double)p[a] - S.lo[a]) * S.inv_side;
        v = floor(v);
        if (!(v >= 0.0)) v = 0.0;  // negative or NaN
        const double mx = (double)(S.n[a] - 1);
        if (v > mx) v = mx;
        ix[a] = (int64_t)v;
    }
    return (uint64_t)((ix[0] * S.n[1] + ix[1]) * S.n[2] + ix[2]);
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// Open-addressing cell hash; an entry whose epoch differs from the current
// one counts as empty, so the table never needs clearing between slides.
__device__ __forceinline__ void hash_insert(const StreamDev &S, uint64_t ck, int c, uint32_t epoch) {
    const uint64_t tag = ((uint64_t)epoch << kKeyBits) | ck;
    uint32_t h = (uint32_t)mix64(ck) & S.hmask;
    while (true) {
        uint64_t cur = __hip_atomic_load(S.htab + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)(cur >> kKeyBits) != epoch) {
            uint64_t prev = atomicCAS((unsigned long long *)(S.htab + h), (unsigned long long)cur,
                                      (unsigned long long)tag);
            if (prev == cur) {
                S.hval[h] = c;
                return;
            }
            continue;
        }
        h = (h + 1) & S.hmask;
    }
}

__device__ __forceinline__ int hash_find(const StreamDev &S, uint64_t ck, uint32_t epoch) {
    const uint64_t tag = ((uint64_t)epoch << kKeyBits) | ck;
    uint32_t h = (uint32_t)mix64(ck) & S.hmask;
    while (true) {
        uint64_t cur = S.htab[h];
        if (cur == tag) return S.hval[h];
        if ((uint32_t)(cur >> kKeyBits) != epoch) return -1;
        h = (h + 1) & S.hmask;
    }
}

// ---------------------------------------------------------------------------
// sorting: LDS bitonic tile sort + merge-path merges

__device__ __forceinline__ void bitonic_sort_tile(uint64_t *sk) {
    for (int k = 2; k <= kTile; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = threadIdx.x; q < kTile / 2; q += kTileThreads) {
                const int i = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                const int ixj = i + j;
                const uint64_t a = sk[i], b = sk[ixj];
                const bool up = (i & k) == 0;
                if ((a > b) == up) {
                    sk[i] = b;
                    sk[ixj] = a;
                }
            }
            __syncthreads();
        }
    }
}

// Cell keys of up to kTile points of a stride (or of the first window),
// sorted by (cell, slot) in LDS and written with their coordinates.
template <bool INIT>
__global__ __launch_bounds__(kTileThreads) void k_stride_tiles(BatchArgs a) {
    __shared__ uint64_t sk[kTile];
    const int g = blockIdx.x;
    if (g == 0 && threadIdx.x < kNumCounters) a.counters[threadIdx.x] = 0;
    const int32_t *pre = INIT ? a.pre_wtile : a.pre_stile;
    const int s = find_stream(pre, a.n, g);
    const StreamDev &S = a.s[s];
    const IO io = a.use_io1 ? a.io1 : a.io[s];
    const int len = INIT ? S.W : S.S;
    const int tile0 = (g - pre[s]) * kTile;
    const int slot0 = INIT ? 0 : S.mut->u * S.S;
    const int dim = S.dim;
    for (int i = threadIdx.x; i < kTile; i += kTileThreads) {
        const int gi = tile0 + i;
        uint64_t key = ~0ull;
        if (gi < len)
            key = (cell_key_of(S, io.in + (size_t)gi * dim) << S.slot_bits) | (uint64_t)(slot0 + gi);
        sk[i] = key;
    }
    __syncthreads();
    bitonic_sort_tile(sk);
    const int nvalid = min(kTile, len - tile0);
    uint64_t *dk = INIT ? S.akey : S.bkey0;
    float *dx = INIT ? S.ax : S.bx0;
    float *dy = INIT ? S.ay : S.by0;
    float *dz = INIT ? S.az : S.bz0;
    const uint32_t m = slot_mask(S);
    for (int i = threadIdx.x; i < nvalid; i += kTileThreads) {
        const uint64_t key = sk[i];
        const int gi = (int)((uint32_t)key & m) - slot0;
        const float *p = io.in + (size_t)gi * dim;
        dk[tile0 + i] = key;
        dx[tile0 + i] = p[0];
        if (dim > 1) dy[tile0 + i] = p[1];
        if (dim > 2) dz[tile0 + i] = p[2];
    }
}

// Number of elements of A among the first d of merge(A, B); keys are unique.
__device__ __forceinline__ int merge_path(const uint64_t *A, int la, const uint64_t *B, int lb, int d) {
    int lo = max(0, d - lb), hi = min(d, la);
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid] < B[d - 1 - mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

struct MergeLds {
    uint64_t in[kTile];
    uint64_t out[kTile];
    int32_t src[kTile];  // >= 0: index into A; < 0: ~index into B
    int32_t sp[2];
    uint64_t pred;
};

// Merges outputs [d0, d1) of merge(A, B) into m.out / m.src (block-wide).
__device__ __forceinline__ void merge_tile(const uint64_t *A, int la, const uint64_t *B, int lb, int d0,
                                           int d1, MergeLds &m, int *a0_out, int *b0_out) {
    const int tid = threadIdx.x;
    if (tid < 2) m.sp[tid] = merge_path(A, la, B, lb, tid ? d1 : d0);
    __syncthreads();
    const int a0 = m.sp[0], a1 = m.sp[1];
    const int b0 = d0 - a0, b1 = d1 - a1;
    const int na = a1 - a0, nb = b1 - b0, n = na + nb;
    for (int i = tid; i < na; i += kTileThreads) m.in[i] = A[a0 + i];
    for (int i = tid; i < nb; i += kTileThreads) m.in[na + i] = B[b0 + i];
    __syncthreads();
    constexpr int per = kTile / kTileThreads;
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

// One pass of the bottom-up merge sort: merges runs of length L pairwise.
// INIT: first window, ping-pong between akey (sel 0: a->s) and skey (sel 1:
// s->a). Otherwise the stride, between bkey0 (sel 0: b0->b1) and bkey1.
template <bool INIT>
__global__ __launch_bounds__(kTileThreads) void k_merge_pass(BatchArgs a, int L, int sel) {
    __shared__ MergeLds m;
    const int g = blockIdx.x;
    const int32_t *pre = INIT ? a.pre_wtile : a.pre_stile;
    const int s = find_stream(pre, a.n, g);
    const StreamDev &S = a.s[s];
    const int len = INIT ? S.W : S.S;
    const int o0 = (g - pre[s]) * kTile;
    const int o1 = min(o0 + kTile, len);
    const uint64_t *sk;
    const float *sx, *sy, *sz;
    uint64_t *dk;
    float *dx, *dy, *dz;
    if (INIT) {
        if (sel == 0) {
            sk = S.akey; sx = S.ax; sy = S.ay; sz = S.az;
            dk = S.skey; dx = S.sx; dy = S.sy; dz = S.sz;
        } else {
            sk = S.skey; sx = S.sx; sy = S.sy; sz = S.sz;
            dk = S.akey; dx = S.ax; dy = S.ay; dz = S.az;
        }
    } else {
        if (sel == 0) {
            sk = S.bkey0; sx = S.bx0; sy = S.by0; sz = S.bz0;
            dk = S.bkey1; dx = S.bx1; dy = S.by1; dz = S.bz1;
        } else {
            sk = S.bkey1; sx = S.bx1; sy = S.by1; sz = S.bz1;
            dk = S.bkey0; dx = S.bx0; dy = S.by0; dz = S.bz0;
        }
    }
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
// window update: drop the expiring stride, merge the new one

__device__ __forceinline__ bool survives(const StreamDev &S, uint64_t key, int lo_slot) {
    const int slot = (int)((uint32_t)key & slot_mask(S));
    return slot < lo_slot || slot >= lo_slot + S.S;
}

// Survivors per window tile (the reference's stable_partition,
// thrust_helper.cu:37-39, split into count + compact).
__global__ __launch_bounds__(kTileThreads) void k_expire_count(BatchArgs a) {
    __shared__ int lds4[4];
    const int g = blockIdx.x;
    const int s = find_stream(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - a.pre_wtile[s];
    const int p0 = tile * kTile, p1 = min(p0 + kTile, S.W);
    const int lo_slot = S.mut->u * S.S;
    int cnt = 0;
    for (int p = p0 + threadIdx.x; p < p1; p += kTileThreads) cnt += survives(S, S.skey[p], lo_slot);
    cnt = block_sum(cnt, lds4);
    if (threadIdx.x == 0) S.tile_a[tile] = cnt;
}

// Ordered compaction of the survivors into the A arrays. Each tile's offset
// is the sum of the preceding tiles' counts (no spin-waiting look-back).
__global__ __launch_bounds__(kTileThreads) void k_expire_compact(BatchArgs a) {
    __shared__ int lds4[4];
    const int g = blockIdx.x;
    const int s = find_stream(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - a.pre_wtile[s];
    const int p0 = tile * kTile;
    const int lo_slot = S.mut->u * S.S;
    int part = 0;
    for (int t = threadIdx.x; t < tile; t += kTileThreads) part += S.tile_a[t];
    int run = block_sum(part, lds4);
    const int dim = S.dim;
    for (int r = 0; r < kTile / kTileThreads; r++) {
        const int p = p0 + r * kTileThreads + threadIdx.x;
        uint64_t key = 0;
        bool f = false;
        if (p < S.W) {
            key = S.skey[p];
            f = survives(S, key, lo_slot);
        }
        int tot;
        const int rank = block_rank_flags(f, &tot, lds4);
        if (f) {
            const int d = run + rank;
            S.akey[d] = key;
            S.ax[d] = S.sx[p];
            if (dim > 1) S.ay[d] = S.sy[p];
            if (dim > 2) S.az[d] = S.sz[p];
        }
        run += tot;
    }
}

// merge(survivors, sorted stride) -> window, plus cell starts per tile
// (thrust::merge at thrust_helper.cu:47-50 and the run starts of
// compute_offsets_of_cells, aabb.cu:377-395).
__global__ __launch_bounds__(kTileThreads) void k_merge_final(BatchArgs a, int bsel) {
    __shared__ MergeLds m;
    __shared__ int lds4[4];
    const int g = blockIdx.x;
    const int s = find_stream(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - a.pre_wtile[s];
    const int o0 = tile * kTile, o1 = min(o0 + kTile, S.W);
    const int la = S.W - S.S, lb = S.S;
    const uint64_t *B = bsel ? S.bkey1 : S.bkey0;
    const float *bx = bsel ? S.bx1 : S.bx0;
    const float *by = bsel ? S.by1 : S.by0;
    const float *bz = bsel ? S.bz1 : S.bz0;
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
    int starts = 0;
    for (int i = threadIdx.x; i < o1 - o0; i += kTileThreads) {
        const uint64_t key = m.out[i];
        const int sidx = m.src[i];
        S.skey[o0 + i] = key;
        if (sidx >= 0) {
            S.sx[o0 + i] = S.ax[sidx];
            if (dim > 1) S.sy[o0 + i] = S.ay[sidx];
            if (dim > 2) S.sz[o0 + i] = S.az[sidx];
        } else {
            const int j = ~sidx;
            S.sx[o0 + i] = bx[j];
            if (dim > 1) S.sy[o0 + i] = by[j];
            if (dim > 2) S.sz[o0 + i] = bz[j];
        }
        const uint64_t prev = i > 0 ? m.out[i - 1] : m.pred;
        starts += (o0 + i == 0) || ((key >> sb) != (prev >> sb));
    }
    starts = block_sum(starts, lds4);
    if (threadIdx.x == 0) S.tile_b[tile] = starts;
}

// Cell table: ordered compaction of run starts -> cstart/ckey, position ->
// cell, and the cell hash. Replaces compute_offsets_of_cells + thrust_sort
// (aabb.cu:377-410, thrust_helper.cu:16-19): the order comes from the scan,
// so no sort and no host round trip (dbscan.cpp:549) are needed.
__global__ __launch_bounds__(kTileThreads) void k_cells(BatchArgs a) {
    __shared__ int lds4[4];
    const int g = blockIdx.x;
    const int s = find_stream(a.pre_wtile, a.n, g);
    const StreamDev &S = a.s[s];
    const int tile = g - a.pre_wtile[s];
    const int ntile = a.pre_wtile[s + 1] - a.pre_wtile[s];
    const int p0 = tile * kTile;
    const uint32_t epoch = S.mut->epoch;
    const int sb = S.slot_bits;
    int part = 0;
    for (int t = threadIdx.x; t < tile; t += kTileThreads) part += S.tile_b[t];
    const int base = block_sum(part, lds4);
    int run = 0;
    for (int r = 0; r < kTile / kTileThreads; r++) {
        const int p = p0 + r * kTileThreads + threadIdx.x;
        const bool valid = p < S.W;
        uint64_t ck = 0;
        bool f = false;
        if (valid) {
            ck = S.skey[p] >> sb;
            f = (p == 0) || (ck != (S.skey[p - 1] >> sb));
        }
        int tot;
        const int rank = block_rank_flags(f, &tot, lds4);
        if (valid) S.pcell[p] = base + run + rank + (f ? 1 : 0) - 1;
        if (f) {
            const int c = base + run + rank;
            S.cstart[c] = p;
            S.ckey[c] = ck;
            hash_insert(S, ck, c, epoch);
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
// only the fastest axis, so its cells are consecutive in the sorted table.

__device__ __forceinline__ int2 cell_row(const StreamDev &S, int c, int r, uint32_t epoch) {
    const uint64_t ck = S.ckey[c];
    int64_t ix[3];
    ix[2] = (int64_t)(ck % (uint64_t)S.n[2]);
    const uint64_t rem = ck / (uint64_t)S.n[2];
    ix[1] = (int64_t)(rem % (uint64_t)S.n[1]);
    ix[0] = (int64_t)(rem / (uint64_t)S.n[1]);
    const int k = S.k, last = S.dim - 1, w = 2 * k + 1;
    const int center = (S.R - 1) / 2;
    int rr = r == 0 ? center : (r == center ? 0 : r);  // centre row first
    int64_t jx[3] = {ix[0], ix[1], ix[2]};
    for (int ax = S.dim - 2; ax >= 0; ax--) {
        const int off = rr % w - k;
        rr /= w;
        jx[ax] = ix[ax] + off;
        if (jx[ax] < 0 || jx[ax] >= S.n[ax]) return make_int2(0, 0);
    }
    const int64_t zlo = ix[last] - k < 0 ? 0 : ix[last] - k;
    const int64_t zhi = ix[last] + k >= S.n[last] ? S.n[last] - 1 : ix[last] + k;
    int first = -1, lastc = -1;
    for (int64_t z = zlo; z <= zhi; z++) {
        jx[last] = z;
        const uint64_t key = (uint64_t)((jx[0] * S.n[1] + jx[1]) * S.n[2] + jx[2]);
        const int c2 = hash_find(S, key, epoch);
        if (c2 >= 0) {
            if (first < 0) first = c2;
            lastc = c2;
        }
    }
    if (first < 0) return make_int2(0, 0);
    return make_int2(S.cstart[first], S.cstart[lastc + 1]);
}

// A cell whose >= min_pts points are all core without counting: the grid is
// dense-safe (k*k > dim: any two points of one cell are closer than eps) and
// the cell is interior. Border cells may hold points clamped in from outside
// the grid bounds, which the side-length argument does not cover.
__device__ __forceinline__ bool cell_is_dense(const StreamDev &S, int c) {
    const int cnt = S.cstart[c + 1] - S.cstart[c];
    if (!S.dense_ok || cnt < S.min_pts) return false;
    const uint64_t ck = S.ckey[c];
    int64_t ix[3];
    ix[2] = (int64_t)(ck % (uint64_t)S.n[2]);
    const uint64_t rem = ck / (uint64_t)S.n[2];
    ix[1] = (int64_t)(rem % (uint64_t)S.n[1]);
    ix[0] = (int64_t)(rem / (uint64_t)S.n[1]);
    for (int a = 0; a < S.dim; a++)
        if (ix[a] <= 0 || ix[a] >= S.n[a] - 1) return false;
    return true;
}

// Persistent: walks the device-side list of (stream, cell, row) items. The
// row-0 item of each cell also records whether the cell is dense.
__global__ __launch_bounds__(kTileThreads) void k_rows(BatchArgs a) {
    __shared__ long long pre[kMaxBatch + 1];
    __shared__ long long tmp4[4];
    block_stream_prefix(a.n, pre, tmp4,
                        [&](int i) { return (long long)a.s[i].mut->ncell * a.s[i].R; });
    const long long total = pre[a.n];
    for (long long gi = (long long)blockIdx.x * kTileThreads + threadIdx.x; gi < total;
         gi += (long long)gridDim.x * kTileThreads) {
        const int s = find_stream(pre, a.n, gi);
        const StreamDev &S = a.s[s];
        const long long local = gi - pre[s];
        const int c = (int)(local / S.R), r = (int)(local % S.R);
        S.rows[(size_t)c * S.R + r] = cell_row(S, c, r, S.mut->epoch);
        if (r == 0) S.cdense[c] = cell_is_dense(S, c) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// neighbourhood iteration: a 64-thread block (one wave) stages the stencil
// candidates of one cell into LDS, kCap at a time, and calls body(fill) when
// at least flush_at candidates are staged (and at the end). body returns
// false to stop early.

template <class Stage, class Body>
__device__ __forceinline__ void for_candidates(const StreamDev &S, int c, int flush_at, Stage stage,
                                               Body body) {
    const int lane = lane_id();
    int2 myrow = make_int2(0, 0);
    if (lane < S.R) myrow = S.rows[(size_t)c * S.R + lane];
    int fill = 0;
    for (int r = 0; r < S.R; r++) {
        int rb = __shfl(myrow.x, r);
        const int re = __shfl(myrow.y, r);
        while (rb < re) {
            const int take = min(re - rb, kCap - fill);
            for (int j = lane; j < take; j += 64) stage(fill + j, rb + j);
            fill += take;
            rb += take;
            if (fill == kCap) {
                __syncthreads();
                const bool more = body(fill);
                __syncthreads();
                fill = 0;
                if (!more) return;
            }
        }
        if (fill >= flush_at) {
            __syncthreads();
            const bool more = body(fill);
            __syncthreads();
            fill = 0;
            if (!more) return;
        }
    }
    if (fill > 0) {
        __syncthreads();
        body(fill);
        __syncthreads();
    }
}

struct ChunkCtx {
    int s, p0, p1;
};

__device__ __forceinline__ ChunkCtx chunk_of(const BatchArgs &a, int g) {
    ChunkCtx x;
    x.s = find_stream(a.pre_chunk, a.n, g);
    const int W = a.s[x.s].W;
    x.p0 = (g - a.pre_chunk[x.s]) * kChunk;
    x.p1 = min(x.p0 + kChunk, W);
    return x;
}

__device__ __forceinline__ bool is_dense(const StreamDev &S, int c) { return S.cdense[c] != 0; }

// Range count = core identification (the reference's __raygen__grid +
// __intersection__grid, dbscan.cu:155-194,250-270, and set_label,
// aabb.cu:575-587). Per cell segment of the chunk, candidates are staged in
// LDS and each query's neighbours are counted 64 at a time with
// ballot/popcount, stopping at min_pts. Cells of a dense-safe grid holding
// >= min_pts points are core without counting (all pairs within eps).
__global__ __launch_bounds__(64) void k_count(BatchArgs a) {
    __shared__ float cx[kCap], cy[kCap], cz[kCap];
    const ChunkCtx x = chunk_of(a, blockIdx.x);
    const StreamDev &S = a.s[x.s];
    const int lane = lane_id();
    const int p = x.p0 + lane;
    const bool mine = p < x.p1;
    const int dim = S.dim, minpts = S.min_pts;
    const float e2 = S.eps2;
    const int cfirst = S.pcell[x.p0], clast = S.pcell[x.p1 - 1];
    int mycnt = 0;
    for (int c = cfirst; c <= clast; c++) {
        const int cs = S.cstart[c], ce = S.cstart[c + 1];
        const int q0 = max(cs, x.p0), q1 = min(ce, x.p1);
        const bool inseg = mine && p >= q0 && p < q1;
        if (is_dense(S, c)) {
            if (inseg) {
                S.flags[p] = kFlagCore | kFlagDense;
                S.parent[p] = cs;  // first position of the cell holds its minimum slot
            }
            continue;
        }
        mycnt = 0;
        for_candidates(
            S, c, 128,
            [&](int i, int pos) {
                cx[i] = S.sx[pos];
                if (dim > 1) cy[i] = S.sy[pos];
                if (dim > 2) cz[i] = S.sz[pos];
            },
            [&](int fill) {
                for (int jl = q0 - x.p0; jl < q1 - x.p0; jl++) {
                    int cj = __shfl(mycnt, jl);
                    if (cj >= minpts) continue;
                    const int pq = x.p0 + jl;
                    const float qx = S.sx[pq];
                    const float qy = dim > 1 ? S.sy[pq] : 0.f;
                    const float qz = dim > 2 ? S.sz[pq] : 0.f;
                    for (int i = 0; i < fill && cj < minpts; i += 64) {
                        const int idx = i + lane;
                        bool hit = false;
                        if (idx < fill) hit = sqd(dim, qx, qy, qz, cx[idx], cy[idx], cz[idx]) < e2;
                        cj += __popcll(__ballot(hit));
                    }
                    if (lane == jl) mycnt = cj;
                }
                return __ballot(inseg && mycnt < minpts) != 0ull;
            });
        if (inseg) {
            S.flags[p] = mycnt >= minpts ? kFlagCore : 0;
            S.parent[p] = p;
        }
    }
}

// Cluster merge, point level (the reference's __raygen__cluster +
// __intersection__hybrid_radius_sphere core branch, dbscan.cu:279-321,
// 354-394): every core of a sparse cell unites with every core within eps
// that lies in a dense cell or at a smaller position (each sparse-sparse pair
// once). Pairs inside a dense cell are pre-linked; dense-dense pairs are
// handled per cell pair by k_union_dense.
__global__ __launch_bounds__(64) void k_union_sparse(BatchArgs a) {
    __shared__ float cx[kCap], cy[kCap], cz[kCap];
    __shared__ int32_t cpos[kCap];
    __shared__ uint8_t cfl[kCap];
    const ChunkCtx x = chunk_of(a, blockIdx.x);
    const StreamDev &S = a.s[x.s];
    const int lane = lane_id();
    const int p = x.p0 + lane;
    const bool mine = p < x.p1;
    const int dim = S.dim;
    const float e2 = S.eps2;
    const uint32_t sm = slot_mask(S);
    const bool mycore = mine && (S.flags[p] & kFlagCore);
    const int cfirst = S.pcell[x.p0], clast = S.pcell[x.p1 - 1];
    for (int c = cfirst; c <= clast; c++) {
        const int cs = S.cstart[c], ce = S.cstart[c + 1];
        if (is_dense(S, c)) continue;
        const int q0 = max(cs, x.p0), q1 = min(ce, x.p1);
        const bool qcore = mycore && p >= q0 && p < q1;
        if (__ballot(qcore) == 0ull) continue;
        for_candidates(
            S, c, kCap,
            [&](int i, int pos) {
                cx[i] = S.sx[pos];
                if (dim > 1) cy[i] = S.sy[pos];
                if (dim > 2) cz[i] = S.sz[pos];
                cpos[i] = pos;
                cfl[i] = S.flags[pos];
            },
            [&](int fill) {
                const uint64_t qm = __ballot(qcore);
                for (int jl = q0 - x.p0; jl < q1 - x.p0; jl++) {
                    if (!((qm >> jl) & 1ull)) continue;
                    const int pq = x.p0 + jl;
                    const float qx = S.sx[pq];
                    const float qy = dim > 1 ? S.sy[pq] : 0.f;
                    const float qz = dim > 2 ? S.sz[pq] : 0.f;
                    for (int i = 0; i < fill; i += 64) {
                        const int idx = i + lane;
                        if (idx < fill) {
                            const uint8_t f = cfl[idx];
                            const int cp = cpos[idx];
                            if ((f & kFlagCore) && ((f & kFlagDense) || cp < pq) &&
                                sqd(dim, qx, qy, qz, cx[idx], cy[idx], cz[idx]) < e2)
                                uf_unite(S.parent, S.skey, sm, pq, cp);
                        }
                    }
                }
                return true;
            });
    }
}

// Cluster merge, cell level: two dense cells (all core, pre-linked) join as
// soon as one pair of their points is within eps - the analogue of the
// reference's dense-cell spheres, where a ray unites with the first point of
// the cell within eps and stops (dbscan.cu:378-393). Persistent over the
// device-side cell lists; one wave per cell, each unordered pair once.
__global__ __launch_bounds__(64) void k_union_dense(BatchArgs a) {
    __shared__ long long pre[kMaxBatch + 1];
    __shared__ long long tmp4[4];
    // block_stream_prefix assumes 256 threads; emulate with a serial prefix
    // over at most kMaxBatch streams (one wave, 16 per lane).
    {
        const int lane = lane_id();
        long long local = 0;
        long long v[kMaxBatch / 64];
        for (int q = 0; q < kMaxBatch / 64; q++) {
            const int i = lane * (kMaxBatch / 64) + q;
            v[q] = i < a.n ? (long long)a.s[i].mut->ncell : 0;
            local += v[q];
        }
        long long xs = local;
        for (int o = 1; o < 64; o <<= 1) {
            long long y = __shfl_up(xs, o);
            if (lane >= o) xs += y;
        }
        long long run = xs - local;
        for (int q = 0; q < kMaxBatch / 64; q++) {
            const int i = lane * (kMaxBatch / 64) + q;
            run += v[q];
            if (i < a.n) pre[i + 1] = run;
        }
        if (lane == 0) pre[0] = 0;
        (void)tmp4;
        __syncthreads();
    }
    const long long total = pre[a.n];
    const int lane = lane_id();
    for (long long gc = blockIdx.x; gc < total; gc += gridDim.x) {
        const int s = find_stream(pre, a.n, gc);
        const StreamDev &S = a.s[s];
        const int c = (int)(gc - pre[s]);
        const int cs = S.cstart[c], ce = S.cstart[c + 1];
        if (!is_dense(S, c)) continue;
        const int dim = S.dim;
        const float e2 = S.eps2;
        const uint32_t sm = slot_mask(S);
        int2 myrow = make_int2(0, 0);
        if (lane < S.R) myrow = S.rows[(size_t)c * S.R + lane];
        for (int r = 0; r < S.R; r++) {
            const int rb = __shfl(myrow.x, r), re = __shfl(myrow.y, r);
            if (rb >= re) continue;
            const int c2lo = S.pcell[rb];
            const int c2hi = min(S.pcell[re - 1], c - 1);
            for (int c2 = c2lo; c2 <= c2hi; c2++) {
                const int cs2 = S.cstart[c2], ce2 = S.cstart[c2 + 1];
                if (!is_dense(S, c2)) continue;
                int same = 0;
                if (lane == 0) same = uf_find(S.parent, cs) == uf_find(S.parent, cs2);
                if (__shfl(same, 0)) continue;
                bool found = false;
                for (int b0 = cs2; b0 < ce2 && !found; b0 += 64) {
                    const int pb = b0 + lane;
                    const bool valid = pb < ce2;
                    float bx = 0.f, by = 0.f, bz = 0.f;
                    if (valid) {
                        bx = S.sx[pb];
                        if (dim > 1) by = S.sy[pb];
                        if (dim > 2) bz = S.sz[pb];
                    }
                    for (int pa = cs; pa < ce; pa++) {
                        const float qx = S.sx[pa];
                        const float qy = dim > 1 ? S.sy[pa] : 0.f;
                        const float qz = dim > 2 ? S.sz[pa] : 0.f;
                        const bool hit = valid && sqd(dim, qx, qy, qz, bx, by, bz) < e2;
                        if (__ballot(hit) != 0ull) {
                            found = true;
                            break;
                        }
                    }
                }
                if (found && lane == 0) uf_unite(S.parent, S.skey, sm, cs, cs2);
            }
        }
    }
}

// Canonical ids of cores (post_cluster_t, aabb.cu:173-196): the root's slot.
__global__ __launch_bounds__(64) void k_flatten(BatchArgs a) {
    const ChunkCtx x = chunk_of(a, blockIdx.x);
    const StreamDev &S = a.s[x.s];
    const IO io = a.use_io1 ? a.io1 : a.io[x.s];
    const int p = x.p0 + lane_id();
    if (p >= x.p1) return;
    if (!(S.flags[p] & kFlagCore)) return;
    const uint32_t sm = slot_mask(S);
    const int root = uf_find(S.parent, p);
    const int v = (int)((uint32_t)S.skey[root] & sm);
    S.canon[p] = v;
    const int slot = (int)((uint32_t)S.skey[p] & sm);
    io.label[slot] = 0;
    io.cid[slot] = v;
}

// Border / noise (the reference's border claim, dbscan.cu:370-377, with the
// deterministic rule: the minimum canonical id among cores within eps).
// Block 0 of each stream also advances the ring cursor and hash epoch; no
// kernel of this launch reads them.
__global__ __launch_bounds__(64) void k_border(BatchArgs a) {
    __shared__ float cx[kCap], cy[kCap], cz[kCap];
    __shared__ int32_t cmeta[kCap];
    const int g = blockIdx.x;
    const ChunkCtx x = chunk_of(a, g);
    const StreamDev &S = a.s[x.s];
    const IO io = a.use_io1 ? a.io1 : a.io[x.s];
    const int lane = lane_id();
    const int p = x.p0 + lane;
    const bool mine = p < x.p1;
    const int dim = S.dim;
    const float e2 = S.eps2;
    const uint32_t sm = slot_mask(S);
    const bool mynon = mine && !(S.flags[p] & kFlagCore);
    const int cfirst = S.pcell[x.p0], clast = S.pcell[x.p1 - 1];
    int best = INT_MAX;
    for (int c = cfirst; c <= clast; c++) {
        const int cs = S.cstart[c], ce = S.cstart[c + 1];
        if (is_dense(S, c)) continue;
        const int q0 = max(cs, x.p0), q1 = min(ce, x.p1);
        const bool qnon = mynon && p >= q0 && p < q1;
        if (__ballot(qnon) == 0ull) continue;
        for_candidates(
            S, c, kCap,
            [&](int i, int pos) {
                cx[i] = S.sx[pos];
                if (dim > 1) cy[i] = S.sy[pos];
                if (dim > 2) cz[i] = S.sz[pos];
                cmeta[i] = (S.flags[pos] & kFlagCore) ? S.canon[pos] : INT_MAX;
            },
            [&](int fill) {
                const uint64_t qm = __ballot(qnon);
                for (int jl = q0 - x.p0; jl < q1 - x.p0; jl++) {
                    if (!((qm >> jl) & 1ull)) continue;
                    const int pq = x.p0 + jl;
                    const float qx = S.sx[pq];
                    const float qy = dim > 1 ? S.sy[pq] : 0.f;
                    const float qz = dim > 2 ? S.sz[pq] : 0.f;
                    int bj = __shfl(best, jl);
                    for (int i = 0; i < fill; i += 64) {
                        const int idx = i + lane;
                        int v = INT_MAX;
                        if (idx < fill) {
                            const int m = cmeta[idx];
                            if (m != INT_MAX && sqd(dim, qx, qy, qz, cx[idx], cy[idx], cz[idx]) < e2) v = m;
                        }
                        if (__ballot(v != INT_MAX) != 0ull) bj = min(bj, wave_min(v));
                    }
                    if (lane == jl) best = bj;
                }
                return true;
            });
    }
    if (mynon) {
        const int slot = (int)((uint32_t)S.skey[p] & sm);
        io.label[slot] = best != INT_MAX ? 1 : 2;
        io.cid[slot] = best != INT_MAX ? best : slot;
    }
    if (g == a.pre_chunk[x.s] && lane == 0) {
        StreamMut *m = S.mut;
        m->t = m->t + 1;
        m->u = m->u + 1 == S.units ? 0 : m->u + 1;
        m->epoch = m->epoch >= kEpochMax ? 1u : m->epoch + 1u;
    }
}

__global__ __launch_bounds__(kTileThreads) void k_clear_hash(BatchArgs a) {
    const StreamDev &S = a.s[blockIdx.y];
    const size_t sz = (size_t)S.hmask + 1;
    for (size_t i = (size_t)blockIdx.x * kTileThreads + threadIdx.x; i < sz;
         i += (size_t)gridDim.x * kTileThreads)
        S.htab[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) S.mut->epoch = 1u;
}

// ---------------------------------------------------------------------------
// launchers

#define RTDBS_LAUNCH(kernel, grid, block, ...)                        \
    do {                                                              \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, st, __VA_ARGS__); \
        hipError_t e_ = hipGetLastError();                            \
        if (e_ != hipSuccess) return e_;                              \
    } while (0)

static inline hipError_t mark(hipEvent_t *ev, int idx, hipStream_t st) {
    if (!ev) return hipSuccess;
    return hipEventRecord(ev[idx], st);
}

hipError_t launch_slide(const BatchArgs &a, const LaunchShape &sh, hipStream_t st, hipEvent_t *ev) {
    hipError_t e;
#define PH_BEGIN(ph) if ((e = mark(ev, 2 * (ph), st)) != hipSuccess) return e
#define PH_END(ph) if ((e = mark(ev, 2 * (ph) + 1, st)) != hipSuccess) return e
    PH_BEGIN(kPhStrideSort);
    RTDBS_LAUNCH(k_stride_tiles<false>, sh.stiles, kTileThreads, a);
    PH_END(kPhStrideSort);
    PH_BEGIN(kPhMerge);
    for (int p = 0; p < sh.passes; p++) RTDBS_LAUNCH(k_merge_pass<false>, sh.stiles, kTileThreads, a, kTile << p, p & 1);
    PH_END(kPhMerge);
    PH_BEGIN(kPhExpire);
    RTDBS_LAUNCH(k_expire_count, sh.wtiles, kTileThreads, a);
    RTDBS_LAUNCH(k_expire_compact, sh.wtiles, kTileThreads, a);
    PH_END(kPhExpire);
    PH_BEGIN(kPhMergeFinal);
    RTDBS_LAUNCH(k_merge_final, sh.wtiles, kTileThreads, a, sh.passes & 1);
    PH_END(kPhMergeFinal);
    PH_BEGIN(kPhCells);
    RTDBS_LAUNCH(k_cells, sh.wtiles, kTileThreads, a);
    PH_END(kPhCells);
    PH_BEGIN(kPhRows);
    RTDBS_LAUNCH(k_rows, sh.persist_blocks, kTileThreads, a);
    PH_END(kPhRows);
    PH_BEGIN(kPhCount);
    RTDBS_LAUNCH(k_count, sh.chunks, 64, a);
    PH_END(kPhCount);
    PH_BEGIN(kPhUnion);
    RTDBS_LAUNCH(k_union_sparse, sh.chunks, 64, a);
    RTDBS_LAUNCH(k_union_dense, sh.persist_blocks * 4, 64, a);
    PH_END(kPhUnion);
    PH_BEGIN(kPhFlatten);
    RTDBS_LAUNCH(k_flatten, sh.chunks, 64, a);
    PH_END(kPhFlatten);
    PH_BEGIN(kPhBorder);
    RTDBS_LAUNCH(k_border, sh.chunks, 64, a);
    PH_END(kPhBorder);
#undef PH_BEGIN
#undef PH_END
    return hipSuccess;
}

hipError_t launch_init(const BatchArgs &a, int32_t itiles, int32_t passes, hipStream_t st) {
    RTDBS_LAUNCH(k_stride_tiles<true>, itiles, kTileThreads, a);
    for (int p = 0; p < passes; p++) RTDBS_LAUNCH(k_merge_pass<true>, itiles, kTileThreads, a, kTile << p, p & 1);
    return hipSuccess;
}

hipError_t launch_clear_hash(const BatchArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_hash, dim3(256, a.n), dim3(kTileThreads), 0, st, a);
    return hipGetLastError();
}

}  // namespace rtdbs
