// Device helpers shared by the slide kernels and the brute-force validator.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtdbs {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// A value every lane of the wave holds alike, moved to a scalar register (the
// compiler cannot always prove it uniform; loops and addresses built from it
// then stay scalar).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Global-memory view of a device pointer: per-lane accesses through it are
// global loads and stores with a scalar base instead of flat accesses.
template <class T>
using gptr = __attribute__((address_space(1))) T *;

template <class T>
__device__ __forceinline__ gptr<T> gl(T *p) {
    return (gptr<T>)p;
}

// Read of a field of a descriptor that no kernel writes during a launch, as
// constant memory: a scalar load when the address is uniform.
template <class T>
__device__ __forceinline__ T ldc(const T &f) {
    return *(const __attribute__((address_space(4))) T *)&f;
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

// v of lane (this lane ^ m), for m a power of two known at compile time: DPP
// moves within 16-lane rows for m <= 8 (quad permutes, row shifts, a row
// rotate) and the row-pair / half-wave swaps for 16 and 32 - no LDS round
// trip, unlike __shfl_xor (ds_bpermute).
__device__ __forceinline__ int xor_lane(int v, int m) {
    const int lane = lane_id();
    switch (m) {
    case 1:
        return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    case 2:
        return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    case 4: {
        const int up = __builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0xF, false);  // row_shl:4, from lane + 4
        const int dn = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4, from lane - 4
        return (lane & 4) ? dn : up;
    }
    case 8:
        return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);  // row_ror:8
    case 16: {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (lane & 16) ? r[0] : r[1];
    }
    case 32: {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (lane & 32) ? r[0] : r[1];
    }
    default:
        return __shfl_xor(v, m);
    }
}

// v of lane (this lane - 1), lane 0 gets 0: one DPP move across the whole
// wave (wave_shr:1), no LDS round trip, unlike __shfl_up (ds_bpermute).
__device__ __forceinline__ int prev_lane(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, xor_lane(v, o));
    return v;
}

// Union-find parent reads may be stale (served by a CU's L1 after another CU
// lowered an entry): a stale value is still an ancestor, and hooks go through
// L2 atomics, so stale reads cost extra steps, never correctness. Cacheable
// loads keep the hot root entries of large clusters from funnelling every walk
// into one L2 line.
__device__ __forceinline__ int uf_load(const int32_t *p) { return *p; }

// Union-find over ring slots, parent[x] <= x for every slot, so the root of a
// component is its minimum slot - the reference's canonical cluster id (it
// hooks the larger root under the smaller, dbscan.cu:22-43).
//
// Hooking uses atomicMin, which never fails: when b turned out not to be a
// root any more, min(old, a) is stored and the caller goes on uniting a with
// the displaced parent `old`, so no link is lost and b strictly decreases.
// (A compare-and-swap hook fails and retries whenever another lane hooked the
// same root first, which serialises every edge of a large cluster on its
// root.) Path halving and compression also lower entries with atomicMin, so
// every entry only ever decreases and always points at an ancestor.
__device__ __forceinline__ int uf_find(int32_t *parent, int x) {
    int p = uf_load(parent + x);
    while (p != x) {
        const int gp = uf_load(parent + p);
        if (gp != p) atomicMin(parent + x, gp);
        x = p;
        p = gp;
    }
    return x;
}

// The roots of a and b, walked together so that the loads of both walks are
// in flight at once. The walks only load; afterwards each starting entry is
// pointed at the root found (one atomicMin each). Atomics inside a walk would
// stall it: the vector memory counter also counts atomics, in issue order, so
// the wait for the next step's load would include the atomic issued before
// it. The starting entries are the elements the merge unites (cell
// representatives, core slots), so their next walks take one step.
__device__ __forceinline__ void uf_find2(int32_t *parent, int &a, int &b) {
    const int a0 = a, b0 = b;
    int pa = uf_load(parent + a), pb = uf_load(parent + b);
    const int pa0 = pa, pb0 = pb;
    while (pa != a || pb != b) {
        const int ga = pa != a ? uf_load(parent + pa) : a;
        const int gb = pb != b ? uf_load(parent + pb) : b;
        if (pa != a) {
            a = pa;
            pa = ga;
        }
        if (pb != b) {
            b = pb;
            pb = gb;
        }
    }
    if (a != pa0) atomicMin(parent + a0, a);
    if (b != pb0) atomicMin(parent + b0, b);
}

__device__ __forceinline__ void uf_unite(int32_t *parent, int a, int b) {
    while (true) {
        uf_find2(parent, a, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        const int old = atomicMin(parent + b, a);
        if (old == b) return;
        b = old;
    }
}

}  // namespace rtdbs
