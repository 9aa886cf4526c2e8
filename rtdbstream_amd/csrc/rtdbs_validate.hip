// On-device brute-force DBSCAN of one window: every pair of points is tested.
// It is the large-window validator of the slide step. It plays the role of the
// reference's own GPU check path: cluster_with_cuda and find_neighbors_t /
// find_cores_t / set_cluster_id_op_t / post_cluster_t (util.cpp:357-381,
// aabb.cu:150-261).
// It uses the same FP32 distance predicate as the slide kernels (sqd: the
// reference's un-contracted order, d2 < fl32(eps*eps)) and the same output
// rules. It shares no grid, sorting or cell logic with them: its candidates
// are all n points, staged through LDS in tiles.
//   nn    - neighbours within eps, the point itself included (find_neighbors_t)
//   label - 0 core (nn >= min_pts), 1 border, 2 noise (find_cores_t)
//   cid   - cores: the minimum index of their cluster's cores (union-find
//           hooking the larger root under the smaller, dbscan.cu:22-43, then
//           post_cluster_t); borders: the minimum cluster id among the cores
//           within eps; noise: its own index.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>
#include <cmath>
#include <mutex>
#include <vector>

#include "rtdbs.h"
#include "rtdbs_devfn.h"
#include "rtdbs_host.h"

namespace rtdbs {

// Scratch of rtdbs_bruteforce_noalloc: union-find parents [0, kNoAllocMax)
// and, when the caller passes no nn array, neighbour counts [kNoAllocMax,
// 2 * kNoAllocMax). One copy per GPU, held until the process ends (8 MiB).
constexpr int64_t kNoAllocMax = int64_t(1) << 20;
static __device__ __align__(256) int32_t scratch[2 * kNoAllocMax];

namespace {

constexpr int kBfTile = 256;  // candidates staged per LDS tile = threads per block
constexpr int kMaxDevices = 64;

struct BfArgs {
    const float *pts;  // n points, dim floats each, point-major
    int n, dim, min_pts;
    float e2;
    int32_t *nn, *parent, *label, *cid;
};

// Coordinates of point i (0 on the axes past dim; those are never read).
__device__ __forceinline__ void bf_point(const BfArgs &a, int i, float &x, float &y, float &z) {
    const float *p = a.pts + (size_t)i * a.dim;
    x = p[0];
    y = a.dim > 1 ? p[1] : 0.f;
    z = a.dim > 2 ? p[2] : 0.f;
}

// Stages candidate j (or zeros past n) into the block's LDS tile.
__device__ __forceinline__ void bf_stage(const BfArgs &a, int j, float *tx, float *ty, float *tz) {
    float x = 0.f, y = 0.f, z = 0.f;
    if (j < a.n) bf_point(a, j, x, y, z);
    tx[threadIdx.x] = x;
    ty[threadIdx.x] = y;
    tz[threadIdx.x] = z;
}

// Neighbour counts: one query per thread, candidates in LDS tiles.
__global__ __launch_bounds__(kBfTile) void k_bf_count(BfArgs a) {
    __shared__ float tx[kBfTile], ty[kBfTile], tz[kBfTile];
    const int i = blockIdx.x * kBfTile + threadIdx.x;
    const bool q = i < a.n;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    if (q) bf_point(a, i, qx, qy, qz);
    int cnt = 0;
    for (int t0 = 0; t0 < a.n; t0 += kBfTile) {
        __syncthreads();
        bf_stage(a, t0 + threadIdx.x, tx, ty, tz);
        __syncthreads();
        const int m = min(kBfTile, a.n - t0);
        for (int k = 0; k < m; k++) cnt += sqd(a.dim, qx, qy, qz, tx[k], ty[k], tz[k]) < a.e2;
    }
    if (q) {
        a.nn[i] = cnt;
        a.label[i] = cnt >= a.min_pts ? 0 : 2;
        a.parent[i] = i;
    }
}

// Core-core unions: each core unites with every core j < i within eps, so
// every pair is tested once.
__global__ __launch_bounds__(kBfTile) void k_bf_union(BfArgs a) {
    __shared__ float tx[kBfTile], ty[kBfTile], tz[kBfTile];
    __shared__ int tc[kBfTile];
    const int i = blockIdx.x * kBfTile + threadIdx.x;
    const bool q = i < a.n && a.label[i] == 0;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    if (q) bf_point(a, i, qx, qy, qz);
    const int last = min(a.n, (int)(blockIdx.x + 1) * kBfTile);  // candidates below the block's last query
    for (int t0 = 0; t0 < last; t0 += kBfTile) {
        __syncthreads();
        const int j = t0 + threadIdx.x;
        const bool v = j < a.n;
        bf_stage(a, j, tx, ty, tz);
        tc[threadIdx.x] = v && a.label[j] == 0;
        __syncthreads();
        if (!q) continue;
        const int m = min(kBfTile, i - t0);
        for (int k = 0; k < m; k++)
            if (tc[k] && sqd(a.dim, qx, qy, qz, tx[k], ty[k], tz[k]) < a.e2) uf_unite(a.parent, i, t0 + k);
    }
}

__global__ __launch_bounds__(kBfTile) void k_bf_flatten(BfArgs a) {
    const int i = blockIdx.x * kBfTile + threadIdx.x;
    if (i >= a.n) return;
    a.cid[i] = a.label[i] == 0 ? uf_find(a.parent, i) : i;
}

// Borders: the minimum cluster id among the cores within eps.
__global__ __launch_bounds__(kBfTile) void k_bf_border(BfArgs a) {
    __shared__ float tx[kBfTile], ty[kBfTile], tz[kBfTile];
    __shared__ int tcid[kBfTile];
    const int i = blockIdx.x * kBfTile + threadIdx.x;
    const bool q = i < a.n && a.label[i] != 0;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    if (q) bf_point(a, i, qx, qy, qz);
    int best = INT_MAX;
    for (int t0 = 0; t0 < a.n; t0 += kBfTile) {
        __syncthreads();
        const int j = t0 + threadIdx.x;
        const bool v = j < a.n;
        bf_stage(a, j, tx, ty, tz);
        tcid[threadIdx.x] = v && a.label[j] == 0 ? a.cid[j] : INT_MAX;
        __syncthreads();
        if (!q) continue;
        const int m = min(kBfTile, a.n - t0);
        for (int k = 0; k < m; k++)
            if (tcid[k] < best && sqd(a.dim, qx, qy, qz, tx[k], ty[k], tz[k]) < a.e2) best = tcid[k];
    }
    if (q && best != INT_MAX) {
        a.label[i] = 1;
        a.cid[i] = best;
    }
}

// The four launches of one validation, in order, on `st`.
hipError_t launch_bf(const BfArgs &a, hipStream_t st) {
    const unsigned blocks = (unsigned)((a.n + kBfTile - 1) / kBfTile);
    hipLaunchKernelGGL(k_bf_count, dim3(blocks), dim3(kBfTile), 0, st, a);
    hipLaunchKernelGGL(k_bf_union, dim3(blocks), dim3(kBfTile), 0, st, a);
    hipLaunchKernelGGL(k_bf_flatten, dim3(blocks), dim3(kBfTile), 0, st, a);
    hipLaunchKernelGGL(k_bf_border, dim3(blocks), dim3(kBfTile), 0, st, a);
    return hipGetLastError();
}

int check_bf_args(int32_t dim, float eps, int32_t min_pts, int64_t n) {
    if (dim < 1 || dim > 3) return fail(RTDBS_EINVAL, "dim must be 1, 2 or 3 (got %d)", dim);
    if (!(eps > 0.0f) || !(eps < INFINITY)) return fail(RTDBS_EINVAL, "eps must be finite and > 0");
    if (min_pts < 1) return fail(RTDBS_EINVAL, "min_pts must be >= 1 (got %d)", min_pts);
    if (n < 1 || n > (int64_t)INT_MAX - kBfTile) return fail(RTDBS_EINVAL, "n out of range (got %lld)", (long long)n);
    return RTDBS_OK;
}

#define BFCHK(expr)                                                                  \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            rc = fail(RTDBS_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
            goto done;                                                               \
        }                                                                            \
    } while (0)

// Device address of this GPU's scratch array, looked up once per device with
// the device current. Host code uses only this address, never the array's
// name (which is a host-side stand-in, not a device address).
int scratch_ptr(int32_t **out) {
    static std::mutex mu;
    static int32_t *ptr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(RTDBS_EHIP, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(RTDBS_EHIP, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(RTDBS_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        ptr[dev] = (int32_t *)p;
    }
    *out = ptr[dev];
    return RTDBS_OK;
}

}  // namespace
}  // namespace rtdbs

using namespace rtdbs;

extern "C" int rtdbs_bruteforce(int32_t dim, float eps, int32_t min_pts, const float *points, int64_t n,
                                int32_t *label, int32_t *cid, int32_t *nn) {
    int rc = check_bf_args(dim, eps, min_pts, n);
    if (rc) return rc;
    if (!points || !label || !cid) return fail(RTDBS_EINVAL, "NULL argument");
    const size_t N = (size_t)n;
    char *mem = nullptr;
    BfArgs a{};
    BFCHK(hipMalloc((void **)&mem, N * (size_t)dim * 4 + 4 * N * 4));
    a.pts = (const float *)mem;
    a.nn = (int32_t *)(mem + N * (size_t)dim * 4);
    a.parent = a.nn + N;
    a.label = a.parent + N;
    a.cid = a.label + N;
    a.n = (int)n;
    a.dim = dim;
    a.min_pts = min_pts;
    a.e2 = eps * eps;  // fl32(eps*eps), as the reference's radius2 (dbscan.cpp:41)
    BFCHK(hipMemcpy(mem, points, N * (size_t)dim * 4, hipMemcpyHostToDevice));
    BFCHK(launch_bf(a, 0));
    BFCHK(hipMemcpy(label, a.label, N * 4, hipMemcpyDeviceToHost));
    BFCHK(hipMemcpy(cid, a.cid, N * 4, hipMemcpyDeviceToHost));
    if (nn) BFCHK(hipMemcpy(nn, a.nn, N * 4, hipMemcpyDeviceToHost));
done:
    if (mem) (void)hipFree(mem);
    return rc;
}

extern "C" int rtdbs_bruteforce_noalloc(int32_t dim, float eps, int32_t min_pts, const float *points_dev, int64_t n,
                                        int32_t *label_dev, int32_t *cid_dev, int32_t *nn_dev) {
    int rc = check_bf_args(dim, eps, min_pts, n);
    if (rc) return rc;
    if (n > kNoAllocMax)
        return fail(RTDBS_EINVAL, "n = %lld exceeds the %lld points the allocation-free validator's scratch holds",
                    (long long)n, (long long)kNoAllocMax);
    if (!points_dev || !label_dev || !cid_dev) return fail(RTDBS_EINVAL, "NULL argument");
    int32_t *scr = nullptr;
    rc = scratch_ptr(&scr);
    if (rc) return rc;
    BfArgs a{};
    a.pts = points_dev;
    a.n = (int)n;
    a.dim = dim;
    a.min_pts = min_pts;
    a.e2 = eps * eps;  // fl32(eps*eps), as the reference's radius2 (dbscan.cpp:41)
    a.parent = scr;  // every entry < n is written by the count kernel before it is read
    a.nn = nn_dev ? nn_dev : scr + kNoAllocMax;
    a.label = label_dev;
    a.cid = cid_dev;
    hipError_t e = launch_bf(a, 0);
    if (e == hipSuccess) e = hipStreamSynchronize(0);
    if (e != hipSuccess) return fail(RTDBS_EHIP, "validator launch failed: %s", hipGetErrorString(e));
    return RTDBS_OK;
}
