// Host engine of the MI355X streaming DBSCAN slide step: stream state,
// plain (one stream, eager launches) and batched (many streams, one launch per
// phase, optional HIP graph) slides, and the C ABI declared in include/rtdbs.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "rtdbs.h"
#include "rtdbs_dev.h"
#include "rtdbs_host.h"

using namespace rtdbs;

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;

namespace rtdbs {
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
}  // namespace rtdbs

extern "C" const char *rtdbs_last_error(void) { return g_err.c_str(); }
extern "C" const char *rtdbs_version(void) { return "rtdbs-mi355x 0.1 (gfx950)"; }

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(RTDBS_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_));  \
    } while (0)

namespace rtdbs {

int device_cus() {
    static int cus = 0;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    });
    return cus;
}

// Unique across all streams, so a batch never mistakes a new stream that
// reuses a freed stream's address for the old one.
static uint64_t next_generation() {
    static std::atomic<uint64_t> g{1};
    return g.fetch_add(1);
}

static uint32_t next_pow2(uint64_t v) {
    uint64_t p = 64;
    while (p < v) p <<= 1;
    return (uint32_t)p;
}

static int ceil_log2(int64_t v) {
    int b = 0;
    while ((int64_t(1) << b) < v) b++;
    return b;
}

static int default_cell_div(const rtdbs_params &p) {
    // k = 2: cells of side eps/2 are dense-safe for dim <= 3 (k*k > dim), so a
    // cell holding >= min_pts points is all core without counting.
    (void)p;
    return 2;
}

int validate_params(const rtdbs_params *p) {
    if (!p) return fail(RTDBS_EINVAL, "params is NULL");
    if (p->dim < 1 || p->dim > 3) return fail(RTDBS_EINVAL, "dim must be 1, 2 or 3 (got %d)", p->dim);
    if (p->window < 1 || p->window > (1 << kMaxSlotBits))
        return fail(RTDBS_EINVAL, "window must be in [1, %d] (got %d)", 1 << kMaxSlotBits, p->window);
    if (p->stride < 1 || p->stride > p->window)
        return fail(RTDBS_EINVAL, "stride must be in [1, window] (got %d)", p->stride);
    if (p->window % p->stride != 0)
        return fail(RTDBS_EINVAL, "stride %d must divide window %d (the ring of dbscan.cpp:2260,2289)",
                    p->stride, p->window);
    if (!(p->eps > 0.0f) || !std::isfinite(p->eps)) return fail(RTDBS_EINVAL, "eps must be finite and > 0");
    if (p->min_pts < 1) return fail(RTDBS_EINVAL, "min_pts must be >= 1 (got %d)", p->min_pts);
    if (p->cell_div < 0 || p->cell_div > kMaxK)
        return fail(RTDBS_EINVAL, "cell_div must be 0 (auto) or 1..%d (got %d)", kMaxK, p->cell_div);
    if (!(p->lanes_per_query == 0 || p->lanes_per_query == 1 || p->lanes_per_query == 2 || p->lanes_per_query == 4))
        return fail(RTDBS_EINVAL, "lanes_per_query must be 0 (auto), 1, 2 or 4 (got %d)", p->lanes_per_query);
    if (p->neighbor_mode < 0 || p->neighbor_mode > 2)
        return fail(RTDBS_EINVAL, "neighbor_mode must be 0 (auto), 1 or 2 (got %d)", p->neighbor_mode);
    if (p->cellpar_per_cu < 0 || p->cellpar_per_cu > kMaxPerCu)
        return fail(RTDBS_EINVAL, "cellpar_per_cu must be 0 (auto) or 1..%d (got %d)", kMaxPerCu, p->cellpar_per_cu);
    if (p->rows_per_cu < 0 || p->rows_per_cu > kMaxPerCu)
        return fail(RTDBS_EINVAL, "rows_per_cu must be 0 (auto) or 1..%d (got %d)", kMaxPerCu, p->rows_per_cu);
    for (int a = 0; a < p->dim; a++)
        if (!std::isfinite(p->lo[a]) || !std::isfinite(p->hi[a]))
            return fail(RTDBS_EINVAL, "grid bounds must be finite");
    return RTDBS_OK;
}

// Buffer layout of one stream's device arena.
struct Layout {
    size_t off_mut, off_skey, off_sx, off_sy, off_sz, off_akey, off_ax, off_ay, off_az;
    size_t off_bkey0, off_bx0, off_by0, off_bz0, off_bkey1, off_bx1, off_by1, off_bz1;
    size_t off_cstart, off_ckey, off_pcell, off_rows, off_cflag, off_fcore, off_flags, off_parent, off_rep,
        off_tile_a, off_tile_b, off_in, off_label, off_cid;
    size_t total;
};

static Layout make_layout(const rtdbs_params &p, int R) {
    Layout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += (bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t W = (size_t)p.window, S = (size_t)p.stride, d = (size_t)p.dim;
    const size_t ntile = (W + kWTile - 1) / kWTile;
    L.off_mut = take(sizeof(StreamMut));
    L.off_skey = take(W * 8);
    // window coordinates (both buffers of the ping-pong) with kCandPad
    // readable floats past the end, a margin: the candidate walks read real
    // candidates only
    const size_t wc = (W + kCandPad) * 4;
    L.off_sx = take(wc);
    L.off_sy = d > 1 ? take(wc) : 0;
    L.off_sz = d > 2 ? take(wc) : 0;
    L.off_akey = take(W * 8);
    L.off_ax = take(wc);
    L.off_ay = d > 1 ? take(wc) : 0;
    L.off_az = d > 2 ? take(wc) : 0;
    L.off_bkey0 = take(S * 8);
    L.off_bx0 = take(S * 4);
    L.off_by0 = d > 1 ? take(S * 4) : 0;
    L.off_bz0 = d > 2 ? take(S * 4) : 0;
    L.off_bkey1 = take(S * 8);
    L.off_bx1 = take(S * 4);
    L.off_by1 = d > 1 ? take(S * 4) : 0;
    L.off_bz1 = d > 2 ? take(S * 4) : 0;
    L.off_cstart = take((W + 1) * 4);
    L.off_ckey = take(W * 8);
    L.off_pcell = take(W * 4);
    L.off_rows = take(W * (size_t)R * 8);
    L.off_cflag = take(W);
    L.off_fcore = take(W * 4);
    L.off_flags = take(W);
    L.off_parent = take(W * 4);
    L.off_rep = take(W * 4);
    L.off_tile_a = take(ntile * 4);
    L.off_tile_b = take(ntile * 4);
    L.off_in = take(W * d * 4);
    // label and cid adjacent: a host-buffer slide copies both down at once
    // when the caller's arrays are adjacent too
    L.off_label = take(W * 8);
    L.off_cid = L.off_label + W * 4;
    L.total = o;
    return L;
}

static void stream_shape(const rtdbs_params &p, int *k, int *R) {
    *k = p.cell_div > 0 ? p.cell_div : default_cell_div(p);
    int r = 1;
    for (int a = 0; a < p.dim - 1; a++) r *= 2 * (*k) + 1;
    *R = r;
}

size_t stream_device_bytes(const rtdbs_params &p) {
    int k, R;
    stream_shape(p, &k, &R);
    return make_layout(p, R).total;
}

}  // namespace rtdbs

// ---------------------------------------------------------------------------
// streams

extern "C" int rtdbs_stream_device_bytes(const rtdbs_params *params, size_t *bytes) {
    int rc = validate_params(params);
    if (rc) return rc;
    if (!bytes) return fail(RTDBS_EINVAL, "bytes is NULL");
    *bytes = stream_device_bytes(*params);
    return RTDBS_OK;
}

extern "C" int rtdbs_stream_create(const rtdbs_params *params, rtdbs_stream **out) {
    int rc = validate_params(params);
    if (rc) return rc;
    if (!out) return fail(RTDBS_EINVAL, "out is NULL");
    *out = nullptr;
    rtdbs_stream *s = new (std::nothrow) rtdbs_stream();
    if (!s) return fail(RTDBS_ENOMEM, "host allocation failed");
    s->p = *params;
    stream_shape(s->p, &s->k, &s->R);
    s->dense_ok = (s->k * s->k > s->p.dim) ? 1 : 0;
    s->slot_bits = std::max(1, ceil_log2(s->p.window));
    const Layout L = make_layout(s->p, s->R);
    s->bytes = L.total;
    if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
    if (alloc_hook_should_fail(L.total)) {
        delete s;
        return fail(RTDBS_ENOMEM, "device allocation of %zu bytes refused (allocation hook)", L.total);
    }
    hipError_t e = hipMalloc((void **)&s->arena, L.total);
    if (e != hipSuccess) {
        delete s;
        (void)hipGetLastError();
        return fail(RTDBS_ENOMEM, "hipMalloc(%zu) failed: %s", L.total, hipGetErrorString(e));
    }
    // zeroed state
    e = hipMemset(s->arena, 0, L.total);
    if (e != hipSuccess) {
        (void)hipFree(s->arena);
        delete s;
        return fail(RTDBS_EHIP, "hipMemset failed: %s", hipGetErrorString(e));
    }
    char *A = s->arena;
    StreamDev &D = s->dev;
    std::memset(&D, 0, sizeof(D));
    D.dim = s->p.dim;
    D.W = s->p.window;
    D.S = s->p.stride;
    D.min_pts = s->p.min_pts;
    D.k = s->k;
    D.R = s->R;
    D.slot_bits = s->slot_bits;
    D.units = s->p.window / s->p.stride;
    D.dense_ok = s->dense_ok;
    {
        volatile float e_ = s->p.eps;
        D.eps2 = e_ * e_;  // fl32(eps*eps), dbscan.cpp:41
    }
    D.mut = (StreamMut *)(A + L.off_mut);
    auto f = [&](size_t off, bool used) { return used ? (float *)(A + off) : (float *)nullptr; };
    const int d = s->p.dim;
    D.skey = (uint64_t *)(A + L.off_skey);
    D.sx = f(L.off_sx, true);
    D.sy = f(L.off_sy, d > 1);
    D.sz = f(L.off_sz, d > 2);
    D.akey = (uint64_t *)(A + L.off_akey);
    D.ax = f(L.off_ax, true);
    D.ay = f(L.off_ay, d > 1);
    D.az = f(L.off_az, d > 2);
    D.bkey0 = (uint64_t *)(A + L.off_bkey0);
    D.bx0 = f(L.off_bx0, true);
    D.by0 = f(L.off_by0, d > 1);
    D.bz0 = f(L.off_bz0, d > 2);
    D.bkey1 = (uint64_t *)(A + L.off_bkey1);
    D.bx1 = f(L.off_bx1, true);
    D.by1 = f(L.off_by1, d > 1);
    D.bz1 = f(L.off_bz1, d > 2);
    D.cstart = (int32_t *)(A + L.off_cstart);
    D.ckey = (uint64_t *)(A + L.off_ckey);
    D.pcell = (int32_t *)(A + L.off_pcell);
    D.rows = (int2 *)(A + L.off_rows);
    D.cflag = (uint8_t *)(A + L.off_cflag);
    D.fcore = (int32_t *)(A + L.off_fcore);
    D.flags = (uint8_t *)(A + L.off_flags);
    D.parent = (int32_t *)(A + L.off_parent);
    D.rep = (int32_t *)(A + L.off_rep);
    D.tile_a = (int32_t *)(A + L.off_tile_a);
    D.tile_b = (int32_t *)(A + L.off_tile_b);
    s->d_in = (float *)(A + L.off_in);
    s->d_label = (int32_t *)(A + L.off_label);
    s->d_cid = (int32_t *)(A + L.off_cid);
    *out = s;
    return RTDBS_OK;
}

namespace rtdbs {

// Grid geometry: side = eps/k * (1 + 2^-16). A pair accepted by the FP32 test
// has true distance below eps*(1 + 2^-21), so on every axis the exact cell
// indices differ by at most k; the double evaluation of an index adds at most
// 2^-22 cells for grids of up to 2^30 cells per axis.
//
// The packed cell key (i0 * n1 + i1) * n2 + i2 has no per-axis cap: each axis
// gets the cells its extent needs, as long as their product fits the key bits
// the window leaves (64 - slot bits, at most 52 so that a key is exact in a
// double). Without given bounds the grid spans the first window with margins
// for drift: twice its extent on each side (at least eps), or less where the
// key bits run out. Only beyond that are the axes shrunk, proportionally, and
// points outside clamped into the border cells, which keeps the search exact
// (never dense or safe) but slower.
static void set_geometry(rtdbs_stream *s, const float *pts) {
    const rtdbs_params &p = s->p;
    const int d = p.dim;
    s->side = (double)p.eps / (double)s->k * (1.0 + 1.0 / 65536.0);
    const double inv = 1.0 / s->side;
    const int key_bits = std::min(52, 64 - s->slot_bits);
    const double key_cells = std::ldexp(1.0, key_bits);
    const double axis_cap = std::ldexp(1.0, 30);
    bool given = true;
    for (int a = 0; a < d; a++)
        if (!(p.hi[a] > p.lo[a])) given = false;
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int a = 0; a < d; a++) {
        if (given) {
            mn[a] = p.lo[a];
            mx[a] = p.hi[a];
        } else {
            float lo = INFINITY, hi = -INFINITY;
            for (int i = 0; i < p.window; i++) {
                const float v = pts[(size_t)i * d + a];
                if (v < lo) lo = v;
                if (v > hi) hi = v;
            }
            if (!(hi >= lo)) lo = hi = 0.0f;  // all NaN
            mn[a] = lo;
            mx[a] = hi;
        }
    }
    // cells per axis for bounds [mn - m*ext - pad, mx + m*ext + pad]
    auto cells_for = [&](double m, double *lo_out, double *n_out) {
        double prod = 1.0;
        for (int a = 0; a < d; a++) {
            const double ext = mx[a] - mn[a];
            const double pad = given ? 0.0 : (double)p.eps;
            double lo = mn[a] - m * ext - pad;
            const double hi = mx[a] + m * ext + pad;
            if (!given) lo = (double)(float)lo;
            double cells = std::floor((hi - lo) * inv) + 1.0;
            if (!(cells >= 1.0)) cells = 1.0;
            cells = std::min(cells, axis_cap);
            lo_out[a] = lo;
            n_out[a] = cells;
            prod *= cells;
        }
        return prod;
    };
    double lo[3] = {0, 0, 0}, n[3] = {1, 1, 1};
    double m = given ? 0.0 : 2.0;
    double prod = cells_for(m, lo, n);
    while (!given && prod > key_cells && m > 0.25) {
        m *= 0.5;
        prod = cells_for(m, lo, n);
    }
    if (prod > key_cells) {
        // still too many cells: shrink every axis by the same factor (the
        // outermost points are then clamped into the border cells)
        const double f = std::pow(key_cells / prod, 1.0 / d) * (1.0 - 1e-9);
        for (int a = 0; a < d; a++) n[a] = std::max(1.0, std::floor(n[a] * f));
    }
    for (int a = 0; a < 3; a++) {
        s->lo[a] = a < d ? lo[a] : 0.0;
        s->n[a] = a < d ? (int64_t)n[a] : 1;
    }
    StreamDev &D = s->dev;
    D.inv_side = inv;
    for (int a = 0; a < 3; a++) {
        D.lo[a] = s->lo[a];
        D.n[a] = s->n[a];
    }
}

}  // namespace rtdbs

static int stream_load(rtdbs_stream *s, const float *points, int64_t slides_done) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    if (!points) return fail(RTDBS_EINVAL, "points is NULL");
    if (slides_done < 0) return fail(RTDBS_EINVAL, "slides_done must be >= 0");
    HIPCHK(hipSetDevice(s->device));
    if (s->solo) {
        int rc = batch_sync(s->solo);
        if (rc) return rc;
    }
    set_geometry(s, points);
    s->gen = next_generation();
    int rc = ensure_solo(s);
    if (rc) return rc;
    rtdbs_batch *b = s->solo;
    rc = refresh_batch(b);
    if (rc) return rc;
    const size_t W = (size_t)s->p.window, d = (size_t)s->p.dim;
    const int units = s->p.window / s->p.stride;
    HIPCHK(hipMemcpyAsync(s->d_in, points, W * d * 4, hipMemcpyHostToDevice, b->hs));
    StreamMut m0 = {(int32_t)(slides_done % (1 << 30)), (int32_t)(slides_done % units), 0, {0, 0, 0, 0, 0}};
    HIPCHK(hipMemcpyAsync(s->dev.mut, &m0, sizeof(m0), hipMemcpyHostToDevice, b->hs));
    BatchArgs a = b->args;
    a.use_io1 = 1;
    a.io1.in = s->d_in;
    a.io1.label = s->d_label;
    a.io1.cid = s->d_cid;
    const int itiles = (int)((W + kTile - 1) / kTile);
    int passes = ceil_log2(itiles);
    if (passes % 2 == 0) passes++;  // result must land in the window arrays
    HIPCHK(launch_init(a, itiles, passes, b->hs));
    HIPCHK(hipStreamSynchronize(b->hs));
    s->initialized = true;
    s->slides = slides_done;
    return RTDBS_OK;
}

extern "C" int rtdbs_stream_init(rtdbs_stream *s, const float *points) { return stream_load(s, points, 0); }

extern "C" int rtdbs_stream_restore(rtdbs_stream *s, const float *window_points, int64_t slides_done) {
    return stream_load(s, window_points, slides_done);
}

extern "C" int rtdbs_stream_slide(rtdbs_stream *s, const float *stride_points, int32_t *label, int32_t *cid) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    if (!stride_points || !label || !cid) return fail(RTDBS_EINVAL, "NULL buffer");
    if (!s->initialized) return fail(RTDBS_ESTATE, "slide before rtdbs_stream_init");
    HIPCHK(hipSetDevice(s->device));
    return batch_slide_host(s->solo, &stride_points, &label, &cid);
}

extern "C" int rtdbs_stream_slide_async(rtdbs_stream *s, const float *stride_points_dev, int32_t *label_dev,
                                        int32_t *cid_dev) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    if (!stride_points_dev || !label_dev || !cid_dev) return fail(RTDBS_EINVAL, "NULL buffer");
    if (!s->initialized) return fail(RTDBS_ESTATE, "slide before rtdbs_stream_init");
    HIPCHK(hipSetDevice(s->device));
    return batch_launch_solo(s->solo, stride_points_dev, label_dev, cid_dev);
}

extern "C" int rtdbs_stream_sync(rtdbs_stream *s) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    if (!s->solo) return RTDBS_OK;
    HIPCHK(hipSetDevice(s->device));
    return batch_sync(s->solo);
}

extern "C" int rtdbs_stream_get_info(const rtdbs_stream *s, rtdbs_stream_info *info) {
    if (!s || !info) return fail(RTDBS_EINVAL, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    info->slides = s->slides;
    info->cell_div = s->k;
    info->rows = s->R;
    info->dense_shortcut = s->dense_ok;
    for (int a = 0; a < 3; a++) {
        info->grid[a] = s->n[a];
        info->lo[a] = s->lo[a];
    }
    info->cell_side = s->side;
    info->device_bytes = s->bytes;
    if (s->initialized) {
        StreamMut m;
        if (hipMemcpy(&m, s->dev.mut, sizeof(m), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(RTDBS_EHIP, "reading stream state failed");
        info->cells = m.ncell;
    }
    return RTDBS_OK;
}

extern "C" int rtdbs_stream_get_window(rtdbs_stream *s, float *points) {
    if (!s || !points) return fail(RTDBS_EINVAL, "NULL argument");
    if (!s->initialized) return fail(RTDBS_ESTATE, "stream not initialised");
    HIPCHK(hipSetDevice(s->device));
    if (s->solo) {
        int rc = batch_sync(s->solo);
        if (rc) return rc;
    }
    const size_t W = (size_t)s->p.window;
    const int d = s->p.dim;
    std::vector<uint64_t> key(W);
    std::vector<float> c[3];
    HIPCHK(hipMemcpy(key.data(), s->dev.skey, W * 8, hipMemcpyDeviceToHost));
    const float *src[3] = {s->dev.sx, s->dev.sy, s->dev.sz};
    for (int a = 0; a < d; a++) {
        c[a].resize(W);
        HIPCHK(hipMemcpy(c[a].data(), src[a], W * 4, hipMemcpyDeviceToHost));
    }
    const uint64_t m = (1ull << s->slot_bits) - 1ull;
    for (size_t i = 0; i < W; i++) {
        const size_t slot = (size_t)(key[i] & m);
        if (slot >= W) return fail(RTDBS_EHIP, "corrupt window state");
        for (int a = 0; a < d; a++) points[slot * d + a] = c[a][i];
    }
    return RTDBS_OK;
}

extern "C" void rtdbs_stream_destroy(rtdbs_stream *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->solo) {
        batch_free(s->solo);
        s->solo = nullptr;
    }
    if (s->arena) (void)hipFree(s->arena);
    delete s;
}

// ---------------------------------------------------------------------------
// batches

namespace rtdbs {

static int batch_alloc(rtdbs_batch *b) {
    const int n = b->capacity;
    const size_t desc_b = ((size_t)n * sizeof(StreamDev) + 255) & ~size_t(255);
    const size_t io_b = ((size_t)n * sizeof(IO) + 255) & ~size_t(255);
    const size_t pre_b = ((size_t)4 * (n + 1) * 4 + 255) & ~size_t(255);
    const size_t cnt_b = ((size_t)n * kNumStats * 8 + 255) & ~size_t(255);
    const size_t cpre_b = ((size_t)(n + 1) * 4 + 255) & ~size_t(255);
    HIPCHK(hipMalloc((void **)&b->dmem, desc_b + io_b + pre_b + cnt_b + cpre_b));
    b->d_cell_pre = (int32_t *)(b->dmem + desc_b + io_b + pre_b + cnt_b);
    b->d_desc = (StreamDev *)b->dmem;
    b->d_io = (IO *)(b->dmem + desc_b);
    b->d_pre = (int32_t *)(b->dmem + desc_b + io_b);
    b->d_stats = (unsigned long long *)(b->dmem + desc_b + io_b + pre_b);
    HIPCHK(hipHostMalloc((void **)&b->h_io, (size_t)n * sizeof(IO), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&b->h_desc, (size_t)n * sizeof(StreamDev), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&b->h_pre, (size_t)4 * (n + 1) * 4, hipHostMallocDefault));
    HIPCHK(hipStreamCreateWithFlags(&b->hs, hipStreamNonBlocking));
    return RTDBS_OK;
}

rtdbs_batch *batch_new(int capacity, bool use_graph, int *rc) {
    rtdbs_batch *b = new (std::nothrow) rtdbs_batch();
    if (!b) {
        *rc = fail(RTDBS_ENOMEM, "host allocation failed");
        return nullptr;
    }
    b->capacity = capacity;
    b->use_graph = use_graph;
    *rc = batch_alloc(b);
    if (*rc) {
        batch_free(b);
        return nullptr;
    }
    return b;
}

int batch_set_members(rtdbs_batch *b, rtdbs_stream *const *streams, int n) {
    if (n < 1 || n > b->capacity) return fail(RTDBS_EINVAL, "batch size %d outside [1, %d]", n, b->capacity);
    bool same = n == b->n;
    for (int i = 0; i < n && same; i++) same = b->streams[i] == streams[i] && b->gens[i] == streams[i]->gen;
    if (same) return RTDBS_OK;
    // the device may still read the old tables
    if (b->hs) HIPCHK(hipStreamSynchronize(b->hs));
    b->n = n;
    b->streams.assign(streams, streams + n);
    b->gens.assign(n, 0);
    b->stage_io.resize(n);
    for (int i = 0; i < n; i++) b->stage_io[i] = IO{streams[i]->d_in, streams[i]->d_label, streams[i]->d_cid};
    b->xfer_ready = false;  // host-buffer transfer layout follows the members
    return refresh_batch(b);
}

static void pipe_drop_graphs(rtdbs_batch *b) {
    for (int p = 0; p < 2; p++) {
        if (b->pipe.gexec[p]) (void)hipGraphExecDestroy(b->pipe.gexec[p]);
        if (b->pipe.graph[p]) (void)hipGraphDestroy(b->pipe.graph[p]);
        b->pipe.gexec[p] = nullptr;
        b->pipe.graph[p] = nullptr;
    }
}

void batch_free(rtdbs_batch *b) {
    if (!b) return;
    (void)pipe_drain(b);
    if (b->pipe.up) (void)hipStreamSynchronize(b->pipe.up);
    if (b->pipe.down) (void)hipStreamSynchronize(b->pipe.down);
    if (b->hs) (void)hipStreamSynchronize(b->hs);
    pipe_drop_graphs(b);
    for (int p = 0; p < 2; p++) {
        if (b->pipe.area[p]) (void)hipFree(b->pipe.area[p]);
        for (hipEvent_t e : {b->pipe.in_done[p], b->pipe.kern_done[p], b->pipe.out_done[p]})
            if (e) (void)hipEventDestroy(e);
    }
    if (b->pipe.d_io) (void)hipFree(b->pipe.d_io);
    if (b->pipe.h_io) (void)hipHostFree(b->pipe.h_io);
    if (b->pipe.up) (void)hipStreamDestroy(b->pipe.up);
    if (b->pipe.down) (void)hipStreamDestroy(b->pipe.down);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t e : b->ev_ring)
        if (e) (void)hipEventDestroy(e);
    if (b->hs) (void)hipStreamDestroy(b->hs);
    if (b->h_io) (void)hipHostFree(b->h_io);
    if (b->h_desc) (void)hipHostFree(b->h_desc);
    if (b->h_pre) (void)hipHostFree(b->h_pre);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->d_xfer) (void)hipFree(b->d_xfer);
    if (b->dmem) (void)hipFree(b->dmem);
    delete b;
}

// Integer launch-shape override from the environment, `def` when unset or
// outside [lo, hi].
static int env_int(const char *name, int def, int lo, int hi) {
    const char *e = std::getenv(name);
    if (!e || !*e) return def;
    const int x = std::atoi(e);
    return (x >= lo && x <= hi) ? x : def;
}

// Windows up to this many points in one launch use the cell-parallel count
// and border kernels.
constexpr long long kCellParMaxPoints = 1000000;

// One-wave workgroups per CU for the persistent cell-level merge
// (RTDBS_WAVES_PER_CU overrides, 1..64).
static int waves_per_cu() { return env_int("RTDBS_WAVES_PER_CU", 16, 1, 64); }

// Uploads descriptors and prefix tables; drops a captured graph. Called when
// the batch is built and whenever a member stream was re-initialised.
int refresh_batch(rtdbs_batch *b) {
    const int n = b->n;
    // the device may still read the tables (and the copy from their pinned
    // mirrors may be pending)
    int rc0 = pipe_drain(b);
    if (rc0) return rc0;
    if (b->hs) HIPCHK(hipStreamSynchronize(b->hs));
    StreamDev *desc = b->h_desc;
    int32_t *pre = b->h_pre;
    std::memset(pre, 0, (size_t)4 * (n + 1) * 4);
    long long rows_total = 0, cells_total = 0;
    for (int i = 0; i < n; i++) {
        rtdbs_stream *s = b->streams[i];
        desc[i] = s->dev;
        const int st = (s->p.stride + kTile - 1) / kTile;
        const int wt = (s->p.window + kWTile - 1) / kWTile;
        const int ch = (s->p.window + kChunk - 1) / kChunk;
        const int it = (s->p.window + kTile - 1) / kTile;
        pre[0 * (n + 1) + i + 1] = pre[0 * (n + 1) + i] + st;
        pre[1 * (n + 1) + i + 1] = pre[1 * (n + 1) + i] + wt;
        pre[2 * (n + 1) + i + 1] = pre[2 * (n + 1) + i] + ch;
        pre[3 * (n + 1) + i + 1] = pre[3 * (n + 1) + i] + it;
        rows_total += (long long)s->p.window * s->R;
        cells_total += s->p.window;
        b->gens[i] = s->gen;
    }
    // asynchronous uploads from the pinned mirrors, ordered before the next
    // launch on the batch's stream
    HIPCHK(hipMemcpyAsync(b->d_desc, desc, (size_t)n * sizeof(StreamDev), hipMemcpyHostToDevice, b->hs));
    HIPCHK(hipMemcpyAsync(b->d_pre, pre, (size_t)4 * (n + 1) * 4, hipMemcpyHostToDevice, b->hs));
    b->shape.stiles = pre[0 * (n + 1) + n];
    b->shape.wtiles = pre[1 * (n + 1) + n];
    b->shape.chunks = pre[2 * (n + 1) + n];
    // stride rank: 64-key blocks when 256-key blocks would not cover the device
    // (RTDBS_RANK_SPLIT: 0 forces 256-key blocks, 1 the 64-key form)
    const int rsplit = env_int("RTDBS_RANK_SPLIT", -1, 0, 1);
    b->shape.rank_split =
        rsplit >= 0 ? rsplit : ((long long)b->shape.stiles * (kTile / kTileThreads) < device_cus() ? 1 : 0);
    const long long pb = (rows_total + kTileThreads - 1) / kTileThreads;
    // a single stream's own tuned grid sizes (0: by launch size)
    const rtdbs_params *p1 = n == 1 ? &b->streams[0]->p : nullptr;
    // rows: a persistent grid of blocks per CU (measured best: 8 for a
    // latency-bound launch, 32 for a large batch); RTDBS_ROWS_PER_CU overrides
    const int r_def = p1 && p1->rows_per_cu > 0 ? p1->rows_per_cu : (cells_total <= kCellParMaxPoints ? 8 : 32);
    const int rcap = env_int("RTDBS_ROWS_PER_CU", r_def, 1, kMaxPerCu);
    b->shape.rows_blocks = (int)std::max(1ll, std::min((long long)device_cus() * rcap, pb));
    // Count and border: cell-parallel (coalesced candidate loads, more
    // instructions per query) while a launch is small and latency-bound,
    // point-parallel once its total window saturates the device.
    // RTDBS_NEIGHBOR_MODE: 1 forces cell-parallel, 2 point-parallel (over a
    // single stream's own neighbor_mode).
    const int nmode = env_int("RTDBS_NEIGHBOR_MODE", 0, 0, 2);
    const int pmode = n == 1 ? b->streams[0]->p.neighbor_mode : 0;  // a single stream's own setting
    const int mode = nmode ? nmode : pmode;
    b->shape.cellpar = mode == 1 || (mode == 0 && cells_total <= kCellParMaxPoints);
    // the range count's own form (RTDBS_COUNT_MODE: 1 cell-parallel, 2
    // point-parallel; otherwise as the output kernel)
    const int cmode = env_int("RTDBS_COUNT_MODE", 0, 0, 2);
    b->shape.count_cell = cmode ? (cmode == 1) : b->shape.cellpar;
    // the cell-parallel merge also wins in large batches (RTDBS_UNION_CELLS=1
    // selects the older cell-level merge kernels there)
    b->shape.union_cells = env_int("RTDBS_UNION_CELLS", 0, 0, 1);
    // cell-parallel grid: 16 blocks per CU while the launch is small; 32 for
    // large batches, where it carries the merge alone (64 streams: 455 ->
    // ~400 us). RTDBS_CELLPAR_PER_CU overrides.
    const int cp_def =
        p1 && p1->cellpar_per_cu > 0 ? p1->cellpar_per_cu : (cells_total <= kCellParMaxPoints ? 16 : 32);
    b->shape.cellpar_blocks = (int)std::max(
        1ll, std::min((long long)device_cus() * env_int("RTDBS_CELLPAR_PER_CU", cp_def, 1, kMaxPerCu),
                      (cells_total + 3) / 4));
    b->shape.cell_blocks = (int)std::max(
        1ll, std::min((long long)device_cus() * waves_per_cu(), (cells_total + kUGCells - 1) / kUGCells));
    // lanes per query: enough threads to cover the device several times over,
    // unless a single stream asks for a tuned setting
    const long long lanes_wanted = (long long)device_cus() * 32 * 64 * 4;
    b->shape.lpq = cells_total * 4 <= lanes_wanted ? 4 : (cells_total * 2 <= lanes_wanted ? 2 : 1);
    // the output kernel walks every stencil row of a non-core point (no early
    // exit), so its per-lane chains are longer: 4 lanes per query, also where
    // the count takes 1 (64 streams: 238 vs 284 us; RTDBS_BORDER_LPQ=1|2|4
    // overrides)
    b->shape.lpq_border = env_int("RTDBS_BORDER_LPQ", 4, 1, 4);
    // point-parallel outputs: one lane per core and four per non-core point
    // (RTDBS_BORDER_REGROUP=0: lpq_border lanes for every position)
    b->shape.border_regroup = env_int("RTDBS_BORDER_REGROUP", 1, 0, 1);
    if (n == 1 && b->streams[0]->p.lanes_per_query > 0) {
        // a single stream's own setting (the autotuner's point-parallel forms)
        b->shape.lpq = b->shape.lpq_border = b->streams[0]->p.lanes_per_query;
        b->shape.border_regroup = 0;
    }
    (void)cells_total;
    BatchArgs &a = b->args;
    std::memset(&a, 0, sizeof(a));
    a.s = b->d_desc;
    a.io = b->d_io;
    a.n = n;
    a.pre_stile = b->d_pre;
    a.pre_wtile = b->d_pre + (n + 1);
    a.pre_chunk = b->d_pre + 2 * (n + 1);
    a.pre_itile = b->d_pre + 3 * (n + 1);
    a.stats = b->stats_on ? b->d_stats : nullptr;
    a.cell_pre = b->d_cell_pre;
    // XCD-grouped work order: in large batches for the merge and the
    // point-parallel output kernel, where it pays (RTDBS_XCD_GROUPS sets the
    // kXcd* bits)
    a.xcd_groups = env_int("RTDBS_XCD_GROUPS", cells_total > kCellParMaxPoints ? kXcdUnion | kXcdBorderPoints : 0, 0, 3);
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    pipe_drop_graphs(b);
    b->pipe.ready = false;  // the areas' layout and tables follow the members
    b->io_valid = false;
    return RTDBS_OK;
}

static int batch_check_current(rtdbs_batch *b) {
    for (int i = 0; i < b->n; i++) {
        if (!b->streams[i]->initialized) return fail(RTDBS_ESTATE, "stream %d of the batch is not initialised", i);
        if (b->gens[i] != b->streams[i]->gen) return refresh_batch(b);
    }
    return RTDBS_OK;
}

int ensure_solo(rtdbs_stream *s) {
    if (s->solo) return RTDBS_OK;
    int rc = 0;
    rtdbs_batch *b = batch_new(1, false, &rc);
    if (!b) return rc;
    b->solo = true;
    rtdbs_stream *one[1] = {s};
    rc = batch_set_members(b, one, 1);
    if (rc) {
        batch_free(b);
        return rc;
    }
    s->solo = b;
    return RTDBS_OK;
}

// Event set for the next timed slide (reads the full ring first).
static int timing_events(rtdbs_batch *b, hipEvent_t **ev) {
    *ev = nullptr;
    if (!b->timing) return RTDBS_OK;
    if (b->ring_used == kTimingRing) {
        int rc = batch_collect_timing(b);
        if (rc) return rc;
    }
    *ev = b->ev_ring.data() + (size_t)b->ring_used * 2 * kNumPhases;
    b->ring_used++;
    return RTDBS_OK;
}

int batch_launch_solo(rtdbs_batch *b, const float *in, int32_t *label, int32_t *cid) {
    int rc = batch_check_current(b);
    if (rc) return rc;
    BatchArgs a = b->args;
    a.use_io1 = 1;
    a.io1.in = in;
    a.io1.label = label;
    a.io1.cid = cid;
    hipEvent_t *ev = nullptr;
    rc = timing_events(b, &ev);
    if (rc) return rc;
    HIPCHK(launch_slide(a, b->shape, b->hs, ev, b->timing_mask));
    for (auto *s : b->streams) s->slides++;
    if (b->stats_on) b->stats_slides++;
    return RTDBS_OK;
}

// Adds the elapsed times of every recorded slide to the totals.
int batch_collect_timing(rtdbs_batch *b) {
    if (b->ring_used == 0) return RTDBS_OK;
    HIPCHK(hipStreamSynchronize(b->hs));
    for (int k = 0; k < b->ring_used; k++) {
        const hipEvent_t *ev = b->ev_ring.data() + (size_t)k * 2 * kNumPhases;
        for (int ph = 0; ph < kNumPhases; ph++) {
            if (!((b->timing_mask >> ph) & 1u)) continue;
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev[2 * ph], ev[2 * ph + 1]));
            b->acc_ms[ph] += ms;
            b->acc_n[ph]++;
        }
    }
    b->ring_used = 0;
    return RTDBS_OK;
}

int batch_sync(rtdbs_batch *b) {
    int rc = pipe_drain(b);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->hs));
    return RTDBS_OK;
}

int pipe_drain(rtdbs_batch *b) {
    if (!b->pipe.waiting) return RTDBS_OK;
    b->pipe.waiting = false;
    HIPCHK(hipEventSynchronize(b->pipe.out_done[(b->pipe.issued - 1) & 1]));
    return RTDBS_OK;
}

// One slide of all member streams with device I/O (table in d_io).
int batch_launch(rtdbs_batch *b, const IO *io) {
    int rc = pipe_drain(b);
    if (rc) return rc;
    rc = batch_check_current(b);
    if (rc) return rc;
    bool changed = !b->io_valid;
    for (int i = 0; i < b->n && !changed; i++)
        if (b->h_io[i].in != io[i].in || b->h_io[i].label != io[i].label || b->h_io[i].cid != io[i].cid)
            changed = true;
    if (changed) {
        // the previous slide may still read the table
        HIPCHK(hipStreamSynchronize(b->hs));
        for (int i = 0; i < b->n; i++) b->h_io[i] = io[i];
        HIPCHK(hipMemcpyAsync(b->d_io, b->h_io, (size_t)b->n * sizeof(IO), hipMemcpyHostToDevice, b->hs));
        b->io_valid = true;
    }
    BatchArgs a = b->args;
    a.use_io1 = 0;
    if (b->use_graph && !b->timing) {
        if (!b->gexec) {
            HIPCHK(hipStreamBeginCapture(b->hs, hipStreamCaptureModeThreadLocal));
            hipError_t le = launch_slide(a, b->shape, b->hs, nullptr, 0u);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(b->hs, &g);
            if (le != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(RTDBS_EHIP, "graph capture launch failed: %s", hipGetErrorString(le));
            }
            if (ce != hipSuccess) return fail(RTDBS_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ce));
            b->graph = g;
            HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        HIPCHK(hipGraphLaunch(b->gexec, b->hs));
    } else {
        hipEvent_t *ev = nullptr;
        rc = timing_events(b, &ev);
        if (rc) return rc;
        HIPCHK(launch_slide(a, b->shape, b->hs, ev, b->timing_mask));
    }
    for (auto *s : b->streams) s->slides++;
    if (b->stats_on) b->stats_slides++;
    return RTDBS_OK;
}

// ---------------------------------------------------------------------------
// host-buffer slides

struct Copy {
    void *dst;
    const void *src;
    size_t bytes;
};

// Appends a copy, merged into the previous one when both ranges abut.
static void add_copy(std::vector<Copy> &v, void *dst, const void *src, size_t bytes) {
    if (!v.empty()) {
        Copy &c = v.back();
        if ((char *)c.dst + c.bytes == (char *)dst && (const char *)c.src + c.bytes == (const char *)src) {
            c.bytes += bytes;
            return;
        }
    }
    v.push_back(Copy{dst, src, bytes});
}

// Offsets of every member's stride and outputs in the transfer area and the
// pinned staging (same layout); (re)allocates both when they are too small.
static int ensure_xfer(rtdbs_batch *b) {
    if (b->xfer_ready) return RTDBS_OK;
    const int n = b->n;
    b->in_off.assign(n, 0);
    b->out_off.assign(n, 0);
    size_t o = 0;
    for (int i = 0; i < n; i++) {
        const rtdbs_stream *s = b->streams[i];
        b->in_off[i] = o;
        o += (size_t)s->p.stride * s->p.dim * 4;
    }
    o = (o + 255) & ~size_t(255);
    for (int i = 0; i < n; i++) {
        b->out_off[i] = o;
        o += (size_t)b->streams[i]->p.window * 8;
    }
    const size_t total = o;
    // grown geometrically: a server lane's members change with every batch
    const size_t grow = std::max(total, total + total / 2);
    if (!b->solo && total > b->d_xfer_bytes) {
        HIPCHK(hipStreamSynchronize(b->hs));
        const size_t want = std::max(grow, b->d_xfer_bytes * 2);
        if (b->d_xfer) (void)hipFree(b->d_xfer);
        b->d_xfer = nullptr;
        b->d_xfer_bytes = 0;
        hipError_t e = hipMalloc((void **)&b->d_xfer, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(RTDBS_ENOMEM, "transfer area of %zu bytes: %s", want, hipGetErrorString(e));
        }
        b->d_xfer_bytes = want;
    }
    if (total > b->h_stage_bytes) {
        HIPCHK(hipStreamSynchronize(b->hs));
        if (b->h_stage) (void)hipHostFree(b->h_stage);
        b->h_stage = nullptr;
        b->h_stage_bytes = 0;
        const size_t want = b->solo ? total : grow;
        hipError_t e = hipHostMalloc((void **)&b->h_stage, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(RTDBS_ENOMEM, "pinned staging of %zu bytes: %s", want, hipGetErrorString(e));
        }
        b->h_stage_bytes = want;
    }
    if (!b->solo) {
        for (int i = 0; i < n; i++) {
            const size_t W = (size_t)b->streams[i]->p.window;
            b->stage_io[i] = IO{(const float *)(b->d_xfer + b->in_off[i]), (int32_t *)(b->d_xfer + b->out_off[i]),
                                (int32_t *)(b->d_xfer + b->out_off[i] + W * 4)};
        }
    }
    b->xfer_ready = true;
    return RTDBS_OK;
}

int batch_slide_host(rtdbs_batch *b, const float *const *in, int32_t *const *labels, int32_t *const *cids) {
    int rc = pipe_drain(b);
    if (rc) return rc;
    rc = batch_check_current(b);
    if (rc) return rc;
    rc = ensure_xfer(b);
    if (rc) return rc;
    std::vector<Copy> up, down, post;
    for (int i = 0; i < b->n; i++) {
        const rtdbs_stream *s = b->streams[i];
        const size_t nb = (size_t)s->p.stride * s->p.dim * 4;
        void *dev = (void *)b->stage_io[i].in;
        if (host_pinned(in[i], nb)) {
            add_copy(up, dev, in[i], nb);
        } else {
            char *h = b->h_stage + b->in_off[i];
            std::memcpy(h, in[i], nb);
            add_copy(up, dev, h, nb);
        }
    }
    for (const Copy &c : up) HIPCHK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, b->hs));
    if (b->solo)
        rc = batch_launch_solo(b, b->stage_io[0].in, b->stage_io[0].label, b->stage_io[0].cid);
    else
        rc = batch_launch(b, b->stage_io.data());
    if (rc) return rc;
    for (int i = 0; i < b->n; i++) {
        const size_t nb = (size_t)b->streams[i]->p.window * 4;
        const IO &io = b->stage_io[i];
        int32_t *dst[2] = {labels[i], cids[i]};
        const int32_t *src[2] = {io.label, io.cid};
        for (int k = 0; k < 2; k++) {
            if (host_pinned(dst[k], nb)) {
                add_copy(down, dst[k], src[k], nb);
            } else {
                char *h = b->h_stage + b->out_off[i] + (size_t)k * nb;
                add_copy(down, h, src[k], nb);
                post.push_back(Copy{dst[k], h, nb});
            }
        }
    }
    for (const Copy &c : down) HIPCHK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, b->hs));
    rc = batch_sync(b);
    if (rc) return rc;
    for (const Copy &c : post) std::memcpy(c.dst, c.src, c.bytes);
    return RTDBS_OK;
}

// Areas, tables, streams and events of the pipelined slides, for the current
// members (the caller has drained the pipe).
static int pipe_setup(rtdbs_batch *b) {
    auto &P = b->pipe;
    if (P.ready) return RTDBS_OK;
    int rc = ensure_xfer(b);  // the transfer area's layout (in_off, out_off)
    if (rc) return rc;
    const int n = b->n;
    const size_t total = b->out_off[n - 1] + (size_t)b->streams[n - 1]->p.window * 8;
    for (hipStream_t s : {P.up, P.down})
        if (s) HIPCHK(hipStreamSynchronize(s));
    if (total > P.area_bytes) {
        for (int p = 0; p < 2; p++) {
            if (P.area[p]) (void)hipFree(P.area[p]);
            P.area[p] = nullptr;
        }
        P.area_bytes = 0;
        const size_t want = std::max(total, total + total / 2);
        for (int p = 0; p < 2; p++) {
            hipError_t e = hipMalloc((void **)&P.area[p], want);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                return fail(RTDBS_ENOMEM, "pipelined transfer area of %zu bytes: %s", want, hipGetErrorString(e));
            }
        }
        P.area_bytes = want;
    }
    if (!P.d_io) {
        HIPCHK(hipMalloc((void **)&P.d_io, (size_t)2 * b->capacity * sizeof(IO)));
        HIPCHK(hipHostMalloc((void **)&P.h_io, (size_t)2 * b->capacity * sizeof(IO), hipHostMallocDefault));
    }
    if (!P.up) {
        HIPCHK(hipStreamCreateWithFlags(&P.up, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&P.down, hipStreamNonBlocking));
        for (int p = 0; p < 2; p++)
            for (hipEvent_t *e : {&P.in_done[p], &P.kern_done[p], &P.out_done[p]})
                HIPCHK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    for (int p = 0; p < 2; p++)
        for (int i = 0; i < n; i++) {
            char *area = P.area[p];
            int32_t *label = (int32_t *)(area + b->out_off[i]);
            P.h_io[p * b->capacity + i] =
                IO{(const float *)(area + b->in_off[i]), label, label + b->streams[i]->p.window};
        }
    HIPCHK(hipMemcpyAsync(P.d_io, P.h_io, (size_t)2 * b->capacity * sizeof(IO), hipMemcpyHostToDevice, b->hs));
    HIPCHK(hipStreamSynchronize(b->hs));
    pipe_drop_graphs(b);
    P.issued = 0;
    P.ready = true;
    return RTDBS_OK;
}

// One pipelined slide (rtdbs_batch_slide_pipelined). Area p = slide index & 1:
//   up stream:      [wait: kernels of slide t-2 done]  strides -> area p      -> in_done[p]
//   batch's stream: [wait: in_done[p], outputs of slide t-2 copied]  slide    -> kern_done[p]
//   down stream:    [wait: kern_done[p]]  labels, ids of area p -> host       -> out_done[p]
// then waits for out_done of slide t-1. So slide t's kernels run while slide
// t-1's outputs go down and slide t+1's strides go up.
int batch_slide_pipelined(rtdbs_batch *b, const float *const *in, int32_t *const *labels, int32_t *const *cids) {
    if (b->solo) return fail(RTDBS_EINVAL, "pipelined slides are for batches");
    int rc = batch_check_current(b);  // may refresh (which drains the pipe)
    if (rc) return rc;
    for (int i = 0; i < b->n; i++) {
        const rtdbs_stream *s = b->streams[i];
        const size_t ni = (size_t)s->p.stride * s->p.dim * 4, no = (size_t)s->p.window * 4;
        if (!host_pinned(in[i], ni) || !host_pinned(labels[i], no) || !host_pinned(cids[i], no))
            return fail(RTDBS_EINVAL, "pipelined slides take rtdbs_host_alloc buffers only (stream %d)", i);
    }
    if (!b->pipe.ready) {
        rc = pipe_drain(b);
        if (rc) return rc;
        rc = pipe_setup(b);
        if (rc) return rc;
    }
    auto &P = b->pipe;
    const int p = (int)(P.issued & 1);
    char *area = P.area[p];
    std::vector<Copy> up, down;
    for (int i = 0; i < b->n; i++) {
        const rtdbs_stream *s = b->streams[i];
        add_copy(up, area + b->in_off[i], in[i], (size_t)s->p.stride * s->p.dim * 4);
    }
    for (int i = 0; i < b->n; i++) {
        const size_t nb = (size_t)b->streams[i]->p.window * 4;
        add_copy(down, labels[i], area + b->out_off[i], nb);
        add_copy(down, cids[i], area + b->out_off[i] + nb, nb);
    }
    if (P.issued >= 2) HIPCHK(hipStreamWaitEvent(P.up, P.kern_done[p], 0));
    for (const Copy &c : up) HIPCHK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, P.up));
    HIPCHK(hipEventRecord(P.in_done[p], P.up));
    HIPCHK(hipStreamWaitEvent(b->hs, P.in_done[p], 0));
    if (P.issued >= 2) HIPCHK(hipStreamWaitEvent(b->hs, P.out_done[p], 0));
    BatchArgs a = b->args;
    a.use_io1 = 0;
    a.io = P.d_io + (size_t)p * b->capacity;
    if (b->use_graph) {
        if (!P.gexec[p]) {
            HIPCHK(hipStreamBeginCapture(b->hs, hipStreamCaptureModeThreadLocal));
            hipError_t le = launch_slide(a, b->shape, b->hs, nullptr, 0u);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(b->hs, &g);
            if (le != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(RTDBS_EHIP, "graph capture launch failed: %s", hipGetErrorString(le));
            }
            if (ce != hipSuccess) return fail(RTDBS_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ce));
            P.graph[p] = g;
            HIPCHK(hipGraphInstantiate(&P.gexec[p], g, nullptr, nullptr, 0));
        }
        HIPCHK(hipGraphLaunch(P.gexec[p], b->hs));
    } else {
        HIPCHK(launch_slide(a, b->shape, b->hs, nullptr, 0u));
    }
    HIPCHK(hipEventRecord(P.kern_done[p], b->hs));
    HIPCHK(hipStreamWaitEvent(P.down, P.kern_done[p], 0));
    for (const Copy &c : down) HIPCHK(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, P.down));
    HIPCHK(hipEventRecord(P.out_done[p], P.down));
    for (auto *s : b->streams) s->slides++;
    if (b->stats_on) b->stats_slides++;
    const bool prev = P.waiting;
    P.issued++;
    P.waiting = true;
    if (prev) HIPCHK(hipEventSynchronize(P.out_done[p ^ 1]));  // the previous slide's outputs
    return RTDBS_OK;
}

}  // namespace rtdbs

extern "C" int rtdbs_batch_create(rtdbs_stream *const *streams, int n, int use_graph, rtdbs_batch **out) {
    if (!out) return fail(RTDBS_EINVAL, "out is NULL");
    *out = nullptr;
    if (!streams || n < 1 || n > kMaxBatch) return fail(RTDBS_EINVAL, "batch size must be in [1, %d]", kMaxBatch);
    for (int i = 0; i < n; i++) {
        if (!streams[i]) return fail(RTDBS_EINVAL, "stream %d is NULL", i);
        if (!streams[i]->initialized) return fail(RTDBS_ESTATE, "stream %d is not initialised", i);
        for (int j = 0; j < i; j++)
            if (streams[j] == streams[i]) return fail(RTDBS_EINVAL, "stream %d appears twice", i);
    }
    int rc = 0;
    rtdbs_batch *b = batch_new(n, use_graph != 0, &rc);
    if (!b) return rc;
    rc = batch_set_members(b, streams, n);
    if (rc) {
        batch_free(b);
        return rc;
    }
    *out = b;
    return RTDBS_OK;
}

extern "C" int rtdbs_batch_slide(rtdbs_batch *b, const float *const *stride_points, int32_t *const *labels,
                                 int32_t *const *cids) {
    if (!b || !stride_points || !labels || !cids) return fail(RTDBS_EINVAL, "NULL argument");
    for (int i = 0; i < b->n; i++)
        if (!stride_points[i] || !labels[i] || !cids[i]) return fail(RTDBS_EINVAL, "NULL buffer for stream %d", i);
    return batch_slide_host(b, stride_points, labels, cids);
}

extern "C" int rtdbs_batch_slide_async(rtdbs_batch *b, const float *const *stride_points_dev,
                                       int32_t *const *labels_dev, int32_t *const *cids_dev) {
    if (!b || !stride_points_dev || !labels_dev || !cids_dev) return fail(RTDBS_EINVAL, "NULL argument");
    std::vector<IO> io(b->n);
    for (int i = 0; i < b->n; i++) {
        if (!stride_points_dev[i] || !labels_dev[i] || !cids_dev[i])
            return fail(RTDBS_EINVAL, "NULL buffer for stream %d", i);
        io[i] = IO{stride_points_dev[i], labels_dev[i], cids_dev[i]};
    }
    return batch_launch(b, io.data());
}

extern "C" int rtdbs_batch_slide_pipelined(rtdbs_batch *b, const float *const *stride_points, int32_t *const *labels,
                                           int32_t *const *cids) {
    if (!b || !stride_points || !labels || !cids) return fail(RTDBS_EINVAL, "NULL argument");
    for (int i = 0; i < b->n; i++)
        if (!stride_points[i] || !labels[i] || !cids[i]) return fail(RTDBS_EINVAL, "NULL buffer for stream %d", i);
    return batch_slide_pipelined(b, stride_points, labels, cids);
}

extern "C" int rtdbs_batch_drain(rtdbs_batch *b) {
    if (!b) return fail(RTDBS_EINVAL, "batch is NULL");
    return pipe_drain(b);
}

extern "C" int rtdbs_batch_sync(rtdbs_batch *b) {
    if (!b) return fail(RTDBS_EINVAL, "batch is NULL");
    return batch_sync(b);
}

extern "C" void rtdbs_batch_destroy(rtdbs_batch *b) { batch_free(b); }

// ---------------------------------------------------------------------------
// work counters

extern "C" int rtdbs_batch_set_stats(rtdbs_batch *b, int enable) {
    if (!b) return fail(RTDBS_EINVAL, "batch is NULL");
    int rc = batch_sync(b);
    if (rc) return rc;
    HIPCHK(hipMemset(b->d_stats, 0, (size_t)b->capacity * kNumStats * 8));
    b->stats_on = enable != 0;
    b->stats_slides = 0;
    b->args.stats = b->stats_on ? b->d_stats : nullptr;
    // a captured graph holds the previous argument
    pipe_drop_graphs(b);
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    return RTDBS_OK;
}

extern "C" int rtdbs_batch_get_stats(rtdbs_batch *b, int index, rtdbs_stats *out) {
    if (!b || !out) return fail(RTDBS_EINVAL, "NULL argument");
    if (index < 0 || index >= b->n) return fail(RTDBS_EINVAL, "stream index %d outside [0, %d)", index, b->n);
    int rc = batch_sync(b);
    if (rc) return rc;
    unsigned long long v[kNumStats];
    HIPCHK(hipMemcpy(v, b->d_stats + (size_t)index * kNumStats, sizeof(v), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    out->slides = b->stats_slides;
    out->cells = (int64_t)v[kStatCells];
    out->dense_points = (int64_t)v[kStatDensePoints];
    out->count_queries = (int64_t)v[kStatQueries];
    out->count_tests = (int64_t)v[kStatCountTests];
    out->merge_tests = (int64_t)v[kStatMergeTests];
    out->border_tests = (int64_t)v[kStatBorderTests];
    out->unites = (int64_t)v[kStatUnites];
    out->merge_pairs = (int64_t)v[kStatMergePairs];
    out->cell_parallel = b->shape.cellpar && b->shape.count_cell;
    return RTDBS_OK;
}

extern "C" int rtdbs_stream_set_stats(rtdbs_stream *s, int enable) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    int rc = ensure_solo(s);
    if (rc) return rc;
    return rtdbs_batch_set_stats(s->solo, enable);
}

extern "C" int rtdbs_stream_get_stats(rtdbs_stream *s, rtdbs_stats *out) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    int rc = ensure_solo(s);
    if (rc) return rc;
    return rtdbs_batch_get_stats(s->solo, 0, out);
}

// ---------------------------------------------------------------------------
// timing

extern "C" int rtdbs_batch_set_timing(rtdbs_batch *b, int enable) {
    if (!b) return fail(RTDBS_EINVAL, "batch is NULL");
    int rc = batch_sync(b);
    if (rc) return rc;
    b->ring_used = 0;  // recorded slides of the previous setting are dropped
    if (enable && b->ev_ring.empty()) {
        b->ev_ring.assign((size_t)kTimingRing * 2 * kNumPhases, nullptr);
        // timing-only events: no system-scope fence (no cache writeback and
        // invalidate around the timed kernel); they are read after the stream
        // is synchronised
        for (hipEvent_t &e : b->ev_ring) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    }
    // 1: every phase; 2: the range-count kernel only (two events per slide)
    b->timing = enable != 0;
    b->timing_mask = enable == 2 ? (1u << kPhCount) : (enable ? (1u << kNumPhases) - 1u : 0u);
    for (int i = 0; i < kNumPhases; i++) {
        b->acc_ms[i] = 0.0;
        b->acc_n[i] = 0;
    }
    return RTDBS_OK;
}

extern "C" int rtdbs_stream_set_timing(rtdbs_stream *s, int enable) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    int rc = ensure_solo(s);
    if (rc) return rc;
    return rtdbs_batch_set_timing(s->solo, enable);
}

extern "C" int rtdbs_batch_get_timing(rtdbs_batch *b, double *ms, int64_t *launches, int max_phases, char *names,
                                      size_t names_len) {
    if (!b) return fail(RTDBS_EINVAL, "batch is NULL");
    int rc = batch_collect_timing(b);
    if (rc) return rc;
    const int np = std::min(max_phases, (int)kNumPhases);
    for (int i = 0; i < np; i++) {
        if (ms) ms[i] = b->acc_ms[i];
        if (launches) launches[i] = b->acc_n[i];
    }
    if (names && names_len > 0) {
        std::string all;
        for (int i = 0; i < kNumPhases; i++) {
            if (i) all += ",";
            all += kPhaseNames[i];
        }
        std::snprintf(names, names_len, "%s", all.c_str());
    }
    return kNumPhases;
}

extern "C" int rtdbs_stream_get_timing(rtdbs_stream *s, double *ms, int64_t *launches, int max_phases, char *names,
                                       size_t names_len) {
    if (!s) return fail(RTDBS_EINVAL, "stream is NULL");
    int rc = ensure_solo(s);
    if (rc) return rc;
    return rtdbs_batch_get_timing(s->solo, ms, launches, max_phases, names, names_len);
}

// ---------------------------------------------------------------------------
// pinned host memory

#include <map>

static std::mutex g_pin_mu;
static std::map<uintptr_t, size_t> g_pinned;  // start -> bytes of each rtdbs_host_alloc block

namespace rtdbs {
bool host_pinned(const void *p, size_t bytes) {
    const uintptr_t a = (uintptr_t)p;
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pinned.upper_bound(a);
    if (it == g_pinned.begin()) return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}
}  // namespace rtdbs

extern "C" int rtdbs_host_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(RTDBS_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(RTDBS_ENOMEM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pinned[(uintptr_t)*ptr] = bytes;
    return RTDBS_OK;
}

extern "C" int rtdbs_host_free(void *ptr) {
    if (!ptr) return RTDBS_OK;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        auto it = g_pinned.find((uintptr_t)ptr);
        if (it == g_pinned.end()) return fail(RTDBS_EINVAL, "pointer was not returned by rtdbs_host_alloc");
        g_pinned.erase(it);
    }
    HIPCHK(hipHostFree(ptr));
    return RTDBS_OK;
}

// ---------------------------------------------------------------------------
// device helpers

extern "C" int rtdbs_device_alloc(size_t bytes, void **ptr) {
    if (!ptr) return fail(RTDBS_EINVAL, "ptr is NULL");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(RTDBS_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return RTDBS_OK;
}

extern "C" int rtdbs_device_free(void *ptr) {
    HIPCHK(hipFree(ptr));
    return RTDBS_OK;
}

extern "C" int rtdbs_memcpy_h2d(void *dst_dev, const void *src_host, size_t bytes) {
    HIPCHK(hipMemcpy(dst_dev, src_host, bytes, hipMemcpyHostToDevice));
    return RTDBS_OK;
}

extern "C" int rtdbs_memcpy_d2h(void *dst_host, const void *src_dev, size_t bytes) {
    HIPCHK(hipMemcpy(dst_host, src_dev, bytes, hipMemcpyDeviceToHost));
    return RTDBS_OK;
}

extern "C" int rtdbs_device_copy(void *dst_dev, const void *src_dev, size_t bytes) {
    if (!dst_dev || !src_dev || bytes % 16 != 0) return fail(RTDBS_EINVAL, "device copy needs 16-byte multiples");
    HIPCHK(launch_copy16(src_dev, dst_dev, bytes / 16, 0));
    HIPCHK(hipDeviceSynchronize());
    return RTDBS_OK;
}

extern "C" int rtdbs_device_synchronize(void) {
    HIPCHK(hipDeviceSynchronize());
    return RTDBS_OK;
}

extern "C" int rtdbs_device_count(int *count) {
    if (!count) return fail(RTDBS_EINVAL, "count is NULL");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
    }
    return RTDBS_OK;
}

// ---------------------------------------------------------------------------
// test hook: host-forced allocation failures

#include <atomic>

static std::atomic<int> g_fail_allocs{0};

namespace rtdbs {
bool alloc_hook_should_fail(size_t bytes) {
    (void)bytes;
    int v = g_fail_allocs.load();
    while (v > 0) {
        if (g_fail_allocs.compare_exchange_weak(v, v - 1)) return true;
    }
    return false;
}
}  // namespace rtdbs

extern "C" int rtdbs_test_fail_allocations(int count) {
    g_fail_allocs.store(count > 0 ? count : 0);
    return RTDBS_OK;
}
