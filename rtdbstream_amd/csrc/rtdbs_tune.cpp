// Autotuner: picks the grid side (stencil reach k, side = Eps/k), the form
// of the neighbourhood kernels (cell-parallel, or point-parallel with 1, 2 or
// 4 lanes per query) and the grid sizes of the persistent kernels by timing
// trial slides on a sample of the stream, and remembers the choice by a
// workload signature.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include <sys/stat.h>

#include "rtdbs.h"
#include "rtdbs_host.h"

using namespace rtdbs;

namespace {

std::mutex g_tune_mu;

std::string cache_path() {
    const char *e = std::getenv("RTDBS_TUNE_CACHE");
    if (e && *e) return e;
    const char *h = std::getenv("HOME");
    return std::string(h && *h ? h : "/tmp") + "/.cache/rtdbs_autotune.txt";
}

// Point-weighted mean occupancy of Eps-side cells of the first window: the
// density a query sees, which decides between fine and coarse grids.
double occupancy(const rtdbs_params &p, const float *pts) {
    const int d = p.dim;
    std::unordered_map<uint64_t, int> cnt;
    cnt.reserve((size_t)p.window * 2);
    const double inv = 1.0 / (double)p.eps;
    for (int i = 0; i < p.window; i++) {
        uint64_t key = 1469598103934665603ull;
        for (int a = 0; a < d; a++) {
            const double v = std::floor((double)pts[(size_t)i * d + a] * inv);
            const int64_t iv = std::isfinite(v) ? (int64_t)v : 0;
            key = (key ^ (uint64_t)iv) * 1099511628211ull;
        }
        cnt[key]++;
    }
    double s2 = 0.0;
    for (auto &kv : cnt) s2 += (double)kv.second * kv.second;
    return s2 / (double)p.window;
}

std::string make_key(const rtdbs_params &p, const float *pts) {
    const double occ = occupancy(p, pts);
    char buf[96];
    std::snprintf(buf, sizeof(buf), "d%d_w%d_s%d_m%d_o%d", p.dim, (int)std::lround(std::log2((double)p.window)),
                  (int)std::lround(std::log2((double)p.window / p.stride)), p.min_pts,
                  (int)std::lround(2.0 * std::log2(std::max(occ, 1e-3))));
    return buf;
}

struct Choice {
    int k = 2, lpq = 0, mode = 0, cp = 0, rows = 0;
    double ms = 0.0;
};

// Cache lines: "key k lanes_per_query neighbor_mode ms cellpar_per_cu
// rows_per_cu"; the last complete entry of a key wins.
bool cache_lookup(const std::string &key, Choice *c) {
    FILE *f = std::fopen(cache_path().c_str(), "r");
    if (!f) return false;
    char line[256], kb[128];
    bool hit = false;
    while (std::fgets(line, sizeof(line), f)) {
        Choice x;
        if (std::sscanf(line, "%127s %d %d %d %lf %d %d", kb, &x.k, &x.lpq, &x.mode, &x.ms, &x.cp, &x.rows) == 7 &&
            key == kb) {
            *c = x;
            hit = true;
        }
    }
    std::fclose(f);
    return hit;
}

void cache_store(const std::string &key, const Choice &c) {
    const std::string path = cache_path();
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) {
        const std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); i++)
            if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
    }
    FILE *f = std::fopen(path.c_str(), "a");
    if (!f) return;  // read-only home: tuning still returns its result
    std::fprintf(f, "%s %d %d %d %.6f %d %d\n", key.c_str(), c.k, c.lpq, c.mode, c.ms, c.cp, c.rows);
    std::fclose(f);
}

// Milliseconds per slide of `slides` timed slides (after one warm-up slide).
// The stride points are on the device first and the slides are enqueued back
// to back, as a serving loop runs them: the host copies of the synchronous
// call (the whole window's labels, from pageable memory) would otherwise hide
// the differences between the kernel forms.
int trial(const rtdbs_params &p, const float *pts, int slides, double *ms) {
    rtdbs_stream *s = nullptr;
    int rc = rtdbs_stream_create(&p, &s);
    if (rc) return rc;
    const size_t W = (size_t)p.window, S = (size_t)p.stride, d = (size_t)p.dim;
    const size_t n_in = (size_t)(slides + 1) * S * d;
    float *d_in = nullptr;
    int32_t *d_label = nullptr, *d_cid = nullptr;
    rc = rtdbs_stream_init(s, pts);
    if (!rc && (hipMalloc((void **)&d_in, n_in * 4) != hipSuccess ||
                hipMalloc((void **)&d_label, W * 4) != hipSuccess || hipMalloc((void **)&d_cid, W * 4) != hipSuccess))
        rc = fail(RTDBS_ENOMEM, "autotune: trial buffers (%zu bytes)", (n_in + 2 * W) * 4);
    if (!rc && hipMemcpy(d_in, pts + W * d, n_in * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(RTDBS_EHIP, "autotune: uploading the sample failed");
    if (!rc) rc = rtdbs_stream_slide_async(s, d_in, d_label, d_cid);
    if (!rc) rc = rtdbs_stream_sync(s);
    if (!rc) {
        auto t0 = std::chrono::steady_clock::now();
        for (int t = 1; t <= slides && !rc; t++)
            rc = rtdbs_stream_slide_async(s, d_in + (size_t)t * S * d, d_label, d_cid);
        if (!rc) rc = rtdbs_stream_sync(s);
        auto t1 = std::chrono::steady_clock::now();
        *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / slides;
    }
    if (d_in) (void)hipFree(d_in);
    if (d_label) (void)hipFree(d_label);
    if (d_cid) (void)hipFree(d_cid);
    rtdbs_stream_destroy(s);
    return rc;
}

}  // namespace

extern "C" int rtdbs_autotune(const rtdbs_params *params, const float *points, int64_t n_points, int32_t slides,
                              rtdbs_tune_result *out) {
    int rc = validate_params(params);
    if (rc) return rc;
    if (!points || !out) return fail(RTDBS_EINVAL, "NULL argument");
    if (slides < 1) return fail(RTDBS_EINVAL, "slides must be >= 1");
    const rtdbs_params &p0 = *params;
    if (n_points < (int64_t)p0.window + (int64_t)(slides + 1) * p0.stride)
        return fail(RTDBS_EINVAL, "the sample needs W + (slides+1)*S = %lld points (got %lld)",
                    (long long)p0.window + (long long)(slides + 1) * p0.stride, (long long)n_points);
    std::memset(out, 0, sizeof(*out));
    const std::string key = make_key(p0, points);
    std::snprintf(out->key, sizeof(out->key), "%s", key.c_str());
    std::lock_guard<std::mutex> lk(g_tune_mu);
    Choice c;
    if (cache_lookup(key, &c)) {
        out->cell_div = c.k;
        out->lanes_per_query = c.lpq;
        out->neighbor_mode = c.mode;
        out->cellpar_per_cu = c.cp;
        out->rows_per_cu = c.rows;
        out->ms_per_slide = c.ms;
        out->from_cache = 1;
        return RTDBS_OK;
    }
    // 1. stencil reach, with the launch-size defaults for the rest; like the
    // forms below, each reach timed twice in interleaved rounds, its faster
    // trial kept
    Choice best;
    double reach_ms[kMaxK + 1];
    for (int kk = 1; kk <= kMaxK; kk++) reach_ms[kk] = 1e300;
    for (int round = 0; round < 2; round++) {
        for (int kk = 1; kk <= kMaxK; kk++) {
            rtdbs_params p = p0;
            p.cell_div = kk;
            p.lanes_per_query = 0;
            p.neighbor_mode = 0;
            double m = 0.0;
            rc = trial(p, points, slides, &m);
            if (rc) return rc;
            out->trials++;
            reach_ms[kk] = std::min(reach_ms[kk], m);
        }
    }
    best.ms = 1e300;
    for (int kk = 1; kk <= kMaxK; kk++) {
        if (reach_ms[kk] < best.ms) {
            best.ms = reach_ms[kk];
            best.k = kk;
        }
    }
    // 2. kernel form at that reach: cell-parallel, then point-parallel with
    // 1, 2 and 4 lanes per query; the result always names one explicitly.
    // Each form is timed twice, the rounds interleaved, and keeps its faster
    // trial: the merge's time varies from run to run by more than the forms
    // differ on some streams, and one slow trial must not decide.
    const int forms[4][2] = {{1, 0}, {2, 1}, {2, 2}, {2, 4}};
    double form_ms[4] = {1e300, 1e300, 1e300, 1e300};
    for (int round = 0; round < 2; round++) {
        for (int f = 0; f < 4; f++) {
            rtdbs_params p = p0;
            p.cell_div = best.k;
            p.neighbor_mode = forms[f][0];
            p.lanes_per_query = forms[f][1];
            double m = 0.0;
            rc = trial(p, points, slides, &m);
            if (rc) return rc;
            out->trials++;
            form_ms[f] = std::min(form_ms[f], m);
        }
    }
    best.ms = 1e300;
    for (int f = 0; f < 4; f++) {
        if (form_ms[f] < best.ms) {
            best.ms = form_ms[f];
            best.mode = forms[f][0];
            best.lpq = forms[f][1];
        }
    }
    // 3. workgroups per CU of the persistent cell-parallel kernels (the merge
    // always runs on that grid), then of the stencil-rows kernel, each with the
    // settings chosen so far
    auto pick = [&](const int *vals, int nv, int Choice::*field) -> int {
        double bm = 1e300;
        int bv = vals[0];
        for (int i = 0; i < nv; i++) {
            Choice t = best;
            t.*field = vals[i];
            rtdbs_params p = p0;
            p.cell_div = t.k;
            p.neighbor_mode = t.mode;
            p.lanes_per_query = t.lpq;
            p.cellpar_per_cu = t.cp;
            p.rows_per_cu = t.rows;
            double m = 0.0;
            const int r = trial(p, points, slides, &m);
            if (r) return r;
            out->trials++;
            if (m < bm) {
                bm = m;
                bv = vals[i];
            }
        }
        best.*field = bv;
        best.ms = bm;
        return RTDBS_OK;
    };
    static const int kCellparPerCu[] = {8, 12, 16, 24, 32, 48};
    static const int kRowsPerCu[] = {4, 8, 16, 32};
    rc = pick(kCellparPerCu, 6, &Choice::cp);
    if (rc) return rc;
    rc = pick(kRowsPerCu, 4, &Choice::rows);
    if (rc) return rc;
    out->cell_div = best.k;
    out->lanes_per_query = best.lpq;
    out->neighbor_mode = best.mode;
    out->cellpar_per_cu = best.cp;
    out->rows_per_cu = best.rows;
    out->ms_per_slide = best.ms;
    cache_store(key, best);
    return RTDBS_OK;
}
