// rtdbs_run: the command-line runner of the slide engine, with the reference
// runner's options and output (main / parse_args / dbscan(),
// dbscan.cpp:2258-2364,2450-2549, util.cpp:454-500, timer.h:212).
//
//   rtdbs_run -f <dataset> -W <window> -S <stride> -K <minPts> -R <radius> [-n <records>]
//
// It loads the dataset (format from the file name, as the reference), runs
// warm-up passes over the whole stream, then one timed pass, and prints the
// reference's lines: "DIM[i]: min, max", "[Step] Warmup i" and
// "[Time] total: <ms per window> ms". On datasets the reference checks (tao,
// RBF, stock) every window of the first warm-up pass is compared with the
// on-device brute-force validator, with the reference's rules (check(),
// util.cpp:383-442): labels equal, core cluster ids equal, every border
// within eps of a point of its cluster; the first mismatch is printed as the
// reference prints it and the run exits with status 1.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rtdbs.h"

namespace {

struct Options {
    std::string file;
    long long n = 0;
    int window = 0, stride = 0, min_pts = 0;
    double radius = 0.0;
    int format = RTDBS_FMT_AUTO;
    int warmup = 10;
    int check = -1;  // -1: as the reference for the format
    int cell_div = 0;
    bool phases = false;
    bool stats = false;
};

[[noreturn]] void usage(const char *argv0) {
    std::fprintf(stderr,
                 "Usage  : %s [options]\n"
                 "Options: --file | -f <filename>      Specify file for data input\n"
                 "         --help | -h                 Print this usage message\n"
                 "         -n <int>                    Records to read; defaults to the whole file\n"
                 "         --window | -W <int>         Window size\n"
                 "         --stride | -S <int>         Stride size (must divide the window size)\n"
                 "         --minPts | -K <int>         MinPts\n"
                 "         --radius | -R <float>       Eps\n"
                 "         --format <name>             tao | geolife | rbf | eds | stock (default: from the file name)\n"
                 "         --warmup <int>              Warm-up passes over the stream; defaults to 10\n"
                 "         --check | --nc              Check the first warm-up pass against the brute-force\n"
                 "                                     validator, or not (default: tao, RBF and stock are checked)\n"
                 "         --cell-div <1|2|3>          Grid side Eps/k (default: automatic)\n"
                 "         --phases                    Also print the time of every kernel phase\n"
                 "         --stats                     Also print work counters per window (extra pass)\n",
                 argv0);
    std::exit(1);
}

int parse_format(const std::string &s) {
    if (s == "tao") return RTDBS_FMT_TAO;
    if (s == "geolife") return RTDBS_FMT_GEOLIFE;
    if (s == "rbf" || s == "RBF") return RTDBS_FMT_RBF;
    if (s == "eds" || s == "EDS") return RTDBS_FMT_EDS;
    if (s == "stock") return RTDBS_FMT_STOCK;
    return -1;
}

Options parse_args(int argc, char **argv) {
    Options o;
    auto value = [&](int &i) -> const char * {
        if (i >= argc - 1) usage(argv[0]);
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string arg(argv[i]);
        if (arg == "--help" || arg == "-h") usage(argv[0]);
        else if (arg == "--file" || arg == "-f") o.file = value(i);
        else if (arg == "-n") o.n = std::atoll(value(i));
        else if (arg == "--window" || arg == "-W") o.window = std::atoi(value(i));
        else if (arg == "--stride" || arg == "-S") o.stride = std::atoi(value(i));
        else if (arg == "--minPts" || arg == "-K") o.min_pts = std::atoi(value(i));
        else if (arg == "--radius" || arg == "-R") o.radius = std::atof(value(i));
        else if (arg == "--warmup") o.warmup = std::atoi(value(i));
        else if (arg == "--check") o.check = 1;
        else if (arg == "--nc") o.check = 0;
        else if (arg == "--cell-div") o.cell_div = std::atoi(value(i));
        else if (arg == "--phases") o.phases = true;
        else if (arg == "--stats") o.stats = true;
        else if (arg == "--format") {
            o.format = parse_format(value(i));
            if (o.format < 0) usage(argv[0]);
        } else {
            std::fprintf(stderr, "Unknown option '%s'\n", arg.c_str());
            usage(argv[0]);
        }
    }
    if (o.file.empty() || o.window <= 0 || o.stride <= 0 || o.min_pts <= 0 || !(o.radius > 0.0) || o.warmup < 0)
        usage(argv[0]);
    return o;
}

void die(const char *what, int rc) {
    std::fprintf(stderr, "[Error] %s: %s (%d)\n", what, rtdbs_last_error(), rc);
    std::exit(1);
}

// The reference's acceptance rules for one window (util.cpp:383-442).
bool check_window(int window_id, const float *win, int dim, int W, float radius2, const int32_t *label,
                  const int32_t *cid, const int32_t *ref_label, const int32_t *ref_cid, const int32_t *ref_nn) {
    for (int i = 0; i < W; i++)
        if (label[i] != ref_label[i]) {
            std::printf("Error on window %d: label[%d] = %d, check_label[%d] = %d; check_nn[%d] = %d\n", window_id, i,
                        label[i], i, ref_label[i], i, ref_nn[i]);
            return false;
        }
    for (int i = 0; i < W; i++) {
        if (label[i] == 0 && cid[i] != ref_cid[i]) {
            std::printf("Error on window %d: cid[%d] = %d, check_cid[%d] = %d; label[%d] = %d, check_nn[%d] = %d\n",
                        window_id, i, cid[i], i, ref_cid[i], i, label[i], i, ref_nn[i]);
            return false;
        }
        if (label[i] == 1 && cid[i] != ref_cid[i]) {
            // any point of the cluster within eps makes the border's id valid
            bool ok = false;
            for (int j = 0; j < W && !ok; j++) {
                if (j == i || cid[j] != cid[i]) continue;
                float d = 0.0f;
                for (int a = 0; a < dim; a++) {
                    const float o = win[(size_t)i * dim + a] - win[(size_t)j * dim + a];
                    d = d + o * o;
                }
                ok = d < radius2;
            }
            if (!ok) {
                std::printf("Error on window %d: cid[%d] = %d, but border[%d] doesn't have a core belonging to "
                            "cluster %d\n",
                            window_id, i, cid[i], i, cid[i]);
                return false;
            }
        }
    }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    setbuf(stdout, nullptr);
    const Options o = parse_args(argc, argv);
    rtdbs_dataset ds;
    int rc = rtdbs_dataset_load(o.file.c_str(), o.format, o.n, &ds);
    if (rc) {
        std::printf("[Error] Dataset %s: %s\n", o.file.c_str(), rtdbs_last_error());
        return 1;
    }
    for (int a = 0; a < ds.dim; a++) std::printf("DIM[%d]: %g, %g\n", a, ds.lo[a], ds.hi[a]);
    const bool check = o.check >= 0 ? o.check == 1
                                    : (ds.format == RTDBS_FMT_TAO || ds.format == RTDBS_FMT_RBF ||
                                       ds.format == RTDBS_FMT_STOCK);
    const int dim = ds.dim, W = o.window, S = o.stride;
    if (ds.n < W + S) {
        std::printf("[Error] %lld records hold no slide of W = %d, S = %d\n", (long long)ds.n, W, S);
        return 1;
    }
    rtdbs_params p;
    std::memset(&p, 0, sizeof(p));
    p.dim = dim;
    p.window = W;
    p.stride = S;
    p.eps = (float)o.radius;
    p.min_pts = o.min_pts;
    p.cell_div = o.cell_div;
    for (int a = 0; a < dim; a++) {
        p.lo[a] = ds.lo[a];
        p.hi[a] = ds.hi[a];
    }
    rtdbs_stream *s = nullptr;
    if ((rc = rtdbs_stream_create(&p, &s))) die("rtdbs_stream_create", rc);
    const long long slides = (ds.n - W) / S;  // the reference's loop: remaining >= stride
    std::vector<int32_t> label(W), cid(W), ref_label(W), ref_cid(W), ref_nn(W);
    std::vector<float> win((size_t)W * dim);
    const float radius2 = p.eps * p.eps;

    auto pass = [&](bool checked, bool timed, double *total_ms) {
        if ((rc = rtdbs_stream_init(s, ds.points))) die("rtdbs_stream_init", rc);
        if (checked) std::memcpy(win.data(), ds.points, win.size() * sizeof(float));
        double acc = 0.0;
        for (long long t = 0; t < slides; t++) {
            const float *stride = ds.points + (size_t)(W + t * S) * dim;
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = rtdbs_stream_slide(s, stride, label.data(), cid.data()))) die("rtdbs_stream_slide", rc);
            const auto t1 = std::chrono::steady_clock::now();
            acc += std::chrono::duration<double, std::milli>(t1 - t0).count();
            if (checked) {
                const int u = (int)(t % (W / S));
                std::memcpy(win.data() + (size_t)u * S * dim, stride, (size_t)S * dim * sizeof(float));
                if ((rc = rtdbs_bruteforce(dim, p.eps, p.min_pts, win.data(), W, ref_label.data(), ref_cid.data(),
                                           ref_nn.data())))
                    die("rtdbs_bruteforce", rc);
                if (!check_window((int)t + 1, win.data(), dim, W, radius2, label.data(), cid.data(),
                                  ref_label.data(), ref_cid.data(), ref_nn.data()))
                    std::exit(1);
            }
        }
        (void)timed;
        if (total_ms) *total_ms = acc;
    };

    for (int i = 0; i < o.warmup; i++) {
        pass(check && i == 0, false, nullptr);
        std::printf("[Step] Warmup %d\n", i);
    }
    if (check && o.warmup == 0) pass(true, false, nullptr);
    double total = 0.0;
    pass(false, true, &total);
    if (slides > 0) std::printf("[Time] total: %g ms\n\n", total / (double)slides);
    if (o.phases && slides > 0) {
        if ((rc = rtdbs_stream_set_timing(s, 1))) die("rtdbs_stream_set_timing", rc);
        pass(false, true, nullptr);
        double ms[32];
        int64_t n[32];
        char names[1024];
        const int np = rtdbs_stream_get_timing(s, ms, n, 32, names, sizeof(names));
        if (np < 0) die("rtdbs_stream_get_timing", np);
        char *save = nullptr;
        char *name = strtok_r(names, ",", &save);
        for (int i = 0; i < np && name; i++, name = strtok_r(nullptr, ",", &save))
            std::printf("[Time] %s: %g ms\n", name, n[i] ? ms[i] / (double)n[i] : 0.0);
    }
    if (o.stats && slides > 0) {
        // the reference's DEBUG_INFO summary lines (print_overall_rt_info)
        if ((rc = rtdbs_stream_set_timing(s, 0))) die("rtdbs_stream_set_timing", rc);
        if ((rc = rtdbs_stream_set_stats(s, 1))) die("rtdbs_stream_set_stats", rc);
        pass(false, false, nullptr);
        rtdbs_stats st;
        if ((rc = rtdbs_stream_get_stats(s, &st))) die("rtdbs_stream_get_stats", rc);
        const double w = (double)st.slides;
        std::printf("avg num spheres for each window: %lf\n", st.cells / w);
        std::printf("avg num points in dense cells for each window: %lf\n", st.dense_points / w);
        std::printf("avg num dist calculations for each window: %lf\n",
                    (st.count_tests + st.merge_tests + st.border_tests) / w);
        std::printf("avg counted points for each window: %lf\n", st.count_queries / w);
        std::printf("avg dist calculations for each window: count %lf, merge %lf, border %lf\n", st.count_tests / w,
                    st.merge_tests / w, st.border_tests / w);
        std::printf("avg unions for each window: %lf (touching pairs %lf)\n", st.unites / w, st.merge_pairs / w);
    }
    if (check) std::printf("[Check] %lld windows equal to the brute-force validator\n", slides);
    rtdbs_stream_destroy(s);
    rtdbs_dataset_free(&ds);
    return 0;
}
