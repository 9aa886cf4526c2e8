// Dataset loaders for the five on-disk formats of the reference's runner
// (util.cpp:12-254), picked as its main does (dbscan.cpp:2454-2469): by a
// substring of the file name. Points are read as double and stored as FP32,
// as the reference stores them in its float3 window; min/max per used axis
// are the reference's grid bounds (state.min_value / max_value).
//   tao      CSV "x,y,z", no header                      (read_data_from_tao)
//   geolife  raw little-endian double x,y,z records      (read_data_from_geolife)
//   RBF      one header line, then CSV "x,y"              (read_data_from_rbf)
//   EDS      one header line, then "id x y cid"           (read_data_from_eds)
//   stock    one header line, then one value per line     (read_data_from_stk)
// Differences: a line that does not parse, or a file with fewer than the
// requested records, is an error here (the reference reuses the previous
// values or reads past the end); n <= 0 reads every record.
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rtdbs.h"
#include "rtdbs_host.h"

using rtdbs::fail;

namespace {

int format_dim(int fmt) {
    switch (fmt) {
        case RTDBS_FMT_TAO:
        case RTDBS_FMT_GEOLIFE: return 3;
        case RTDBS_FMT_RBF:
        case RTDBS_FMT_EDS: return 2;
        case RTDBS_FMT_STOCK: return 1;
        default: return 0;
    }
}

// Parses one text record into v[0..dim); false if the line does not parse.
bool parse_line(int fmt, const char *line, double v[3]) {
    switch (fmt) {
        case RTDBS_FMT_TAO: return std::sscanf(line, "%lf,%lf,%lf", &v[0], &v[1], &v[2]) == 3;
        case RTDBS_FMT_RBF: return std::sscanf(line, "%lf,%lf", &v[0], &v[1]) == 2;
        case RTDBS_FMT_EDS: {
            int id, cid;
            return std::sscanf(line, "%d %lf %lf %d", &id, &v[0], &v[1], &cid) == 4;
        }
        case RTDBS_FMT_STOCK: return std::sscanf(line, "%lf", &v[0]) == 1;
        default: return false;
    }
}

bool read_line(FILE *f, std::string &line) {
    line.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '\n') return true;
        line.push_back((char)c);
    }
    return !line.empty();
}

}  // namespace

extern "C" int rtdbs_dataset_format_of(const char *path) {
    if (!path) return RTDBS_FMT_AUTO;
    // the reference's order of tests (dbscan.cpp:2454-2469)
    if (std::strstr(path, "tao")) return RTDBS_FMT_TAO;
    if (std::strstr(path, "geolife")) return RTDBS_FMT_GEOLIFE;
    if (std::strstr(path, "RBF")) return RTDBS_FMT_RBF;
    if (std::strstr(path, "EDS")) return RTDBS_FMT_EDS;
    if (std::strstr(path, "stock")) return RTDBS_FMT_STOCK;
    return RTDBS_FMT_AUTO;
}

extern "C" int rtdbs_dataset_load(const char *path, int32_t format, int64_t n, rtdbs_dataset *out) {
    if (!path || !out) return fail(RTDBS_EINVAL, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    const int fmt = format == RTDBS_FMT_AUTO ? rtdbs_dataset_format_of(path) : format;
    const int dim = format_dim(fmt);
    if (dim == 0) return fail(RTDBS_EINVAL, "dataset %s: unknown format (name holds none of tao, geolife, RBF, EDS, stock)", path);
    FILE *f = std::fopen(path, fmt == RTDBS_FMT_GEOLIFE ? "rb" : "r");
    if (!f) return fail(RTDBS_ENOTFOUND, "Fail to open [%s]!", path);
    std::vector<float> pts;
    if (n > 0) pts.reserve((size_t)n * dim);
    int rc = RTDBS_OK;
    int64_t got = 0;
    if (fmt == RTDBS_FMT_GEOLIFE) {
        double rec[3];
        while ((n <= 0 || got < n) && std::fread(rec, sizeof(double), 3, f) == 3) {
            for (int a = 0; a < 3; a++) pts.push_back((float)rec[a]);
            got++;
        }
    } else {
        std::string line;
        if (fmt != RTDBS_FMT_TAO) read_line(f, line);  // header line, unused
        int64_t lineno = fmt != RTDBS_FMT_TAO ? 1 : 0;
        while ((n <= 0 || got < n) && read_line(f, line)) {
            lineno++;
            double v[3] = {0, 0, 0};
            if (!parse_line(fmt, line.c_str(), v)) {
                if (n <= 0 && line.find_first_not_of(" \t\r") == std::string::npos) continue;  // trailing blank
                rc = fail(RTDBS_EINVAL, "dataset %s line %lld does not parse as this format: '%.60s'", path,
                          (long long)lineno, line.c_str());
                break;
            }
            for (int a = 0; a < dim; a++) pts.push_back((float)v[a]);
            got++;
        }
    }
    std::fclose(f);
    if (rc) return rc;
    if (n > 0 && got < n)
        return fail(RTDBS_EINVAL, "dataset %s holds %lld records, %lld requested", path, (long long)got, (long long)n);
    if (got == 0) return fail(RTDBS_EINVAL, "dataset %s holds no records", path);
    float *buf = (float *)std::malloc(pts.size() * sizeof(float));
    if (!buf) return fail(RTDBS_ENOMEM, "host allocation of %zu bytes failed", pts.size() * sizeof(float));
    std::memcpy(buf, pts.data(), pts.size() * sizeof(float));
    out->format = fmt;
    out->dim = dim;
    out->n = got;
    out->points = buf;
    for (int a = 0; a < 3; a++) {
        out->lo[a] = a < dim ? FLT_MAX : 0.0f;
        out->hi[a] = a < dim ? -FLT_MAX : 0.0f;
    }
    for (int64_t i = 0; i < got; i++)
        for (int a = 0; a < dim; a++) {
            const float v = buf[i * dim + a];
            if (v < out->lo[a]) out->lo[a] = v;
            if (v > out->hi[a]) out->hi[a] = v;
        }
    return RTDBS_OK;
}

extern "C" void rtdbs_dataset_free(rtdbs_dataset *ds) {
    if (!ds) return;
    std::free(ds->points);
    ds->points = nullptr;
    ds->n = 0;
}
