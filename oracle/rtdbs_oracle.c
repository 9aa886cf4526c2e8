/*
 * CPU oracle for the sliding-window DBSCAN slide step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke
 * check and bench.py's CPU-baseline leg may load it. The product path (the HIP
 * engine in rtdbstream_amd/) never links, loads or calls it.
 *
 * It restates, in plain C, the semantics the reference checks its own GPU
 * output against (RTDBStream, src/dbscan/util.cpp:285-442 and
 * src/dbscan/aabb.cu:150-261):
 *
 *  - distances are FP32 with the evaluation order ((dx*dx + dy*dy) + dz*dz),
 *    no fused multiply-add (build with -ffp-contract=off); missing axes are 0
 *    in the reference (util.cpp:146,194,238-239) and contribute +0, so the
 *    1-D/2-D forms below are bit-identical to the reference's float3 form
 *    (util.cpp:310-313, aabb.cu:203-204, dbscan.cu:48-49);
 *  - eps2 = fl32(eps * eps) (dbscan.cpp:41, eps is float: state.h:68);
 *  - the neighbour test is strict: d2 < eps2 (util.cpp:314, aabb.cu:205);
 *  - the neighbour count includes the point itself (util.cpp:308,
 *    aabb.cu:202-207); core iff count >= min_pts (util.cpp:321, aabb.cu:153);
 *  - label 0 = core, 1 = border, 2 = noise (dbscan.h:27);
 *  - clusters are discovered by BFS seeded at ascending core slot
 *    (util.cpp:325-351), so a core's cid is the minimum core slot of its
 *    cluster (equal to the reference's union-find root, which always hooks the
 *    larger root under the smaller: dbscan.cu:22-43, util.cpp:289-297);
 *  - a border point takes the id of the first cluster (in seed order) that
 *    reaches it, i.e. the minimum canonical id among clusters with a core
 *    within eps; this satisfies the reference's border rule
 *    (util.cpp:420-438) deterministically;
 *  - noise keeps cid = its own slot (aabb.cu:160-170, dbscan.cpp:2321);
 *  - the window is a ring of W slots; slide t overwrites slots
 *    [u*S, u*S+S) with u = t mod (W/S) (dbscan.cpp:2259-2353); the first
 *    window is loaded but never clustered (dbscan.cpp:2268-2276).
 *
 * Two implementations with identical output:
 *  - oracle_bruteforce: O(n^2) count + O(n * cores) BFS, the literal
 *    restatement of cluster_with_cpu (util.cpp:299-354);
 *  - oracle_grid: the same BFS with candidates taken from a uniform grid of
 *    side eps*(1+2^-16) and per-cell lists of unvisited points; the count
 *    phase is OpenMP-parallel. Used at full window sizes and as the timed CPU
 *    baseline.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_OK 0
#define ORACLE_EINVAL -1
#define ORACLE_ENOMEM -2

static inline float sqdist(const float *a, const float *b, int dim) {
    /* util.cpp:310-313: O = p - q; d = O.x*O.x + O.y*O.y + O.z*O.z */
    float dx = a[0] - b[0];
    float d = dx * dx;
    if (dim > 1) {
        float dy = a[1] - b[1];
        d = d + dy * dy;
    }
    if (dim > 2) {
        float dz = a[2] - b[2];
        d = d + dz * dz;
    }
    return d;
}

float oracle_eps2(float eps) {
    volatile float e = eps; /* dbscan.cpp:41: radius2 = radius * radius in FP32 */
    return e * e;
}

/* Self-inclusive neighbour count, uncapped (util.cpp:306-318). */
int oracle_neighbor_counts(const float *pts, int n, int dim, float eps, int32_t *nn) {
    if (n < 0 || dim < 1 || dim > 3 || !(eps > 0.0f)) return ORACLE_EINVAL;
    const float e2 = oracle_eps2(eps);
    for (int i = 0; i < n; i++) nn[i] = 1; /* the point itself */
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            if (sqdist(pts + (size_t)i * dim, pts + (size_t)j * dim, dim) < e2) {
                nn[i]++;
                nn[j]++;
            }
        }
    }
    return ORACLE_OK;
}

/* Literal restatement of cluster_with_cpu (util.cpp:299-354) with noise
 * cid = own slot (aabb.cu:168). */
int oracle_bruteforce(const float *pts, int n, int dim, float eps, int min_pts,
                      int32_t *label, int32_t *cid) {
    if (n < 0 || dim < 1 || dim > 3 || !(eps > 0.0f) || min_pts < 1) return ORACLE_EINVAL;
    const float e2 = oracle_eps2(eps);
    int32_t *nn = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    char *vis = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!nn || !vis || !queue) {
        free(nn); free(vis); free(queue);
        return ORACLE_ENOMEM;
    }
    oracle_neighbor_counts(pts, n, dim, eps, nn);
    for (int i = 0; i < n; i++) {
        label[i] = nn[i] >= min_pts ? 0 : 2;
        cid[i] = i;
    }
    for (int i = 0; i < n; i++) {
        if (vis[i] || label[i] != 0) continue;
        vis[i] = 1;
        cid[i] = i;
        int qh = 0, qt = 0;
        queue[qt++] = i;
        while (qh < qt) {
            const float *p = pts + (size_t)queue[qh++] * dim;
            for (int j = 0; j < n; j++) {
                if (vis[j]) continue;
                if (sqdist(p, pts + (size_t)j * dim, dim) < e2) {
                    vis[j] = 1;
                    cid[j] = i;
                    if (label[j] == 0) queue[qt++] = j;
                    else label[j] = 1;
                }
            }
        }
    }
    free(nn); free(vis); free(queue);
    return ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Grid-accelerated restatement with identical output.                       */

typedef struct {
    uint64_t key;
    int32_t idx;
} keyed_t;

static int cmp_keyed(const void *a, const void *b) {
    const keyed_t *x = (const keyed_t *)a, *y = (const keyed_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

typedef struct {
    int dim;
    int64_t nc[3];
    double lo[3];
    double inv_side;
    int ncells;
    uint64_t *cell_key;  /* sorted, unique */
    int32_t *cell_start; /* ncells + 1 */
    int32_t *order;      /* point indices sorted by (key, idx) */
    int32_t *cell_of;    /* point -> cell */
    int64_t *coord;      /* point -> packed axis cell indices (3 x 21 bits) */
} grid_t;

static void grid_free(grid_t *g) {
    free(g->cell_key); free(g->cell_start); free(g->order); free(g->cell_of); free(g->coord);
}

/* Cell side is eps*(1+2^-16): a pair accepted by the FP32 test has true
 * distance below eps*(1+2^-21), so the exact (double) cell indices of the two
 * points differ by at most 1 on every axis. */
static int grid_build(grid_t *g, const float *pts, int n, int dim, float eps) {
    memset(g, 0, sizeof(*g));
    g->dim = dim;
    const double side = (double)eps * (1.0 + 1.0 / 65536.0);
    g->inv_side = 1.0 / side;
    double hi[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) { g->lo[a] = 0.0; g->nc[a] = 1; }
    for (int a = 0; a < dim; a++) {
        double mn = INFINITY, mx = -INFINITY;
        for (int i = 0; i < n; i++) {
            double v = pts[(size_t)i * dim + a];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        if (n == 0) { mn = 0; mx = 0; }
        g->lo[a] = mn;
        hi[a] = mx;
        int64_t c = (int64_t)floor((mx - mn) * g->inv_side) + 1;
        if (c > (1 << 20)) return ORACLE_EINVAL; /* extent/eps too large for the oracle */
        g->nc[a] = c;
    }
    keyed_t *kv = (keyed_t *)malloc(sizeof(keyed_t) * (size_t)(n > 0 ? n : 1));
    g->order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    g->cell_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    g->coord = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    g->cell_key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    g->cell_start = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    if (!kv || !g->order || !g->cell_of || !g->coord || !g->cell_key || !g->cell_start) {
        free(kv);
        grid_free(g);
        return ORACLE_ENOMEM;
    }
    for (int i = 0; i < n; i++) {
        int64_t ix[3] = {0, 0, 0};
        for (int a = 0; a < dim; a++) {
            double v = ((double)pts[(size_t)i * dim + a] - g->lo[a]) * g->inv_side;
            int64_t c = (int64_t)floor(v);
            if (c < 0) c = 0;
            if (c >= g->nc[a]) c = g->nc[a] - 1;
            ix[a] = c;
        }
        kv[i].key = (uint64_t)((ix[0] * g->nc[1] + ix[1]) * g->nc[2] + ix[2]);
        kv[i].idx = i;
        g->coord[i] = (ix[0] << 42) | (ix[1] << 21) | ix[2];
    }
    qsort(kv, (size_t)n, sizeof(keyed_t), cmp_keyed);
    int nc = 0;
    for (int i = 0; i < n; i++) {
        if (i == 0 || kv[i].key != kv[i - 1].key) {
            g->cell_key[nc] = kv[i].key;
            g->cell_start[nc] = i;
            nc++;
        }
        g->order[i] = kv[i].idx;
        g->cell_of[kv[i].idx] = nc - 1;
    }
    g->cell_start[nc] = n;
    g->ncells = nc;
    free(kv);
    return ORACLE_OK;
}

static int grid_find(const grid_t *g, uint64_t key) {
    int lo = 0, hi = g->ncells;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (g->cell_key[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < g->ncells && g->cell_key[lo] == key) ? lo : -1;
}

/* Neighbour cells (3^dim stencil) of point i; returns the count written. */
static int grid_stencil(const grid_t *g, int i, int32_t *out) {
    int64_t c = g->coord[i];
    int64_t ix[3] = {(c >> 42) & 0x1FFFFF, (c >> 21) & 0x1FFFFF, c & 0x1FFFFF};
    int cnt = 0;
    int rx = 1, ry = g->dim > 1 ? 1 : 0, rz = g->dim > 2 ? 1 : 0;
    for (int dx = -rx; dx <= rx; dx++) {
        int64_t x = ix[0] + dx;
        if (x < 0 || x >= g->nc[0]) continue;
        for (int dy = -ry; dy <= ry; dy++) {
            int64_t y = ix[1] + dy;
            if (y < 0 || y >= g->nc[1]) continue;
            for (int dz = -rz; dz <= rz; dz++) {
                int64_t z = ix[2] + dz;
                if (z < 0 || z >= g->nc[2]) continue;
                int cc = grid_find(g, (uint64_t)((x * g->nc[1] + y) * g->nc[2] + z));
                if (cc >= 0) out[cnt++] = cc;
            }
        }
    }
    return cnt;
}

int oracle_grid(const float *pts, int n, int dim, float eps, int min_pts, int threads,
                int32_t *label, int32_t *cid) {
    if (n < 0 || dim < 1 || dim > 3 || !(eps > 0.0f) || min_pts < 1) return ORACLE_EINVAL;
    const float e2 = oracle_eps2(eps);
    grid_t g;
    int rc = grid_build(&g, pts, n, dim, eps);
    if (rc) return rc;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    /* Phase 1: core flags, count capped at min_pts (the cap cannot change
     * the flag: core iff count >= min_pts). */
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        int32_t st[27];
        int ns = grid_stencil(&g, i, st);
        const float *p = pts + (size_t)i * dim;
        int cnt = 0;
        for (int s = 0; s < ns && cnt < min_pts; s++) {
            for (int k = g.cell_start[st[s]]; k < g.cell_start[st[s] + 1]; k++) {
                if (sqdist(p, pts + (size_t)g.order[k] * dim, dim) < e2) {
                    if (++cnt >= min_pts) break;
                }
            }
        }
        label[i] = cnt >= min_pts ? 0 : 2;
        cid[i] = i;
    }
    /* Phase 2: BFS in ascending seed order (util.cpp:325-351) over per-cell
     * lists of unvisited points. */
    int32_t *unv = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *unv_cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(g.ncells > 0 ? g.ncells : 1));
    int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    char *vis = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int32_t *pos_in_cell = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!unv || !unv_cnt || !queue || !vis || !pos_in_cell) {
        free(unv); free(unv_cnt); free(queue); free(vis); free(pos_in_cell);
        grid_free(&g);
        return ORACLE_ENOMEM;
    }
    for (int k = 0; k < n; k++) {
        unv[k] = g.order[k];
        pos_in_cell[g.order[k]] = k;
    }
    for (int c = 0; c < g.ncells; c++) unv_cnt[c] = g.cell_start[c + 1] - g.cell_start[c];
#define REMOVE_UNVISITED(j)                                                   \
    do {                                                                      \
        int cc_ = g.cell_of[j];                                               \
        int base_ = g.cell_start[cc_];                                        \
        int last_ = base_ + unv_cnt[cc_] - 1;                                 \
        int at_ = pos_in_cell[j];                                             \
        int moved_ = unv[last_];                                              \
        unv[at_] = moved_;                                                    \
        pos_in_cell[moved_] = at_;                                            \
        unv[last_] = (j);                                                     \
        pos_in_cell[j] = last_;                                               \
        unv_cnt[cc_]--;                                                       \
    } while (0)
    for (int i = 0; i < n; i++) {
        if (vis[i] || label[i] != 0) continue;
        vis[i] = 1;
        REMOVE_UNVISITED(i);
        cid[i] = i;
        int qh = 0, qt = 0;
        queue[qt++] = i;
        while (qh < qt) {
            int pi = queue[qh++];
            const float *p = pts + (size_t)pi * dim;
            int32_t st[27];
            int ns = grid_stencil(&g, pi, st);
            for (int s = 0; s < ns; s++) {
                int c = st[s];
                int base = g.cell_start[c];
                int k = 0;
                while (k < unv_cnt[c]) {
                    int j = unv[base + k];
                    if (sqdist(p, pts + (size_t)j * dim, dim) < e2) {
                        vis[j] = 1;
                        REMOVE_UNVISITED(j); /* moves another point into slot k */
                        cid[j] = i;
                        if (label[j] == 0) queue[qt++] = j;
                        else label[j] = 1;
                    } else {
                        k++;
                    }
                }
            }
        }
    }
#undef REMOVE_UNVISITED
    free(unv); free(unv_cnt); free(queue); free(vis); free(pos_in_cell);
    grid_free(&g);
    return ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Sliding driver: the reference's ring (dbscan.cpp:2259-2353).              */

/* data: n_total points (dim floats each). Runs n_slides slides; slide t
 * writes labels/cids of all W slots to labels[t*W..], cids[t*W..].
 * method 0 = brute force, 1 = grid. Returns 0 or a negative error code. */
int oracle_stream(const float *data, int64_t n_total, int dim, int W, int S, float eps,
                  int min_pts, int n_slides, int method, int threads,
                  int32_t *labels, int32_t *cids) {
    if (W <= 0 || S <= 0 || W % S != 0 || dim < 1 || dim > 3 || n_slides < 0) return ORACLE_EINVAL;
    if ((int64_t)W + (int64_t)n_slides * S > n_total) return ORACLE_EINVAL;
    float *win = (float *)malloc(sizeof(float) * (size_t)W * dim);
    if (!win) return ORACLE_ENOMEM;
    memcpy(win, data, sizeof(float) * (size_t)W * dim);
    const int units = W / S;
    for (int t = 0; t < n_slides; t++) {
        int u = t % units;
        memcpy(win + (size_t)u * S * dim, data + ((size_t)W + (size_t)t * S) * dim,
               sizeof(float) * (size_t)S * dim);
        int rc = method == 0
                     ? oracle_bruteforce(win, W, dim, eps, min_pts, labels + (size_t)t * W,
                                         cids + (size_t)t * W)
                     : oracle_grid(win, W, dim, eps, min_pts, threads, labels + (size_t)t * W,
                                   cids + (size_t)t * W);
        if (rc) {
            free(win);
            return rc;
        }
    }
    free(win);
    return ORACLE_OK;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
