/*
 * rtdbs.h - C ABI of the MI355X streaming DBSCAN slide engine.
 *
 * Drop-in for the per-slide step of RTDBStream's sliding-window DBSCAN
 * (reference: src/dbscan/dbscan.cpp:2258-2364, the OPTIMIZATION_LEVEL 9
 * pipeline). The window-slide API keeps the reference's parameters
 * (window W, stride S, Eps, MinPts) and its per-slide output: two slot-indexed
 * int32 arrays of length W,
 *   label[slot] in {0 core, 1 border, 2 noise}   (reference dbscan.h:27)
 *   cid[slot]   = cluster id: for cores the minimum core slot of the cluster,
 *                 for borders the minimum such id among clusters with a core
 *                 within Eps, for noise the point's own slot.
 *
 * Conventions
 *  - Plain pointers and sizes only. Points are FP32, point-major, `dim`
 *    floats per point (dim 1, 2 or 3; the reference pads to 3 with zeros,
 *    util.cpp:146,194,238-239, which gives identical distances).
 *  - Return codes: 0 = OK, negative = error (RTDBS_E*); nothing is thrown
 *    across the ABI. rtdbs_last_error() returns a thread-local message.
 *  - The caller owns host buffers; the engine owns all device state.
 *  - A stream or batch handle is used by one host thread at a time. The
 *    server object is thread-safe.
 *  - "_async" entry points take DEVICE pointers and enqueue work on the
 *    handle's own HIP stream; call the matching _sync before reading results
 *    or reusing the inputs.
 */
#ifndef RTDBS_H
#define RTDBS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTDBS_OK 0
#define RTDBS_EINVAL -1     /* invalid argument (S does not divide W, eps <= 0, ...) */
#define RTDBS_ENOMEM -2     /* device or host allocation failed */
#define RTDBS_EHIP -3       /* HIP runtime error */
#define RTDBS_ESTATE -4     /* call not valid in the handle's state (e.g. slide before init) */
#define RTDBS_EBUDGET -5    /* server: admission refused by the HBM budget */
#define RTDBS_ETIMEOUT -6   /* server: wait timed out */
#define RTDBS_ECANCELLED -7 /* server: request cancelled */
#define RTDBS_ENOTFOUND -8  /* server: unknown stream or ticket */

#define RTDBS_LABEL_CORE 0
#define RTDBS_LABEL_BORDER 1
#define RTDBS_LABEL_NOISE 2

/* Stream parameters.
 * Replaces the reference's ScanState/Params set-up (state.h:23-129,
 * dbscan.h:15-90, initialize_params at dbscan.cpp:30-132). */
typedef struct rtdbs_params {
    int32_t dim;      /* 1, 2 or 3 coordinates per point */
    int32_t window;   /* W: points in the window (reference -W, util.cpp:463) */
    int32_t stride;   /* S: points per slide, must divide W (ring: dbscan.cpp:2260,2289) */
    float eps;        /* Eps (reference -R, narrowed to float: util.cpp:491, state.h:68) */
    int32_t min_pts;  /* MinPts, self included (reference -K, util.cpp:480; aabb.cu:153) */
    float lo[3];      /* grid bounds (reference: whole-file min/max, util.cpp:23-55);   */
    float hi[3];      /* if hi[a] <= lo[a] on a used axis they come from the first window */
    int32_t cell_div; /* stencil reach k (grid side = Eps/k), 1..3; 0 = default (2) */
    int32_t lanes_per_query; /* threads per query of the point-parallel kernels: 1, 2, 4; 0 = auto */
    int32_t neighbor_mode;   /* count/border kernels of a single-stream launch: 1 cell-parallel,
                                2 point-parallel; 0 = by launch size */
    int32_t cellpar_per_cu;  /* workgroups per CU of the persistent cell-parallel kernels (count,
                                merge, outputs) in a single-stream launch, 1..1024; 0 = by launch size */
    int32_t rows_per_cu;     /* workgroups per CU of the stencil-rows kernel in a single-stream
                                launch, 1..1024; 0 = by launch size */
    int32_t reserved[3];
} rtdbs_params;

typedef struct rtdbs_stream rtdbs_stream;
typedef struct rtdbs_batch rtdbs_batch;
typedef struct rtdbs_server rtdbs_server;

typedef struct rtdbs_stream_info {
    int64_t slides;        /* slides completed */
    int32_t cells;         /* non-empty grid cells in the current window */
    int32_t cell_div;      /* stencil reach k in use */
    int32_t rows;          /* stencil rows per cell */
    int32_t dense_shortcut;/* 1 if cells are small enough that >= MinPts points => all core */
    int64_t grid[3];       /* grid cells per axis */
    double cell_side;      /* grid side length */
    double lo[3];          /* grid origin */
    size_t device_bytes;   /* device memory held by the stream */
} rtdbs_stream_info;

/* ---- single stream ("plain" path) --------------------------------------- */

/* Validates parameters and allocates the stream's device state.
 * Replaces initialize_params (dbscan.cpp:30-132). */
int rtdbs_stream_create(const rtdbs_params *params, rtdbs_stream **out);

/* Device bytes a stream with these parameters holds (for HBM budgeting). */
int rtdbs_stream_device_bytes(const rtdbs_params *params, size_t *bytes);

/* Loads the first window (W points, host memory). No labels are produced,
 * as in the reference (dbscan.cpp:2268-2272: the first window is loaded and
 * gridded but not clustered). Resets the ring cursor. */
int rtdbs_stream_init(rtdbs_stream *s, const float *points);

/* Restores a stream from a checkpoint: the window in slot order (W*dim
 * floats, as rtdbs_stream_get_window returns it) and the number of slides
 * already done, which fixes the ring cursor u = slides mod (W/S). The
 * reference keeps no checkpoint; a window is fully determined by its W
 * points and the cursor, so the labels that follow are identical. */
int rtdbs_stream_restore(rtdbs_stream *s, const float *window_points, int64_t slides_done);

/* One slide, synchronous, host buffers: inserts S points (host) into the
 * ring slots [u*S, u*S+S), expires the S oldest, clusters the window and
 * writes label[W], cid[W] (host). Replaces the loop body
 * dbscan.cpp:2276-2362 (H2D of the stride at :2289 through D2H at
 * :2344-2345). */
int rtdbs_stream_slide(rtdbs_stream *s, const float *stride_points, int32_t *label, int32_t *cid);

/* Same slide with device buffers, enqueued on the stream's HIP stream. */
int rtdbs_stream_slide_async(rtdbs_stream *s, const float *stride_points_dev, int32_t *label_dev,
                             int32_t *cid_dev);
int rtdbs_stream_sync(rtdbs_stream *s);
int rtdbs_stream_get_info(const rtdbs_stream *s, rtdbs_stream_info *info);
/* Copies the current window back to the host in slot order (W*dim floats). */
int rtdbs_stream_get_window(rtdbs_stream *s, float *points);
void rtdbs_stream_destroy(rtdbs_stream *s);

/* ---- batched streams (one launch per phase, optional HIP graph) ---------- */
/* No reference counterpart: the reference slides one stream at a time. */

/* Groups n initialised streams. With use_graph != 0 the slide step is
 * captured once into a HIP graph and replayed. A stream may belong to
 * several batches but must not be slid by two of them concurrently. */
int rtdbs_batch_create(rtdbs_stream *const *streams, int n, int use_graph, rtdbs_batch **out);
/* One slide of every stream in the batch; host buffers, synchronous. */
int rtdbs_batch_slide(rtdbs_batch *b, const float *const *stride_points, int32_t *const *labels,
                      int32_t *const *cids);
/* Device buffers, enqueued on the batch's HIP stream. */
int rtdbs_batch_slide_async(rtdbs_batch *b, const float *const *stride_points_dev,
                            int32_t *const *labels_dev, int32_t *const *cids_dev);
int rtdbs_batch_sync(rtdbs_batch *b);
/* Pipelined host-buffer slide. No reference counterpart: the reference's loop
 * copies the stride up, clusters and copies the labels down one after the
 * other (dbscan.cpp:2289, :2344-2345). This call enqueues one slide of every
 * member - the strides go up on one copy stream, the slide runs on the batch's
 * stream, labels and ids come down on another copy stream, through two device
 * areas used by alternate slides - and returns once the PREVIOUS pipelined
 * slide's labels and ids are in the buffers given with that call. So one
 * slide's transfers overlap the next slide's kernels. The buffers of this call
 * must stay untouched until the next pipelined call, rtdbs_batch_drain or any
 * other slide or sync of the batch returns; all of them must come from
 * rtdbs_host_alloc (RTDBS_EINVAL otherwise). */
int rtdbs_batch_slide_pipelined(rtdbs_batch *b, const float *const *stride_points, int32_t *const *labels,
                                int32_t *const *cids);
/* Waits until the last pipelined slide's labels and ids are in its buffers. */
int rtdbs_batch_drain(rtdbs_batch *b);
void rtdbs_batch_destroy(rtdbs_batch *b);

/* ---- autotuner ------------------------------------------------------------ */
/* No reference counterpart (the reference fixes the grid side to Eps/sqrt(3),
 * dbscan.cpp:34). Times trial slides of the given stream sample for every
 * stencil reach k (grid side Eps/k), then for each form of the neighbourhood
 * kernels (cell-parallel; point-parallel with 1, 2 or 4 lanes per query) at
 * the best reach (each reach and each form timed twice in interleaved
 * rounds, its faster trial kept), then for the
 * workgroups per CU of the cell-parallel kernels
 * and then of the stencil-rows kernel, and returns the fastest setting.
 * `points` holds W + (slides+1)*S points (slides >= 1: one warm-up slide, then
 * `slides` timed ones per trial). Results are cached by (dim, W, S, MinPts,
 * density) in the file named by the RTDBS_TUNE_CACHE environment variable
 * (default $HOME/.cache/rtdbs_autotune.txt); a cache hit runs no trials. */
typedef struct rtdbs_tune_result {
    int32_t cell_div;
    int32_t lanes_per_query;
    double ms_per_slide;   /* of the chosen setting (from the trials or the cache) */
    int32_t from_cache;
    int32_t trials;
    char key[96];
    int32_t neighbor_mode; /* 1 cell-parallel, 2 point-parallel (with lanes_per_query) */
    int32_t cellpar_per_cu;/* workgroups per CU of the cell-parallel kernels */
    int32_t rows_per_cu;   /* workgroups per CU of the stencil-rows kernel */
    int32_t pad;
} rtdbs_tune_result;

int rtdbs_autotune(const rtdbs_params *params, const float *points, int64_t n_points, int32_t slides,
                   rtdbs_tune_result *out);

/* ---- request server (continuous batching, HBM-budget admission) ---------- */
/* No reference counterpart. */

typedef struct rtdbs_server_config {
    size_t hbm_budget_bytes; /* admission limit on the sum of stream device bytes */
    int32_t max_batch;       /* most streams slid together in one launch */
    int32_t lanes;           /* batches that may be on the device at once (HIP streams); 3 serves the
                                most on one MI355X (one lane's transfers and host work overlap the
                                others' kernels; a fourth stream shares one of the process's 4
                                hardware queues) */
    int32_t use_graph;       /* capture recurring batch shapes as HIP graphs */
    int32_t admit_wait_ms;   /* open_stream waits this long for budget to free up */
    int32_t swap;            /* 1: streams beyond the budget live as host checkpoints and are
                                swapped in when a request of theirs is ready, least recently
                                used idle streams out (open_stream then never waits) */
    int32_t reserved[3];
} rtdbs_server_config;

typedef struct rtdbs_server_stats {
    int64_t completed;
    int64_t failed;
    int64_t batches;
    double mean_batch;
    double p50_ms;
    double p99_ms;
    double max_ms;
    size_t bytes_in_use;
    int32_t streams;
    int32_t resident;      /* streams whose state is on the device */
    int64_t swap_ins;      /* streams rebuilt from a checkpoint (readmits included) */
    int64_t swap_outs;     /* streams checkpointed to host memory (evictions included) */
} rtdbs_server_stats;

int rtdbs_server_create(const rtdbs_server_config *cfg, rtdbs_server **out);
/* Admits a stream (and loads its first window, host memory) if its device
 * bytes fit the budget, else RTDBS_EBUDGET. With cfg.swap, a stream that does
 * not fit now is opened as a host checkpoint of its first window and swapped
 * in when its first request is ready (RTDBS_EBUDGET only if it exceeds the
 * whole budget). */
int rtdbs_server_open_stream(rtdbs_server *srv, const rtdbs_params *params, const float *first_window,
                             int32_t *stream_id);
/* Releases the stream's device memory. Pending requests are cancelled. */
int rtdbs_server_close_stream(rtdbs_server *srv, int32_t stream_id);
/* Checkpoints a stream to host memory and frees its device state; its
 * requests wait. rtdbs_server_readmit rebuilds it (subject to the budget). */
int rtdbs_server_evict_stream(rtdbs_server *srv, int32_t stream_id);
int rtdbs_server_readmit_stream(rtdbs_server *srv, int32_t stream_id);
/* Queues one slide of a stream. stride_points (S*dim floats), label and cid
 * (W ints each) are host buffers that must stay valid until the ticket
 * completes. Slides of one stream run in submission order. */
int rtdbs_server_submit(rtdbs_server *srv, int32_t stream_id, const float *stride_points,
                        int32_t *label, int32_t *cid, int64_t *ticket);
/* Waits for a ticket (timeout_ms < 0: forever). Returns the request's status. */
int rtdbs_server_wait(rtdbs_server *srv, int64_t ticket, int32_t timeout_ms);
/* Non-blocking: 1 done, 0 pending, negative error. */
int rtdbs_server_poll(rtdbs_server *srv, int64_t ticket);
/* Cancels a queued request (RTDBS_ESTATE once it is on the device). */
int rtdbs_server_cancel(rtdbs_server *srv, int64_t ticket);
int rtdbs_server_get_stats(rtdbs_server *srv, rtdbs_server_stats *stats);
int rtdbs_server_reset_stats(rtdbs_server *srv);
void rtdbs_server_destroy(rtdbs_server *srv);

/* ---- pinned host memory --------------------------------------------------
 * Page-locked host memory. The host-buffer slides (rtdbs_stream_slide,
 * rtdbs_batch_slide, rtdbs_server_submit) copy strides and outputs that lie in
 * such memory by DMA directly, as the reference does with its pinned
 * h_label / h_cluster_id (dbscan.cpp:54-55); other host buffers go through the
 * engine's own pinned staging (one extra host memcpy each way). A batch's
 * copies merge when consecutive members' buffers abut: strides one after the
 * other, and each member's label then cid, members in batch order. */
int rtdbs_host_alloc(size_t bytes, void **ptr);
/* Frees memory from rtdbs_host_alloc (any other pointer: RTDBS_EINVAL). */
int rtdbs_host_free(void *ptr);

/* ---- device memory helpers (so callers need no other GPU runtime) -------- */
int rtdbs_device_alloc(size_t bytes, void **ptr);
int rtdbs_device_free(void *ptr);
int rtdbs_memcpy_h2d(void *dst_dev, const void *src_host, size_t bytes);
int rtdbs_memcpy_d2h(void *dst_host, const void *src_dev, size_t bytes);
/* ---- datasets ------------------------------------------------------------
 * The five on-disk formats of the reference's runner (util.cpp:12-254). */
enum {
    RTDBS_FMT_AUTO = 0,    /* from the file name, as the reference's main (dbscan.cpp:2454-2469) */
    RTDBS_FMT_TAO = 1,     /* CSV "x,y,z", no header; 3-D              (read_data_from_tao) */
    RTDBS_FMT_GEOLIFE = 2, /* raw little-endian double x,y,z records   (read_data_from_geolife) */
    RTDBS_FMT_RBF = 3,     /* header line, CSV "x,y"; 2-D              (read_data_from_rbf) */
    RTDBS_FMT_EDS = 4,     /* header line, "id x y cid"; 2-D           (read_data_from_eds) */
    RTDBS_FMT_STOCK = 5    /* header line, one value per line; 1-D     (read_data_from_stk) */
};

typedef struct {
    int32_t format;  /* RTDBS_FMT_* actually read */
    int32_t dim;
    int64_t n;       /* records */
    float lo[3];     /* per-axis min / max of the stored FP32 values (the */
    float hi[3];     /* reference's state.min_value / max_value) */
    float *points;   /* n * dim floats, point-major; owned, rtdbs_dataset_free */
} rtdbs_dataset;

/* Format named by the file name (substrings tao, geolife, RBF, EDS, stock, in
 * the reference's order), RTDBS_FMT_AUTO if none. */
int rtdbs_dataset_format_of(const char *path);
/* Reads the first n records (n <= 0: all) of `path`. Fails on a record that
 * does not parse or when fewer than n records exist.
 * Replaces: read_data_from_tao / _geolife / _rbf / _eds / _stk (util.cpp:12-254). */
int rtdbs_dataset_load(const char *path, int32_t format, int64_t n, rtdbs_dataset *out);
void rtdbs_dataset_free(rtdbs_dataset *ds);

/* On-device brute-force DBSCAN of one window of n points (slot i = point i):
 * every pair is tested, with the slide step's FP32 predicate and output rules
 * (label 0 core / 1 border / 2 noise; cid = minimum core slot of the cluster,
 * a border's = the minimum cluster id among cores within eps, noise = own
 * slot); nn (optional) receives the self-inclusive neighbour counts. A
 * validator for windows too large for a CPU check. Synchronous; allocates
 * 32 bytes of device memory per point.
 * Replaces: cluster_with_cuda (util.cpp:357-381) and its kernels
 * find_neighbors_t / find_cores_t / set_cluster_id_op_t / post_cluster_t
 * (aabb.cu:150-261), the reference's own check path. */
int rtdbs_bruteforce(int32_t dim, float eps, int32_t min_pts, const float *points, int64_t n, int32_t *label,
                     int32_t *cid, int32_t *nn);

/* rtdbs_bruteforce without any allocation: no reference counterpart. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Inputs and outputs
 * are DEVICE buffers: points_dev (n * dim floats, point-major), label_dev,
 * cid_dev and the optional nn_dev (n ints each). The union-find parents (and
 * the neighbour counts when nn_dev is NULL) live in a fixed per-GPU scratch
 * array of the library, so n is limited to 2^20 points (RTDBS_EINVAL beyond);
 * every call on one GPU shares that array, so calls that overlap on one GPU
 * race. Synchronous (null stream); same outputs as rtdbs_bruteforce. */
int rtdbs_bruteforce_noalloc(int32_t dim, float eps, int32_t min_pts, const float *points_dev, int64_t n,
                             int32_t *label_dev, int32_t *cid_dev, int32_t *nn_dev);

/* Device-to-device copy by a plain kernel (bytes a multiple of 16): a kernel
 * of known traffic for scaling profiler byte counters. Synchronous. */
int rtdbs_device_copy(void *dst_dev, const void *src_dev, size_t bytes);
int rtdbs_device_synchronize(void);
int rtdbs_device_count(int *count);

/* ---- kernel timing (HIP events on the engine's own streams) -------------- */
/* Work counters of one stream, summed over the slides since counting was
 * enabled - the analogue of the reference's DEBUG_INFO statistics
 * (print_rt_info / print_overall_rt_info, dbscan.cpp:145-212). */
typedef struct {
    int64_t slides;        /* slides counted */
    int64_t cells;         /* non-empty grid cells (the reference's spheres, num_spheres) */
    int64_t dense_points;  /* points of dense cells: core without counting (num_points_in_dense_cells) */
    int64_t count_queries; /* points whose neighbours were counted (rays of the core pass) */
    int64_t count_tests;   /* distance tests of the range count (intersections; num_dist_calculations) */
    int64_t merge_tests;   /* distance tests of the cluster merge (cluster-pass intersections) */
    int64_t border_tests;  /* distance tests of the border pass */
    int64_t unites;        /* union-find unite calls */
    int64_t merge_pairs;   /* touching element pairs found by the merge */
    int32_t cell_parallel; /* 1: count and border ran cell-parallel; 0: either ran point-parallel (large
                              batches). The merge counters come from the cell-parallel merge,
                              which both paths use by default */
    int32_t pad;
    int64_t reserved[6];
} rtdbs_stats;

/* Enables (1) or disables (0) the counters of every stream of the batch and
 * zeroes them. Counting adds one atomic per counter per wave, so it is off by
 * default and timings taken with it on are not representative. */
int rtdbs_batch_set_stats(rtdbs_batch *b, int enable);
/* Counters of member `index` of the batch (synchronises). */
int rtdbs_batch_get_stats(rtdbs_batch *b, int index, rtdbs_stats *out);
int rtdbs_stream_set_stats(rtdbs_stream *s, int enable);
int rtdbs_stream_get_stats(rtdbs_stream *s, rtdbs_stats *out);

/* Event timing on a batch (or a stream's solo batch): enable = 0 off,
 * 1 every phase, 2 only the range-count kernel (two events per slide). */
int rtdbs_batch_set_timing(rtdbs_batch *b, int enable);
int rtdbs_stream_set_timing(rtdbs_stream *s, int enable);
/* Accumulated milliseconds and launch counts per phase since the last reset.
 * names: comma-separated phase names written to a caller buffer. */
int rtdbs_batch_get_timing(rtdbs_batch *b, double *ms, int64_t *launches, int max_phases,
                           char *names, size_t names_len);
int rtdbs_stream_get_timing(rtdbs_stream *s, double *ms, int64_t *launches, int max_phases,
                            char *names, size_t names_len);

/* Test hook: the next `count` stream creations fail with RTDBS_ENOMEM before
 * touching the device (host-forced allocation failure). */
int rtdbs_test_fail_allocations(int count);

/* Thread-local description of the last error on this thread. */
const char *rtdbs_last_error(void);
const char *rtdbs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RTDBS_H */
