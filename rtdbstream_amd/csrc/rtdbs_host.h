// Host-side objects behind the opaque handles of include/rtdbs.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "rtdbs.h"
#include "rtdbs_dev.h"

struct rtdbs_batch;

namespace rtdbs {
constexpr int kTimingRing = 64;  // timed slides recorded before their events are read
}

struct rtdbs_stream {
    rtdbs_params p{};
    int k = 0, R = 0, slot_bits = 0, dense_ok = 0;
    double side = 0.0;
    double lo[3] = {0, 0, 0};
    int64_t n[3] = {1, 1, 1};
    size_t bytes = 0;
    char *arena = nullptr;
    rtdbs::StreamDev dev{};   // host copy of the device descriptor
    float *d_in = nullptr;     // stride / first-window staging (device)
    int32_t *d_label = nullptr;  // output staging for host-buffer slides
    int32_t *d_cid = nullptr;
    bool initialized = false;
    int64_t slides = 0;
    uint64_t gen = 0;          // bumped by every init: batches re-read descriptors
    rtdbs_batch *solo = nullptr;  // single-stream batch of the plain path
    int device = 0;
};

struct rtdbs_batch {
    int n = 0;
    int capacity = 0;
    bool solo = false;
    bool use_graph = false;
    std::vector<rtdbs_stream *> streams;
    std::vector<uint64_t> gens;
    std::vector<rtdbs::IO> stage_io;
    char *dmem = nullptr;
    rtdbs::StreamDev *d_desc = nullptr;
    rtdbs::IO *d_io = nullptr;
    int32_t *d_pre = nullptr;
    unsigned long long *d_stats = nullptr;  // capacity * kNumStats work counters
    bool stats_on = false;
    int64_t stats_slides = 0;
    int32_t *d_cell_pre = nullptr;
    rtdbs::IO *h_io = nullptr;  // pinned mirror of d_io
    bool io_valid = false;
    rtdbs::StreamDev *h_desc = nullptr;  // pinned mirrors of the descriptor and prefix tables
    int32_t *h_pre = nullptr;            // (uploaded asynchronously on the batch's stream)
    // Host-buffer slides. The device side of each member's I/O (stage_io): the
    // stream's own arena buffers in a solo batch, else one contiguous transfer
    // area holding every member's stride (in_off), then its label and cid
    // next to each other (out_off), so that copies of host buffers laid out the
    // same way merge into one. h_stage is pinned staging in the same layout for
    // host buffers that are not engine-pinned memory.
    bool xfer_ready = false;
    char *d_xfer = nullptr;
    size_t d_xfer_bytes = 0;
    char *h_stage = nullptr;
    size_t h_stage_bytes = 0;
    std::vector<size_t> in_off, out_off;
    // Pipelined host-buffer slides (rtdbs_batch_slide_pipelined): two device
    // areas in the transfer area's layout, used by alternate slides, an I/O
    // table and a captured graph for each, a copy stream each way beside the
    // batch's stream, and per-area events ordering the three.
    struct Pipe {
        char *area[2] = {nullptr, nullptr};
        size_t area_bytes = 0;
        rtdbs::IO *d_io = nullptr;  // 2 * capacity entries: area 0's table, then area 1's
        rtdbs::IO *h_io = nullptr;  // pinned staging of the two tables
        hipGraph_t graph[2] = {nullptr, nullptr};
        hipGraphExec_t gexec[2] = {nullptr, nullptr};
        hipStream_t up = nullptr, down = nullptr;
        hipEvent_t in_done[2] = {nullptr, nullptr}, kern_done[2] = {nullptr, nullptr},
                   out_done[2] = {nullptr, nullptr};
        bool ready = false;    // areas and tables match the members
        int64_t issued = 0;    // slides enqueued since the areas were set up
        bool waiting = false;  // the last slide's outputs are not yet waited for
    } pipe;
    rtdbs::BatchArgs args{};
    rtdbs::LaunchShape shape{};
    hipStream_t hs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool timing = false;
    uint32_t timing_mask = 0;  // phases bracketed by events
    // Event sets of the slides recorded since the last read, 2 * kNumPhases
    // per slide, created when timing is first switched on. The elapsed times
    // are read only when the timings are read or the ring is full, so a timed
    // slide costs no host calls beyond its launch.
    std::vector<hipEvent_t> ev_ring;
    int ring_used = 0;
    double acc_ms[rtdbs::kNumPhases] = {};
    int64_t acc_n[rtdbs::kNumPhases] = {};
};

namespace rtdbs {
int fail(int code, const char *fmt, ...);
int validate_params(const rtdbs_params *p);
size_t stream_device_bytes(const rtdbs_params &p);
int device_cus();
int ensure_solo(rtdbs_stream *s);
int refresh_batch(rtdbs_batch *b);
int batch_launch_solo(rtdbs_batch *b, const float *in, int32_t *label, int32_t *cid);
int batch_launch(rtdbs_batch *b, const IO *io);
int batch_sync(rtdbs_batch *b);
// One slide of every member with host buffers (synchronous): strides up,
// slide, labels and ids down, each set of copies merged where the ranges abut.
// Buffers in engine-pinned memory (rtdbs_host_alloc) are copied by DMA
// directly, others through the batch's pinned staging.
int batch_slide_host(rtdbs_batch *b, const float *const *in, int32_t *const *labels, int32_t *const *cids);
// Pipelined host-buffer slide (engine-pinned buffers): enqueues the slide and
// returns once the previous pipelined slide's outputs are in its buffers.
int batch_slide_pipelined(rtdbs_batch *b, const float *const *in, int32_t *const *labels, int32_t *const *cids);
// Waits for the outputs of the last pipelined slide, if any.
int pipe_drain(rtdbs_batch *b);
// True if [p, p + bytes) lies inside one rtdbs_host_alloc allocation.
bool host_pinned(const void *p, size_t bytes);
int batch_collect_timing(rtdbs_batch *b);
void batch_free(rtdbs_batch *b);
// Allocates a batch able to hold up to `capacity` streams (members set later).
rtdbs_batch *batch_new(int capacity, bool use_graph, int *rc);
// Replaces the member streams (n <= capacity); drops any captured graph
// unless the membership is unchanged.
int batch_set_members(rtdbs_batch *b, rtdbs_stream *const *streams, int n);
// Test hook: makes the next stream allocations fail (host-forced ENOMEM).
bool alloc_hook_should_fail(size_t bytes);
}  // namespace rtdbs
