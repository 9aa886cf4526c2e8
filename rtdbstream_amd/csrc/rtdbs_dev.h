// Device-side layout shared by the HIP kernels and the C++ host engine.
//
// One stream's window lives in HBM sorted by the packed 64-bit key
//     (cell_key << slot_bits) | slot
// i.e. by (grid cell, ring slot) - the order of the reference's pos_arr
// (thrust_helper.cu:25-27,42-46). Coordinates travel with the keys in SoA
// arrays so every stencil row is a contiguous, coalesced range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtdbs {

constexpr int kTile = 2048;        // positions per sort tile (stride, first window)
constexpr int kWTile = 512;        // positions per window tile: expiry, merge, cell table
constexpr int kTileThreads = 256;  // threads per tile block
constexpr int kChunk = 64;         // positions per wave in the neighbourhood kernels
constexpr int kCap = 512;          // candidates staged in LDS per wave
constexpr int kMaxK = 3;           // largest stencil reach
constexpr int kMaxRows = 49;       // (2k+1)^(dim-1) for k <= 3, dim <= 3
constexpr int kMaxSlotBits = 24;  // slot bits of the packed key; the cell key takes the rest (<= 52)
constexpr int kUGCells = 8;  // query cells per wave of the cell-level merge
constexpr int kMaxBatch = 1024;
constexpr int kMaxPerCu = 1024;  // largest workgroups-per-CU setting of a persistent grid
constexpr int kCandPad = 4;      // floats of margin past the end of each window coordinate array
// Per-stream work counters (optional; rtdbs_stats in the C ABI)
constexpr int kNumStats = 8;
enum Stat {
    kStatCells = 0,      // non-empty cells
    kStatDensePoints,    // points of dense cells: core without counting
    kStatQueries,        // points whose neighbours were counted
    kStatCountTests,     // distance tests of the range count
    kStatMergeTests,     // distance tests of the cluster merge
    kStatBorderTests,    // distance tests of the border pass
    kStatUnites,         // union-find unite calls
    kStatMergePairs      // touching (element, element) pairs found by the merge
};

constexpr uint8_t kFlagCore = 1;
constexpr uint8_t kFlagDense = 2;
constexpr uint8_t kCellSafe = 1;
constexpr uint8_t kCellDense = 2;
constexpr int kNoCore = 0x7fffffff;  // fcore of a cell without cores

// Per-stream state that the device itself advances every slide.
struct StreamMut {
    int32_t t;       // slides completed
    int32_t u;       // ring unit the next stride overwrites: slots [u*S, u*S+S)
    int32_t ncell;   // non-empty cells of the current window
    int32_t pad[5];
};

// Per-slide I/O (device pointers).
struct IO {
    const float *in;  // S*dim floats (W*dim for the first window)
    int32_t *label;   // W, slot-indexed
    int32_t *cid;     // W, slot-indexed
};

struct StreamDev {
    int32_t dim, W, S, min_pts;
    int32_t k, R, slot_bits, units;
    int32_t dense_ok;  // side eps/k with k*k > dim: points of one interior cell are within eps
    float eps2;        // fl32(eps*eps)
    double inv_side;
    double lo[3];
    int64_t n[3];  // grid cells per axis (1 for unused axes)
    StreamMut *mut;
    // sorted window (W)
    uint64_t *skey;
    float *sx, *sy, *sz;
    // compacted survivors of the previous window (W), also init ping-pong
    uint64_t *akey;
    float *ax, *ay, *az;
    // stride: sorted tiles (keys, bkey0), then the sorted stride (B1)
    uint64_t *bkey0;
    float *bx0, *by0, *bz0;
    uint64_t *bkey1;
    float *bx1, *by1, *bz1;
    // cell table
    int32_t *cstart;  // W+1
    uint64_t *ckey;   // W
    int32_t *pcell;   // W: position -> cell
    int2 *rows;       // W*R: candidate position ranges per cell
    uint8_t *cflag;   // W: kCellSafe | kCellDense per cell
    int32_t *fcore;   // W: first (minimum-slot) core position of each cell, kNoCore if none
    // clustering
    uint8_t *flags;   // W: kFlagCore | kFlagDense
    int32_t *parent;  // W: union-find over positions, root = minimum slot
    int32_t *rep;     // W: union-find element of each position (cell-parallel path)
    int32_t *tile_a;  // ceil(W/kWTile): survivors per tile
    int32_t *tile_b;  // ceil(W/kWTile): cell starts per tile
};

struct BatchArgs {
    const StreamDev *s;  // n descriptors
    const IO *io;        // n I/O entries (unless use_io1)
    IO io1;              // single-stream I/O passed by value
    int32_t use_io1;
    int32_t n;
    const int32_t *pre_stile;  // n+1 prefix of ceil(S/kTile)
    const int32_t *pre_wtile;  // n+1 prefix of ceil(W/kWTile)
    const int32_t *pre_itile;  // n+1 prefix of ceil(W/kTile) (first-window sort)
    const int32_t *pre_chunk;  // n+1 prefix of ceil(W/kChunk)
    unsigned long long *stats; // n * kNumStats per-stream counters, null when off
    int32_t *cell_pre;         // n+1 prefix of cells per stream (written by k_rows)
    int32_t xcd_groups;        // kXcd* bits: kernels whose XCDs each take one contiguous range of work
};

// BatchArgs::xcd_groups bits
constexpr int32_t kXcdUnion = 1;         // k_union
constexpr int32_t kXcdBorderPoints = 2;  // k_border_points

// Kernel launchers (rtdbs_kernels.hip). All enqueue on `st` and return the
// hipError_t of the launch.
struct LaunchShape {
    int32_t stiles;   // total stride tiles
    int32_t wtiles;   // total window tiles
    int32_t chunks;   // total chunks
    int32_t rows_blocks;     // 256-thread blocks for the row kernel
    int32_t cell_blocks;     // one-wave blocks for the persistent cell kernels
    int32_t lpq;             // lanes per query of the point-parallel count (1, 2 or 4)
    int32_t lpq_border;      // lanes per query of the point-parallel output kernel
    int32_t cellpar_blocks;  // blocks of the cell-parallel count and border kernels
    int32_t cellpar;         // 1: cell-parallel output kernel, 0: point-parallel
    int32_t count_cell;      // 1: cell-parallel range count (writes fcore and rep), 0: point-parallel
    int32_t union_cells;     // point-parallel path: 1 cell-level merge kernels, 0 the cell-parallel merge
    int32_t rank_split;      // 1: stride rank with 64-key blocks (small strides), 0: 256-key blocks
    int32_t border_regroup;  // point-parallel outputs: 1 one lane per core, non-cores regrouped onto 4 lanes
};

enum Phase {
    kPhStrideSort = 0,  // stride tile sort + survivor counts
    kPhMerge,           // stride rank merge + survivor compaction
    kPhMergeFinal,
    kPhCells,
    kPhRows,
    kPhCount,
    kPhUnion,
    kPhBorder,
    kNumPhases
};
extern const char *const kPhaseNames[kNumPhases];

// Enqueues one full slide for every stream of the batch. `ev` (optional,
// 2*kNumPhases events) brackets each phase selected by phase_mask.
hipError_t launch_slide(const BatchArgs &a, const LaunchShape &sh, hipStream_t st, hipEvent_t *ev,
                        uint32_t phase_mask);
// Enqueues the first-window load (sort of W points into the window arrays).
hipError_t launch_init(const BatchArgs &a, int32_t itiles, int32_t passes, hipStream_t st);
// Copies n16 16-byte words (profiler byte-counter calibration).
hipError_t launch_copy16(const void *src, void *dst, size_t n16, hipStream_t st);

}  // namespace rtdbs
