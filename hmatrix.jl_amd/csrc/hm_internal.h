// Internal types shared by the plan builder (hm_plan.hip) and the kernels
// (hm_kernels.hip).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

namespace hm {

enum : int { KIND_DENSE = 0, KIND_RK = 1 };

// Rows of one work tile in the row-owner phase: one wavefront lane per row.
constexpr int TILE_ROWS = 64;
// Default rows of V (or A for the transposed product) one projection item
// streams; the plan's configuration may change it.
constexpr int PROJ_CHUNK = 1024;
// Rank columns one projection item accumulates in registers.
constexpr int PROJ_KMAX = 16;
// Waves per workgroup in the generic kernels.
constexpr int WAVES = 4;
// Size of the plan's zero buffer (covers one t vector and any 16-byte load).
constexpr int ZERO_BYTES = 4096;

// One leaf as the kernels see it.  Offsets are element offsets into the
// storage arena; for the transposed direction A/B and row/col are swapped
// and dense leaves point at a transposed copy.
struct BlkDev {
    int64_t offA;   // dense: C (m x n col-major, ld m); rk: A (m x k, ld m)
    int64_t offB;   // rk: B (n x k col-major, ld n)
    int64_t row0;
    int64_t col0;
    int64_t tOff;   // rk: offset of t = B' x (k entries per right-hand side)
    int32_t m, n, k, kind;
};

// Projection item: t[j] (+)= sum_{i in rows} B[i, j] x[col0 + i] for
// j in [j0, j0 + kc).
struct ProjItem {
    int32_t blk;
    int32_t row0;     // first row of B
    int32_t rows;
    int32_t j0;
    int32_t kc;
    int32_t pad;
    int64_t out;      // destination: partial buffer or t buffer
};

// Reduction item: t[tOff + j] = sum_c part[pOff + c*kc + j], with part the
// partial area that follows t in the scratch buffer.
struct RedItem {
    int64_t pOff;
    int64_t tOff;
    int32_t nchunk;
    int32_t kc;
};

// Row tile of the owner phase.  The tile's work is a stream of columns
// [c0, c1): column c holds `rows` values at stream[base + (c - c0)*rows] and
// is multiplied by mult(c) = x[id] (id >= 0, dense leaf column) or
// t[~id] (id < 0, low-rank column).
struct TileDev {
    int64_t row0;
    int64_t base;
    int32_t rows;
    int32_t c0, c1;
    int32_t pad;
};

// ---------------------------------------------------------------------------
// Batched multi-right-hand-side path (FP32 / BF16 / split-BF16 storage).
// Right-hand sides are processed in panels of NB columns.
constexpr int NB = 32;
// Columns of X one projection workgroup stages in LDS.
constexpr int PROJ_SUPER = 256;

// k-rows of B one staging batch of the row-owner kernel holds in LDS.
constexpr int KSTAGE = 128;
// K-steps (32 k-rows each) per batch.
constexpr int KSLOTS = KSTAGE / 32;

// K segment of a batched row tile: an R x klen block of A (a dense leaf's
// rows, or the tile's panel of low-rank columns) whose B rows are X rows
// [src, src + klen) (kind 0) or the t vectors tidx[src .. src + klen)
// (kind 1).  It is staged at LDS row `lrow` of its batch.
struct BSeg {
    int64_t src;
    int32_t klen;
    int32_t kind;
    int32_t lrow;
    int32_t pad;
};

// One K-step (32 k values) of A fragments: all M-tiles at `frag`; B rows at
// LDS row `krow`; nOct = number of stored 8-value octets (4, or fewer in a
// segment's last K-step).
struct BKs {
    int64_t frag;
    int32_t krow;
    int32_t nOct;
};

// Staging batch: segments [sb, se) staged together, K-steps [kb, ke).
struct BBatch {
    int32_t sb, se, kb, ke;
};

struct BTile {
    int64_t row0;
    int32_t rows;
    int32_t bb, be;    // batch range
    int32_t pad;
};

// Per-tile descriptor blob read by the row-owner kernel in one round trip:
// BLOB_STRIDE 32-bit words per tile.  Layout (words):
//   0-1 row0, 2 rows, 3 nbatch, 4 nseg, 5 nks, 6 ntoff,
//   7 words of the full blob when it does not fit inline (0 otherwise),
//   8-9 word offset of that full blob in the overflow arena,
//   then from word BLOB_HDR: batches (4 words: sb, se, kb, ke; indices local
//   to the tile), segments (6 words: src lo/hi, klen, kind, lrow, tbase),
//   K-steps (4 words: frag lo/hi, krow, nOct), t offsets (1 word each, in
//   units of NB floats; kind-1 segment row r uses toff[tbase + r]).
constexpr int BLOB_STRIDE = 128;
constexpr int BLOB_HDR = 16;
constexpr int BLOB_MAX = 2048;      // words of the largest blob the kernel stages

// Projection work unit: columns [col0, col0 + rows) of X and the B slices
// of items [ib, ie) (bytes [vbeg, vbeg + vlen) of the slice arena, split
// BF16 planes or FP32), all staged in LDS.  A super-tile whose slices exceed
// the stage or PITEMS_MAX items is split into several units.
constexpr int VSTAGE_BYTES = 32768;
constexpr int PITEMS_MAX = 64;
struct PTile {
    int64_t col0;
    int64_t vbeg;
    int32_t vlen;
    int32_t rows;
    int32_t ib, ie;
};

// Projection item: partial^T = X[slice]^T V[slice] for one low-rank leaf.
// The V slice covers super-tile rows [ls, ls + nrow) (ls and nrow multiples
// of 8, zero rows where the leaf does not reach), column-major, ld nrow; for
// split-BF16 storage the lo plane follows the hi plane (k*nrow elements).
struct PItem {
    int64_t voff;   // V slice base (elements)
    int64_t out;    // float offset of the k x NB result ([j][NB])
    int32_t ls;
    int32_t nrow;
    int32_t k;
    int32_t pad;
};

// Batched reduction: t[tOff + e] = sum_c part[pOff + c*len + e], e < len.
struct BRed {
    int64_t pOff;
    int64_t tOff;
    int32_t nchunk;
    int32_t len;
};

struct Batched {
    bool ready = false;
    int64_t ntiles = 0;
    uint32_t *blobs = nullptr;    // ntiles x BLOB_STRIDE words (multi-batch tiles)
    uint32_t *blobx = nullptr;    // overflow blobs
    PTile *ptiles = nullptr;
    int64_t nptiles = 0;
    PItem *pitems = nullptr;
    BRed *red = nullptr;
    int64_t nred = 0;
    void *frag = nullptr;    // A fragments of every tile segment
    void *vslc = nullptr;    // V slices per super-tile
    float *t = nullptr;      // [t (tLen) | partials (partLen)]
    int64_t tLen = 0, partLen = 0;
};

// Device-side description of one product direction (H or H').
struct Direction {
    int64_t nrows = 0, ncols = 0;
    BlkDev *blk = nullptr;
    int64_t nblk = 0;
    ProjItem *proj = nullptr;
    int64_t nproj = 0;
    RedItem *red = nullptr;
    int64_t nred = 0;
    TileDev *tiles = nullptr;
    int64_t ntiles = 0;
    int32_t *mult = nullptr;          // multiplier index per stream column
    int64_t ncolumns = 0;
    void *stream = nullptr;           // packed columns (storage type; hi plane)
    void *stream_lo = nullptr;        // lo plane (BF16X2)
    int64_t tLen = 0;      // length of t (one right-hand side)
    int64_t partLen = 0;   // length of the chunk partials
    void *t = nullptr;     // scratch [t (tLen) | partials (partLen)]
    Batched bat;
};

// Gradient item: elements [e0, e1) of one gradient block in the exported
// arena layout.  kind 0: dC (m x n), 1: dA (m x k), 2: dB (n x k).
struct GradItem {
    int64_t dst;      // offset of the block in the gradient arena
    int64_t vec0;     // first index into the vector of the block's rows (dC, dA: res; dB: x)
    int64_t aux;      // dC: col0 (index into x); dA/dB: t offset
    int64_t e0, e1;
    int32_t ld;       // rows of the block (m, or n for dB)
    int32_t kind;
};

// Per-phase kernel timing (instrumentation): each product records four
// events on its stream -- before the projection kernel, after it, after the
// reduction and after the row pass -- into a ring of TIMING_RING records;
// reading folds the elapsed times into ms[] and empties the ring.
constexpr int TIMING_RING = 1024;
struct Timing {
    bool on = false;
    std::vector<hipEvent_t> ev;   // TIMING_RING * 4 events, created on enable
    int64_t used = 0;             // records in the ring
    double ms[3] = {0.0, 0.0, 0.0};
    int64_t count = 0;            // records folded into ms
};

struct Plan {
    int device = 0;
    int storage = 0;
    int proj_chunk = PROJ_CHUNK;   // rows per projection item (single right-hand side)
    int row_waves = WAVES;         // wavefronts per row tile (single right-hand side): 2, 4, 8
    int flags = 0;
    int64_t nrows = 0, ncols = 0;
    int64_t ndense = 0, nrk = 0, dense_elems = 0, rk_elems = 0, maxrank = 0;
    int64_t device_bytes = 0;
    void *fac = nullptr;      // storage arena (hi plane for BF16X2)
    void *fac_lo = nullptr;   // lo plane for BF16X2
    // ZERO_BYTES of zeros: inactive lanes load from here instead of taking a
    // branch, so a wavefront's loads all issue before its first wait.
    void *zeros = nullptr;
    Direction fwd, bwd;
    // adjoint: gradient items and the balance exponent of every rank column
    GradItem *grad = nullptr;
    int64_t ngrad = 0;
    int32_t *rexp = nullptr;       // length fwd.tLen
    void *lossbuf = nullptr;       // one double
    int64_t ndata = 0;
    std::vector<void *> allocs;
    Timing timing;
};

// Launchers (hm_kernels.hip).  Vectors are double for HM_F64, float otherwise.
// ev (may be NULL): three events recorded after the projection, the
// reduction and the row pass.
hipError_t launch_matvec(const Plan &p, const Direction &d, const void *x, void *y,
                         double alpha, double beta, hipStream_t s, hipEvent_t *ev = nullptr);
// Adjoint: gradients from res, x and the t vectors the two products left in
// fwd.t (t = B' x) and bwd.t (s = A' res); optional loss into p.lossbuf.
hipError_t launch_adjoint_grads(const Plan &p, const void *x, const void *res, void *grad,
                                bool want_loss, hipStream_t s);
// Batched: one panel of ncol <= NB right-hand sides, float vectors.
hipError_t launch_batched(const Plan &p, const Direction &d, int ncol, const float *X, int64_t ldx,
                          float *Y, int64_t ldy, float alpha, float beta, hipStream_t s,
                          hipEvent_t *ev = nullptr);

}  // namespace hm
