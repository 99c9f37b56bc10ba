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
// Rows of V (or A for the transposed product) one projection item streams.
constexpr int PROJ_CHUNK = 1024;
// Rank columns one projection item accumulates in registers.
constexpr int PROJ_KMAX = 16;
// Waves per workgroup in the generic kernels.
constexpr int WAVES = 4;

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

// Row-owner item of the batched path.
struct BItem {
    int64_t off;    // dense: fragment base (elements); rk: U slice base (elements)
    int64_t aux;    // dense: col0; rk: t offset (floats, layout [j][NB])
    int32_t kind;
    int32_t dim;    // dense: n; rk: k
};

struct BTile {
    int64_t row0;
    int32_t rows;
    int32_t ib, ie;
    int32_t pad;
};

// Projection super-tile: columns [col0, col0 + rows) of X staged in LDS.
struct PTile {
    int64_t col0;
    int32_t rows;
    int32_t ib, ie;
    int32_t pad;
};

// Projection item: partial^T = X[slice]^T V[slice] for one low-rank leaf.
// The V slice covers super-tile rows [ls, ls + nrow) (ls and nrow multiples
// of 8, zero rows where the leaf does not reach), column-major, ld nrow.
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
    BTile *tiles = nullptr;
    int64_t ntiles = 0;
    BItem *items = nullptr;
    PTile *ptiles = nullptr;
    int64_t nptiles = 0;
    PItem *pitems = nullptr;
    BRed *red = nullptr;
    int64_t nred = 0;
    void *frag = nullptr;    // dense fragments
    void *uslc = nullptr;    // U row slices per tile ([j][R4] column-major)
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

struct Plan {
    int device = 0;
    int storage = 0;
    int64_t nrows = 0, ncols = 0;
    int64_t ndense = 0, nrk = 0, dense_elems = 0, rk_elems = 0, maxrank = 0;
    int64_t device_bytes = 0;
    void *fac = nullptr;      // storage arena (hi plane for BF16X2)
    void *fac_lo = nullptr;   // lo plane for BF16X2
    Direction fwd, bwd;
    // adjoint: gradient items and the balance exponent of every rank column
    GradItem *grad = nullptr;
    int64_t ngrad = 0;
    int32_t *rexp = nullptr;       // length fwd.tLen
    void *lossbuf = nullptr;       // one double
    int64_t ndata = 0;
    std::vector<void *> allocs;
};

// Launchers (hm_kernels.hip).  Vectors are double for HM_F64, float otherwise.
hipError_t launch_matvec(const Plan &p, const Direction &d, const void *x, void *y,
                         double alpha, double beta, hipStream_t s);
// Adjoint: gradients from res, x and the t vectors the two products left in
// fwd.t (t = B' x) and bwd.t (s = A' res); optional loss into p.lossbuf.
hipError_t launch_adjoint_grads(const Plan &p, const void *x, const void *res, void *grad,
                                bool want_loss, hipStream_t s);
// Batched: one panel of ncol <= NB right-hand sides, float vectors.
hipError_t launch_batched(const Plan &p, const Direction &d, int ncol, const float *X, int64_t ldx,
                          float *Y, int64_t ldy, float alpha, float beta, hipStream_t s);

}  // namespace hm
