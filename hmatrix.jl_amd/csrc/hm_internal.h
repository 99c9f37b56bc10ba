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
constexpr int PROJ_CHUNK = 2048;
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

// Row tile of the owner phase.
struct TileDev {
    int64_t row0;
    int32_t rows;
    int32_t ib, ie;   // item range in the item list
    int32_t pad;
};

struct TileItem {
    int32_t blk;
    int32_t ro;       // first row of the tile inside the leaf
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
    TileItem *items = nullptr;
    int64_t nitems = 0;
    int64_t tLen = 0;      // length of t (one right-hand side)
    int64_t partLen = 0;   // length of the chunk partials
    void *t = nullptr;     // scratch [t (tLen) | partials (partLen)]
    const void *denseA = nullptr;   // arena for dense leaves (fwd: fac; bwd: transposed copy)
    const void *denseA_lo = nullptr;
    Batched bat;
};

struct Plan {
    int device = 0;
    int storage = 0;
    int64_t nrows = 0, ncols = 0;
    int64_t ndense = 0, nrk = 0, dense_elems = 0, rk_elems = 0, maxrank = 0;
    int64_t device_bytes = 0;
    void *fac = nullptr;      // storage arena (hi plane for BF16X2)
    void *fac_lo = nullptr;   // lo plane for BF16X2
    void *facT = nullptr;     // transposed dense leaves (hi)
    void *facT_lo = nullptr;
    Direction fwd, bwd;
    std::vector<void *> allocs;
};

// Launchers (hm_kernels.hip).  Vectors are double for HM_F64, float otherwise.
hipError_t launch_matvec(const Plan &p, const Direction &d, const void *x, void *y,
                         double alpha, double beta, hipStream_t s);
// Batched: one panel of ncol <= NB right-hand sides, float vectors.
hipError_t launch_batched(const Plan &p, const Direction &d, int ncol, const float *X, int64_t ldx,
                          float *Y, int64_t ldy, float alpha, float beta, hipStream_t s);

}  // namespace hm
