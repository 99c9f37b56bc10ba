// Host layout builder: turns the oriented leaves of one product direction
// into the work lists and packed arenas the kernels read, and checks every
// offset of the result against the arena it indexes.  Plain C++ (no HIP):
// the plan builder (hm_plan.hip) uploads what it returns, and the same
// source builds on the CPU alone, e.g. under AddressSanitizer and
// UndefinedBehaviorSanitizer.  Not part of the C ABI.
#pragma once

#include <stdint.h>

#include <string>
#include <utility>
#include <vector>

#include "hm_types.h"

namespace hm {

// storage: enum hm_storage, strip_shape: enum hm_strip_shape (hmat_amd.h)
struct OB {  // oriented leaf
    int kind;
    int64_t row0, col0, m, n, k, offA, offB;
};

uint16_t f32_to_bf16_rne(float f);
float bf16_to_f32_host(uint16_t h);
size_t storage_size(int st);
size_t acc_size(int st);
// Convert a double array to the storage arena(s) on the host.
void pack_storage(int st, const double *src, int64_t n, std::vector<uint8_t> &hi, std::vector<uint8_t> &lo);
bool contributes(const OB &o);
// Row tiles of at most TILE_ROWS rows, cut at every leaf's first and end
// row and at the rows in `extra` (member boundaries of a many-matrix plan).
std::vector<std::pair<int64_t, int32_t>> make_tiles(int64_t nrows, const std::vector<OB> &obs,
                                                    const std::vector<int64_t> &extra = {});

// Both orientations of a validated block table (caller offsets into data):
// low-rank factors copied into one compact arena (rkv, A then B per leaf)
// with each column pair balanced by powers of two (rexp: the exponents, in
// table order); dense leaves of H stay in data, those of H' are transposed
// into dT.  fw / bw: the leaves of H and H' with offsets into rkv and data
// (fw) or dT (bw).
struct Oriented {
    std::vector<OB> fw, bw;
    std::vector<double> rkv, dT;
    std::vector<int32_t> rexp;
};
void orient(const std::vector<OB> &leaves, const double *data, int64_t rk_elems, int64_t dense_elems,
            Oriented &O);

// Single-right-hand-side layout of one direction.
struct DirectionHost {
    std::vector<BlkDev> blk;
    std::vector<ProjItem> proj;
    std::vector<RedItem> red;
    std::vector<TileDev> tiles;
    std::vector<int32_t> mult;          // multiplier index per stream column
    std::vector<uint8_t> shi, slo;      // packed column stream (storage type)
    int64_t ncolumns = 0, tLen = 0, partLen = 0;
    int nmem = 0;                       // many-matrix one-launch layout: members
    std::vector<uint32_t> need;         // and its counter targets (hm_types.h)
    std::vector<int32_t> redmem;        // member of each reduction item
};
// Projection items over the factor arena `fac` (rkv) and the row tiles'
// column streams built from `fac` (low-rank A) and `dense` (dense leaves).
// mrow (many-matrix one-launch plans, else empty): first row of each member
// in this direction, and nrows last; items and tiles are then ordered by
// member, largest members first.
void layout_direction(int storage, int proj_chunk, int row_waves, int64_t nrows, const std::vector<OB> &obs,
                      const double *fac, const double *dense, const std::vector<int64_t> &mrow,
                      DirectionHost &D);
// Every projection, reduction and tile offset of D inside its arena (fac of
// fac_elems elements, x of ncols, t and partials, the column stream, y of
// nrows), every t entry and partial written exactly once, every row in one
// tile, the counter targets consistent with the items.  Returns "" or the
// first violation.
std::string check_direction(const DirectionHost &D, int storage, int row_waves, int64_t nrows, int64_t ncols,
                            int64_t fac_elems);

// Element arena of the batched layout: double, float, or BF16 (one plane,
// or the split hi / lo planes of each block).
struct Arena {
    int mode = 0;                 // 1 f32, 2 bf16 planes, 3 f64
    std::vector<float> f;
    std::vector<uint16_t> h;
    std::vector<double> d;
    int64_t size() const {
        return mode == 1 ? static_cast<int64_t>(f.size())
                         : mode == 3 ? static_cast<int64_t>(d.size()) : static_cast<int64_t>(h.size());
    }
    void align(int64_t a) {
        while (size() % a) push(0.0);
    }
    void push(double v);
    void push_lo(double v);       // low half of the split of v
};

// Host image of the batched layout of one direction.
struct BatchedHost {
    int64_t nstrips = 0;
    RowShape shape = ROW_NARROW;
    std::vector<uint32_t> soff, recs;
    std::vector<uint32_t> soffw, wpos;  // record offsets in wavefront order, run starts (hm_types.h)
    std::vector<int32_t> wgx;
    std::vector<uint32_t> pcol;        // projection column descriptors, 4 words each
    int64_t nst = 0, gps = 0;
    std::vector<BRed> red;
    Arena fr, vs;
    int64_t tLen = 0, partLen = 0;
};
// Batched (multi-RHS) layout of one direction for storage `mode` (enum
// hm_storage: 0 FP64 fragments and slices, 1 FP32, 2 BF16, 3 split BF16).
// rk factors come from `fac` (balanced FP64 arena), dense leaves from
// `dense` at each leaf's offA.  FP64 plans always take the row shape
// ROW_F64 (strip_shape is ignored).  Returns false when the matrix does not fit
// the layout (the product then runs one right-hand side at a time).
bool layout_batched(int mode, int strip_shape, int64_t nrows, int64_t ncols, const std::vector<OB> &obs,
                    const double *fac, const double *dense, BatchedHost &H);
// Every strip record (header, K-step descriptors, t-row index blocks, the
// 128 words the row pass reads from each record start), every A fragment
// block, every staged X window, every projection column descriptor and B
// slice, every reduction inside its arena, and every t entry and partial
// written exactly once.  Returns "" or the first violation.
std::string check_batched(const BatchedHost &H, int mode, int64_t nrows, int64_t ncols);

}  // namespace hm
