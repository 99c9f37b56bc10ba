// Device-side plan structures shared by the plan builder (hm_plan.hip) and
// the kernels (hm_kernels.hip, hm_batched.hip); the layout types and
// constants they point at are in hm_types.h.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "hm_types.h"

namespace hm {

// 16-byte loads of data a product streams once, where one wavefront
// instruction covers whole cache lines (the row pass's BF16 A fragments:
// lane l reads bytes [16 l, 16 l + 16) of a contiguous block):
// non-temporal, so that they do not displace from L2 the X windows that
// neighbouring bands read again.  (Data whose lines several instructions
// share, such as FP64 fragments read 64 bytes per lane, stays cached: a
// non-temporal load would fetch each line once per instruction.  The
// projection's X tile and B slices measured slower non-temporal.)
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ uint4 ld_stream(const void *p) {
    const u32x4_t v = __builtin_nontemporal_load(static_cast<const u32x4_t *>(p));
    return uint4{v[0], v[1], v[2], v[3]};
}

// Device arrays of the one-launch product (counter layout: hm_types.h).
struct Fused {
    int nmem = 0;
    uint32_t *counts = nullptr;
    uint32_t *need = nullptr;
    int32_t *redmem = nullptr;
};

// Device copies of the batched layout (formats: hm_types.h).
struct Batched {
    bool ready = false;
    int64_t nstrips = 0;
    uint32_t *soffw = nullptr;    // word offsets of the strip records in wavefront runs
    uint32_t *wpos = nullptr;     // run starts per band (formats: hm_types.h)
    uint32_t *recs = nullptr;     // strip records
    RowShape shape = ROW_NARROW;  // row-pass workgroup shape
    int32_t *wgx = nullptr;       // per band: X window (first row, rows; 0 rows: not staged)
    int64_t nunstaged = 0;        // bands whose window is not staged
    uint32_t *ulist = nullptr;    // their indices
    uint4 *pcol = nullptr;        // projection column descriptors [nst][gps][16]
    int64_t nst = 0, gps = 0;     // super-tiles, groups per super-tile
    BRed *red = nullptr;          // largest first
    int64_t nred = 0, nredbig = 0; // items; leading items with more than 32 chunks
    void *frag = nullptr;    // A fragments of every strip K-step
    void *vslc = nullptr;    // B slices per super-tile column
    void *t = nullptr;       // [t (tLen) | partials (partLen) | NB zeros], accumulator type
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
    void *tm = nullptr;    // the same for RM right-hand sides at once, RM values per entry
    Fused fmem;            // many-matrix plans: the one-launch product
    Batched bat;
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
    int strip_shape = 0;           // enum hm_strip_shape requested for the batched row pass
    int flags = 0;
    int64_t nrows = 0, ncols = 0;
    int64_t ndense = 0, nrk = 0, dense_elems = 0, rk_elems = 0, maxrank = 0;
    int64_t device_bytes = 0;
    void *fac = nullptr;      // storage arena (hi plane for BF16X2)
    void *fac_lo = nullptr;   // lo plane for BF16X2
    // ZERO_BYTES of zeros: inactive lanes load from here instead of taking a
    // branch, so a wavefront's loads all issue before its first wait.
    void *zeros = nullptr;
    uint32_t *fault = nullptr;     // raised when a bounded wait of the one-launch product ran out
    Direction fwd, bwd;
    // adjoint: gradient items and the balance exponent of every rank column
    GradItem *grad = nullptr;
    int64_t ngrad = 0;
    int32_t *rexp = nullptr;       // length fwd.tLen
    void *lossbuf = nullptr;       // [loss | LOSS_PARTS partial sums], doubles
    double *losshost = nullptr;    // pinned host copy of the loss
    int64_t ndata = 0;
    std::vector<void *> allocs;
    Timing timing;
};

// Launchers (hm_kernels.hip).  Vectors are double for HM_F64, float otherwise.
// ev (may be NULL): three events recorded after the projection, the
// reduction and the row pass.  scratch (may be NULL): device buffer of at
// least (d.tLen + d.partLen) accumulator elements used instead of d.t.
hipError_t launch_matvec(const Plan &p, const Direction &d, const void *x, void *y,
                         double alpha, double beta, hipStream_t s, hipEvent_t *ev = nullptr,
                         void *scratch = nullptr);
// The projection phase of a single-right-hand-side product alone: t = B' x
// of every low-rank leaf of d into d.t.
hipError_t launch_projection(const Plan &p, const Direction &d, const void *x, hipStream_t s);
// Adjoint: gradients from res, x and the t vectors the two products left in
// fwd.t (t = B' x) and bwd.t (s = A' res); optional loss into p.lossbuf.
hipError_t launch_adjoint_grads(const Plan &p, const void *x, const void *res, void *grad,
                                bool want_loss, hipStream_t s);
// Several right-hand sides on the single-right-hand-side layout: one panel
// of ncol <= MULTI_RHS columns (any storage; vectors of the plan's vector type).
hipError_t launch_matvec_m(const Plan &p, const Direction &d, int ncol, const void *x, int64_t ldx, void *y,
                           int64_t ldy, double alpha, double beta, hipStream_t s, hipEvent_t *ev = nullptr);
// Batched: one panel of ncol <= NB right-hand sides, float vectors.
hipError_t launch_batched(const Plan &p, const Direction &d, int ncol, const float *X, int64_t ldx,
                          float *Y, int64_t ldy, float alpha, float beta, hipStream_t s,
                          hipEvent_t *ev = nullptr);
// Batched FP64 plans (hm_batched64.hip): one panel of ncol <= NB
// right-hand sides, double vectors, on the f64 MFMA.
hipError_t launch_batched64(const Plan &p, const Direction &d, int ncol, const double *X, int64_t ldx,
                            double *Y, int64_t ldy, double alpha, double beta, hipStream_t s,
                            hipEvent_t *ev = nullptr);

}  // namespace hm
