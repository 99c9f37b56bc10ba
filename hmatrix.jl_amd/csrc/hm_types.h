// Plain types and constants of the plan layout, shared by the host layout
// builder (hm_layout.cpp, no HIP) and the kernels.  Not part of the C ABI.
#pragma once

#include <stdint.h>

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
// Right-hand sides one pass of the single-right-hand-side layout serves
// when a product has several (FP64 H*X, or plans without the batched layout).
constexpr int MULTI_RHS = 8;
// Partial sums of the loss reduction (workgroups of its first pass, at most).
constexpr int LOSS_PARTS = 256;
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

// Projection item: out[j] = sum_{i < rows} fac[base + j*ld + i] x[xoff + i]
// for j < kc: rows [row0, row0 + rows) of rank columns [j0, j0 + kc) of one
// leaf's B (col-major, ld = n), with base and xoff already pointing at
// row row0 / column j0, so a wavefront needs this record only.
struct ProjItem {
    int64_t base;     // arena offset of B[row0, j0]
    int64_t xoff;     // index into x of row row0 (col0 + row0)
    int64_t out;      // destination: partial buffer or t buffer
    int32_t ld;
    int32_t rows;
    int32_t kc;       // 1 .. PROJ_KMAX
    int32_t pad;      // member of a many-matrix plan (0 otherwise)
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
    int32_t pad;      // member of a many-matrix plan (0 otherwise)
};

// One-launch single-right-hand-side product of a many-matrix plan (nmem
// members).  Workgroup b of the launch counts its finished items on slot
// b % FUSED_SLOTS (the XCD it runs on) of its member's counters; every
// counter has a 256-byte line of its own.  Counters (lines), zero between
// products: [items (projection + reduction) | row tiles] per (member, slot),
// then one per member (tile slots complete).  need (uint32): items, then
// projection items, then row tiles per (member, slot), then the number of
// slots with row tiles per member.  redmem = member of each reduction item.
constexpr int FUSED_SLOTS = 8;
constexpr int FUSED_STRIDE = 64;

// ---------------------------------------------------------------------------
// Batched multi-right-hand-side path (FP32 / BF16 / split-BF16 storage).
// Right-hand sides are processed in panels of NB columns.
constexpr int NB = 32;
// Columns of X one projection workgroup stages in LDS.
constexpr int PROJ_SUPER = 256;

// Row pass of the batched path: the rows are cut into strips of SROWS rows,
// one wavefront per strip (MFMA M = 16).  A strip is a list of K-steps of
// 32 columns; each K-step multiplies a block of A fragments (pre-packed in
// the MFMA A-operand order) with 32 rows of B that the wavefront loads
// straight from global memory: X rows [src, src + 32) for a dense piece, or
// 32 gathered rows of t for the strip's panel of low-rank columns.
constexpr int SROWS = 16;
// Row-pass workgroup shapes: nw wavefronts (one per strip at a time) take
// spw strips each of a band of nw * spw strips, sharing the band's X window
// of at most xsm rows (all NB columns), staged in LDS as split BF16; a band
// whose dense pieces span more rows reads X straight from memory.  The plan
// takes the narrow shape (two workgroups of 256 rows per CU) unless more
// than a tenth of its bands would not stage their window; then the wide one
// (one workgroup of 512 rows per CU), for leaves that reach far from the
// diagonal (strong admissibility).  xsm / 8 is even: the LDS octet stride
// xsm / 8 + 1 is odd.
// FP64 plans stage the window as doubles (256 bytes per row of X): bands
// of 16 strips as in the narrow shape (8 wavefronts, two strips each), one
// workgroup and a 132 KiB window per CU.
struct RowShape {
    int nw, spw, xsm;
};
constexpr RowShape ROW_NARROW = {8, 2, 528};
constexpr RowShape ROW_WIDE = {16, 2, 1056};
constexpr RowShape ROW_F64 = {8, 2, 528};

// Strip record, 32-bit words, 16-byte aligned:
//   0 row0, 1 rows (<= 16), 2 nd = dense K-steps, 3 np = panel K-steps,
//   4 word offset (from the record start) of the first t-row index block,
//   5 word offset in the record arena of the strip's next record (0: none;
//     a strip with more than SREC_MAXK K-steps continues there), 6-7 zero,
//   then 4 words per K-step, the nd dense ones first:
//     afrag  offset of the K-step's A fragments in 16-byte units;
//     src    dense: first X row; panel: word offset (from the record start)
//            of its 32 t-row indices (entries past the panel point at a
//            zero row of t);
//     info   roff | rows << 4 | nOct << 9 | kind << 12, where the A block
//            covers strip rows [roff, roff + rows) and octets [0, nOct) of
//            the 4 8-column octets (others are zero and not stored); a
//            dense piece whose octets would reach past the last row of X
//            starts up to 7 columns early instead, with zero A columns;
//     zero;
//   then the 32-word t-row index blocks of the panel K-steps (one block of
//   zero-row indices when the record has no panel).  Header and
//   descriptors fit the first 128 words; 128 zero words follow the last
//   record: the row pass loads 128 words from each record start.
// The row pass takes a band's strips in wavefront runs: soffw holds the
// record offsets of band b's strips at [b * band, b * band + n) grouped by
// wavefront, and wpos[b * (nw + 1) + w] is the first slot of wavefront w's
// run (wpos[.. + nw] = n), so wavefront w takes slots [wpos[w], wpos[w+1]).
// The runs balance the strips' K-steps across the wavefronts.
constexpr int SREC_HDR = 8;
constexpr int SREC_MAXK = (128 - SREC_HDR) / 4;
// A fragments of one K-step: `planes` planes of nOct octets x rows x 8
// values (lane (g, r) reads values [8g, 8g + 8) of strip row roff + r).
// FP64 fragments: 4 x rows lanes of 2 nOct values, lane (g, r) holding
// columns 4 m + g (m < 2 nOct) of strip row roff + r; 8 zero values follow
// the last fragment (a lane reads 4 16-byte units from its start).
constexpr int SK_KIND_PANEL = 1;

// Projection: the rank columns (leaf b, column j) crossing a super-tile of
// PROJ_SUPER columns of X form groups of 16; gps groups per super-tile (the
// largest count of the plan; unused groups hold empty descriptors).  Column
// descriptor (uint4): x = offset / 8 of the column's B values in the slice
// arena, y = s0 | s1 << 8 | ngroups << 16 with [s0, s1) the 32-row K-steps
// of the super-tile the leaf reaches (values stored for those rows only,
// zeros around the leaf; split-BF16: hi rows then lo rows), z = float offset
// of the result row t[j][0..NB) (partial or t; -1 for no column).  The
// arena starts with VZERO zero elements, read for K-steps a column lacks;
// element offsets stay below 2^32.
constexpr int VZERO = 16;

// Batched reduction: t[tOff + e] = sum_c part[pOff + c*len + e], e < len.
struct BRed {
    int64_t pOff;
    int64_t tOff;
    int32_t nchunk;
    int32_t len;
};

// Gradient item: rows [i0, i0 + rows) x columns [j0, j0 + ncol) of one
// gradient block in the exported arena layout (col-major, ld rows), one
// wavefront each: lane = row within a 64-row strip, at most GRAD_STRIPS
// strips.  kind 0: dC = res x' (m x n), 1: dA = res t' (m x k),
// 2: dB = x s' (n x k).
constexpr int GRAD_STRIPS = 16;
struct GradItem {
    int64_t dst;      // arena offset of element (i0, j0)
    int64_t vec0;     // index of row i0 in the row vector (dC, dA: res; dB: x)
    int64_t aux;      // index of column j0 in the column vector (dC: x; dA: t; dB: s)
    int32_t ld;       // rows of the block (m, or n for dB)
    int32_t rows;     // <= 64 * GRAD_STRIPS
    int32_t ncol;
    int32_t kind;
};

}  // namespace hm
