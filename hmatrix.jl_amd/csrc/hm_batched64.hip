// FP64 multi-right-hand-side product Y = beta*Y + alpha*H*X for panels of up
// to NB = 32 right-hand sides on v_mfma_f64_16x16x4f64: one pass over the
// FP64 factors serves all 32 columns (the reference multiplies a matrix
// column by column, src/harithm.jl:317-323, each column a full pass of
// hmat_matvec!, src/harithm.jl:340-354).
//
// The layout is the batched one of the FP32 / BF16 plans (strip records,
// A fragments in the register order of the MFMA A operand, projection
// column descriptors over 256-row super-tiles, fixed-order reductions),
// with FP64 fragments and slices.  The f64 MFMA multiplies a 16 x 4 A block
// by a 4 x 16 B block; lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[l / 16 + 4 r][l % 16], r < 4 (checked on
// the device).  The row pass computes Y' = X' C' per 16 right-hand sides:
// A = X' (i = right-hand side), B = C' (j = strip row).  A K-step of 32
// columns is eight such MFMAs: in MFMA m, lane (g, r) supplies column
// 4 m + g of its fragment row (FP64 fragments hold, per lane (g, r), the
// columns 4 m + g of row r) and row src + 4 m + g of X, and the lane ends
// with strip row r of right-hand sides g + 4 r'.  A K-step of nOct octets
// runs MFMAs m < 2 nOct only.  The sum over a K-step is the same as in the
// single-column product, in another order.
//
// Three kernels per panel, as in the other storages:
//   kb_proj64  t_b = B_b' X[cols(b)] per 256-row super-tile (partials of
//              leaves wider than one super-tile), X[super-tile] staged in
//              LDS once as doubles, 16 rank columns per MFMA group;
//   kb_reduce<double> the fixed-order sum of the partials (hm_bred.h);
//   kb_rows64  16-row strips, two per wavefront, 16 per workgroup sharing
//              the band's X window (<= 528 rows x 32 columns of doubles,
//              132 KiB) in LDS; dense and panel K-steps on MFMA; every
//              element of Y written once.
#include "hm_bred.h"
#include "hm_internal.h"

namespace hm {

typedef double doublex4 __attribute__((ext_vector_type(4)));
// 16 bytes with 8-byte alignment (leading dimensions may be odd)
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));

__device__ __forceinline__ doublex4 mfma64(double a, double b, doublex4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t rword64(uint32_t h0, uint32_t h1, int i) {
    const uint32_t a = __builtin_amdgcn_readlane(h0, i & 63);
    const uint32_t b = __builtin_amdgcn_readlane(h1, i & 63);
    return i < 64 ? a : b;
}

// ---------------------------------------------------------------------------
// Projection.  One workgroup (4 wavefronts) per super-tile of PROJ_SUPER
// rows of X: the super-tile is staged in LDS as doubles, element (row rho,
// right-hand side c) at pair rho * 16 + c % 16, half c / 16, so one 16-byte
// read gives a lane the A operands of both halves of the panel.  Wavefront
// w takes groups w, w + 4, ... of 16 rank columns: lane (g, j) loads the
// 8 B values of column j for rows [32 s + 8 g, +8) of every K-step s (zeros
// outside the column's K-step range [s0, s1), read from the slice arena's
// zero head), then runs 8 x 8 MFMAs per half.

// K-steps [h * 4, h * 4 + 4) of this lane's column
__device__ __forceinline__ void load_v64(const double *vslc, const uint4 &cd, int g, int h, double (&v)[4][8]) {
    const int s0 = cd.y & 255, s1 = (cd.y >> 8) & 255;
    const uint32_t base = cd.x * 8u + 8u * static_cast<uint32_t>(g) - 32u * static_cast<uint32_t>(s0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int s = 4 * h + q;
        const bool on = s >= s0 && s < s1;
        const double *p = vslc + (on ? base + 32u * s : 0u);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const d2u w = reinterpret_cast<const d2u *>(p)[e];
            v[q][2 * e] = w[0];
            v[q][2 * e + 1] = w[1];
        }
    }
}

__global__ __launch_bounds__(256) void kb_proj64(int64_t gps, const uint4 *__restrict__ pcol,
                                                 const double *__restrict__ vslc, const double *__restrict__ zeros,
                                                 const double *__restrict__ X, int64_t ldx, int64_t ncols, int ncol,
                                                 double *__restrict__ out) {
    __shared__ d2u xs[PROJ_SUPER * 16];
    const int64_t col0 = static_cast<int64_t>(blockIdx.x) * PROJ_SUPER;
    const int rows = static_cast<int>(min<int64_t>(PROJ_SUPER, ncols - col0));
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const uint4 *gdesc = pcol + static_cast<int64_t>(blockIdx.x) * gps * 16;
    uint4 cd = gps > 0 ? gdesc[min(wave, static_cast<int>(gps) - 1) * 16 + r16] : uint4{0u, 0u, 0xffffffffu, 0u};
    if (wave >= gps) cd = uint4{0u, cd.y & 0xffff0000u, 0xffffffffu, 0u};
    // Stage: unit p = column pair j + 16 * row pair q (rows 2q, 2q + 1 of
    // columns j and j + 16); 4 lanes read 64 contiguous bytes of a column.
    constexpr int NU = PROJ_SUPER / 2 * 16;
    constexpr int UT = NU / 256;
    d2u xa[UT], xb[UT];
#pragma unroll
    for (int u = 0; u < UT; ++u) {
        const int p = threadIdx.x + 256 * u;
        const int j = p & 15, q = p >> 4;
        const int64_t row = col0 + 2 * q;
        const bool full = 2 * q + 2 <= rows;
        const double *s0 = (full && j < ncol) ? X + static_cast<int64_t>(j) * ldx + row : zeros;
        const double *s1 = (full && j + 16 < ncol) ? X + static_cast<int64_t>(j + 16) * ldx + row : zeros;
        xa[u] = *reinterpret_cast<const d2u *>(s0);
        xb[u] = *reinterpret_cast<const d2u *>(s1);
        if (!full && 2 * q < rows) {      // the pair cut by the matrix end
            xa[u] = d2u{j < ncol ? X[static_cast<int64_t>(j) * ldx + row] : 0.0, 0.0};
            xb[u] = d2u{j + 16 < ncol ? X[static_cast<int64_t>(j + 16) * ldx + row] : 0.0, 0.0};
        }
    }
    // B values in two halves of 4 K-steps: the second half (or the next
    // group's first) loads during the MFMAs of the other
    double va[4][8], vb[4][8];
    load_v64(vslc, cd, g, 0, va);
#pragma unroll
    for (int u = 0; u < UT; ++u) {
        const int p = threadIdx.x + 256 * u;
        const int j = p & 15, q = p >> 4;
        xs[(2 * q) * 16 + j] = d2u{xa[u][0], xb[u][0]};
        xs[(2 * q + 1) * 16 + j] = d2u{xa[u][1], xb[u][1]};
    }
    __syncthreads();
    const int ngroups = static_cast<int>(__builtin_amdgcn_readfirstlane(cd.y >> 16));
    doublex4 acc0, acc1;
    // K-steps some column of the group reaches (the others are zero)
    auto kmask = [&]() __attribute__((always_inline)) {
        const int s0 = cd.y & 255, s1 = (cd.y >> 8) & 255;
        uint32_t mk = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (__ballot(q >= s0 && q < s1) != 0) mk |= 1u << q;
        return mk;
    };
    uint32_t km = kmask();
    auto half = [&](int h, const double (&v)[4][8]) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!((km >> (4 * h + q)) & 1u)) continue;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                const d2u a = xs[(32 * (4 * h + q) + 8 * g + m) * 16 + r16];
                acc0 = mfma64(v[q][m], a[0], acc0);
                acc1 = mfma64(v[q][m], a[1], acc1);
            }
        }
    };
    for (int gi = wave;; gi += 4) {
        acc0 = doublex4{0.0, 0.0, 0.0, 0.0};
        acc1 = doublex4{0.0, 0.0, 0.0, 0.0};
        load_v64(vslc, cd, g, 1, vb);
        __builtin_amdgcn_sched_barrier(0);
        half(0, va);
        // D[i = column][j = rhs]: lane (g, r16) holds columns g + 4r of
        // right-hand side r16 (acc0) and 16 + r16 (acc1); column c's output
        // row is in the descriptor of lane c
        int o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) o[r] = __shfl(static_cast<int>(cd.z), g + 4 * r, 64);
        const bool more = gi + 4 < ngroups;
        if (more) {            // the next group's descriptor and first half (km keeps this group's mask)
            cd = gdesc[(gi + 4) * 16 + r16];
            load_v64(vslc, cd, g, 0, va);
        }
        __builtin_amdgcn_sched_barrier(0);
        half(1, vb);
        if (more) km = kmask();
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (o[r] >= 0) {
                out[o[r] + r16] = acc0[r];
                out[o[r] + 16 + r16] = acc1[r];
            }
        if (!more) break;
    }
}

// ---------------------------------------------------------------------------
// Row pass.  Workgroup b (consecutive bands on one XCD, as in the other
// storages) owns a band of NW * SPW strips; wavefront w takes its run of
// the band's strips (balanced by K-steps).  The band's X window [cmin, cmin + W) is staged in LDS as
// doubles in the projection's pair layout (row rho, right-hand side c at
// pair rho * 16 + c % 16, half c / 16); a band whose window does not fit
// (W = 0) runs in the second instance and reads X from memory.  Each
// wavefront walks its records as one stream: the A fragments of K-step
// k + 2 are in flight during the MFMAs of K-step k, and at the end of a
// record the next record's first two dense K-steps and its panel K-step (A
// fragments and 16 gathered t values per lane) are issued as the operand
// sets free up.
template <int NW, int SPW, int XSM, bool STAGED>
__global__ __launch_bounds__(64 * NW) void kb_rows64(int64_t nstrips, int64_t wg_per_xcd,
                                                     const uint32_t *__restrict__ soffw,
                                                     const uint32_t *__restrict__ wpos,
                                                     const uint32_t *__restrict__ recs,
                                                     const int32_t *__restrict__ wgx,
                                                     const uint32_t *__restrict__ ulist,
                                                     const double *__restrict__ frag,
                                                     const double *__restrict__ zeros,
                                                     const double *__restrict__ T,
                                                     const double *__restrict__ X, int64_t ldx, int64_t ncols,
                                                     double *__restrict__ Y, int64_t ldy, int ncol, double alpha,
                                                     double beta, int beta_zero) {
    __shared__ d2u xs[STAGED ? XSM * 16 : 1];
    const int64_t wg = STAGED ? static_cast<int64_t>(blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8
                              : static_cast<int64_t>(ulist[blockIdx.x]);
    constexpr int BAND = NW * SPW;
    if (wg * BAND >= nstrips) return;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const int cmin = wgx[2 * wg], W = wgx[2 * wg + 1];
    if ((W > 0) != STAGED) return;
    // this wavefront's run of the band's strips (slots [s0, s0 + nrun));
    // every lane loads one slot of the band, in parallel with the run bounds
    const uint32_t *wp = wpos + wg * (NW + 1);
    const int s0 = static_cast<int>(wp[wave]);
    const int nrun = static_cast<int>(wp[wave + 1]) - s0;
    const uint32_t sv = soffw[min(wg * BAND + l % BAND, nstrips - 1)];
    constexpr uint32_t NONE = 0xffffffffu;

    // window staging: unit p = column pair j + 16 * row pair q
    constexpr int NU = XSM / 2 * 16;
    constexpr int U = (NU + 64 * NW - 1) / (64 * NW);
    d2u xa[STAGED ? U : 1], xb[STAGED ? U : 1];
    if constexpr (STAGED) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 64 * NW * u;
            const int j = p & 15, q = p >> 4;
            const int64_t row = static_cast<int64_t>(cmin) + 2 * q;
            const bool in = p < NU && 2 * q < W;
            const bool full = in && row + 2 <= ncols;
            const double *s0 = (full && j < ncol) ? X + static_cast<int64_t>(j) * ldx + row : zeros;
            const double *s1 = (full && j + 16 < ncol) ? X + static_cast<int64_t>(j + 16) * ldx + row : zeros;
            xa[u] = *reinterpret_cast<const d2u *>(s0);
            xb[u] = *reinterpret_cast<const d2u *>(s1);
            if (in && !full && row < ncols) {     // the pair cut by the last row of X
                xa[u] = d2u{j < ncol ? X[static_cast<int64_t>(j) * ldx + row] : 0.0, 0.0};
                xb[u] = d2u{j + 16 < ncol ? X[static_cast<int64_t>(j + 16) * ldx + row] : 0.0, 0.0};
            }
        }
    }

    // this lane's fragment values of K-step k of record (h0, h1): columns
    // 4m + g, m < 2 nOct (the 4 units read past them belong to the next
    // lanes and meet only skipped MFMAs)
    auto issue_a = [&](uint32_t h0, uint32_t h1, int k, bool on, double (&a)[8]) __attribute__((always_inline)) {
        const int i = SREC_HDR + 4 * k;
        const uint32_t afrag = rword64(h0, h1, i), info = rword64(h0, h1, i + 2);
        const int roff = info & 15, rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
        const bool lon = on && r16 >= roff && r16 < roff + rows;
        const double *p = lon ? frag + 2 * (static_cast<int64_t>(afrag) + nOct * (g * rows + r16 - roff)) : zeros;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const d2u w = reinterpret_cast<const d2u *>(p)[q];
            a[2 * q] = w[0];
            a[2 * q + 1] = w[1];
        }
    };
    // a panel K-step: A fragments and the t rows of columns 8g..8g+7
    auto issue_panel = [&](uint32_t h0, uint32_t h1, int k, bool on, double (&a)[8], double (&t0)[8],
                           double (&t1)[8]) __attribute__((always_inline)) {
        const int i = SREC_HDR + 4 * k;
        const uint32_t afrag = rword64(h0, h1, i), src = rword64(h0, h1, i + 1), info = rword64(h0, h1, i + 2);
        const int rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
        const bool lon = on && r16 < rows;
        const double *p = lon ? frag + 2 * (static_cast<int64_t>(afrag) + nOct * (g * rows + r16)) : zeros;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const d2u w = reinterpret_cast<const d2u *>(p)[q];
            a[2 * q] = w[0];
            a[2 * q + 1] = w[1];
        }
        uint32_t ix[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int wi = (static_cast<int>(src) + 4 * e + g) & 127;
            const int ad = (wi & 63) << 2;
            const uint32_t i0 = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(ad, static_cast<int>(h0)));
            const uint32_t i1 = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(ad, static_cast<int>(h1)));
            ix[e] = wi < 64 ? i0 : i1;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const double *tr = on ? T + static_cast<int64_t>(ix[e]) * NB + r16 : zeros;
            t0[e] = tr[0];
            t1[e] = tr[16];
        }
    };
    doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    auto compute_dense = [&](uint32_t h0, uint32_t h1, int k, const double (&a)[8]) __attribute__((always_inline)) {
        const int i = SREC_HDR + 4 * k;
        const uint32_t src = rword64(h0, h1, i + 1), info = rword64(h0, h1, i + 2);
        const int nm = 2 * static_cast<int>((info >> 9) & 7);     // MFMAs of this K-step
        if constexpr (STAGED) {
            const int rho = static_cast<int>(src) - cmin + g;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                if (m >= nm) break;
                const d2u b = xs[(rho + 4 * m) * 16 + r16];
                acc0 = mfma64(b[0], a[m], acc0);
                acc1 = mfma64(b[1], a[m], acc1);
            }
        } else {
            const int64_t row = static_cast<int64_t>(src) + g;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                if (m >= nm) break;
                const int64_t rm = row + 4 * m;
                const double b0 = r16 < ncol && rm < ncols ? X[static_cast<int64_t>(r16) * ldx + rm] : 0.0;
                const double b1 = 16 + r16 < ncol && rm < ncols ? X[static_cast<int64_t>(16 + r16) * ldx + rm] : 0.0;
                acc0 = mfma64(b0, a[m], acc0);
                acc1 = mfma64(b1, a[m], acc1);
            }
        }
    };
    auto compute_panel = [&](uint32_t h0, uint32_t h1, int k, const double (&a)[8], const double (&t0)[8],
                             const double (&t1)[8]) __attribute__((always_inline)) {
        const int nm = 2 * static_cast<int>((rword64(h0, h1, SREC_HDR + 4 * k + 2) >> 9) & 7);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            if (m >= nm) break;
            acc0 = mfma64(t0[m], a[m], acc0);
            acc1 = mfma64(t1[m], a[m], acc1);
        }
    };
    auto store = [&](int row0, int R) __attribute__((always_inline)) {
        // D layout: lane (g, r16) holds strip row r16 of right-hand sides
        // 16 h + g + 4 r (acc0: h = 0, acc1: h = 1); 16 lanes write 16
        // consecutive rows of one column
        if (r16 < R) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const doublex4 a = h == 0 ? acc0 : acc1;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = 16 * h + g + 4 * r;
                    if (c >= ncol) continue;
                    double *dst = Y + static_cast<int64_t>(c) * ldy + row0 + r16;
                    *dst = beta_zero ? alpha * a[r] : beta * *dst + alpha * a[r];
                }
            }
        }
        acc0 = doublex4{0.0, 0.0, 0.0, 0.0};
        acc1 = doublex4{0.0, 0.0, 0.0, 0.0};
    };
    auto next_of = [&](uint32_t h0, uint32_t h1, int at, int &nat) __attribute__((always_inline)) -> uint32_t {
        const uint32_t ch = rword64(h0, h1, 5);
        if (ch != 0u) {
            nat = at;
            return ch;
        }
        nat = at + 1;
        return at + 1 < nrun ? __builtin_amdgcn_readlane(sv, (s0 + at + 1) & 63) : NONE;
    };
    auto load_rec = [&](uint32_t off, uint32_t &h0, uint32_t &h1) __attribute__((always_inline)) {
        const uint32_t *rw = recs + (off == NONE ? 0u : off);
        h0 = rw[l];
        h1 = rw[64 + l];
    };

    uint32_t c0, c1, n0, n1;
    load_rec(nrun > 0 ? __builtin_amdgcn_readlane(sv, s0 & 63) : NONE, c0, c1);
    int nsi = 0;
    uint32_t noff = nrun > 0 ? next_of(c0, c1, 0, nsi) : NONE;
    load_rec(noff, n0, n1);
    int nd = nrun > 0 ? static_cast<int>(rword64(c0, c1, 2)) : 0;
    double A[8], B[8], C[8];
    double PA[8], PT0[8], PT1[8];
    issue_a(c0, c1, 0, nd > 0, A);
    issue_a(c0, c1, 1, nd > 1, B);
    if constexpr (STAGED) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 64 * NW * u;
            const int j = p & 15, q = p >> 4;
            if (p < NU) {
                xs[(2 * q) * 16 + j] = d2u{xa[u][0], xb[u][0]};
                xs[(2 * q + 1) * 16 + j] = d2u{xa[u][1], xb[u][1]};
            }
        }
    }
    __syncthreads();
    if (nrun == 0) return;
    issue_panel(c0, c1, nd, rword64(c0, c1, 3) > 0u, PA, PT0, PT1);
    for (;;) {
        // dense K-steps in groups of 3 through the sets A, B, C: the
        // fragments of K-step k + 2 are loaded before the MFMAs of k (steps
        // past nd load the zero buffer and compute nothing)
        const int trips = (max(nd, 1) + 2) / 3;
        for (int q = 0; q + 1 < trips; ++q) {
            const int k = 3 * q;
            issue_a(c0, c1, k + 2, k + 2 < nd, C);
            __builtin_amdgcn_sched_barrier(0);
            compute_dense(c0, c1, k, A);
            issue_a(c0, c1, k + 3, k + 3 < nd, A);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 1 < nd) compute_dense(c0, c1, k + 1, B);
            issue_a(c0, c1, k + 4, k + 4 < nd, B);
            __builtin_amdgcn_sched_barrier(0);
            if (k + 2 < nd) compute_dense(c0, c1, k + 2, C);
        }
        // last group: the record after next is loaded first; the next
        // record's first two dense K-steps go into A and B as those free up,
        // its panel into PA / PT after this record's panel
        const int kl = 3 * (trips - 1);
        const bool more = noff != NONE;
        int nnsi = nsi;
        const uint32_t nnoff = more ? next_of(n0, n1, nsi, nnsi) : NONE;
        uint32_t m0, m1;
        load_rec(nnoff, m0, m1);
        const int nnd = more ? static_cast<int>(rword64(n0, n1, 2)) : 0;
        const bool nnp = more && rword64(n0, n1, 3) > 0u;
        issue_a(c0, c1, kl + 2, kl + 2 < nd, C);
        __builtin_amdgcn_sched_barrier(0);
        if (kl < nd) compute_dense(c0, c1, kl, A);
        issue_a(n0, n1, 0, more && nnd > 0, A);
        __builtin_amdgcn_sched_barrier(0);
        if (kl + 1 < nd) compute_dense(c0, c1, kl + 1, B);
        issue_a(n0, n1, 1, more && nnd > 1, B);
        __builtin_amdgcn_sched_barrier(0);
        if (rword64(c0, c1, 3) > 0u) compute_panel(c0, c1, nd, PA, PT0, PT1);
        issue_panel(n0, n1, nnd, nnp, PA, PT0, PT1);
        __builtin_amdgcn_sched_barrier(0);
        if (kl + 2 < nd) compute_dense(c0, c1, kl + 2, C);
        if (rword64(c0, c1, 5) == 0u) store(static_cast<int>(rword64(c0, c1, 0)), static_cast<int>(rword64(c0, c1, 1)));
        if (!more) break;
        c0 = n0;
        c1 = n1;
        n0 = m0;
        n1 = m1;
        nsi = nnsi;
        noff = nnoff;
        nd = nnd;
    }
}

hipError_t launch_batched64(const Plan &p, const Direction &d, int ncol, const double *X, int64_t ldx, double *Y,
                            int64_t ldy, double alpha, double beta, hipStream_t s, hipEvent_t *ev) {
    const Batched &b = d.bat;
    const double *zd = static_cast<const double *>(p.zeros);
    double *t = static_cast<double *>(b.t);
    hipError_t e = hipSuccess;
    if (b.nst > 0 && b.gps > 0) {
        hipLaunchKernelGGL(kb_proj64, dim3(static_cast<unsigned>(b.nst)), dim3(256), 0, s, b.gps, b.pcol,
                           static_cast<const double *>(b.vslc), zd, X, ldx, d.ncols, ncol, t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[0], s);
    if (b.nst > 0 && b.gps > 0 && b.nred > 0) {
        const int64_t nwg = b.nredbig + (b.nred - b.nredbig + 3) / 4;
        hipLaunchKernelGGL(kb_reduce<double>, dim3(static_cast<unsigned>(nwg)), dim3(256), 0, s, b.red, b.nredbig,
                           b.nred, t + b.tLen, zd, t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (b.nstrips > 0) {
        constexpr RowShape SH = ROW_F64;
        if (b.shape.nw != SH.nw || b.shape.spw != SH.spw || b.shape.xsm != SH.xsm) return hipErrorInvalidValue;
        const int64_t band = static_cast<int64_t>(SH.nw) * SH.spw;
        const int64_t nwg = (b.nstrips + band - 1) / band;
        const int64_t per = (nwg + 7) / 8;
        const dim3 block(static_cast<unsigned>(64 * SH.nw));
        const int bz = beta == 0.0 ? 1 : 0;
        const double *fr = static_cast<const double *>(b.frag);
        hipLaunchKernelGGL((kb_rows64<SH.nw, SH.spw, SH.xsm, true>), dim3(static_cast<unsigned>(8 * per)), block, 0,
                           s, b.nstrips, per, b.soffw, b.wpos, b.recs, b.wgx, b.ulist, fr, zd, t, X, ldx, d.ncols, Y, ldy,
                           ncol, alpha, beta, bz);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (b.nunstaged > 0) {
            hipLaunchKernelGGL((kb_rows64<SH.nw, SH.spw, SH.xsm, false>), dim3(static_cast<unsigned>(b.nunstaged)),
                               block, 0, s, b.nstrips, per, b.soffw, b.wpos, b.recs, b.wgx, b.ulist, fr, zd, t, X, ldx,
                               d.ncols, Y, ldy, ncol, alpha, beta, bz);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    return hipSuccess;
}

}  // namespace hm
