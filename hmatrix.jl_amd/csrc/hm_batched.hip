// Batched multi-right-hand-side product Y = beta*Y + alpha*H*X for panels of
// up to NB = 32 right-hand sides, FP32 accumulation, on MFMA.
//
// Three kernels per panel, each streaming its factors exactly once:
//   kb_proj   t_b = B_b' X[cols(b)] for every low-rank leaf, as partial sums
//             per 256-column super-tile.  One workgroup per super-tile
//             stages X[super-tile] in LDS as split BF16 (hi + lo); the rank
//             columns crossing the super-tile (several leaves, ranks 1-2
//             each here) form groups of 16, the N side of
//             v_mfma_f32_16x16x32_bf16 with X' as the A operand (M =
//             right-hand sides).  Each lane loads its column's B values for
//             all K-steps in one round trip, overlapped with the X stage;
//             X is read from HBM once per product, not once per leaf.
//   kb_reduce sums the per-super-tile partials of leaves wider than one
//             super-tile, in a fixed order (deterministic, no atomics),
//             largest leaves first.
//   kb_rows   row owner: a workgroup owns a band of strips of 16 rows (8
//             or 16 wavefronts, two strips each; the plan's row shape) and
//             stages the band's X window in LDS once, split into BF16 hi/lo
//             in the B-operand order.  A strip is a list of K-steps of 32
//             columns: each dense leaf crossing it (A = its rows, B = X
//             rows of its columns) and a panel holding the A columns of
//             every low-rank leaf crossing it (B = the matching t rows).
//             A operands are pre-packed in MFMA fragment order (zero rows
//             and octets not stored); each wavefront walks the records of
//             its strips as one software-pipelined stream (the A fragments
//             of K-step k + 2 in flight during the MFMAs of K-step k, the
//             next record's panel and first K-steps issued before the
//             current record's last MFMAs).  Each element of Y is written
//             once.  Workgroups are remapped so consecutive bands run on
//             the same XCD and share X rows in its L2.
// Storage modes: HM_BF16X2 (dense leaves as hi+lo BF16 planes, 3 MFMAs per
// product term: hi*hi + hi*lo + lo*hi; low-rank factors FP32), HM_F32 (dense
// leaves FP32, split in registers, same 3 MFMAs), HM_BF16 (dense leaves and
// factors BF16; X still split, 2 MFMAs).
#include <algorithm>
#include <type_traits>

#include "hm_bred.h"
#include "hm_internal.h"

namespace hm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short short8 __attribute__((ext_vector_type(8)));

enum : int { MODE_F32 = 1, MODE_BF16 = 2, MODE_BF16X2 = 3 };

__device__ __forceinline__ uint32_t f2bf_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf_bits2f(uint32_t h) { return __uint_as_float(h << 16); }

// Split 8 floats into BF16 hi and lo = bf16(v - hi).
__device__ __forceinline__ void split8(const float (&v)[8], bf16x8 &hi, bf16x8 &lo) {
    short8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint32_t hb = f2bf_bits(v[e]);
        h[e] = static_cast<short>(hb);
        l[e] = static_cast<short>(f2bf_bits(v[e] - bf_bits2f(hb)));
    }
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
}

typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

// Split two floats into BF16 hi (round to nearest even, v_cvt_pk_bf16_f32)
// and lo = bf16(v - hi); returns the packed pairs.
__device__ __forceinline__ void split2(float a, float b, uint32_t &hi, uint32_t &lo) {
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(float2v{a, b}, bf16x2v));
    const float ha = __uint_as_float(hi << 16);
    const float hb = __uint_as_float(hi & 0xffff0000u);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(float2v{a - ha, b - hb}, bf16x2v));
}

// Split eight floats; hi/lo as 8-element BF16 vectors.
__device__ __forceinline__ void split8hw(const float (&v)[8], bf16x8 &hi, bf16x8 &lo) {
    uint4 h, l;
    split2(v[0], v[1], h.x, l.x);
    split2(v[2], v[3], h.y, l.y);
    split2(v[4], v[5], h.z, l.z);
    split2(v[6], v[7], h.w, l.w);
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
}

// 16-byte vector with 4-byte alignment: global_load/store_dwordx4 at any
// float boundary (the device runs in unaligned-access mode).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// projection
constexpr int XOCT = PROJ_SUPER / 8;     // row octets of a super-tile
constexpr int XO_LD = XOCT + 1;          // [rhs][octet] stride: conflict-free reads and writes
constexpr int PKS = PROJ_SUPER / 32;     // K-steps of a super-tile

// B operand of the K-steps of one 16-column group, this lane's column
// (descriptor cd) and row octet g: 8 values per K-step, zero outside the
// column's K-step range [s0, s1).  Raw values; split by split_v.
template <int MODE>
struct VRaw {
    static constexpr int W = MODE == MODE_F32 ? 2 : (MODE == MODE_BF16X2 ? 2 : 1);   // 16-byte words per K-step
    uint4 w[PKS][W];
};

// The slice arena starts with VZERO zero elements: a lane whose column does
// not reach K-step s loads those instead, so every load is the arena base
// plus a 32-bit element offset (one address register, no branch).
template <int MODE>
__device__ __forceinline__ void load_v(const void *vslc, const uint4 &cd, int g, VRaw<MODE> &v) {
    const int s0 = cd.y & 255, s1 = (cd.y >> 8) & 255;
    const uint32_t base = cd.x * 8u + 8u * static_cast<uint32_t>(g) - 32u * static_cast<uint32_t>(s0);
    const uint32_t len = static_cast<uint32_t>(s1 - s0) * 32u;
#pragma unroll
    for (int s = 0; s < PKS; ++s) {
        const bool on = s >= s0 && s < s1;
        const uint32_t e = on ? base + 32u * s : 0u;
        if constexpr (MODE == MODE_F32) {
            const uint4 *p = reinterpret_cast<const uint4 *>(static_cast<const float *>(vslc) + e);
            v.w[s][0] = p[0];
            v.w[s][1] = p[1];
        } else {
            const uint16_t *vp = static_cast<const uint16_t *>(vslc);
            v.w[s][0] = *reinterpret_cast<const uint4 *>(vp + e);
            if constexpr (MODE == MODE_BF16X2)
                v.w[s][1] = *reinterpret_cast<const uint4 *>(vp + (on ? e + len : 0u));
        }
    }
}

// Projection: t_b = B_b' X[cols(b)] for every low-rank leaf, as partial sums
// per super-tile of PROJ_SUPER columns of X.  One workgroup per super-tile:
// it stages X[super-tile] in LDS as split BF16 (hi + lo); the rank columns
// (leaf, column) crossing the super-tile form groups of 16, the N side of
// v_mfma_f32_16x16x32_bf16 (A operand X', M = right-hand sides), and
// wavefront w takes groups w, w + 4, ...  Each lane loads its column's B
// values for every K-step at once, in the same round trip as the X stage
// (descriptors and X first, then B).  Column descriptors (16 per group,
// gps groups per super-tile): x = B offset / 8, y = s0 | s1 << 8 |
// ngroups << 16, z = float offset of the result row t[j][0..NB) (-1: no
// column).
template <int MODE>
__global__ __launch_bounds__(256) void kb_proj(int64_t gps, const uint4 *__restrict__ pcol,
                                               const void *__restrict__ vslc,
                                               const float *__restrict__ zeros,
                                               const float *__restrict__ X, int64_t ldx,
                                               int64_t ncols, int ncol, float *__restrict__ out) {
    __shared__ uint4 xs[2][NB][XO_LD];              // [hi/lo][rhs][row octet], 8 BF16 each
    const int64_t col0 = static_cast<int64_t>(blockIdx.x) * PROJ_SUPER;
    const int rows = static_cast<int>(min<int64_t>(PROJ_SUPER, ncols - col0));
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const uint4 *gdesc = pcol + static_cast<int64_t>(blockIdx.x) * gps * 16;
    // every wavefront loads a descriptor (the last group's when it has none)
    // and empties its K-step range when the index is past the groups; a
    // plan without low-rank columns (gps == 0) has no descriptors
    uint4 cd = gps > 0 ? gdesc[min(wave, static_cast<int>(gps) - 1) * 16 + r16] : uint4{0u, 0u, 0xffffffffu, 0u};
    if (wave >= gps) cd = uint4{0u, cd.y & 0xffff0000u, 0xffffffffu, 0u};
    // Stage X[col0:col0+rows] in (column, 4-row quad) units: thread t reads
    // quad t % 64 of columns t / 64 + 4u, a wavefront 1 KB of one column per
    // instruction; all loads first.
    constexpr int XQ = PROJ_SUPER / 4;
    static_assert(XQ == 64, "one quad per lane");
    constexpr int XIT = NB * XQ / 256;
    const int q = threadIdx.x & 63;
    const int rhs0 = threadIdx.x >> 6;
    const float *xb = X + col0 + 4 * q + static_cast<int64_t>(rhs0) * ldx;
    const int64_t xstride = 4 * ldx;
    const bool qfull = 4 * q + 4 <= rows;
    f4u xv[XIT];
#pragma unroll
    for (int u = 0; u < XIT; ++u) {
        const float *src = xb + u * xstride;
        const bool full = qfull && rhs0 + 4 * u < ncol;
        xv[u] = *reinterpret_cast<const f4u *>(full ? src : zeros);
    }
    if (rows < PROJ_SUPER) {      // last super-tile: the quad cut by the matrix end
#pragma unroll
        for (int u = 0; u < XIT; ++u) {
            if (rhs0 + 4 * u < ncol && 4 * q < rows && !qfull) {
                const float *src = xb + u * xstride;
                f4u v = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int e = 0; e < 3; ++e)
                    if (4 * q + e < rows) v[e] = src[e];
                xv[u] = v;
            }
        }
    }
    // B values of the wavefront's first group (addresses from its descriptors)
    VRaw<MODE> vr;
    load_v<MODE>(vslc, cd, g, vr);
#pragma unroll
    for (int u = 0; u < XIT; ++u) {
        const int rhs = rhs0 + 4 * u;
        uint2 hi, lo;
        split2(xv[u][0], xv[u][1], hi.x, lo.x);
        split2(xv[u][2], xv[u][3], hi.y, lo.y);
        *(reinterpret_cast<uint2 *>(&xs[0][rhs][q >> 1]) + (q & 1)) = hi;
        *(reinterpret_cast<uint2 *>(&xs[1][rhs][q >> 1]) + (q & 1)) = lo;
    }
    __syncthreads();
    const int ngroups = static_cast<int>(__builtin_amdgcn_readfirstlane(cd.y >> 16));
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
    // one group: MFMAs over the K-steps, then the 32 x 16 result rows
    auto group = [&]() {
        floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < PKS; ++s) {
            bf16x8 vh, vl;
            if constexpr (MODE == MODE_F32) {
                const float v[8] = {__uint_as_float(vr.w[s][0].x), __uint_as_float(vr.w[s][0].y),
                                    __uint_as_float(vr.w[s][0].z), __uint_as_float(vr.w[s][0].w),
                                    __uint_as_float(vr.w[s][1].x), __uint_as_float(vr.w[s][1].y),
                                    __uint_as_float(vr.w[s][1].z), __uint_as_float(vr.w[s][1].w)};
                split8hw(v, vh, vl);
            } else {
                vh = __builtin_bit_cast(bf16x8, vr.w[s][0]);
                vl = MODE == MODE_BF16X2 ? __builtin_bit_cast(bf16x8, vr.w[s][VRaw<MODE>::W - 1]) : zero8;
            }
            const int oct = 4 * s + g;
            const bf16x8 xh0 = __builtin_bit_cast(bf16x8, xs[0][r16][oct]);
            const bf16x8 xl0 = __builtin_bit_cast(bf16x8, xs[1][r16][oct]);
            const bf16x8 xh1 = __builtin_bit_cast(bf16x8, xs[0][16 + r16][oct]);
            const bf16x8 xl1 = __builtin_bit_cast(bf16x8, xs[1][16 + r16][oct]);
            acc0 = mfma(xh0, vh, acc0);
            acc0 = mfma(xl0, vh, acc0);
            acc1 = mfma(xh1, vh, acc1);
            acc1 = mfma(xl1, vh, acc1);
            if (MODE != MODE_BF16) {
                acc0 = mfma(xh0, vl, acc0);
                acc1 = mfma(xh1, vl, acc1);
            }
        }
        // D[i = rhs][j = column]: lane holds rhs 4g..4g+3 (+16 for acc1) of column r16
        const int o = static_cast<int>(cd.z);
        if (o >= 0) {
            float *dst = out + o + 4 * g;
            reinterpret_cast<float4 *>(dst)[0] = make_float4(acc0[0], acc0[1], acc0[2], acc0[3]);
            reinterpret_cast<float4 *>(dst + 16)[0] = make_float4(acc1[0], acc1[1], acc1[2], acc1[3]);
        }
    };
    // The first group runs unconditionally (a wavefront without one has an
    // empty descriptor: zero operands, no store), so its B loads stay ahead
    // of the barrier; later groups load descriptors, then B values.
    group();
    for (int gi = wave + 4; gi < ngroups; gi += 4) {
        cd = gdesc[gi * 16 + r16];
        load_v<MODE>(vslc, cd, g, vr);
        group();
    }
}

// ---------------------------------------------------------------------------
// lane's 8 A values of one K-step as BF16 hi / lo vectors
template <int MODE>
__device__ __forceinline__ void load_a(const void *frag, const float *zeros, bool on, uint32_t afrag,
                                       int lane_unit, int lo_units, bf16x8 &ah, bf16x8 &al) {
    if (MODE == MODE_F32) {
        const float *fp = on ? static_cast<const float *>(frag) + 4 * (static_cast<int64_t>(afrag) + lane_unit)
                             : zeros;
        const float4 a = reinterpret_cast<const float4 *>(fp)[0];
        const float4 b = reinterpret_cast<const float4 *>(fp)[1];
        float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        split8hw(v, ah, al);
    } else {
        const uint16_t *fb = static_cast<const uint16_t *>(frag) + 8 * (static_cast<int64_t>(afrag) + lane_unit);
        const uint16_t *zb = reinterpret_cast<const uint16_t *>(zeros);
        ah = __builtin_bit_cast(bf16x8, ld_stream(on ? fb : zb));
        if (MODE == MODE_BF16X2) al = __builtin_bit_cast(bf16x8, ld_stream(on ? fb + 8 * lo_units : zb));
        else al = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
    }
}

// ---------------------------------------------------------------------------
// Row pass.  Workgroup b (remapped so that consecutive bands run on one XCD
// and share X rows in its L2) owns a band of NW * SPW strips; wavefront w
// takes its run of the band's strips (runs balanced by K-steps at plan
// creation: the workgroup lasts as long as its slowest wavefront).  The
// band's X window
// [cmin, cmin + W) (every dense piece of the band, W <= XSM rows) is staged
// once in LDS, split into BF16 hi/lo planes in the B-operand order; a band
// whose pieces span more rows (W = 0) reads X straight from memory.  Each
// wavefront then walks the records of its strips as one stream.  A record
// has nd >= 0 dense K-steps and at most one panel K-step; the wavefront
// treats it as max(nd, 1) dense K-steps (padded to an even count) and one
// panel K-step, the missing ones read from the zero buffer, so that every
// record issues the same load sequence and the compiler's wait counts stay
// exact:
//   * dense K-steps run in pairs through two A-operand sets: the A
//     fragments of K-step k + 1 are loaded before the MFMAs of K-step k,
//     whose B operand is read from the LDS window;
//   * at the last dense K-step, the registers of the record after next, the
//     next record's panel K-step (A fragments, 16 gathered t values, row
//     indices taken from the record registers by ds_bpermute) and its first
//     dense A fragments are issued first, then the current record's last
//     dense and panel MFMAs run and its Y rows are written (unless the strip
//     continues in a chained record).

__device__ __forceinline__ uint32_t rword(uint32_t h0, uint32_t h1, int i) {
    const uint32_t a = __builtin_amdgcn_readlane(h0, i & 63);
    const uint32_t b = __builtin_amdgcn_readlane(h1, i & 63);
    return i < 64 ? a : b;
}

// A operand of one K-step, and a panel K-step's operands (A, then 16 t
// values: rows 8g..8g+7 of columns r16 and 16 + r16)
struct AOp {
    bf16x8 ah, al;
};
struct POp {
    bf16x8 ah, al;
    float t[16];
};

template <int MODE, int NW, int SPW, int XSM, bool STAGED>
__global__ __launch_bounds__(64 * NW, 4) void kb_rows(int64_t nstrips, int64_t wg_per_xcd,
                                                   const uint32_t *__restrict__ soffw,
                                                   const uint32_t *__restrict__ wpos,
                                                   const uint32_t *__restrict__ recs,
                                                   const int32_t *__restrict__ wgx,
                                                   const uint32_t *__restrict__ ulist,
                                                   const void *__restrict__ frag,
                                                   const float *__restrict__ zeros,
                                                   const float *__restrict__ T,
                                                   const float *__restrict__ X, int64_t ldx, int64_t ncols,
                                                   float *__restrict__ Y, int64_t ldy, int ncol, float alpha,
                                                   float beta, int beta_zero) {
    // window: [hi, lo plane][column of X][octet] vectors of 8 BF16 (rows
    // 8 octet .. of the window); odd octet stride: the B reads of 16
    // consecutive lanes (16 columns) and the staging writes (consecutive
    // rows of one column) fall on distinct banks
    constexpr int NOCT = XSM / 8;
    constexpr int OLD = NOCT + 1;
    static_assert(OLD % 2 == 1, "odd octet stride");
    __shared__ uint4 xs[STAGED ? 2 : 1][STAGED ? NB : 1][STAGED ? OLD : 1];
    // staged instance: band b%8 * per + b/8; the other instance: band list[b]
    const int64_t wg = STAGED ? static_cast<int64_t>(blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8
                              : static_cast<int64_t>(ulist[blockIdx.x]);
    constexpr int BAND = NW * SPW;
    if (wg * BAND >= nstrips) return;          // whole workgroup
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const int cmin = wgx[2 * wg], W = wgx[2 * wg + 1];
    if ((W > 0) != STAGED) return;             // the band is the other instance's
    // this wavefront's run of the band's strips (slots [s0, s0 + nrun));
    // every lane loads one slot of the band, in parallel with the run bounds
    const uint32_t *wp = wpos + wg * (NW + 1);
    const int s0 = static_cast<int>(wp[wave]);
    const int nrun = static_cast<int>(wp[wave + 1]) - s0;
    const uint32_t sv = soffw[min(wg * BAND + l % BAND, nstrips - 1)];
    constexpr int AU = MODE == MODE_F32 ? 2 : 1;     // 16-byte units per lane and plane
    constexpr uint32_t NONE = 0xffffffffu;

    // Stage the window: thread t takes (column, 4-row quad) units along the
    // rows of one column, loads all of them, then splits them into LDS.
    constexpr int NQ = XSM / 4;
    constexpr int U = (NB * NQ + 64 * NW - 1) / (64 * NW);     // units per thread (the last partly)
    // rows of unit u of this thread that lie in the window and in X (0..4)
    auto unit_rows = [&](int u) __attribute__((always_inline)) {
        const int p = threadIdx.x + 64 * NW * u;
        const int c = p / NQ, q = p % NQ;
        const int64_t row = static_cast<int64_t>(cmin) + 4 * q;
        return (p < NB * NQ && c < ncol && 4 * q < W) ? static_cast<int>(min<int64_t>(4, ncols - row)) : 0;
    };
    f4u xv[U];
    if constexpr (STAGED) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 64 * NW * u;
            const int c = p / NQ, q = p % NQ;
            const float *src = unit_rows(u) == 4 ? X + static_cast<int64_t>(c) * ldx + cmin + 4 * q : zeros;
            xv[u] = *reinterpret_cast<const f4u *>(src);
        }
    }

    auto issue_a = [&](uint32_t h0, uint32_t h1, int k, bool on, AOp &o) __attribute__((always_inline)) {
        const int i = SREC_HDR + 4 * k;
        const uint32_t afrag = rword(h0, h1, i), info = rword(h0, h1, i + 2);
        const int roff = info & 15, rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
        load_a<MODE>(frag, zeros, on && g < nOct && r16 >= roff && r16 < roff + rows, afrag,
                     (g * rows + r16 - roff) * AU, nOct * rows, o.ah, o.al);
    };
    auto issue_panel = [&](uint32_t h0, uint32_t h1, int k, bool on, POp &o) __attribute__((always_inline)) {
        const int i = SREC_HDR + 4 * k;
        const uint32_t afrag = rword(h0, h1, i), src = rword(h0, h1, i + 1), info = rword(h0, h1, i + 2);
        const int rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
        load_a<MODE>(frag, zeros, on && g < nOct && r16 < rows, afrag, (g * rows + r16) * AU, nOct * rows, o.ah,
                     o.al);
        uint32_t ix[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int wi = (static_cast<int>(src) + 8 * g + e) & 127;   // record word holding the t-row index
            const int a = (wi & 63) << 2;
            const uint32_t i0 = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(a, static_cast<int>(h0)));
            const uint32_t i1 = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(a, static_cast<int>(h1)));
            ix[e] = wi < 64 ? i0 : i1;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float *tr = on ? T + static_cast<int64_t>(ix[e]) * NB + r16 : zeros;
            o.t[e] = tr[0];
            o.t[8 + e] = tr[16];
        }
    };
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    auto mfmas = [&](const bf16x8 &ah, const bf16x8 &al, const bf16x8 &xh0, const bf16x8 &xl0,
                     const bf16x8 &xh1, const bf16x8 &xl1) __attribute__((always_inline)) {
        acc0 = mfma(ah, xh0, acc0);
        acc0 = mfma(ah, xl0, acc0);
        acc1 = mfma(ah, xh1, acc1);
        acc1 = mfma(ah, xl1, acc1);
        if (MODE != MODE_BF16) {
            acc0 = mfma(al, xh0, acc0);
            acc1 = mfma(al, xh1, acc1);
        }
    };
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
    // dense K-step k of record (h0, h1) with its A operand; B from the LDS
    // window, or from X when the band is not staged
    auto compute_dense = [&](uint32_t h0, uint32_t h1, int k, bool on, const AOp &a,
                             auto staged) __attribute__((always_inline)) {
        constexpr bool FROM_LDS = decltype(staged)::value;
        const int i = SREC_HDR + 4 * k;
        const uint32_t src = rword(h0, h1, i + 1), info = rword(h0, h1, i + 2);
        if (!on) return;                          // a padding step: nothing to add
        const int nOct = (info >> 9) & 7;
        const bool bon = g < nOct;
        bf16x8 xh0 = zero8, xl0 = zero8, xh1 = zero8, xl1 = zero8;
        if constexpr (FROM_LDS) {
            // lanes past the piece's octets hold zero A values: they read
            // the piece's last octet instead of branching (same window row)
            const int o = ((static_cast<int>(src) - cmin) >> 3) + min(g, nOct - 1);
            xh0 = __builtin_bit_cast(bf16x8, xs[0][r16][o]);
            xl0 = __builtin_bit_cast(bf16x8, xs[1][r16][o]);
            xh1 = __builtin_bit_cast(bf16x8, xs[0][16 + r16][o]);
            xl1 = __builtin_bit_cast(bf16x8, xs[1][16 + r16][o]);
        } else {
            float v0[8], v1[8];
            const int64_t row = static_cast<int64_t>(src) + 8 * g;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float *v = h == 0 ? v0 : v1;
                const int c = 16 * h + r16;
                const float *xp = X + static_cast<int64_t>(c) * ldx + row;
                if (bon && c < ncol && row + 8 <= ncols) {
                    const f4u p0 = reinterpret_cast<const f4u *>(xp)[0];
                    const f4u p1 = reinterpret_cast<const f4u *>(xp)[1];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        v[e] = p0[e];
                        v[4 + e] = p1[e];
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = (bon && c < ncol && row + e < ncols) ? xp[e] : 0.0f;
                }
            }
            split8hw(v0, xh0, xl0);
            split8hw(v1, xh1, xl1);
        }
        mfmas(a.ah, a.al, xh0, xl0, xh1, xl1);
    };
    auto compute_panel = [&](const POp &o) __attribute__((always_inline)) {
        const float v0[8] = {o.t[0], o.t[1], o.t[2], o.t[3], o.t[4], o.t[5], o.t[6], o.t[7]};
        const float v1[8] = {o.t[8], o.t[9], o.t[10], o.t[11], o.t[12], o.t[13], o.t[14], o.t[15]};
        bf16x8 xh0, xl0, xh1, xl1;
        split8hw(v0, xh0, xl0);
        split8hw(v1, xh1, xl1);
        mfmas(o.ah, o.al, xh0, xl0, xh1, xl1);
    };
    auto store = [&](int row0, int R) __attribute__((always_inline)) {
        // D layout: lane (g, r16) holds strip rows 4g .. 4g+3 of column 16h + r16
        const int r = 4 * g;
        if (r < R) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = 16 * h + r16;
                if (c >= ncol) continue;
                const floatx4 a = h == 0 ? acc0 : acc1;
                float *dst = Y + static_cast<int64_t>(c) * ldy + row0 + r;
                f4u v = {alpha * a[0], alpha * a[1], alpha * a[2], alpha * a[3]};
                if (r + 4 <= R) {
                    if (!beta_zero) v += beta * *reinterpret_cast<const f4u *>(dst);
                    *reinterpret_cast<f4u *>(dst) = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (r + e < R) dst[e] = beta_zero ? v[e] : beta * dst[e] + v[e];
                }
            }
        }
        acc0 = floatx4{0.f, 0.f, 0.f, 0.f};
        acc1 = floatx4{0.f, 0.f, 0.f, 0.f};
    };
    // record after (h0, h1) of strip position at: the chain's continuation,
    // else the next strip's first record (NONE past the wavefront's strips)
    auto next_of = [&](uint32_t h0, uint32_t h1, int at, int &nat) __attribute__((always_inline)) -> uint32_t {
        const uint32_t ch = rword(h0, h1, 5);
        if (ch != 0u) {
            nat = at;
            return ch;
        }
        nat = at + 1;
        return at + 1 < nrun ? __builtin_amdgcn_readlane(sv, (s0 + at + 1) & 63) : NONE;
    };
    auto load_rec = [&](uint32_t off, uint32_t &h0, uint32_t &h1) __attribute__((always_inline)) {
        const uint32_t *rw = recs + (off == NONE ? 0u : off);     // record 0 stands in for none
        h0 = rw[l];
        h1 = rw[64 + l];
    };

    // the wavefront's first record, the one after it, and the first
    // record's first dense A fragments, in flight with the window (its panel
    // K-step is issued after the barrier, ahead of the first dense MFMAs)
    uint32_t c0, c1, n0, n1;
    load_rec(nrun > 0 ? __builtin_amdgcn_readlane(sv, s0 & 63) : NONE, c0, c1);
    int nsi = 0;
    uint32_t noff = nrun > 0 ? next_of(c0, c1, 0, nsi) : NONE;
    load_rec(noff, n0, n1);
    AOp A, B, C;
    POp PN;
    int nd = nrun > 0 ? static_cast<int>(rword(c0, c1, 2)) : 0;
    issue_a(c0, c1, 0, nd > 0, A);
    issue_a(c0, c1, 1, nd > 1, B);
    if constexpr (STAGED) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 64 * NW * u;
            const int c = p / NQ, q = p % NQ;
            const int part = unit_rows(u);
            if (part > 0 && part < 4) {              // the quad cut by the last row of X
                const float *src = X + static_cast<int64_t>(c) * ldx + cmin + 4 * q;
                f4u w = {0.f, 0.f, 0.f, 0.f};
                for (int e = 0; e < part; ++e) w[e] = src[e];
                xv[u] = w;
            }
            uint32_t hi0, lo0, hi1, lo1;
            split2(xv[u][0], xv[u][1], hi0, lo0);
            split2(xv[u][2], xv[u][3], hi1, lo1);
            if (p < NB * NQ) {
                *(reinterpret_cast<uint2 *>(&xs[0][c][q >> 1]) + (q & 1)) = uint2{hi0, hi1};
                *(reinterpret_cast<uint2 *>(&xs[1][c][q >> 1]) + (q & 1)) = uint2{lo0, lo1};
            }
        }
    }
    __syncthreads();
    if (nrun == 0) return;
    issue_panel(c0, c1, nd, rword(c0, c1, 3) > 0u, PN);

    floatx4 p0 = {0.f, 0.f, 0.f, 0.f}, p1 = {0.f, 0.f, 0.f, 0.f};
    int prow = -1, pR = 0;
    auto walk = [&](auto staged) __attribute__((always_inline)) {
        for (;;) {
            // dense K-steps in groups of 3 through the sets A, B, C: the A
            // fragments of K-step k + 2 are loaded before the MFMAs of k
            const int trips = (max(nd, 1) + 2) / 3;
            for (int q = 0; q + 1 < trips; ++q) {
                const int k = 3 * q;
                issue_a(c0, c1, k + 2, k + 2 < nd, C);
                __builtin_amdgcn_sched_barrier(0);      // C's loads stay ahead of A's MFMAs
                compute_dense(c0, c1, k, true, A, staged);
                issue_a(c0, c1, k + 3, k + 3 < nd, A);
                __builtin_amdgcn_sched_barrier(0);
                compute_dense(c0, c1, k + 1, k + 1 < nd, B, staged);
                issue_a(c0, c1, k + 4, k + 4 < nd, B);
                __builtin_amdgcn_sched_barrier(0);
                compute_dense(c0, c1, k + 2, k + 2 < nd, C, staged);
            }
            // last group: the record after next is loaded first; the next
            // record's first two dense K-steps go into A and B as soon as
            // those are free, its panel into PN after this record's panel
            const int kl = 3 * (trips - 1);
            const bool more = noff != NONE;
            int nnsi = nsi;
            const uint32_t nnoff = more ? next_of(n0, n1, nsi, nnsi) : NONE;
            uint32_t m0, m1;
            load_rec(nnoff, m0, m1);
            const int nnd = more ? static_cast<int>(rword(n0, n1, 2)) : 0;
            const int nnp = more ? static_cast<int>(rword(n0, n1, 3)) : 0;
            issue_a(c0, c1, kl + 2, kl + 2 < nd, C);
            __builtin_amdgcn_sched_barrier(0);
            compute_dense(c0, c1, kl, kl < nd, A, staged);
            issue_a(n0, n1, 0, more && nnd > 0, A);
            __builtin_amdgcn_sched_barrier(0);
            compute_dense(c0, c1, kl + 1, kl + 1 < nd, B, staged);
            issue_a(n0, n1, 1, more && nnd > 1, B);
            __builtin_amdgcn_sched_barrier(0);
            compute_panel(PN);
            issue_panel(n0, n1, nnd, more && nnp > 0, PN);
            __builtin_amdgcn_sched_barrier(0);
            compute_dense(c0, c1, kl + 2, kl + 2 < nd, C, staged);
            if (rword(c0, c1, 5) == 0u) {
                // the wavefront's first finished strip is written at its end,
                // after its last loads; later strips as they finish
                if (prow < 0) {
                    p0 = acc0;
                    p1 = acc1;
                    prow = static_cast<int>(rword(c0, c1, 0));
                    pR = static_cast<int>(rword(c0, c1, 1));
                    acc0 = floatx4{0.f, 0.f, 0.f, 0.f};
                    acc1 = floatx4{0.f, 0.f, 0.f, 0.f};
                } else {
                    store(static_cast<int>(rword(c0, c1, 0)), static_cast<int>(rword(c0, c1, 1)));
                }
            }
            if (!more) break;
            c0 = n0;
            c1 = n1;
            n0 = m0;
            n1 = m1;
            nsi = nnsi;
            noff = nnoff;
            nd = nnd;
        }
    };
    walk(std::integral_constant<bool, STAGED>{});
    if (prow >= 0) {
        acc0 = p0;
        acc1 = p1;
        store(prow, pR);
    }
}

template <int MODE>
static hipError_t run_b(const Plan &p, const Direction &d, int ncol, const float *X, int64_t ldx, float *Y,
                        int64_t ldy, float alpha, float beta, hipStream_t s, hipEvent_t *ev) {
    const Batched &b = d.bat;
    const float *zf = static_cast<const float *>(p.zeros);
    hipError_t e = hipSuccess;
    if (b.nst > 0 && b.gps > 0) {
        hipLaunchKernelGGL(kb_proj<MODE>, dim3(static_cast<unsigned>(b.nst)), dim3(256), 0, s, b.gps, b.pcol,
                           b.vslc, zf, X, ldx, d.ncols, ncol, static_cast<float *>(b.t));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[0], s);
    if (b.nst > 0 && b.gps > 0 && b.nred > 0) {
        const int64_t nwg = b.nredbig + (b.nred - b.nredbig + 3) / 4;
        hipLaunchKernelGGL(kb_reduce<float>, dim3(static_cast<unsigned>(nwg)), dim3(256), 0, s, b.red, b.nredbig, b.nred,
                           static_cast<const float *>(b.t) + b.tLen, zf, static_cast<float *>(b.t));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (b.nstrips > 0) {
        const int64_t band = static_cast<int64_t>(b.shape.nw) * b.shape.spw;
        const int64_t nwg = (b.nstrips + band - 1) / band;
        const int64_t per = (nwg + 7) / 8;        // workgroups per XCD
        const dim3 grid(static_cast<unsigned>(8 * per)), block(static_cast<unsigned>(64 * b.shape.nw));
        const float *tf = static_cast<const float *>(b.t);
        const int bz = beta == 0.0f ? 1 : 0;
        // bands with a staged window, then (if the plan has any) the others
        for (int staged = 1; staged >= 0; --staged) {
            if (!staged && b.nunstaged == 0) break;
            const bool wide = b.shape.nw == ROW_WIDE.nw && b.shape.spw == ROW_WIDE.spw && b.shape.xsm == ROW_WIDE.xsm;
            const bool narrow =
                b.shape.nw == ROW_NARROW.nw && b.shape.spw == ROW_NARROW.spw && b.shape.xsm == ROW_NARROW.xsm;
            if (!wide && !narrow) return hipErrorInvalidValue;
            const dim3 g2 = staged ? grid : dim3(static_cast<unsigned>(b.nunstaged));
#define HM_ROWS(SH, ST)                                                                                          \
    hipLaunchKernelGGL((kb_rows<MODE, SH.nw, SH.spw, SH.xsm, ST>), g2, block, 0, s, b.nstrips, per, b.soffw, b.wpos, b.recs, \
                       b.wgx, b.ulist, b.frag, zf, tf, X, ldx, d.ncols, Y, ldy, ncol, alpha, beta, bz)
            if (wide && staged) HM_ROWS(ROW_WIDE, true);
            else if (wide) HM_ROWS(ROW_WIDE, false);
            else if (staged) HM_ROWS(ROW_NARROW, true);
            else HM_ROWS(ROW_NARROW, false);
#undef HM_ROWS
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    return hipSuccess;
}

hipError_t launch_batched(const Plan &p, const Direction &d, int ncol, const float *X, int64_t ldx,
                          float *Y, int64_t ldy, float alpha, float beta, hipStream_t s,
                          hipEvent_t *ev) {
    switch (p.storage) {
        case MODE_F32: return run_b<MODE_F32>(p, d, ncol, X, ldx, Y, ldy, alpha, beta, s, ev);
        case MODE_BF16: return run_b<MODE_BF16>(p, d, ncol, X, ldx, Y, ldy, alpha, beta, s, ev);
        case MODE_BF16X2: return run_b<MODE_BF16X2>(p, d, ncol, X, ldx, Y, ldy, alpha, beta, s, ev);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace hm
