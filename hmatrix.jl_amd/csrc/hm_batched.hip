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
//   kb_strips row owner: one wavefront per 16-row strip, 8 or 16 per
//             workgroup (the plan's row shape).  A strip is a list of
//             K-steps of 32 columns: each
//             dense leaf crossing it (A = its rows, B = X rows of its
//             columns) and a panel holding the A columns of every low-rank
//             leaf crossing it (B = the matching t rows).  A operands are
//             pre-packed in MFMA fragment order (zero rows and octets not
//             stored) and loaded straight into registers, K-steps in
//             batches; X rows come from a workgroup window staged in LDS.
//             Each element of Y is written once.  Workgroups are remapped
//             so consecutive strips run on the same XCD and share X rows in
//             its L2.
// Storage modes: HM_BF16X2 (dense leaves as hi+lo BF16 planes, 3 MFMAs per
// product term: hi*hi + hi*lo + lo*hi; low-rank factors FP32), HM_F32 (dense
// leaves FP32, split in registers, same 3 MFMAs), HM_BF16 (dense leaves and
// factors BF16; X still split, 2 MFMAs).
#include <algorithm>
#include <type_traits>

#include "hm_internal.h"

namespace hm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short short8 __attribute__((ext_vector_type(8)));

enum : int { MODE_F32 = 1, MODE_BF16 = 2, MODE_BF16X2 = 3 };

// K-steps a row-pass wavefront loads before their MFMAs.
#ifndef HM_SKC
#define HM_SKC 2
#endif
constexpr int SKC = HM_SKC;

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

__device__ __forceinline__ bf16x8 load_bf8(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

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
    // and empties its K-step range when the index is past the groups
    uint4 cd = gdesc[min(wave, static_cast<int>(gps) - 1) * 16 + r16];
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
// t = sum of the per-super-tile partials of one leaf, in ascending chunk
// order (deterministic, no atomics).  Items come largest first: the first
// nbig (more than RED_WAVE chunks) take one workgroup each -- threads split
// into G = 256/len groups, group q sums chunks q, q+G, ... (RED_BATCH loads
// in flight per thread) and the groups are combined in LDS in ascending
// order; the rest take one wavefront each, lanes over the elements, all of
// an element's chunks loaded at once and summed in order.
constexpr int RED_BATCH = 32;
constexpr int RED_WAVE = 32;
__global__ __launch_bounds__(256) void kb_reduce(const BRed *__restrict__ red, int64_t nbig, int64_t nred,
                                                 const float *__restrict__ part,
                                                 const float *__restrict__ zeros,
                                                 float *__restrict__ t) {
    __shared__ float sh[256];
    if (static_cast<int64_t>(blockIdx.x) < nbig) {
        const BRed r = red[blockIdx.x];
        const int per = r.len >= 256 ? 256 : r.len;    // elements per pass
        const int G = 256 / per;
        for (int e0 = 0; e0 < r.len; e0 += per) {
            const int e = e0 + static_cast<int>(threadIdx.x) % per;
            const int q = static_cast<int>(threadIdx.x) / per;
            float acc = 0.0f;
            if (e < r.len && q < G) {
                const float *src = part + r.pOff + e;
                for (int c0 = q; c0 < r.nchunk; c0 += RED_BATCH * G) {
                    float v[RED_BATCH];
#pragma unroll
                    for (int u = 0; u < RED_BATCH; ++u) {
                        const int c = c0 + u * G;
                        v[u] = *(c < r.nchunk ? src + static_cast<int64_t>(c) * r.len : zeros);
                    }
#pragma unroll
                    for (int u = 0; u < RED_BATCH; ++u) acc += v[u];
                }
            }
            sh[threadIdx.x] = acc;
            __syncthreads();
            if (q == 0 && e < r.len) {
                float sum = acc;
                for (int gq = 1; gq < G; ++gq) sum += sh[gq * per + threadIdx.x];
                t[r.tOff + e] = sum;
            }
            __syncthreads();
        }
        return;
    }
    const int64_t item = nbig + (static_cast<int64_t>(blockIdx.x) - nbig) * 4 + (threadIdx.x >> 6);
    if (item >= nred) return;
    const BRed r = red[item];
    const int l = threadIdx.x & 63;
    for (int e = l; e < r.len; e += 64) {
        const float *src = part + r.pOff + e;
        float v[RED_WAVE];
#pragma unroll
        for (int c = 0; c < RED_WAVE; ++c) v[c] = *(c < r.nchunk ? src + static_cast<int64_t>(c) * r.len : zeros);
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < RED_WAVE; ++c) acc += v[c];
        t[r.tOff + e] = acc;
    }
}

// ---------------------------------------------------------------------------
// row pass: one wavefront per strip of SROWS rows x NB right-hand sides,
// SW strips per workgroup.  Per K-step a lane loads its 8 A values of
// strip row r16 (octet g) and, for each right-hand-side block h, the 8 B
// values of k-rows [8g, 8g + 8) in column 16h + r16: X rows from the
// workgroup's X window in LDS (or straight from X when the window is wider
// than XSM), panel rows gathered from t.  KC K-steps are loaded together
// before their MFMAs.  The strip record and the first batch's A fragments
// and t-row indices are loaded before the window's barrier, in flight
// together with the window's X loads.

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
        ah = load_bf8(on ? fb : zb);
        if (MODE == MODE_BF16X2) al = load_bf8(on ? fb + 8 * lo_units : zb);
        else al = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
    }
}

template <int MODE, int KC, int SW, int XSM>
__global__ __launch_bounds__(64 * SW) void kb_strips(int64_t nstrips, int64_t wg_per_xcd,
                                                 const uint32_t *__restrict__ soff,
                                                 const uint32_t *__restrict__ recs,
                                                 const int32_t *__restrict__ wgx,
                                                 const void *__restrict__ frag,
                                                 const float *__restrict__ zeros,
                                                 const float *__restrict__ T,
                                                 const float *__restrict__ X, int64_t ldx, int64_t ncols,
                                                 float *__restrict__ Y, int64_t ldy,
                                                 int ncol, float alpha, float beta, int beta_zero) {
    // X window of the workgroup as FP32 [column of X][row], then 8 zeros
    // that lanes without data read
    constexpr int XS_LD = XSM + 4;
    static_assert((XS_LD / 4) % 2 == 1, "odd 16-byte stride");
    __shared__ float xs[NB * XS_LD + 8];
    // Hardware places workgroup b on XCD b % 8; give each XCD a contiguous
    // range of strips so neighbouring strips share X rows in that L2.
    const int64_t wg = static_cast<int64_t>(blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8;
    if (wg * SW >= nstrips) return;           // whole workgroup
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t s = wg * SW + wave;
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const int cmin = wgx[2 * wg], W = wgx[2 * wg + 1];
    const float *xzero = xs + NB * XS_LD;

    const bool has = s < nstrips;
    constexpr int NQ = XSM / 4;                    // quads per column slot
    constexpr int U = NB * NQ / (64 * SW);         // units per thread
    static_assert(U * 64 * SW == NB * NQ, "whole units per thread");

    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const bool rok0 = r16 < ncol, rok1 = 16 + r16 < ncol;
    const float *xc0 = X + static_cast<int64_t>(r16) * ldx;
    const float *xc1 = X + static_cast<int64_t>(16 + r16) * ldx;
    int row0 = 0, R = 0;
    // Strip record: its first 128 words in two vector registers (lane l holds
    // words l and 64 + l; the record arena is padded so these loads stay in
    // bounds), read with readlane -- one memory round trip for the header
    // and every K-step descriptor.  A wavefront without a strip reads strip
    // 0's record (valid addresses, results unused).
    uint32_t rec = soff[has ? s : 0];
    const uint32_t *rw = recs + rec;
    uint32_t hv0 = rw[l], hv1 = rw[64 + l];
    int nd = 0, np = 0;
    uint32_t pidx = 0;
    auto word = [&](int i) __attribute__((always_inline)) -> uint32_t {
        const uint32_t a = __builtin_amdgcn_readlane(hv0, i & 63);
        const uint32_t b = __builtin_amdgcn_readlane(hv1, i & 63);
        return i < 64 ? a : b;
    };
    auto decode = [&]() __attribute__((always_inline)) {
        row0 = static_cast<int>(word(0));
        R = static_cast<int>(word(1));
        nd = static_cast<int>(word(2));
        np = static_cast<int>(word(3));
        pidx = word(4);
    };
    auto desc = [&](int k) __attribute__((always_inline)) {
        const int i = SREC_HDR + 4 * k;
        return uint4{word(i), word(i + 1), word(i + 2), 0u};
    };
    // MFMAs of one K-step: 8 B values per lane and block (hi/lo split here)
    auto kstep = [&](const bf16x8 &ah, const bf16x8 &al, const f4u (&b)[4]) __attribute__((always_inline)) {
        bf16x8 xh0, xl0, xh1, xl1;
        {
            const float v[8] = {b[0][0], b[0][1], b[0][2], b[0][3], b[1][0], b[1][1], b[1][2], b[1][3]};
            split8hw(v, xh0, xl0);
        }
        {
            const float v[8] = {b[2][0], b[2][1], b[2][2], b[2][3], b[3][0], b[3][1], b[3][2], b[3][3]};
            split8hw(v, xh1, xl1);
        }
        acc0 = mfma(ah, xh0, acc0);
        acc0 = mfma(ah, xl0, acc0);
        acc1 = mfma(ah, xh1, acc1);
        acc1 = mfma(ah, xl1, acc1);
        if (MODE != MODE_BF16) {
            acc0 = mfma(al, xh0, acc0);
            acc1 = mfma(al, xh1, acc1);
        }
    };
    // Panel K-step q: A fragments (on == false: the zero buffer); the t rows
    // of this lane's octet are indices [8g, 8g + 8) of a 32-word index block
    // of the record.  Every strip has at least one index block (zero rows
    // when it has no panel), so the loads issued never depend on the strip.
    auto load_panel_a = [&](int q, bool on, bf16x8 &ah, bf16x8 &al) __attribute__((always_inline)) {
        const uint4 kd = desc(nd + (on ? q : 0));
        const uint32_t afrag = kd.x, info = kd.z;
        const int rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
        load_a<MODE>(frag, zeros, on && g < nOct && r16 < rows, afrag,
                     (g * rows + r16) * (MODE == MODE_F32 ? 2 : 1), nOct * rows, ah, al);
    };
    auto gather_t = [&](const uint4 &q0, const uint4 &q1, f4u (&b)[4]) __attribute__((always_inline)) {
        const uint32_t ix[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float *tr = T + static_cast<int64_t>(ix[e]) * NB + r16;
            b[e >> 2][e & 3] = tr[0];
            b[2 + (e >> 2)][e & 3] = tr[16];
        }
    };
    // A operands of dense K-steps [k0, k0 + KC); K-steps past the strip's end
    // load the zero buffer, and their MFMAs run anyway (a branch around them
    // would let the compiler move their loads behind the first wait)
    auto load_dense_a = [&](int k0, bf16x8 (&ah)[KC], bf16x8 (&al)[KC]) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const uint4 kd = desc(max(0, min(k0 + u, nd - 1)));
            const bool live = k0 + u < nd;
            const uint32_t afrag = kd.x, info = kd.z;
            const int roff = info & 15, rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
            const bool aon = live && g < nOct && r16 >= roff && r16 < roff + rows;
            load_a<MODE>(frag, zeros, aon, afrag, (g * rows + r16 - roff) * (MODE == MODE_F32 ? 2 : 1),
                         nOct * rows, ah[u], al[u]);
        }
    };
    // B operands of dense K-steps [k0, k0 + KC): X rows from the LDS window
    // (16-byte aligned reads) or straight from X
    auto load_dense_b = [&](int k0, auto staged, f4u (&b)[KC][4]) __attribute__((always_inline)) {
        constexpr bool STAGED = decltype(staged)::value;
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const uint4 kd = desc(max(0, min(k0 + u, nd - 1)));
            const bool live = k0 + u < nd;
            const uint32_t src = kd.y, info = kd.z;
            const int nOct = (info >> 9) & 7;
            const bool on = live && g < nOct;
            if constexpr (STAGED) {
                const int off = static_cast<int>(src) - cmin + 8 * g;
                const float *p0 = on ? xs + r16 * XS_LD + off : xzero;
                const float *p1 = on ? xs + (16 + r16) * XS_LD + off : xzero;
                const float4 v0 = reinterpret_cast<const float4 *>(p0)[0];
                const float4 v1 = reinterpret_cast<const float4 *>(p0)[1];
                const float4 v2 = reinterpret_cast<const float4 *>(p1)[0];
                const float4 v3 = reinterpret_cast<const float4 *>(p1)[1];
                b[u][0] = f4u{v0.x, v0.y, v0.z, v0.w};
                b[u][1] = f4u{v1.x, v1.y, v1.z, v1.w};
                b[u][2] = f4u{v2.x, v2.y, v2.z, v2.w};
                b[u][3] = f4u{v3.x, v3.y, v3.z, v3.w};
            } else {
                const int64_t off = static_cast<int64_t>(src) + 8 * g;
                const float *p0 = on && rok0 ? xc0 + off : zeros;
                const float *p1 = on && rok1 ? xc1 + off : zeros;
                b[u][0] = reinterpret_cast<const f4u *>(p0)[0];
                b[u][1] = reinterpret_cast<const f4u *>(p0)[1];
                b[u][2] = reinterpret_cast<const f4u *>(p1)[0];
                b[u][3] = reinterpret_cast<const f4u *>(p1)[1];
            }
        }
    };
    auto batch = [&](int k0, auto staged) __attribute__((always_inline)) {
        bf16x8 ah[KC], al[KC];
        f4u b[KC][4];
        load_dense_a(k0, ah, al);
        load_dense_b(k0, staged, b);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < KC; ++u) kstep(ah[u], al[u], b[u]);
    };
    // rest of a record after its first batch and first panel K-step
    auto rest = [&](auto staged) __attribute__((always_inline)) {
        for (int k0 = KC; k0 < nd; k0 += KC) batch(k0, staged);
        for (int q = 1; q < np; ++q) {
            bf16x8 ah, al;
            f4u b[4];
            load_panel_a(q, true, ah, al);
            const uint4 *ip = reinterpret_cast<const uint4 *>(rw + desc(nd + q).y + 8 * g);
            gather_t(ip[0], ip[1], b);
            kstep(ah, al, b);
        }
    };
    // first batch of a record: loads (issued early), then B reads and MFMAs
    auto first_loads = [&](uint4 &q0, uint4 &q1, bf16x8 &pah, bf16x8 &pal, bf16x8 (&ah)[KC],
                           bf16x8 (&al)[KC]) __attribute__((always_inline)) {
        const uint4 *ip = reinterpret_cast<const uint4 *>(rw + pidx + 8 * g);
        q0 = ip[0];
        q1 = ip[1];
        load_panel_a(0, np > 0, pah, pal);
        load_dense_a(0, ah, al);
    };
    auto first_compute = [&](auto staged, const uint4 &q0, const uint4 &q1, const bf16x8 &pah,
                             const bf16x8 &pal, const bf16x8 (&ah)[KC], const bf16x8 (&al)[KC]) __attribute__((always_inline)) {
        f4u b[KC][4];
        load_dense_b(0, staged, b);
        f4u pb[4];
        __builtin_amdgcn_sched_barrier(0);
        gather_t(q0, q1, pb);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < KC; ++u) kstep(ah[u], al[u], b[u]);
        kstep(pah, pal, pb);
    };
    auto finish = [&](auto staged) __attribute__((always_inline)) {
        rest(staged);
        rec = word(5);
        while (rec != 0u) {                        // chained records
            rw = recs + rec;
            hv0 = rw[l];
            hv1 = rw[64 + l];
            decode();
            uint4 q0, q1;
            bf16x8 pah, pal, ah[KC], al[KC];
            first_loads(q0, q1, pah, pal, ah, al);
            first_compute(staged, q0, q1, pah, pal, ah, al);
            rest(staged);
            rec = word(5);
        }
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
    };

    if (W > 0) {
        // The record loads (issued above), the X window loads, then -- once
        // the record has arrived -- the first batch's A and index loads: all
        // in flight together; the LDS stores wait only for the X loads.
        f4u v[U];
        int part[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 64 * SW * u;
            const int c = p / NQ, q = p % NQ;
            const int64_t row = static_cast<int64_t>(cmin) + 4 * q;
            const bool in = c < ncol && 4 * q < W;
            part[u] = in ? static_cast<int>(min<int64_t>(4, ncols - row)) : 0;
            const float *src = part[u] == 4 ? X + static_cast<int64_t>(c) * ldx + row : zeros;
            v[u] = *reinterpret_cast<const f4u *>(src);
        }
        decode();
        uint4 q0, q1;
        bf16x8 pah, pal, ah[KC], al[KC];
        first_loads(q0, q1, pah, pal, ah, al);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (part[u] > 0 && part[u] < 4) {        // the quad cut by the last row of X
                const int p = threadIdx.x + 64 * SW * u;
                const float *src = X + static_cast<int64_t>(p / NQ) * ldx + cmin + 4 * (p % NQ);
                f4u w = {0.f, 0.f, 0.f, 0.f};
                for (int e = 0; e < part[u]; ++e) w[e] = src[e];
                v[u] = w;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 64 * SW * u;
            *reinterpret_cast<float4 *>(&xs[(p / NQ) * XS_LD + 4 * (p % NQ)]) =
                make_float4(v[u][0], v[u][1], v[u][2], v[u][3]);
        }
        if (threadIdx.x < 8) xs[NB * XS_LD + threadIdx.x] = 0.0f;
        __syncthreads();
        if (!has) return;
        first_compute(std::true_type{}, q0, q1, pah, pal, ah, al);
        finish(std::true_type{});
    } else {
        if (!has) return;
        decode();
        uint4 q0, q1;
        bf16x8 pah, pal, ah[KC], al[KC];
        first_loads(q0, q1, pah, pal, ah, al);
        first_compute(std::false_type{}, q0, q1, pah, pal, ah, al);
        finish(std::false_type{});
    }
}


template <int MODE>
static hipError_t run_b(const Direction &d, const void *zeros, int ncol, const float *X,
                        int64_t ldx, float *Y, int64_t ldy, float alpha, float beta, hipStream_t s,
                        hipEvent_t *ev) {
    const Batched &b = d.bat;
    hipError_t e = hipSuccess;
    if (b.nst > 0 && b.gps > 0) {
        hipLaunchKernelGGL(kb_proj<MODE>, dim3(static_cast<unsigned>(b.nst)), dim3(256), 0, s,
                           b.gps, b.pcol, b.vslc, static_cast<const float *>(zeros), X, ldx,
                           d.ncols, ncol, b.t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[0], s);
    if (b.nst > 0 && b.gps > 0 && b.nred > 0) {
        const int64_t nwg = b.nredbig + (b.nred - b.nredbig + 3) / 4;
        hipLaunchKernelGGL(kb_reduce, dim3(static_cast<unsigned>(nwg)), dim3(256), 0, s, b.red, b.nredbig, b.nred,
                           static_cast<const float *>(b.t) + b.tLen, static_cast<const float *>(zeros), b.t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (b.nstrips > 0) {
        const int sw = b.shape.sw;
        const int64_t nwg = (b.nstrips + sw - 1) / sw;
        const int64_t per = (nwg + 7) / 8;        // workgroups per XCD
        const dim3 grid(static_cast<unsigned>(8 * per)), block(static_cast<unsigned>(64 * sw));
        const float *zf = static_cast<const float *>(zeros);
        const float *tf = static_cast<const float *>(b.t);
        const int bz = beta == 0.0f ? 1 : 0;
        if (sw == ROW_WIDE.sw && b.shape.xsm == ROW_WIDE.xsm)
            hipLaunchKernelGGL((kb_strips<MODE, SKC, ROW_WIDE.sw, ROW_WIDE.xsm>), grid, block, 0, s, b.nstrips,
                               per, b.soff, b.recs, b.wgx, b.frag, zf, tf, X, ldx, d.ncols, Y, ldy, ncol, alpha,
                               beta, bz);
        else if (sw == ROW_NARROW.sw && b.shape.xsm == ROW_NARROW.xsm)
            hipLaunchKernelGGL((kb_strips<MODE, SKC, ROW_NARROW.sw, ROW_NARROW.xsm>), grid, block, 0, s,
                               b.nstrips, per, b.soff, b.recs, b.wgx, b.frag, zf, tf, X, ldx, d.ncols, Y, ldy, ncol,
                               alpha, beta, bz);
        else
            return hipErrorInvalidValue;
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
        case MODE_F32: return run_b<MODE_F32>(d, p.zeros, ncol, X, ldx, Y, ldy, alpha, beta, s, ev);
        case MODE_BF16: return run_b<MODE_BF16>(d, p.zeros, ncol, X, ldx, Y, ldy, alpha, beta, s, ev);
        case MODE_BF16X2: return run_b<MODE_BF16X2>(d, p.zeros, ncol, X, ldx, Y, ldy, alpha, beta, s, ev);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace hm
