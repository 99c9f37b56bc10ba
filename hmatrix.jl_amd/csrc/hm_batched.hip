// Batched multi-right-hand-side product Y = beta*Y + alpha*H*X for panels of
// up to NB = 32 right-hand sides, FP32 accumulation, on MFMA.
//
// Three kernels per panel, each streaming its factors exactly once:
//   kb_proj   t_b = B_b' X[cols(b)] for every low-rank leaf.  One workgroup
//             per 256-column super-tile stages X[super-tile] in LDS as split
//             BF16 (hi + lo); its 4 wavefronts walk the leaves crossing the
//             super-tile, each issuing all B-slice loads of a leaf at once:
//             v_mfma_f32_16x16x32_bf16 with X' as the A operand (M =
//             right-hand sides) and the B slice as the B operand (N = rank).
//             X is read from HBM once per product, not once per leaf.
//   kb_reduce sums the per-super-tile partials of leaves wider than one
//             super-tile, in a fixed order (deterministic, no atomics).
//   kb_rows   row owner: one workgroup per tile of <= 64 rows of Y, one
//             wavefront per 16-row M-tile.  A tile is a list of K segments:
//             each dense leaf crossing it (A = its rows, B = X rows of its
//             columns) and one panel holding the A columns of every low-rank
//             leaf crossing it (B = the matching t vectors).  All A operands
//             are pre-packed in MFMA fragment order (16 rows x 8 k per lane
//             group; zero rows and octets are not stored).  Per batch of
//             K-steps each wave issues all its A-fragment loads while the
//             workgroup stages the B rows (X windows, t vectors) into LDS
//             with coalesced loads: one memory round trip per batch.  Y
//             leaves through LDS as contiguous column segments, each element
//             written once.  Workgroups are remapped so consecutive tiles run
//             on the same XCD and share X rows in its L2.
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

template <int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kb_proj(const PTile *__restrict__ ptiles,
                                               const PItem *__restrict__ pitems,
                                               const void *__restrict__ vslc,
                                               const float *__restrict__ zeros,
                                               const float *__restrict__ X, int64_t ldx,
                                               int64_t ncols, int ncol, float *__restrict__ out) {
    __shared__ uint4 xs[2][NB][XO_LD];              // [hi/lo][rhs][row octet], 8 BF16 each
    const PTile P = ptiles[blockIdx.x];
    // Stage X[col0:col0+rows] as split BF16 in (column, 4-row quad) units: a
    // wavefront reads 1 KB of one column per instruction; all loads first.
    {
        constexpr int XQ = PROJ_SUPER / 4;
        constexpr int XIT = NB * XQ / 256;
        f4u xv[XIT];
#pragma unroll
        for (int u = 0; u < XIT; ++u) {
            const int p = threadIdx.x + 256 * u;
            const int q = p % XQ;
            const int rhs = p / XQ;
            const bool full = rhs < ncol && 4 * q + 4 <= P.rows;
            xv[u] = *reinterpret_cast<const f4u *>(full ? X + static_cast<int64_t>(rhs) * ldx + P.col0 + 4 * q : zeros);
        }
        // the quad cut by the matrix end: element loads after the vector ones
#pragma unroll
        for (int u = 0; u < XIT; ++u) {
            const int p = threadIdx.x + 256 * u;
            const int q = p % XQ;
            const int rhs = p / XQ;
            if (rhs < ncol && 4 * q < P.rows && 4 * q + 4 > P.rows) {
                const float *src = X + static_cast<int64_t>(rhs) * ldx + P.col0 + 4 * q;
                f4u v = {0.f, 0.f, 0.f, 0.f};
                for (int e = 0; 4 * q + e < P.rows; ++e) v[e] = src[e];
                xv[u] = v;
            }
        }
#pragma unroll
        for (int u = 0; u < XIT; ++u) {
            const int p = threadIdx.x + 256 * u;
            const int q = p % XQ;
            const int rhs = p / XQ;
            uint2 hi, lo;
            split2(xv[u][0], xv[u][1], hi.x, lo.x);
            split2(xv[u][2], xv[u][3], hi.y, lo.y);
            *(reinterpret_cast<uint2 *>(&xs[0][rhs][q >> 1]) + (q & 1)) = hi;
            *(reinterpret_cast<uint2 *>(&xs[1][rhs][q >> 1]) + (q & 1)) = lo;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
    constexpr int KS = PROJ_SUPER / 32 + 1;   // K-steps an item can span
    for (int q = P.ib + wave; q < P.ie; q += 4) {
        const PItem it = pitems[q];
        const int nK = (it.nrow + 31) >> 5;
        const bool kon = r16 < it.k;
        // every B-slice load of the item is issued before the first MFMA;
        // lanes without data read the zero buffer
        bf16x8 vh[KS], vl[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int rloc = 32 * s + 8 * g;
            const bool on = s < nK && kon && rloc < it.nrow;
            const int64_t e0 = it.voff + static_cast<int64_t>(r16) * it.nrow + rloc;
            if (MODE == MODE_F32) {
                const float *vp = on ? static_cast<const float *>(vslc) + e0 : zeros;
                const float4 a = reinterpret_cast<const float4 *>(vp)[0];
                const float4 b = reinterpret_cast<const float4 *>(vp)[1];
                float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
                split8hw(v, vh[s], vl[s]);
            } else {
                const uint16_t *vp = static_cast<const uint16_t *>(vslc);
                const uint16_t *zb = reinterpret_cast<const uint16_t *>(zeros);
                vh[s] = load_bf8(on ? vp + e0 : zb);
                vl[s] = MODE == MODE_BF16X2 ? load_bf8(on ? vp + e0 + static_cast<int64_t>(it.k) * it.nrow : zb)
                                            : zero8;
            }
        }
        floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        // MFMAs run for the whole wavefront (wave-uniform condition); lanes
        // past the slice hold zero operands
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            if (s < nK) {
                bf16x8 xh0 = zero8, xl0 = zero8, xh1 = zero8, xl1 = zero8;
                if (32 * s + 8 * g < it.nrow) {
                    const int oct = (it.ls >> 3) + 4 * s + g;
                    xh0 = __builtin_bit_cast(bf16x8, xs[0][r16][oct]);
                    xl0 = __builtin_bit_cast(bf16x8, xs[1][r16][oct]);
                    xh1 = __builtin_bit_cast(bf16x8, xs[0][16 + r16][oct]);
                    xl1 = __builtin_bit_cast(bf16x8, xs[1][16 + r16][oct]);
                }
                acc0 = mfma(xh0, vh[s], acc0);
                acc0 = mfma(xl0, vh[s], acc0);
                acc1 = mfma(xh1, vh[s], acc1);
                acc1 = mfma(xl1, vh[s], acc1);
                if (MODE != MODE_BF16) {
                    acc0 = mfma(xh0, vl[s], acc0);
                    acc1 = mfma(xh1, vl[s], acc1);
                }
            }
        }
        // D[i = rhs][j = rank]: lane holds rhs 4g..4g+3 (+16 for acc1), rank r16
        if (kon) {
            float *dst = out + it.out + static_cast<int64_t>(r16) * NB + 4 * g;
            reinterpret_cast<float4 *>(dst)[0] = make_float4(acc0[0], acc0[1], acc0[2], acc0[3]);
            reinterpret_cast<float4 *>(dst + 16)[0] = make_float4(acc1[0], acc1[1], acc1[2], acc1[3]);
        }
    }
}

// ---------------------------------------------------------------------------
// t = sum of the per-super-tile partials of one leaf.  Threads split into
// G = 256/len groups; group q sums chunks q, q+G, ... (eight loads in flight)
// and the groups are combined in LDS in ascending order: a fixed order, so
// the result is deterministic.
__global__ __launch_bounds__(256) void kb_reduce(const BRed *__restrict__ red,
                                                 const float *__restrict__ part,
                                                 float *__restrict__ t) {
    __shared__ float sh[256];
    const BRed r = red[blockIdx.x];
    const int G = r.len >= 256 ? 1 : 256 / r.len;
    for (int e0 = 0; e0 < r.len; e0 += 256 / G) {
        const int e = e0 + static_cast<int>(threadIdx.x) % (256 / G);
        const int q = static_cast<int>(threadIdx.x) / (256 / G);
        float acc = 0.0f;
        if (e < r.len && q < G) {
            const float *src = part + r.pOff + e;
            int c = q;
            for (; c + 7 * G < r.nchunk; c += 8 * G) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = src[static_cast<int64_t>(c + u * G) * r.len];
#pragma unroll
                for (int u = 0; u < 8; ++u) acc += v[u];
            }
            for (; c < r.nchunk; c += G) acc += src[static_cast<int64_t>(c) * r.len];
        }
        sh[threadIdx.x] = acc;
        __syncthreads();
        if (q == 0 && e < r.len) {
            float s = acc;
            for (int g = 1; g < G; ++g) s += sh[g * (256 / G) + threadIdx.x];
            t[r.tOff + e] = s;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// row pass: one wavefront per strip of SROWS rows x NB right-hand sides.
// No LDS and no barriers: every operand comes from global memory straight
// into the MFMA registers.  Per K-step a lane loads its 8 A values of strip
// row r16 (octet g) and, for each right-hand-side block h, the 8 B values
// of k-rows [8g, 8g + 8) in column 16h + r16: X rows are contiguous in a
// column of X, panel rows are gathered from t.  KC K-steps are loaded
// together before their MFMAs.

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

template <int MODE, int KC>
__global__ __launch_bounds__(256) void kb_strips(int64_t nstrips, int64_t wg_per_xcd,
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
    __shared__ float xs[NB * XS_LD + 8];
    // Hardware places workgroup b on XCD b % 8; give each XCD a contiguous
    // range of strips so neighbouring strips share X rows in that L2.
    const int64_t wg = static_cast<int64_t>(blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8;
    if (wg * SWAVES >= nstrips) return;       // whole workgroup
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t s = wg * SWAVES + wave;
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const int cmin = wgx[2 * wg], W = wgx[2 * wg + 1];
    const float *xzero = xs + NB * XS_LD;

    // Stage X rows [cmin, cmin + W) of all NB columns (zeros past ncols and
    // ncol): (column, 4-row quad) units, a wavefront reading one column's
    // rows contiguously; every load issued before the first LDS store.
    auto stage = [&]() {
        constexpr int NQ = XS_MAX / 4;                 // quads per column slot
        constexpr int U = NB * NQ / 256;               // units per thread
        f4u v[U];
        int part[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 256 * u;
            const int c = p / NQ, q = p % NQ;
            const int64_t row = static_cast<int64_t>(cmin) + 4 * q;
            const bool in = c < ncol && 4 * q < W;
            part[u] = in ? static_cast<int>(min<int64_t>(4, ncols - row)) : 0;
            const float *src = part[u] == 4 ? X + static_cast<int64_t>(c) * ldx + row : zeros;
            v[u] = *reinterpret_cast<const f4u *>(src);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (part[u] > 0 && part[u] < 4) {        // the quad cut by the last row of X
                const int p = threadIdx.x + 256 * u;
                const float *src = X + static_cast<int64_t>(p / NQ) * ldx + cmin + 4 * (p % NQ);
                f4u w = {0.f, 0.f, 0.f, 0.f};
                for (int e = 0; e < part[u]; ++e) w[e] = src[e];
                v[u] = w;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int p = threadIdx.x + 256 * u;
            *reinterpret_cast<float4 *>(&xs[(p / NQ) * XS_LD + 4 * (p % NQ)]) =
                make_float4(v[u][0], v[u][1], v[u][2], v[u][3]);
        }
        if (threadIdx.x < 8) xs[NB * XS_LD + threadIdx.x] = 0.0f;
    };

    auto body = [&](auto staged) {
    constexpr bool STAGED = decltype(staged)::value;
    if (s >= nstrips) return;                 // this wavefront has no strip
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const bool rok0 = r16 < ncol, rok1 = 16 + r16 < ncol;
    const float *xc0 = X + static_cast<int64_t>(r16) * ldx;
    const float *xc1 = X + static_cast<int64_t>(16 + r16) * ldx;
    int row0 = 0, R = 0;
    uint32_t rec = soff[s];
    do {
    const uint32_t *rw = recs + rec;
    // The first 128 words of the record in two vector registers (lane l
    // holds words l and 64 + l; the record arena is padded so these loads
    // stay in bounds): header and descriptors are then read with readlane,
    // one memory round trip for all of them.
    const uint32_t hv0 = rw[l], hv1 = rw[64 + l];
    auto word = [&](int i) -> uint32_t {          // i wave-uniform, < 128
        const uint32_t a = __builtin_amdgcn_readlane(hv0, i & 63);
        const uint32_t b = __builtin_amdgcn_readlane(hv1, i & 63);
        return i < 64 ? a : b;
    };
    row0 = static_cast<int>(word(0));
    R = static_cast<int>(word(1));
    const int nd = static_cast<int>(word(2));        // dense K-steps
    const int np = static_cast<int>(word(3));        // panel K-steps
    const uint32_t pidx = word(4);                   // first t-row index block
    auto desc = [&](int k) {                         // K-step k (dense first, then panel)
        const int i = SREC_HDR + 4 * k;
        return uint4{word(i), word(i + 1), word(i + 2), 0u};
    };

    // MFMAs of one K-step: 8 B values per lane and block (hi/lo split here)
    auto kstep = [&](const bf16x8 &ah, const bf16x8 &al, const f4u (&b)[4]) {
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
    // Panel K-step q: A fragments (on == false: the zero buffer) and the t
    // rows of this lane's octet, indices [8g, 8g + 8) of the 32-word block
    // at word `ib` of the record.  Every strip has at least one index block
    // (rows of zeros when it has no panel), so the loads issued never depend
    // on the strip: the wait counters then never over-wait on a branch.
    auto load_panel_a = [&](int q, bool on, bf16x8 &ah, bf16x8 &al) {
        const uint4 kd = desc(nd + (on ? q : 0));
        const uint32_t afrag = kd.x, info = kd.z;
        const int rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
        load_a<MODE>(frag, zeros, on && g < nOct && r16 < rows, afrag,
                     (g * rows + r16) * (MODE == MODE_F32 ? 2 : 1), nOct * rows, ah, al);
    };
    auto gather_t = [&](const uint4 &q0, const uint4 &q1, f4u (&b)[4]) {
        const uint32_t ix[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float *tr = T + static_cast<int64_t>(ix[e]) * NB + r16;
            b[e >> 2][e & 3] = tr[0];
            b[2 + (e >> 2)][e & 3] = tr[16];
        }
    };

    // One batch of KC dense K-steps (the first batch also carries the first
    // panel K-step): all loads first, then the MFMAs.
    auto batch = [&](int k0, auto first) {
        constexpr bool FIRST = decltype(first)::value;
        uint4 q0, q1;
        bf16x8 pah, pal;
        if constexpr (FIRST) {
            const uint4 *ip = reinterpret_cast<const uint4 *>(rw + pidx + 8 * g);
            q0 = ip[0];
            q1 = ip[1];
            load_panel_a(0, np > 0, pah, pal);
        }
        bf16x8 ah[KC], al[KC];
        f4u b[KC][4];       // [h * 2 + half]: B values 8g .. 8g+7 of block h
        uint4 kd[KC];
#pragma unroll
        for (int u = 0; u < KC; ++u)
            kd[u] = desc(max(0, min(k0 + u, nd - 1)));
#pragma unroll
        for (int u = 0; u < KC; ++u) {
            const bool live = k0 + u < nd;                          // wave-uniform
            const uint32_t afrag = kd[u].x, src = kd[u].y, info = kd[u].z;
            const int roff = info & 15, rows = (info >> 4) & 31, nOct = (info >> 9) & 7;
            const bool aon = live && g < nOct && r16 >= roff && r16 < roff + rows;
            load_a<MODE>(frag, zeros, aon, afrag, (g * rows + r16 - roff) * (MODE == MODE_F32 ? 2 : 1),
                         nOct * rows, ah[u], al[u]);
            const bool on = live && g < nOct;
            if constexpr (STAGED) {
                // B rows from the LDS window (16-byte aligned reads)
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
        f4u pb[4];
        if constexpr (FIRST) {
            // keep the gather behind the loads above: its addresses wait on
            // the index loads
            __builtin_amdgcn_sched_barrier(0);
            gather_t(q0, q1, pb);
        }
        __builtin_amdgcn_sched_barrier(0);
        // K-steps past the strip's end hold zero operands; their MFMAs run
        // anyway (a branch around them would let the compiler move their
        // loads behind the first wait)
#pragma unroll
        for (int u = 0; u < KC; ++u) kstep(ah[u], al[u], b[u]);
        if constexpr (FIRST) kstep(pah, pal, pb);
    };
    batch(0, std::true_type{});
    for (int k0 = KC; k0 < nd; k0 += KC) batch(k0, std::false_type{});
    for (int q = 1; q < np; ++q) {                 // further panel K-steps
        bf16x8 ah, al;
        f4u b[4];
        load_panel_a(q, true, ah, al);
        const uint4 *ip = reinterpret_cast<const uint4 *>(rw + desc(nd + q).y + 8 * g);
        gather_t(ip[0], ip[1], b);
        kstep(ah, al, b);
    }
    rec = word(5);
    } while (rec != 0u);

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
        stage();
        __syncthreads();
        body(std::true_type{});
    } else {
        body(std::false_type{});
    }
}


template <int MODE>
static hipError_t run_b(const Direction &d, const void *zeros, int ncol, const float *X,
                        int64_t ldx, float *Y, int64_t ldy, float alpha, float beta, hipStream_t s,
                        hipEvent_t *ev) {
    const Batched &b = d.bat;
    hipError_t e = hipSuccess;
    if (b.nptiles > 0) {
        hipLaunchKernelGGL(kb_proj<MODE>, dim3(static_cast<unsigned>(b.nptiles)), dim3(256), 0, s,
                           b.ptiles, b.pitems, b.vslc, static_cast<const float *>(zeros), X, ldx,
                           d.ncols, ncol, b.t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[0], s);
    if (b.nptiles > 0 && b.nred > 0) {
        hipLaunchKernelGGL(kb_reduce, dim3(static_cast<unsigned>(b.nred)), dim3(256), 0, s, b.red,
                           static_cast<const float *>(b.t) + b.tLen, b.t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (b.nstrips > 0) {
        const int64_t nwg = (b.nstrips + SWAVES - 1) / SWAVES;
        const int64_t per = (nwg + 7) / 8;        // workgroups per XCD
        hipLaunchKernelGGL((kb_strips<MODE, SKC>), dim3(static_cast<unsigned>(8 * per)), dim3(64 * SWAVES), 0, s,
                           b.nstrips, per, b.soff, b.recs, b.wgx, b.frag, static_cast<const float *>(zeros),
                           static_cast<const float *>(b.t), X, ldx, d.ncols,
                           Y, ldy, ncol, alpha, beta, beta == 0.0f ? 1 : 0);
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
