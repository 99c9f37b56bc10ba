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
// row owner
constexpr int BK_LD = KSTAGE + 8;   // [plane][rhs][k] BF16 stride: conflict-free b128 reads

__device__ __forceinline__ int64_t w64(const uint32_t *w) {
    return static_cast<int64_t>(static_cast<uint64_t>(w[0]) | (static_cast<uint64_t>(w[1]) << 32));
}

template <int MODE>
__global__ __launch_bounds__(256) void kb_rows(int64_t ntiles, int64_t wg_per_xcd,
                                               const uint32_t *__restrict__ blobs,
                                               const uint32_t *__restrict__ blobx,
                                               const void *__restrict__ frag,
                                               const float *__restrict__ zeros,
                                               const float *__restrict__ T,
                                               const float *__restrict__ X, int64_t ldx,
                                               float *__restrict__ Y, int64_t ldy,
                                               int ncol, float alpha, float beta, int beta_zero) {
    // B rows of one staging batch as split BF16, [hi/lo][rhs][k row]; the
    // same storage holds the tile's Y block (as floats) at the end.
    __shared__ uint32_t bsm[2 * NB * BK_LD / 2];
    __shared__ uint4 blob4[BLOB_MAX / 4];
    uint16_t(*bs)[NB][BK_LD] = reinterpret_cast<uint16_t(*)[NB][BK_LD]>(bsm);
    const uint32_t *blob = reinterpret_cast<const uint32_t *>(blob4);
    // Hardware places workgroup b on XCD b % 8; give each XCD a contiguous
    // range of tiles so neighbouring tiles share X rows in that L2.
    const int64_t tile = static_cast<int64_t>(blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8;
    if (tile >= ntiles) return;                       // uniform across the workgroup
    // the tile's descriptor blob: one round trip (two when it overflows)
    if (threadIdx.x < BLOB_STRIDE / 4)
        blob4[threadIdx.x] = reinterpret_cast<const uint4 *>(blobs + tile * BLOB_STRIDE)[threadIdx.x];
    __syncthreads();
    if (blob[7] != 0u) {
        const int nw = static_cast<int>(blob[7]);
        const uint4 *src = reinterpret_cast<const uint4 *>(blobx + w64(blob + 8));
        __syncthreads();
        for (int i = threadIdx.x; i < nw / 4; i += 256) blob4[i] = src[i];
        __syncthreads();
    }
    const int64_t row0 = w64(blob);
    const int R = static_cast<int>(blob[2]);
    const int nbatch = static_cast<int>(blob[3]);
    const int nseg = static_cast<int>(blob[4]);
    const int nks = static_cast<int>(blob[5]);
    const uint32_t *bat = blob + BLOB_HDR;
    const uint32_t *sgw = bat + 4 * nbatch;
    const uint32_t *ksw = sgw + 6 * nseg;
    const uint32_t *tof = ksw + 4 * nks;

    const int wave = threadIdx.x >> 6;               // M-tile of this wavefront
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const int Rm = min(16, R - 16 * wave);            // <= 0: wave has no rows
    constexpr int planes = MODE == MODE_BF16X2 ? 2 : 1;
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};

    for (int bi = 0; bi < nbatch; ++bi) {
        const int sb = static_cast<int>(bat[4 * bi]), se = static_cast<int>(bat[4 * bi + 1]);
        const int kb = static_cast<int>(bat[4 * bi + 2]), ke = static_cast<int>(bat[4 * bi + 3]);
        // (1) this wave's A fragments for every K-step of the batch.  Every
        //     lane issues its loads unconditionally (lanes without data read
        //     the zero buffer), so all of them are in flight together.
        bf16x8 ah[KSLOTS], al[KSLOTS];
        int krow[KSLOTS];
#pragma unroll
        for (int u = 0; u < KSLOTS; ++u) {
            const bool slot = kb + u < ke;
            const uint32_t *kw = ksw + 4 * (slot ? kb + u : 0);
            krow[u] = slot ? static_cast<int>(kw[2]) : 0;
            const int nOct = slot ? static_cast<int>(kw[3]) : 0;
            const bool on = slot && Rm > 0 && r16 < Rm && g < nOct;
            const int64_t la = w64(kw) + static_cast<int64_t>(planes) * 128 * nOct * wave +
                               static_cast<int64_t>(g * Rm + r16) * 8;
            if (MODE == MODE_F32) {
                const float *fp = on ? static_cast<const float *>(frag) + la : zeros;
                const float4 a = reinterpret_cast<const float4 *>(fp)[0];
                const float4 b = reinterpret_cast<const float4 *>(fp)[1];
                float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
                split8hw(v, ah[u], al[u]);
            } else {
                const uint16_t *fb = static_cast<const uint16_t *>(frag);
                const uint16_t *zb = reinterpret_cast<const uint16_t *>(zeros);
                ah[u] = load_bf8(on ? fb + la : zb);
                if (MODE == MODE_BF16X2)
                    al[u] = load_bf8(on ? fb + la + static_cast<int64_t>(nOct) * Rm * 8 : zb);
                else
                    al[u] = zero8;
            }
        }
        // (2) stage the batch's B rows as split BF16 into bs[plane][rhs][k].
        //     X rows (column-major X): (column, 4-row quad) units, one 16-byte
        //     load each, a wavefront reading contiguous column segments.
        //     t vectors: (row pair, 4-rhs quad) units, two 16-byte loads each.
        //     Every load of a thread is issued before its first LDS store.
        {
            constexpr int NQ = KSTAGE / 4;                 // row quads of the stage
            constexpr int UX = NQ * NB / 256;              // X units per thread
            constexpr int UT = (KSTAGE / 2) * (NB / 4) / 256;   // t units per thread
            f4u xv[UX];
            int xq[UX], xpart[UX];
            const float *xsrc[UX];
#pragma unroll
            for (int u = 0; u < UX; ++u) {
                const int p = threadIdx.x + 256 * u;
                const int q = p % NQ;
                const int c = p / NQ;
                const int r = 4 * q;
                xq[u] = -1;
                xpart[u] = 0;
                const float *src = zeros;
                for (int qq = sb; qq < se; ++qq) {
                    const uint32_t *sw = sgw + 6 * qq;
                    const int rr = r - static_cast<int>(sw[4]);
                    const int klen = static_cast<int>(sw[2]);
                    if (sw[3] == 0u && rr >= 0 && rr < ((klen + 31) & ~31)) {
                        xq[u] = 1;
                        if (c < ncol && rr < klen) {
                            src = X + static_cast<int64_t>(c) * ldx + w64(sw) + rr;
                            xpart[u] = min(4, klen - rr);    // valid rows of the quad
                        }
                    }
                }
                xsrc[u] = xpart[u] == 4 ? src : zeros;
                xv[u] = *reinterpret_cast<const f4u *>(xsrc[u]);
                xsrc[u] = src;
            }
            float4 ta[UT], tb[UT];
            int tr[UT];
#pragma unroll
            for (int u = 0; u < UT; ++u) {
                const int p = threadIdx.x + 256 * u;
                const int cq = p % (NB / 4);
                const int r = 2 * (p / (NB / 4));
                tr[u] = -1;
                const float *sa = zeros, *sb2 = zeros;
                for (int qq = sb; qq < se; ++qq) {
                    const uint32_t *sw = sgw + 6 * qq;
                    const int rr = r - static_cast<int>(sw[4]);
                    const int klen = static_cast<int>(sw[2]);
                    if (sw[3] == 1u && rr >= 0 && rr < ((klen + 31) & ~31)) {
                        tr[u] = 1;
                        const uint32_t *tt = tof + sw[5] + rr;
                        if (rr < klen) sa = T + static_cast<int64_t>(tt[0]) * NB + 4 * cq;
                        if (rr + 1 < klen) sb2 = T + static_cast<int64_t>(tt[1]) * NB + 4 * cq;
                    }
                }
                ta[u] = *reinterpret_cast<const float4 *>(sa);
                tb[u] = *reinterpret_cast<const float4 *>(sb2);
            }
            // quads cut by a segment end: element loads, after every vector
            // load above has been issued
#pragma unroll
            for (int u = 0; u < UX; ++u) {
                if (xpart[u] > 0 && xpart[u] < 4) {
                    f4u v = {0.f, 0.f, 0.f, 0.f};
                    for (int e = 0; e < xpart[u]; ++e) v[e] = xsrc[u][e];
                    xv[u] = v;
                }
            }
#pragma unroll
            for (int u = 0; u < UX; ++u) {
                const int p = threadIdx.x + 256 * u;
                const int q = p % NQ;
                const int c = p / NQ;
                uint2 hi, lo;
                split2(xv[u][0], xv[u][1], hi.x, lo.x);
                split2(xv[u][2], xv[u][3], hi.y, lo.y);
                if (xq[u] > 0) {
                    *reinterpret_cast<uint2 *>(&bs[0][c][4 * q]) = hi;
                    *reinterpret_cast<uint2 *>(&bs[1][c][4 * q]) = lo;
                }
            }
#pragma unroll
            for (int u = 0; u < UT; ++u) {
                const int p = threadIdx.x + 256 * u;
                const int cq = p % (NB / 4);
                const int r = 2 * (p / (NB / 4));
                if (tr[u] > 0) {
                    const float a4[4] = {ta[u].x, ta[u].y, ta[u].z, ta[u].w};
                    const float b4[4] = {tb[u].x, tb[u].y, tb[u].z, tb[u].w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        uint32_t hi, lo;
                        split2(a4[e], b4[e], hi, lo);
                        *reinterpret_cast<uint32_t *>(&bs[0][4 * cq + e][r]) = hi;
                        *reinterpret_cast<uint32_t *>(&bs[1][4 * cq + e][r]) = lo;
                    }
                }
            }
        }
        __syncthreads();
        // (3) MFMA over the batch's K-steps, B fragments straight from LDS
        if (Rm > 0) {
#pragma unroll
            for (int u = 0; u < KSLOTS; ++u) {
                if (kb + u < ke) {
                    const int k0 = krow[u] + 8 * g;
                    const bf16x8 xh0 = load_bf8(&bs[0][r16][k0]);
                    const bf16x8 xl0 = load_bf8(&bs[1][r16][k0]);
                    const bf16x8 xh1 = load_bf8(&bs[0][16 + r16][k0]);
                    const bf16x8 xl1 = load_bf8(&bs[1][16 + r16][k0]);
                    acc0 = mfma(ah[u], xh0, acc0);
                    acc0 = mfma(ah[u], xl0, acc0);
                    acc1 = mfma(ah[u], xh1, acc1);
                    acc1 = mfma(ah[u], xl1, acc1);
                    if (MODE != MODE_BF16) {
                        acc0 = mfma(al[u], xh0, acc0);
                        acc1 = mfma(al[u], xh1, acc1);
                    }
                }
            }
        }
        __syncthreads();
    }
    // Y block through LDS: [rhs][row] floats, then (column, 4-row quad)
    // units written with one 16-byte store each.
    float(*ys)[68] = reinterpret_cast<float(*)[68]>(bsm);
    if (Rm > 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ys[r16][16 * wave + 4 * g + i] = acc0[i];
            ys[16 + r16][16 * wave + 4 * g + i] = acc1[i];
        }
    }
    __syncthreads();
    {
        const int q = threadIdx.x & 15;      // row quad
        for (int c = threadIdx.x >> 4; c < ncol; c += 16) {
            const int r = 4 * q;
            if (r < R) {
                float *dst = Y + static_cast<int64_t>(c) * ldy + row0 + r;
                const float4 a = *reinterpret_cast<const float4 *>(&ys[c][r]);
                f4u v = {alpha * a.x, alpha * a.y, alpha * a.z, alpha * a.w};
                if (r + 4 <= R) {
                    if (!beta_zero) {
                        const f4u o = *reinterpret_cast<const f4u *>(dst);
                        v += beta * o;
                    }
                    *reinterpret_cast<f4u *>(dst) = v;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (r + e < R) dst[e] = beta_zero ? v[e] : beta * dst[e] + v[e];
                }
            }
        }
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
    if (b.ntiles > 0) {
        const int64_t per = (b.ntiles + 7) / 8;
        hipLaunchKernelGGL(kb_rows<MODE>, dim3(static_cast<unsigned>(8 * per)), dim3(256), 0, s,
                           b.ntiles, per, b.blobs, b.blobx, b.frag, static_cast<const float *>(zeros),
                           static_cast<const float *>(b.t),
                           X, ldx, Y, ldy, ncol, alpha, beta, beta == 0.0f ? 1 : 0);
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
