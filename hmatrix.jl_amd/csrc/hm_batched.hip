// Batched multi-right-hand-side product Y = beta*Y + alpha*H*X for panels of
// up to NB = 32 right-hand sides, FP32 accumulation, on MFMA.
//
// Three kernels per panel, each streaming its factors exactly once:
//   kb_proj   t_b = B_b' X[cols(b)] for every low-rank leaf.  One workgroup
//             per 256-column super-tile stages X[super-tile] in LDS as split
//             BF16 (hi + lo), then its 4 wavefronts walk the leaves that
//             cross the super-tile: v_mfma_f32_16x16x32_bf16 with X' as the
//             A operand (M = right-hand sides) and the B slice as the B
//             operand (N = rank).  X is read from HBM once per product, not
//             once per leaf.
//   kb_reduce sums the per-super-tile partials of leaves wider than one
//             super-tile, in a fixed order (deterministic, no atomics).
//   kb_rows   row owner: one wavefront per tile of <= 64 rows of Y.  Dense
//             leaves are pre-packed in MFMA A-fragment order (16 rows x 8
//             columns per lane-group, zero rows and columns are not stored)
//             and multiplied with X fragments on MFMA; low-rank leaves add
//             A[rows] t_b on VALU (rank is small).  The accumulator stays in
//             registers and each Y element is written once.  Workgroups are
//             remapped so that consecutive tiles run on the same XCD and
//             share the X segments of their dense leaves through that XCD's
//             L2.
// Storage modes: HM_BF16X2 (dense leaves as hi+lo BF16 planes, 3 MFMAs per
// product term: hi*hi + hi*lo + lo*hi; low-rank factors FP32), HM_F32 (dense
// leaves FP32, split in registers, same 3 MFMAs), HM_BF16 (dense leaves and
// factors BF16; X still split, 2 MFMAs).
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

__device__ __forceinline__ bf16x8 load_bf8(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// projection
template <int MODE>
__global__ __launch_bounds__(256) void kb_proj(const PTile *__restrict__ ptiles,
                                               const PItem *__restrict__ pitems,
                                               const void *__restrict__ vslc,
                                               const float *__restrict__ X, int64_t ldx,
                                               int64_t ncols, int ncol, float *__restrict__ out) {
    __shared__ uint4 xs[2][PROJ_SUPER / 8][NB];   // [hi/lo][row octet][rhs], 8 BF16 each
    const PTile P = ptiles[blockIdx.x];
    // Stage X[col0 : col0+rows, 0:NB] as split BF16.  Thread p reads 8
    // consecutive rows of one column (adjacent threads: adjacent octets).
    for (int p = threadIdx.x; p < NB * (PROJ_SUPER / 8); p += 256) {
        const int oct = p % (PROJ_SUPER / 8);
        const int rhs = p / (PROJ_SUPER / 8);
        float v[8];
        const int64_t r0 = P.col0 + 8 * oct;
        const float *src = X + static_cast<int64_t>(rhs) * ldx + r0;
        const bool colok = rhs < ncol;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const bool ok = colok && (8 * oct + e < P.rows) && (r0 + e < ncols);
            v[e] = ok ? src[e] : 0.0f;
        }
        bf16x8 hi, lo;
        split8(v, hi, lo);
        xs[0][oct][rhs] = __builtin_bit_cast(uint4, hi);
        xs[1][oct][rhs] = __builtin_bit_cast(uint4, lo);
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    for (int q = P.ib + wave; q < P.ie; q += 4) {
        const PItem it = pitems[q];
        floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        const int nK = (it.nrow + 31) >> 5;
        for (int s = 0; s < nK; ++s) {
            const int rloc = 32 * s + 8 * g;            // row inside the slice
            const bool valid = rloc < it.nrow;
            bf16x8 vh, vl;
            if (MODE == MODE_BF16) {
                if (valid && r16 < it.k)
                    vh = load_bf8(static_cast<const uint16_t *>(vslc) + it.voff +
                                  static_cast<int64_t>(r16) * it.nrow + rloc);
                else
                    vh = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
                vl = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
            } else {
                float v[8];
                if (valid && r16 < it.k) {
                    const float *src = static_cast<const float *>(vslc) + it.voff +
                                       static_cast<int64_t>(r16) * it.nrow + rloc;
                    const float4 a = reinterpret_cast<const float4 *>(src)[0];
                    const float4 b = reinterpret_cast<const float4 *>(src)[1];
                    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
                    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = 0.0f;
                }
                split8(v, vh, vl);
            }
            const int oct = (it.ls >> 3) + 4 * s + g;
            bf16x8 xh0, xl0, xh1, xl1;
            if (valid) {
                xh0 = __builtin_bit_cast(bf16x8, xs[0][oct][r16]);
                xl0 = __builtin_bit_cast(bf16x8, xs[1][oct][r16]);
                xh1 = __builtin_bit_cast(bf16x8, xs[0][oct][16 + r16]);
                xl1 = __builtin_bit_cast(bf16x8, xs[1][oct][16 + r16]);
            } else {
                xh0 = xl0 = xh1 = xl1 = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
            }
            acc0 = mfma(xh0, vh, acc0);
            acc0 = mfma(xl0, vh, acc0);
            acc1 = mfma(xh1, vh, acc1);
            acc1 = mfma(xl1, vh, acc1);
            if (MODE != MODE_BF16) {
                acc0 = mfma(xh0, vl, acc0);
                acc1 = mfma(xh1, vl, acc1);
            }
        }
        // D[i = rhs][j = rank]: lane holds rhs 4g..4g+3 (+16 for acc1), rank r16
        if (r16 < it.k) {
            float *dst = out + it.out + static_cast<int64_t>(r16) * NB + 4 * g;
            reinterpret_cast<float4 *>(dst)[0] = make_float4(acc0[0], acc0[1], acc0[2], acc0[3]);
            reinterpret_cast<float4 *>(dst + 16)[0] = make_float4(acc1[0], acc1[1], acc1[2], acc1[3]);
        }
    }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kb_reduce(const BRed *__restrict__ red,
                                                 const float *__restrict__ part,
                                                 float *__restrict__ t) {
    const BRed r = red[blockIdx.x];
    for (int e = threadIdx.x; e < r.len; e += 256) {
        float s = 0.0f;
        for (int c = 0; c < r.nchunk; ++c) s += part[r.pOff + static_cast<int64_t>(c) * r.len + e];
        t[r.tOff + e] = s;
    }
}

// ---------------------------------------------------------------------------
// row owner
template <int MODE>
__global__ __launch_bounds__(256) void kb_rows(const BTile *__restrict__ tiles, int64_t ntiles,
                                               int64_t wg_per_xcd, const BItem *__restrict__ items,
                                               const void *__restrict__ frag,
                                               const void *__restrict__ uslc,
                                               const float *__restrict__ T,
                                               const float *__restrict__ X, int64_t ldx,
                                               int64_t ncols, float *__restrict__ Y, int64_t ldy,
                                               int ncol, float alpha, float beta, int beta_zero) {
    // Hardware places workgroup b on XCD b % 8; give each XCD a contiguous
    // range of tiles.
    const int64_t L = static_cast<int64_t>(blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8;
    const int64_t tile = L * 4 + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    const int l = threadIdx.x & 63;
    const int r16 = l & 15;
    const int g = l >> 4;
    const BTile Tt = tiles[tile];
    const int R = Tt.rows;
    const int nM = (R + 15) >> 4;
    constexpr int planes = MODE == MODE_BF16X2 ? 2 : 1;
    floatx4 acc[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        acc[mt][0] = floatx4{0.f, 0.f, 0.f, 0.f};
        acc[mt][1] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, short8{0, 0, 0, 0, 0, 0, 0, 0});
    for (int q = Tt.ib; q < Tt.ie; ++q) {
        const BItem it = items[q];
        if (it.kind == KIND_DENSE) {
            const int n = it.dim;
            const int nK = (n + 31) >> 5;
            for (int s = 0; s < nK; ++s) {
                const int rem = n - 32 * s;
                const int nOct = rem >= 32 ? 4 : (rem + 7) >> 3;
                const int vcnt = min(8, max(0, rem - 8 * g));
                bf16x8 xh[2], xl[2];
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) {
                    const int col = 16 * nt + r16;
                    float v[8];
                    const float *src = X + static_cast<int64_t>(col) * ldx + it.aux + 32 * s + 8 * g;
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[e] = (e < vcnt && col < ncol) ? src[e] : 0.0f;
                    split8(v, xh[nt], xl[nt]);
                }
                const int64_t kbase = it.off + static_cast<int64_t>(s) * planes * 32 * R;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    if (mt < nM) {
                        const int Rm = min(16, R - 16 * mt);
                        const int64_t mbase = kbase + static_cast<int64_t>(planes) * 128 * nOct * mt;
                        const bool on = r16 < Rm && g < nOct;
                        const int64_t lo_off = static_cast<int64_t>(nOct) * Rm * 8;
                        const int64_t la = mbase + static_cast<int64_t>(g * Rm + r16) * 8;
                        bf16x8 ah, al;
                        if (MODE == MODE_F32) {
                            float v[8];
                            if (on) {
                                const float *src = static_cast<const float *>(frag) + la;
                                const float4 a = reinterpret_cast<const float4 *>(src)[0];
                                const float4 b = reinterpret_cast<const float4 *>(src)[1];
                                v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
                                v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
                            } else {
#pragma unroll
                                for (int e = 0; e < 8; ++e) v[e] = 0.0f;
                            }
                            split8(v, ah, al);
                        } else {
                            const uint16_t *fb = static_cast<const uint16_t *>(frag);
                            ah = on ? load_bf8(fb + la) : zero8;
                            al = (MODE == MODE_BF16X2 && on) ? load_bf8(fb + la + lo_off) : zero8;
                        }
#pragma unroll
                        for (int nt = 0; nt < 2; ++nt) {
                            acc[mt][nt] = mfma(ah, xh[nt], acc[mt][nt]);
                            acc[mt][nt] = mfma(ah, xl[nt], acc[mt][nt]);
                            if (MODE != MODE_BF16) acc[mt][nt] = mfma(al, xh[nt], acc[mt][nt]);
                        }
                    }
                }
            }
        } else {
            const int k = it.dim;
            const int R4 = (R + 3) & ~3;
            for (int j = 0; j < k; ++j) {
                const float t0 = T[it.aux + static_cast<int64_t>(j) * NB + r16];
                const float t1 = T[it.aux + static_cast<int64_t>(j) * NB + 16 + r16];
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    if (mt < nM) {
                        const int row = 16 * mt + 4 * g;
                        float u[4] = {0.f, 0.f, 0.f, 0.f};
                        if (row < R) {
                            const int64_t ua = it.off + static_cast<int64_t>(j) * R4 + row;
                            if (MODE == MODE_BF16) {
                                const uint2 w = *reinterpret_cast<const uint2 *>(
                                    static_cast<const uint16_t *>(uslc) + ua);
                                u[0] = bf_bits2f(w.x & 0xffffu);
                                u[1] = bf_bits2f(w.x >> 16);
                                u[2] = bf_bits2f(w.y & 0xffffu);
                                u[3] = bf_bits2f(w.y >> 16);
                            } else {
                                const float4 w = *reinterpret_cast<const float4 *>(
                                    static_cast<const float *>(uslc) + ua);
                                u[0] = w.x; u[1] = w.y; u[2] = w.z; u[3] = w.w;
                            }
                        }
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            acc[mt][0][i] += u[i] * t0;
                            acc[mt][1][i] += u[i] * t1;
                        }
                    }
                }
            }
        }
    }
    (void)ncols;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        if (mt < nM) {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const int col = 16 * nt + r16;
                if (col >= ncol) continue;
                float *dst = Y + static_cast<int64_t>(col) * ldy + Tt.row0 + 16 * mt + 4 * g;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (16 * mt + 4 * g + i < R)
                        dst[i] = beta_zero ? alpha * acc[mt][nt][i] : beta * dst[i] + alpha * acc[mt][nt][i];
                }
            }
        }
    }
}

template <int MODE>
static hipError_t run_b(const Direction &d, int ncol, const float *X, int64_t ldx, float *Y,
                        int64_t ldy, float alpha, float beta, hipStream_t s) {
    const Batched &b = d.bat;
    if (b.nptiles > 0) {
        hipLaunchKernelGGL(kb_proj<MODE>, dim3(static_cast<unsigned>(b.nptiles)), dim3(256), 0, s,
                           b.ptiles, b.pitems, b.vslc, X, ldx, d.ncols, ncol, b.t);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (b.nred > 0) {
            hipLaunchKernelGGL(kb_reduce, dim3(static_cast<unsigned>(b.nred)), dim3(256), 0, s, b.red,
                               static_cast<const float *>(b.t) + b.tLen, b.t);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    if (b.ntiles > 0) {
        const int64_t nwg = (b.ntiles + 3) / 4;
        const int64_t per = (nwg + 7) / 8;
        hipLaunchKernelGGL(kb_rows<MODE>, dim3(static_cast<unsigned>(8 * per)), dim3(256), 0, s,
                           b.tiles, b.ntiles, per, b.items, b.frag, b.uslc,
                           static_cast<const float *>(b.t), X, ldx, d.ncols, Y, ldy, ncol, alpha,
                           beta, beta == 0.0f ? 1 : 0);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_batched(const Plan &p, const Direction &d, int ncol, const float *X, int64_t ldx,
                          float *Y, int64_t ldy, float alpha, float beta, hipStream_t s) {
    switch (p.storage) {
        case MODE_F32: return run_b<MODE_F32>(d, ncol, X, ldx, Y, ldy, alpha, beta, s);
        case MODE_BF16: return run_b<MODE_BF16>(d, ncol, X, ldx, Y, ldy, alpha, beta, s);
        case MODE_BF16X2: return run_b<MODE_BF16X2>(d, ncol, X, ldx, Y, ldy, alpha, beta, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace hm
