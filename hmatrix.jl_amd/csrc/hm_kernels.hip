// Generic product kernels: y = beta*y + alpha*H*x for one right-hand side,
// any storage precision.
//
// A product is two passes over the factors, each streamed exactly once:
//   1. projection: t_b = B_b' x[cols(b)] for every low-rank leaf b, one
//      wavefront per (leaf, row chunk, rank group); long leaves are cut into
//      chunks whose partial sums a tiny reduction kernel adds in a fixed
//      order (deterministic, no atomics);
//   2. row owner: one workgroup per tile of <= 64 rows of y; each lane owns
//      one row and walks every leaf that intersects the tile: dense C rows
//      (C column-major, so a column of the tile is one contiguous 512-byte
//      read per wavefront for FP64) and low-rank A rows times t_b.  The 4
//      wavefronts split the tile's leaves and combine in LDS; exactly one
//      thread writes each y element.
// This is the arithmetic of hmat_matvec! (src/harithm.jl:340-354): dense
// leaves r += s*C*v (line 342), low-rank leaves r += s*A*(B'*v) (line 344).
#include "hm_internal.h"

namespace hm {

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// Storage readers: element i of the arena, widened to the accumulation type.
struct StF64 {
    using acc = double;
    static __device__ __forceinline__ double ld(const void *hi, const void *, int64_t i) {
        return static_cast<const double *>(hi)[i];
    }
};
struct StF32 {
    using acc = float;
    static __device__ __forceinline__ float ld(const void *hi, const void *, int64_t i) {
        return static_cast<const float *>(hi)[i];
    }
};
struct StBF16 {
    using acc = float;
    static __device__ __forceinline__ float ld(const void *hi, const void *, int64_t i) {
        return bf16_to_f32(static_cast<const uint16_t *>(hi)[i]);
    }
};
struct StBF16x2 {
    using acc = float;
    static __device__ __forceinline__ float ld(const void *hi, const void *lo, int64_t i) {
        return bf16_to_f32(static_cast<const uint16_t *>(hi)[i]) +
               bf16_to_f32(static_cast<const uint16_t *>(lo)[i]);
    }
};

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Pass 1: projection partials.  One wavefront per item; lanes stride the
// rows of B, each lane keeps up to PROJ_KMAX rank accumulators.
template <class St>
__global__ __launch_bounds__(256) void k_proj(const BlkDev *__restrict__ blk,
                                              const ProjItem *__restrict__ items, int64_t nitems,
                                              const void *__restrict__ fac,
                                              const void *__restrict__ fac_lo,
                                              const typename St::acc *__restrict__ x,
                                              typename St::acc *__restrict__ out) {
    using A = typename St::acc;
    const int lane = threadIdx.x & 63;
    const int64_t it = static_cast<int64_t>(blockIdx.x) * WAVES + (threadIdx.x >> 6);
    if (it >= nitems) return;
    const ProjItem w = items[it];
    const BlkDev b = blk[w.blk];
    const int64_t ld = b.n;
    const int64_t base = b.offB + static_cast<int64_t>(w.j0) * ld;
    const A *xs = x + b.col0;
    A acc[PROJ_KMAX];
#pragma unroll
    for (int j = 0; j < PROJ_KMAX; ++j) acc[j] = A(0);
    for (int r = lane; r < w.rows; r += 64) {
        const int64_t i = w.row0 + r;
        const A xv = xs[i];
#pragma unroll
        for (int j = 0; j < PROJ_KMAX; ++j)
            if (j < w.kc) acc[j] += St::ld(fac, fac_lo, base + i + j * ld) * xv;
    }
#pragma unroll
    for (int j = 0; j < PROJ_KMAX; ++j) {
        if (j < w.kc) {
            const A v = wave_sum(acc[j]);
            if (lane == 0) out[w.out + j] = v;
        }
    }
}

// Pass 1b: t = sum of chunk partials, chunks in ascending order.
template <class A>
__global__ __launch_bounds__(256) void k_reduce(const RedItem *__restrict__ items, int64_t nitems,
                                                const A *__restrict__ part, A *__restrict__ t) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t it = g / PROJ_KMAX;
    const int j = static_cast<int>(g % PROJ_KMAX);
    if (it >= nitems) return;
    const RedItem r = items[it];
    if (j >= r.kc) return;
    A s = A(0);
    for (int c = 0; c < r.nchunk; ++c) s += part[r.pOff + static_cast<int64_t>(c) * r.kc + j];
    t[r.tOff + j] = s;
}

// Pass 2: row owner.  One workgroup per tile of <= 64 rows.
template <class St>
__global__ __launch_bounds__(256) void k_rows(const BlkDev *__restrict__ blk,
                                              const TileDev *__restrict__ tiles,
                                              const TileItem *__restrict__ items,
                                              const void *__restrict__ fac,
                                              const void *__restrict__ fac_lo,
                                              const void *__restrict__ dfac,
                                              const void *__restrict__ dfac_lo,
                                              const typename St::acc *__restrict__ x,
                                              const typename St::acc *__restrict__ t,
                                              typename St::acc *__restrict__ y,
                                              typename St::acc alpha, typename St::acc beta,
                                              int beta_zero) {
    using A = typename St::acc;
    __shared__ A red[WAVES][64];
    const TileDev T = tiles[blockIdx.x];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const bool act = lane < T.rows;
    A a0 = A(0), a1 = A(0), a2 = A(0), a3 = A(0);
    for (int q = T.ib + wave; q < T.ie; q += WAVES) {
        const TileItem it = items[q];
        const BlkDev b = blk[it.blk];
        if (!act) continue;
        const int64_t m = b.m;
        int64_t p = b.offA + it.ro + lane;
        if (b.kind == KIND_DENSE) {
            const A *xs = x + b.col0;
            int j = 0;
            for (; j + 4 <= b.n; j += 4, p += 4 * m) {
                a0 += St::ld(dfac, dfac_lo, p) * xs[j];
                a1 += St::ld(dfac, dfac_lo, p + m) * xs[j + 1];
                a2 += St::ld(dfac, dfac_lo, p + 2 * m) * xs[j + 2];
                a3 += St::ld(dfac, dfac_lo, p + 3 * m) * xs[j + 3];
            }
            for (; j < b.n; ++j, p += m) a0 += St::ld(dfac, dfac_lo, p) * xs[j];
        } else {
            const A *ts = t + b.tOff;
            for (int j = 0; j < b.k; ++j, p += m) a1 += St::ld(fac, fac_lo, p) * ts[j];
        }
    }
    red[wave][lane] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (wave == 0 && act) {
        A s = red[0][lane];
#pragma unroll
        for (int w = 1; w < WAVES; ++w) s += red[w][lane];
        const int64_t r = T.row0 + lane;
        y[r] = beta_zero ? alpha * s : beta * y[r] + alpha * s;
    }
}

template <class St>
static hipError_t run(const Plan &p, const Direction &d, const void *xv, void *yv, double alpha,
                      double beta, hipStream_t s) {
    using A = typename St::acc;
    const A *x = static_cast<const A *>(xv);
    A *y = static_cast<A *>(yv);
    A *t = static_cast<A *>(d.t);
    if (d.nproj > 0) {
        // The scratch buffer is [t (tLen) | partials (partLen)]: items that
        // cover a whole leaf write t directly, chunked items write partials
        // (their offsets already include tLen) and k_reduce folds them.
        const unsigned g = static_cast<unsigned>((d.nproj + WAVES - 1) / WAVES);
        hipLaunchKernelGGL(k_proj<St>, dim3(g), dim3(64 * WAVES), 0, s, d.blk, d.proj, d.nproj,
                           p.fac, p.fac_lo, x, t);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (d.nred > 0) {
            const int64_t thr = d.nred * PROJ_KMAX;
            const unsigned g2 = static_cast<unsigned>((thr + 255) / 256);
            hipLaunchKernelGGL(k_reduce<A>, dim3(g2), dim3(256), 0, s, d.red, d.nred,
                               static_cast<const A *>(t) + d.tLen, t);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    if (d.ntiles > 0) {
        hipLaunchKernelGGL(k_rows<St>, dim3(static_cast<unsigned>(d.ntiles)), dim3(64 * WAVES), 0,
                           s, d.blk, d.tiles, d.items, p.fac, p.fac_lo, d.denseA, d.denseA_lo, x,
                           static_cast<const A *>(t), y, static_cast<A>(alpha), static_cast<A>(beta),
                           beta == 0.0 ? 1 : 0);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_matvec(const Plan &p, const Direction &d, const void *x, void *y, double alpha,
                         double beta, hipStream_t s) {
    switch (p.storage) {
        case 0: return run<StF64>(p, d, x, y, alpha, beta, s);
        case 1: return run<StF32>(p, d, x, y, alpha, beta, s);
        case 2: return run<StBF16>(p, d, x, y, alpha, beta, s);
        case 3: return run<StBF16x2>(p, d, x, y, alpha, beta, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace hm
