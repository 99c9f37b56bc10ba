// Generic product kernels: y = beta*y + alpha*H*x for one right-hand side,
// any storage precision.
//
// A product is two passes over the factors, each streamed exactly once:
//   1. projection: t_b = B_b' x[cols(b)] for every low-rank leaf b, one
//      wavefront per (leaf, row chunk, rank group); long leaves are cut into
//      chunks whose partial sums a tiny reduction kernel adds in a fixed
//      order (deterministic, no atomics);
//   2. row owner: one workgroup per tile of <= 64 rows of y; each lane owns
//      one row.  Every leaf that crosses the tile contributes columns: a
//      dense leaf its n columns (multiplier x[col]), a low-rank leaf the k
//      columns of A (multiplier t_b[j]).  The plan packs these columns
//      tile-major, so a tile is one contiguous stream; the 4 wavefronts
//      split it and combine in LDS; exactly one thread writes each y
//      element.
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

// Broadcast lane `src` (wave-uniform) of v to the whole wavefront.
__device__ __forceinline__ float bcast(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ __forceinline__ double bcast(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), src);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<unsigned int>(lo));
}

// Pass 1: projection partials.  One wavefront per item, compiled for the
// item's exact rank-column count KC so that every load of a round is
// unconditional and all of them issue before the first wait: a round covers
// U 64-row strips of all KC columns (about 32 loads per lane), lanes stride
// the rows, each lane keeps KC accumulators, one wave reduction per column.
// Device-scope stores and loads (coherent across XCDs) for values that one
// workgroup of a launch writes and another reads in the same launch.
template <class T>
__device__ __forceinline__ void store_dev(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T load_dev(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class St, int KC, bool DEV = false>
__device__ __forceinline__ void proj_item(const ProjItem &w, const void *__restrict__ fac,
                                          const void *__restrict__ fac_lo,
                                          const typename St::acc *__restrict__ x,
                                          typename St::acc *__restrict__ out, int lane) {
    using A = typename St::acc;
    // The fused kernel (DEV) shares one register budget with the row tiles:
    // eight loads per lane and column group there keeps it at 8 waves/SIMD.
    constexpr int U0 = KC >= 32 ? 1 : (32 / KC > PROJ_CHUNK / 64 ? PROJ_CHUNK / 64 : 32 / KC);
    constexpr int U = DEV && U0 > 8 / KC ? (8 / KC > 0 ? 8 / KC : 1) : U0;
    const int rows = w.rows;
    const int64_t ld = w.ld;
    const A *xs = x + w.xoff;
    A acc[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) acc[j] = A(0);
    // Lanes past the item's rows load a clamped (valid) row and multiply by
    // zero.
    for (int r0 = 0; r0 < rows; r0 += 64 * U) {
        A xv[U];
        int rc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int r = r0 + 64 * u + lane;
            rc[u] = min(r, rows - 1);
            const A v = xs[rc[u]];
            xv[u] = r < rows ? v : A(0);
        }
        A bv[KC][U];
#pragma unroll
        for (int j = 0; j < KC; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u) bv[j][u] = St::ld(fac, fac_lo, w.base + j * ld + rc[u]);
#pragma unroll
        for (int j = 0; j < KC; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u) acc[j] += bv[j][u] * xv[u];
    }
    // butterfly over all columns at once: KC independent exchanges per step
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < KC; ++j) acc[j] += __shfl_xor(acc[j], o, 64);
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            if constexpr (DEV) store_dev(out + w.out + j, acc[j]);
            else out[w.out + j] = acc[j];
        }
    }
}

template <class St>
__global__ __launch_bounds__(256) void k_proj(const ProjItem *__restrict__ items, int64_t nitems,
                                              const void *__restrict__ fac,
                                              const void *__restrict__ fac_lo,
                                              const typename St::acc *__restrict__ x,
                                              typename St::acc *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t it = static_cast<int64_t>(blockIdx.x) * WAVES + wave;
    if (it >= nitems) return;
    const ProjItem w = items[it];
    static_assert(PROJ_KMAX == 16, "one case per rank-column count");
    switch (w.kc) {
        case 1: proj_item<St, 1>(w, fac, fac_lo, x, out, lane); break;
        case 2: proj_item<St, 2>(w, fac, fac_lo, x, out, lane); break;
        case 3: proj_item<St, 3>(w, fac, fac_lo, x, out, lane); break;
        case 4: proj_item<St, 4>(w, fac, fac_lo, x, out, lane); break;
        case 5: proj_item<St, 5>(w, fac, fac_lo, x, out, lane); break;
        case 6: proj_item<St, 6>(w, fac, fac_lo, x, out, lane); break;
        case 7: proj_item<St, 7>(w, fac, fac_lo, x, out, lane); break;
        case 8: proj_item<St, 8>(w, fac, fac_lo, x, out, lane); break;
        case 9: proj_item<St, 9>(w, fac, fac_lo, x, out, lane); break;
        case 10: proj_item<St, 10>(w, fac, fac_lo, x, out, lane); break;
        case 11: proj_item<St, 11>(w, fac, fac_lo, x, out, lane); break;
        case 12: proj_item<St, 12>(w, fac, fac_lo, x, out, lane); break;
        case 13: proj_item<St, 13>(w, fac, fac_lo, x, out, lane); break;
        case 14: proj_item<St, 14>(w, fac, fac_lo, x, out, lane); break;
        case 15: proj_item<St, 15>(w, fac, fac_lo, x, out, lane); break;
        case 16: proj_item<St, 16>(w, fac, fac_lo, x, out, lane); break;
        default: break;
    }
}

// Pass 1b: t = sum of chunk partials, chunks in ascending order.  The
// partials are loaded RED_ROUND at a time (chunks past the last one load
// chunk 0 and are not added), so a round is one memory round trip.
constexpr int RED_ROUND = 16;
constexpr int RM = MULTI_RHS;
template <class A>
__global__ __launch_bounds__(256) void k_reduce(const RedItem *__restrict__ items, int64_t nitems,
                                                const A *__restrict__ part, A *__restrict__ t) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t it = g / PROJ_KMAX;
    const int j = static_cast<int>(g % PROJ_KMAX);
    if (it >= nitems) return;
    const RedItem r = items[it];
    if (j >= r.kc) return;
    const A *src = part + r.pOff + j;
    A s = A(0);
    for (int c0 = 0; c0 < r.nchunk; c0 += RED_ROUND) {
        A v[RED_ROUND];
#pragma unroll
        for (int u = 0; u < RED_ROUND; ++u) {
            const int c = c0 + u < r.nchunk ? c0 + u : 0;
            v[u] = src[static_cast<int64_t>(c) * r.kc];
        }
#pragma unroll
        for (int u = 0; u < RED_ROUND; ++u)
            if (c0 + u < r.nchunk) s += v[u];
    }
    t[r.tOff + j] = s;
}

// Pass 2, one row tile (the body of k_rows and of the fused kernel below).
// DEV: t is written earlier in the same launch; each wavefront issues
// ready.start(lane) (the first readiness check), its first multipliers and
// stream loads, then ready.finish(lane, started), which returns once the t
// values the tile reads are complete, then reads t with device-scope loads.
struct NoWait {
    __device__ uint32_t start(int) const { return 0; }
    __device__ void finish(int, uint32_t) const {}
};

template <class St, int NW, bool DEV, class Ready = NoWait>
__device__ __forceinline__ void rows_tile(const TileDev &T, const int32_t *__restrict__ mult,
                                          const void *__restrict__ stream, const void *__restrict__ stream_lo,
                                          const typename St::acc *__restrict__ x,
                                          const typename St::acc *__restrict__ t, typename St::acc *__restrict__ y,
                                          typename St::acc alpha, typename St::acc beta, int beta_zero,
                                          typename St::acc (*red)[64], Ready ready = Ready()) {
    using A = typename St::acc;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int R = T.rows;
    const bool act = lane < R;
    const int ncol = T.c1 - T.c0;
    const int per = (ncol + NW - 1) / NW;
    const int cb = T.c0 + min(ncol, wave * per);
    const int ce = T.c0 + min(ncol, (wave + 1) * per);
    A a0 = A(0), a1 = A(0), a2 = A(0), a3 = A(0);
    // Lanes past the tile's rows read row R-1 (a valid address) and their
    // sums are never written: every load is unconditional.
    const int lr = min(lane, R - 1);
    if constexpr (DEV) {
        // groups of eight columns; the last group of a chunk loads column
        // nc-1 in place of the missing ones and skips them in the sums
        bool first = true;
        const uint32_t started = cb < ce ? ready.start(lane) : 0u;
        for (int c = cb; c < ce; c += 64) {
            const int nc = min(64, ce - c);
            const int32_t id = mult[c + min(lane, nc - 1)];
            const int64_t p0 = T.base + static_cast<int64_t>(c - T.c0) * R + lr;
            A v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = St::ld(stream, stream_lo, p0 + static_cast<int64_t>(min(u, nc - 1)) * R);
            if (first) {
                ready.finish(lane, started);
                first = false;
            }
            const A mv = id >= 0 ? x[id] : load_dev(t + ~id);
            for (int j = 0; j < nc; j += 8) {
                if (j > 0) {
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        v[u] = St::ld(stream, stream_lo, p0 + static_cast<int64_t>(min(j + u, nc - 1)) * R);
                }
                if (j + 8 <= nc) {
                    a0 += v[0] * bcast(mv, j);
                    a1 += v[1] * bcast(mv, j + 1);
                    a2 += v[2] * bcast(mv, j + 2);
                    a3 += v[3] * bcast(mv, j + 3);
                    a0 += v[4] * bcast(mv, j + 4);
                    a1 += v[5] * bcast(mv, j + 5);
                    a2 += v[6] * bcast(mv, j + 6);
                    a3 += v[7] * bcast(mv, j + 7);
                } else {
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (j + u < nc) a0 += v[u] * bcast(mv, j + u);
                }
            }
        }
    } else {
        for (int c = cb; c < ce; c += 64) {
            const int nc = min(64, ce - c);
            const int32_t id = mult[c + min(lane, nc - 1)];
            const A mv = *(id >= 0 ? x + id : t + ~id);
            const int64_t p0 = T.base + static_cast<int64_t>(c - T.c0) * R + lr;
            int j = 0;
            for (; j + 8 <= nc; j += 8) {
                A v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = St::ld(stream, stream_lo, p0 + static_cast<int64_t>(j + u) * R);
                a0 += v[0] * bcast(mv, j);
                a1 += v[1] * bcast(mv, j + 1);
                a2 += v[2] * bcast(mv, j + 2);
                a3 += v[3] * bcast(mv, j + 3);
                a0 += v[4] * bcast(mv, j + 4);
                a1 += v[5] * bcast(mv, j + 5);
                a2 += v[6] * bcast(mv, j + 6);
                a3 += v[7] * bcast(mv, j + 7);
            }
            for (; j < nc; ++j) {
                const A v = St::ld(stream, stream_lo, p0 + static_cast<int64_t>(j) * R);
                a0 += v * bcast(mv, j);
            }
        }
    }
    red[wave][lane] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if (wave == 0 && act) {
        A s = red[0][lane];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += red[w][lane];
        const int64_t r = T.row0 + lane;
        y[r] = beta_zero ? alpha * s : beta * y[r] + alpha * s;
    }
}

// Pass 2: row owner over the tile's column stream.  One workgroup per tile;
// its NW wavefronts take contiguous parts of the columns, lane = row.
// Multipliers are fetched 64 at a time (one per lane) and broadcast with
// readlane; eight columns are in flight per wavefront.
template <class St, int NW>
__global__ __launch_bounds__(64 * NW) void k_rows(const TileDev *__restrict__ tiles,
                                              const int32_t *__restrict__ mult,
                                              const void *__restrict__ stream,
                                              const void *__restrict__ stream_lo,
                                              const typename St::acc *__restrict__ x,
                                              const typename St::acc *__restrict__ t,
                                              typename St::acc *__restrict__ y,
                                              typename St::acc alpha, typename St::acc beta,
                                              int beta_zero) {
    using A = typename St::acc;
    __shared__ A red[NW][64];
    rows_tile<St, NW, false>(tiles[blockIdx.x], mult, stream, stream_lo, x, t, y, alpha, beta, beta_zero, red);
}

// ---------------------------------------------------------------------------
// Several right-hand sides on the single-right-hand-side layout (any storage;
// the FP64 H*X path): panels of RM columns of X share one pass over the
// factors.  t holds RM values per rank column (t_m[q*RM + r]); columns of a
// panel past ncol load column 0 (valid addresses) and are not written.

// Projection of item w, rank columns [j0, j0 + KC): each lane keeps KC x RM
// accumulators over its rows, then one butterfly per accumulator.
template <class St, int KC>
__device__ __forceinline__ void proj_item_m(const ProjItem &w, int j0, const void *__restrict__ fac,
                                            const void *__restrict__ fac_lo,
                                            const typename St::acc *__restrict__ x, int64_t ldx, int ncol,
                                            typename St::acc *__restrict__ out, int lane) {
    using A = typename St::acc;
    constexpr int U = KC >= 8 ? 1 : 8 / KC;
    const int rows = w.rows;
    const int64_t ld = w.ld;
    const int64_t base = w.base + static_cast<int64_t>(j0) * ld;
    int64_t xo[RM];
#pragma unroll
    for (int r = 0; r < RM; ++r) xo[r] = w.xoff + static_cast<int64_t>(r < ncol ? r : 0) * ldx;
    A acc[KC][RM];
#pragma unroll
    for (int j = 0; j < KC; ++j)
#pragma unroll
        for (int r = 0; r < RM; ++r) acc[j][r] = A(0);
    for (int r0 = 0; r0 < rows; r0 += 64 * U) {
        A xv[U][RM];
        A bv[KC][U];
        int rc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int row = r0 + 64 * u + lane;
            rc[u] = min(row, rows - 1);
#pragma unroll
            for (int r = 0; r < RM; ++r) {
                const A v = x[xo[r] + rc[u]];
                xv[u][r] = row < rows ? v : A(0);
            }
        }
#pragma unroll
        for (int j = 0; j < KC; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u) bv[j][u] = St::ld(fac, fac_lo, base + j * ld + rc[u]);
#pragma unroll
        for (int j = 0; j < KC; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int r = 0; r < RM; ++r) acc[j][r] += bv[j][u] * xv[u][r];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < KC; ++j)
#pragma unroll
            for (int r = 0; r < RM; ++r) acc[j][r] += __shfl_xor(acc[j][r], o, 64);
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < KC; ++j)
#pragma unroll
            for (int r = 0; r < RM; ++r) out[(w.out + j0 + j) * RM + r] = acc[j][r];
    }
}

template <class St>
__global__ __launch_bounds__(256) void k_proj_m(const ProjItem *__restrict__ items, int64_t nitems,
                                                const void *__restrict__ fac, const void *__restrict__ fac_lo,
                                                const typename St::acc *__restrict__ x, int64_t ldx, int ncol,
                                                typename St::acc *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t it = static_cast<int64_t>(blockIdx.x) * WAVES + wave;
    if (it >= nitems) return;
    const ProjItem w = items[it];
    // columns in groups of at most PGROUP (KC x RM <= 32 accumulators per lane)
    auto group = [&](int j0, int kc) {
        switch (kc) {
            case 1: proj_item_m<St, 1>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            case 2: proj_item_m<St, 2>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            case 3: proj_item_m<St, 3>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            case 4: proj_item_m<St, 4>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            case 5: proj_item_m<St, 5>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            case 6: proj_item_m<St, 6>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            case 7: proj_item_m<St, 7>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            case 8: proj_item_m<St, 8>(w, j0, fac, fac_lo, x, ldx, ncol, out, lane); break;
            default: break;
        }
    };
    constexpr int PGROUP = 32 / RM;
    static_assert(PGROUP >= 1 && PGROUP <= 8, "group sizes have compiled cases");
    for (int j0 = 0; j0 < w.kc; j0 += PGROUP) group(j0, min(w.kc - j0, PGROUP));
}

// t_m[(tOff + j)*RM + r] = sum of the chunk partials, chunks in ascending order.
template <class A>
__global__ __launch_bounds__(256) void k_reduce_m(const RedItem *__restrict__ items, int64_t nitems,
                                                  const A *__restrict__ part, A *__restrict__ t) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t it = g / (PROJ_KMAX * RM);
    const int jr = static_cast<int>(g % (PROJ_KMAX * RM));
    const int j = jr / RM, r = jr % RM;
    if (it >= nitems) return;
    const RedItem ri = items[it];
    if (j >= ri.kc) return;
    const A *src = part + (ri.pOff + j) * RM + r;
    const int64_t stride = static_cast<int64_t>(ri.kc) * RM;
    A s = A(0);
    for (int c0 = 0; c0 < ri.nchunk; c0 += RED_ROUND) {
        A v[RED_ROUND];
#pragma unroll
        for (int u = 0; u < RED_ROUND; ++u) {
            const int c = c0 + u < ri.nchunk ? c0 + u : 0;
            v[u] = src[static_cast<int64_t>(c) * stride];
        }
#pragma unroll
        for (int u = 0; u < RED_ROUND; ++u)
            if (c0 + u < ri.nchunk) s += v[u];
    }
    t[(ri.tOff + j) * RM + r] = s;
}

// Row owner over the tile's column stream for RM right-hand sides: each
// stream value meets RM multipliers (x[id + r*ldx] or t_m[~id*RM + r]).
template <class St, int NW>
__global__ __launch_bounds__(64 * NW) void k_rows_m(const TileDev *__restrict__ tiles,
                                                    const int32_t *__restrict__ mult,
                                                    const void *__restrict__ stream,
                                                    const void *__restrict__ stream_lo,
                                                    const typename St::acc *__restrict__ x, int64_t ldx,
                                                    const typename St::acc *__restrict__ t,
                                                    typename St::acc *__restrict__ y, int64_t ldy, int ncol,
                                                    typename St::acc alpha, typename St::acc beta,
                                                    int beta_zero) {
    using A = typename St::acc;
    __shared__ A red[NW][RM][64];
    const TileDev T = tiles[blockIdx.x];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int R = T.rows;
    const int ncols = T.c1 - T.c0;
    const int per = (ncols + NW - 1) / NW;
    const int cb = T.c0 + min(ncols, wave * per);
    const int ce = T.c0 + min(ncols, (wave + 1) * per);
    A a0[RM], a1[RM];
#pragma unroll
    for (int r = 0; r < RM; ++r) a0[r] = a1[r] = A(0);
    const int lr = min(lane, R - 1);
    int64_t xo[RM];
#pragma unroll
    for (int r = 0; r < RM; ++r) xo[r] = static_cast<int64_t>(r < ncol ? r : 0) * ldx;
    for (int c = cb; c < ce; c += 64) {
        const int nc = min(64, ce - c);
        const int32_t id = mult[c + min(lane, nc - 1)];
        A mv[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r)
            mv[r] = *(id >= 0 ? x + id + xo[r] : t + static_cast<int64_t>(~id) * RM + (r < ncol ? r : 0));
        const int64_t p0 = T.base + static_cast<int64_t>(c - T.c0) * R + lr;
        int j = 0;
        for (; j + 4 <= nc; j += 4) {
            A v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = St::ld(stream, stream_lo, p0 + static_cast<int64_t>(j + u) * R);
#pragma unroll
            for (int r = 0; r < RM; ++r) {
                a0[r] += v[0] * bcast(mv[r], j);
                a1[r] += v[1] * bcast(mv[r], j + 1);
                a0[r] += v[2] * bcast(mv[r], j + 2);
                a1[r] += v[3] * bcast(mv[r], j + 3);
            }
        }
        for (; j < nc; ++j) {
            const A v = St::ld(stream, stream_lo, p0 + static_cast<int64_t>(j) * R);
#pragma unroll
            for (int r = 0; r < RM; ++r) a0[r] += v * bcast(mv[r], j);
        }
    }
#pragma unroll
    for (int r = 0; r < RM; ++r) red[wave][r][lane] = a0[r] + a1[r];
    __syncthreads();
    if (wave == 0 && lane < R) {
        const int64_t row = T.row0 + lane;
#pragma unroll
        for (int r = 0; r < RM; ++r) {
            if (r >= ncol) continue;
            A s = red[0][r][lane];
#pragma unroll
            for (int w = 1; w < NW; ++w) s += red[w][r][lane];
            A *yp = y + row + static_cast<int64_t>(r) * ldy;
            *yp = beta_zero ? alpha * s : beta * *yp + alpha * s;
        }
    }
}

// Adjoint gradients of L = 0.5*||H x - b||^2 in the exported arena layout:
// dC = res x', dA = res t' and dB = x s', with t = B'x and s = A'res taken
// from the projection passes of the two products (scaled back by the
// power-of-two balance of each rank column, exactly).  One wavefront per
// item: each lane loads its rows' vector values once, then walks the
// item's columns, reading the column multiplier with a wave-uniform (scalar)
// load and storing one 64-row strip per store instruction.  The loop issues
// no vector load, so nothing waits for the stores in flight.
template <class A>
__global__ __launch_bounds__(256) void k_grad(const GradItem *__restrict__ items, int64_t nitems,
                                              const int32_t *__restrict__ rexp,
                                              const A *__restrict__ x, const A *__restrict__ res,
                                              const A *__restrict__ t, const A *__restrict__ s,
                                              A *__restrict__ grad) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int64_t id = static_cast<int64_t>(blockIdx.x) * WAVES + wave;
    if (id >= nitems) return;
    const GradItem it = items[id];
    const A *uv = (it.kind == 2 ? x : res) + it.vec0;
    const A *wv = (it.kind == 0 ? x : (it.kind == 1 ? t : s)) + it.aux;
    const int32_t *re = rexp + it.aux;
    const int sg = it.kind == 2 ? -1 : 1;
    const int rows = it.rows;
    const int64_t ld = it.ld;
    A u[GRAD_STRIPS];
#pragma unroll
    for (int q = 0; q < GRAD_STRIPS; ++q) u[q] = uv[min(64 * q + lane, rows - 1)];   // clamped: valid
    A *g = grad + it.dst;
    for (int j = 0; j < it.ncol; ++j) {
        A w = wv[j];
        if (it.kind != 0) w = static_cast<A>(ldexp(static_cast<double>(w), sg * re[j]));
        A *gc = g + j * ld;
#pragma unroll
        for (int q = 0; q < GRAD_STRIPS; ++q) {
            const int r = 64 * q + lane;
            if (64 * q < rows && r < rows) gc[r] = u[q] * w;
        }
    }
}

// 0.5 * sum(res^2) into *out (double) in two fixed-order passes: LOSS_PARTS
// workgroups (at most) sum contiguous ranges into part[], one workgroup
// adds them.
template <class A>
__global__ __launch_bounds__(256) void k_loss_part(const A *__restrict__ res, int64_t n, int64_t per,
                                                   double *__restrict__ part) {
    __shared__ double sh[256];
    const int64_t lo = static_cast<int64_t>(blockIdx.x) * per;
    const int64_t hi = min(n, lo + per);
    double acc = 0.0;
    for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += 256 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int64_t i = i0 + 256 * u;
            v[u] = i < hi ? static_cast<double>(res[i]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u] * v[u];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_loss_final(const double *__restrict__ part, int nparts,
                                                    double *__restrict__ out) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) acc += part[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = 0.5 * sh[0];
}

// ---------------------------------------------------------------------------
// One launch for a single-right-hand-side product of a many-matrix plan:
// the workgroups take, in this order, the projection items (NW per
// workgroup), the reduction items (64 NW / PROJ_KMAX per workgroup) and the
// row tiles, with the members in decreasing order of bytes (largest first)
// inside each group.  A member's reductions wait until its projection items
// are done and its row tiles until its projection and reduction items are
// done, by counters in device memory; every item a workgroup waits for has a
// lower workgroup index, so with workgroups dispatched in index order the
// waits end.  A wait is bounded all the same (FUSED_SPIN polls): past it the
// workgroup goes on and raises *fault, so that the kernel always ends.  The
// member's last row tile clears its counters for the next product.  t and
// the partial sums cross workgroups (and XCDs) with device-scope stores and
// loads; each wavefront waits for its stores to complete, then one thread
// per workgroup adds the workgroup's items to the counters (one add per
// member).  Each member's counters are split into FUSED_SLOTS slots, one per
// XCD, each on its own 256-byte line, and a waiter polls all slots at once
// (one lane each): one line per member took one add per item from every
// XCD and serialised the whole launch on a few lines.
constexpr int FUSED_SPIN = 1 << 22;

// Lanes 0 .. FUSED_SLOTS-1 each wait until counter slot `lane` (lines of
// FUSED_STRIDE words from cnt) reaches need[lane]; the other lanes pass.
// start() issues the first check, finish() completes it and polls on.
struct SlotWait {
    const uint32_t *cnt;
    const uint32_t *need;
    uint32_t *fault;
    __device__ uint32_t start(int lane) const { return lane < FUSED_SLOTS ? load_dev(cnt + lane * FUSED_STRIDE) : 0u; }
    __device__ void finish(int lane, uint32_t seen) const {
        if (lane >= FUSED_SLOTS) return;
        const uint32_t target = need[lane];
        if (seen >= target) return;
        const uint32_t *c = cnt + lane * FUSED_STRIDE;
        for (int it = 0; it < FUSED_SPIN; ++it) {
            __builtin_amdgcn_s_sleep(2);
            if (load_dev(c) >= target) return;
        }
        store_dev(fault, 1u);
    }
    __device__ void wait(int lane) const { finish(lane, start(lane)); }
};

// One thread: for each run of equal members in mem[0 .. n) (-1: no item),
// add the run's length to that member's item counter on `slot`.
__device__ __forceinline__ void count_runs(uint32_t *counts, const int *mem, int n, int slot) {
    for (int i = 0; i < n;) {
        const int m = mem[i];
        int j = i + 1;
        while (j < n && mem[j] == m) ++j;
        if (m >= 0)
            __hip_atomic_fetch_add(counts + (m * FUSED_SLOTS + slot) * FUSED_STRIDE, static_cast<uint32_t>(j - i),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        i = j;
    }
}

// Counter lines and need entries are laid out as described with struct Fused.
template <class St, int NW>
__global__ __launch_bounds__(64 * NW) void k_fused(const ProjItem *__restrict__ proj, int64_t nproj,
                                               const RedItem *__restrict__ redi, const int32_t *__restrict__ redmem,
                                               int64_t nred, const TileDev *__restrict__ tiles,
                                               const int32_t *__restrict__ mult, const void *__restrict__ stream,
                                               const void *__restrict__ stream_lo, const void *__restrict__ fac,
                                               const void *__restrict__ fac_lo,
                                               const typename St::acc *__restrict__ x,
                                               typename St::acc *__restrict__ t, int64_t tLen,
                                               typename St::acc *__restrict__ y, typename St::acc alpha,
                                               typename St::acc beta, int beta_zero, uint32_t *__restrict__ counts,
                                               const uint32_t *__restrict__ need, int nmem,
                                               uint32_t *__restrict__ fault) {
    using A = typename St::acc;
    constexpr int S = FUSED_SLOTS;
    constexpr int RPW = 64 * NW / PROJ_KMAX;               // reduction items per workgroup
    __shared__ A red[NW][64];
    __shared__ int wm[RPW > NW ? RPW : NW];
    const int64_t npw = (nproj + NW - 1) / NW;
    const int64_t nrw = (nred + RPW - 1) / RPW;
    const int64_t b = blockIdx.x;
    const int slot = static_cast<int>(b % S);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int64_t MS = static_cast<int64_t>(nmem) * S;
    if (b < npw) {                                         // projection items, one per wavefront
        const int64_t it = b * NW + wave;
        int m = -1;
        if (it < nproj) {
            const ProjItem w = proj[it];
            static_assert(PROJ_KMAX == 16, "one case per rank-column count");
            switch (w.kc) {
#define HM_PCASE(K) case K: proj_item<St, K, true>(w, fac, fac_lo, x, t, lane); break;
                HM_PCASE(1) HM_PCASE(2) HM_PCASE(3) HM_PCASE(4) HM_PCASE(5) HM_PCASE(6) HM_PCASE(7) HM_PCASE(8)
                HM_PCASE(9) HM_PCASE(10) HM_PCASE(11) HM_PCASE(12) HM_PCASE(13) HM_PCASE(14) HM_PCASE(15)
                HM_PCASE(16)
#undef HM_PCASE
                default: break;
            }
            m = w.pad;
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wavefront's stores are done
        if (lane == 0) wm[wave] = m;
        __syncthreads();
        if (threadIdx.x == 0) count_runs(counts, wm, NW, slot);
        return;
    }
    if (b < npw + nrw) {                                   // reduction items, PROJ_KMAX lanes each
        const int g = static_cast<int>(threadIdx.x / PROJ_KMAX);
        const int j = static_cast<int>(threadIdx.x % PROJ_KMAX);
        const int64_t it = (b - npw) * RPW + g;
        int m = -1;
        if (it < nred) {
            const RedItem r = redi[it];
            m = redmem[it];
            SlotWait{counts + m * S * FUSED_STRIDE, need + MS + m * S, fault}.wait(j);   // its projections
            if (j < r.kc) {
                const A *src = t + tLen + r.pOff + j;
                A sum = A(0);
                for (int c0 = 0; c0 < r.nchunk; c0 += RED_ROUND) {
                    A v[RED_ROUND];
#pragma unroll
                    for (int u = 0; u < RED_ROUND; ++u) {
                        const int c = c0 + u < r.nchunk ? c0 + u : 0;
                        v[u] = load_dev(src + static_cast<int64_t>(c) * r.kc);
                    }
#pragma unroll
                    for (int u = 0; u < RED_ROUND; ++u)
                        if (c0 + u < r.nchunk) sum += v[u];
                }
                store_dev(t + r.tOff + j, sum);
            }
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (j == 0) wm[g] = m;
        __syncthreads();
        if (threadIdx.x == 0) count_runs(counts, wm, RPW, slot);
        return;
    }
    const TileDev T = tiles[b - npw - nrw];               // row tile
    const int m = T.pad;
    rows_tile<St, NW, true>(T, mult, stream, stream_lo, x, t, y, alpha, beta, beta_zero, red,
                            SlotWait{counts + m * S * FUSED_STRIDE, need + m * S, fault});
    if (threadIdx.x == 0) {
        // the member's last row tile (last on its slot, then last slot)
        // clears the member's counters for the next product
        uint32_t *tc = counts + (MS + m * S + slot) * FUSED_STRIDE;
        const uint32_t p1 = __hip_atomic_fetch_add(tc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p1 + 1 == need[2 * MS + m * S + slot]) {
            uint32_t *mc = counts + (2 * MS + m) * FUSED_STRIDE;
            const uint32_t p2 = __hip_atomic_fetch_add(mc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (p2 + 1 == need[3 * MS + m]) {
                for (int q = 0; q < S; ++q) {
                    store_dev(counts + (m * S + q) * FUSED_STRIDE, 0u);
                    store_dev(counts + (MS + m * S + q) * FUSED_STRIDE, 0u);
                }
                store_dev(mc, 0u);
            }
        }
    }
}

template <class St>
static hipError_t run_fused(const Plan &p, const Direction &d, const void *xv, void *yv, double alpha,
                            double beta, hipStream_t s, hipEvent_t *ev, void *scratch) {
    using A = typename St::acc;
    const A *x = static_cast<const A *>(xv);
    A *y = static_cast<A *>(yv);
    A *t = static_cast<A *>(scratch ? scratch : d.t);
    if (ev) {
        (void)hipEventRecord(ev[0], s);
        (void)hipEventRecord(ev[1], s);
    }
    const int NW = p.row_waves;
    const int64_t npw = (d.nproj + NW - 1) / NW;
    const int64_t rpw = 64 * NW / PROJ_KMAX;
    const int64_t nrw = (d.nred + rpw - 1) / rpw;
    const dim3 grid(static_cast<unsigned>(npw + nrw + d.ntiles));
    const A al = static_cast<A>(alpha), be = static_cast<A>(beta);
    const int bz = beta == 0.0 ? 1 : 0;
#define HM_FUSED(W)                                                                                             \
    hipLaunchKernelGGL((k_fused<St, W>), grid, dim3(64 * W), 0, s, d.proj, d.nproj, d.red, d.fmem.redmem, d.nred, \
                       d.tiles, d.mult, d.stream, d.stream_lo, p.fac, p.fac_lo, x, t, d.tLen, y, al, be, bz,      \
                       d.fmem.counts, d.fmem.need, d.fmem.nmem, p.fault)
    if (NW == 2) HM_FUSED(2);
    else if (NW == 8) HM_FUSED(8);
    else HM_FUSED(4);
#undef HM_FUSED
    const hipError_t e = hipGetLastError();
    if (ev) (void)hipEventRecord(ev[2], s);
    return e;
}

template <class A>
static hipError_t run_grads(const Plan &p, const void *x, const void *res, void *grad,
                            bool want_loss, hipStream_t s) {
    if (p.ngrad > 0 && grad) {
        hipLaunchKernelGGL(k_grad<A>, dim3(static_cast<unsigned>((p.ngrad + WAVES - 1) / WAVES)), dim3(64 * WAVES),
                           0, s, p.grad, p.ngrad, p.rexp, static_cast<const A *>(x), static_cast<const A *>(res),
                           static_cast<const A *>(p.fwd.t), static_cast<const A *>(p.bwd.t),
                           static_cast<A *>(grad));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (want_loss) {
        // lossbuf: [loss | LOSS_PARTS partial sums]; the split depends on n only
        double *lb = static_cast<double *>(p.lossbuf);
        const int64_t n = p.nrows;
        const int64_t parts = n > 0 ? std::min<int64_t>(LOSS_PARTS, (n + 2047) / 2048) : 0;
        if (parts > 0) {
            const int64_t per = (n + parts - 1) / parts;
            hipLaunchKernelGGL(k_loss_part<A>, dim3(static_cast<unsigned>(parts)), dim3(256), 0, s,
                               static_cast<const A *>(res), n, per, lb + 1);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, s, static_cast<const double *>(lb + 1),
                           static_cast<int>(parts), lb);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_adjoint_grads(const Plan &p, const void *x, const void *res, void *grad,
                                bool want_loss, hipStream_t s) {
    return p.storage == 0 ? run_grads<double>(p, x, res, grad, want_loss, s)
                          : run_grads<float>(p, x, res, grad, want_loss, s);
}

template <class St>
static hipError_t run(const Plan &p, const Direction &d, const void *xv, void *yv, double alpha,
                      double beta, hipStream_t s, hipEvent_t *ev, void *scratch) {
    using A = typename St::acc;
    const A *x = static_cast<const A *>(xv);
    A *y = static_cast<A *>(yv);
    A *t = static_cast<A *>(scratch ? scratch : d.t);
    hipError_t e = hipSuccess;
    if (d.nproj > 0) {
        // The scratch buffer is [t (tLen) | partials (partLen)]: items that
        // cover a whole leaf write t directly, chunked items write partials
        // (their offsets already include tLen) and k_reduce folds them.
        const unsigned g = static_cast<unsigned>((d.nproj + WAVES - 1) / WAVES);
        hipLaunchKernelGGL(k_proj<St>, dim3(g), dim3(64 * WAVES), 0, s, d.proj, d.nproj,
                           p.fac, p.fac_lo, x, t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[0], s);
    if (d.nproj > 0 && d.nred > 0) {
        const int64_t thr = d.nred * PROJ_KMAX;
        const unsigned g2 = static_cast<unsigned>((thr + 255) / 256);
        hipLaunchKernelGGL(k_reduce<A>, dim3(g2), dim3(256), 0, s, d.red, d.nred,
                           static_cast<const A *>(t) + d.tLen, t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (d.ntiles > 0) {
        const dim3 grid(static_cast<unsigned>(d.ntiles));
        const A *tt = static_cast<const A *>(t);
        const A al = static_cast<A>(alpha), be = static_cast<A>(beta);
        const int bz = beta == 0.0 ? 1 : 0;
        if (p.row_waves == 2)
            hipLaunchKernelGGL((k_rows<St, 2>), grid, dim3(128), 0, s, d.tiles, d.mult, d.stream, d.stream_lo,
                               x, tt, y, al, be, bz);
        else if (p.row_waves == 8)
            hipLaunchKernelGGL((k_rows<St, 8>), grid, dim3(512), 0, s, d.tiles, d.mult, d.stream, d.stream_lo,
                               x, tt, y, al, be, bz);
        else
            hipLaunchKernelGGL((k_rows<St, 4>), grid, dim3(256), 0, s, d.tiles, d.mult, d.stream, d.stream_lo,
                               x, tt, y, al, be, bz);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    return hipSuccess;
}

template <class St>
static hipError_t run_m(const Plan &p, const Direction &d, int ncol, const void *xv, int64_t ldx, void *yv,
                        int64_t ldy, double alpha, double beta, hipStream_t s, hipEvent_t *ev) {
    using A = typename St::acc;
    const A *x = static_cast<const A *>(xv);
    A *y = static_cast<A *>(yv);
    A *t = static_cast<A *>(d.tm);
    hipError_t e = hipSuccess;
    if (d.nproj > 0) {
        const unsigned g = static_cast<unsigned>((d.nproj + WAVES - 1) / WAVES);
        hipLaunchKernelGGL(k_proj_m<St>, dim3(g), dim3(64 * WAVES), 0, s, d.proj, d.nproj, p.fac, p.fac_lo, x, ldx,
                           ncol, t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[0], s);
    if (d.nproj > 0 && d.nred > 0) {
        const int64_t thr = d.nred * PROJ_KMAX * RM;
        hipLaunchKernelGGL(k_reduce_m<A>, dim3(static_cast<unsigned>((thr + 255) / 256)), dim3(256), 0, s, d.red,
                           d.nred, static_cast<const A *>(t) + d.tLen * RM, t);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], s);
    if (d.ntiles > 0) {
        const dim3 grid(static_cast<unsigned>(d.ntiles));
        const A al = static_cast<A>(alpha), be = static_cast<A>(beta);
        const int bz = beta == 0.0 ? 1 : 0;
        // the plan's row_waves wavefronts per tile, as in the single-column pass
#define HM_ROWS_M(W)                                                                                            \
    hipLaunchKernelGGL((k_rows_m<St, W>), grid, dim3(64 * W), 0, s, d.tiles, d.mult, d.stream, d.stream_lo, x, \
                       ldx, static_cast<const A *>(t), y, ldy, ncol, al, be, bz)
        if (p.row_waves == 2) HM_ROWS_M(2);
        else if (p.row_waves == 8) HM_ROWS_M(8);
        else HM_ROWS_M(4);
#undef HM_ROWS_M
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[2], s);
    return hipSuccess;
}

hipError_t launch_matvec_m(const Plan &p, const Direction &d, int ncol, const void *x, int64_t ldx, void *y,
                           int64_t ldy, double alpha, double beta, hipStream_t s, hipEvent_t *ev) {
    switch (p.storage) {
        case 0: return run_m<StF64>(p, d, ncol, x, ldx, y, ldy, alpha, beta, s, ev);
        case 1: return run_m<StF32>(p, d, ncol, x, ldx, y, ldy, alpha, beta, s, ev);
        case 2: return run_m<StBF16>(p, d, ncol, x, ldx, y, ldy, alpha, beta, s, ev);
        case 3: return run_m<StBF16x2>(p, d, ncol, x, ldx, y, ldy, alpha, beta, s, ev);
        default: return hipErrorInvalidValue;
    }
}

// The projection phase alone (t = B' x of every low-rank leaf into d.t).
template <class St>
static hipError_t run_proj(const Plan &p, const Direction &d, const void *xv, hipStream_t s) {
    using A = typename St::acc;
    const A *x = static_cast<const A *>(xv);
    A *t = static_cast<A *>(d.t);
    if (d.nproj == 0) return hipSuccess;
    const unsigned g = static_cast<unsigned>((d.nproj + WAVES - 1) / WAVES);
    hipLaunchKernelGGL(k_proj<St>, dim3(g), dim3(64 * WAVES), 0, s, d.proj, d.nproj, p.fac, p.fac_lo, x, t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || d.nred == 0) return e;
    const int64_t thr = d.nred * PROJ_KMAX;
    hipLaunchKernelGGL(k_reduce<A>, dim3(static_cast<unsigned>((thr + 255) / 256)), dim3(256), 0, s, d.red, d.nred,
                       static_cast<const A *>(t) + d.tLen, t);
    return hipGetLastError();
}

hipError_t launch_projection(const Plan &p, const Direction &d, const void *x, hipStream_t s) {
    switch (p.storage) {
        case 0: return run_proj<StF64>(p, d, x, s);
        case 1: return run_proj<StF32>(p, d, x, s);
        case 2: return run_proj<StBF16>(p, d, x, s);
        case 3: return run_proj<StBF16x2>(p, d, x, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_matvec(const Plan &p, const Direction &d, const void *x, void *y, double alpha,
                         double beta, hipStream_t s, hipEvent_t *ev, void *scratch) {
    if (d.fmem.nmem > 0 && d.ntiles > 0) {
        switch (p.storage) {
            case 0: return run_fused<StF64>(p, d, x, y, alpha, beta, s, ev, scratch);
            case 1: return run_fused<StF32>(p, d, x, y, alpha, beta, s, ev, scratch);
            case 2: return run_fused<StBF16>(p, d, x, y, alpha, beta, s, ev, scratch);
            case 3: return run_fused<StBF16x2>(p, d, x, y, alpha, beta, s, ev, scratch);
            default: return hipErrorInvalidValue;
        }
    }
    switch (p.storage) {
        case 0: return run<StF64>(p, d, x, y, alpha, beta, s, ev, scratch);
        case 1: return run<StF32>(p, d, x, y, alpha, beta, s, ev, scratch);
        case 2: return run<StBF16>(p, d, x, y, alpha, beta, s, ev, scratch);
        case 3: return run<StBF16x2>(p, d, x, y, alpha, beta, s, ev, scratch);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace hm
