// Fixed-order reduction of the batched projection's per-super-tile partial
// sums (float for FP32 / BF16 / split-BF16 plans, double for FP64 plans).
// Device code shared by hm_batched.hip and hm_batched64.hip.  Not part of
// the C ABI.
#pragma once

#include "hm_internal.h"

namespace hm {

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
template <class T>
__global__ __launch_bounds__(256) void kb_reduce(const BRed *__restrict__ red, int64_t nbig, int64_t nred,
                                                 const T *__restrict__ part, const T *__restrict__ zeros,
                                                 T *__restrict__ t) {
    __shared__ T sh[256];
    if (static_cast<int64_t>(blockIdx.x) < nbig) {
        const BRed r = red[blockIdx.x];
        const int per = r.len >= 256 ? 256 : r.len;    // elements per pass
        const int G = 256 / per;
        for (int e0 = 0; e0 < r.len; e0 += per) {
            const int e = e0 + static_cast<int>(threadIdx.x) % per;
            const int q = static_cast<int>(threadIdx.x) / per;
            T acc = T(0);
            if (e < r.len && q < G) {
                const T *src = part + r.pOff + e;
                for (int c0 = q; c0 < r.nchunk; c0 += RED_BATCH * G) {
                    T v[RED_BATCH];
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
                T sum = acc;
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
        const T *src = part + r.pOff + e;
        T v[RED_WAVE];
#pragma unroll
        for (int c = 0; c < RED_WAVE; ++c) v[c] = *(c < r.nchunk ? src + static_cast<int64_t>(c) * r.len : zeros);
        T acc = T(0);
#pragma unroll
        for (int c = 0; c < RED_WAVE; ++c) acc += v[c];
        t[r.tOff + e] = acc;
    }
}

}  // namespace hm
