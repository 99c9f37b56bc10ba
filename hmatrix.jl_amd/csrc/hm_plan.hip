// Plan construction and the C ABI (include/hmat_amd.h).
//
// hm_plan_create turns the exported block table into device work lists:
//   * low-rank factors go to one compact arena in the storage precision
//     (column pairs balanced by powers of two), read by the projection pass;
//   * rows are cut at every leaf boundary and then into tiles of <= 64 rows;
//     every tile's dense columns and A columns are packed into one
//     contiguous column stream (row-owner pass);
//   * the transposed product H' gets the same structures built from the
//     transposed leaves (dense leaves transposed on the host, A and B
//     swapping roles), so H' x runs at the same bandwidth as H x;
//   * low-rank leaves get projection items of <= PROJ_CHUNK rows and
//     <= PROJ_KMAX rank columns;
//   * FP32/BF16/split-BF16 plans also get the batched multi-right-hand-side
//     layout (hm_batched.hip): dense leaves in MFMA fragment order per row
//     tile, A rows per tile, B slices per 256-column super-tile.
#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>

#include "../../include/hmat_amd.h"
#include "hm_internal.h"
#include "hm_layout.h"

using namespace hm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

struct HipFail {
    hipError_t e;
    std::string where;
};

void hip_check(hipError_t e, const char *where) {
    if (e != hipSuccess) throw HipFail{e, where};
}




template <class T>
T *dev_upload(hm::Plan &p, const std::vector<T> &v) {
    if (v.empty()) return nullptr;
    void *d = nullptr;
    const size_t bytes = v.size() * sizeof(T);
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) throw HipFail{e, "hipMalloc"};
    p.allocs.push_back(d);
    p.device_bytes += static_cast<int64_t>(bytes);
    hip_check(hipMemcpy(d, v.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy");
    return static_cast<T *>(d);
}

void *dev_alloc(hm::Plan &p, size_t bytes) {
    if (bytes == 0) return nullptr;
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) throw HipFail{e, "hipMalloc"};
    p.allocs.push_back(d);
    p.device_bytes += static_cast<int64_t>(bytes);
    return d;
}








template <class T>
void *upload_vec(hm::Plan &p, const std::vector<T> &v) {
    return v.empty() ? nullptr : static_cast<void *>(dev_upload(p, v));
}



void build_batched(hm::Plan &p, hm::Direction &d, int64_t nrows, int64_t ncols,
                   const std::vector<OB> &obs, const double *fac, const double *dense) {
    using namespace hm;
    BatchedHost H;
    if (!layout_batched(p.storage, p.strip_shape, nrows, ncols, obs, fac, dense, H)) return;
    const std::string bad = check_batched(H, p.storage, nrows, ncols);
    if (!bad.empty()) throw std::logic_error(bad);
    Batched &B = d.bat;
    B.nstrips = H.nstrips;
    B.soffw = dev_upload(p, H.soffw);
    B.wpos = dev_upload(p, H.wpos);
    B.recs = dev_upload(p, H.recs);
    B.shape = H.shape;
    B.wgx = dev_upload(p, H.wgx);
    {
        std::vector<uint32_t> ul;
        for (size_t w = 0; 2 * w + 1 < H.wgx.size(); ++w)
            if (H.wgx[2 * w + 1] == 0) ul.push_back(static_cast<uint32_t>(w));
        B.nunstaged = static_cast<int64_t>(ul.size());
        B.ulist = dev_upload(p, ul);
    }
    B.pcol = H.pcol.empty() ? nullptr : reinterpret_cast<uint4 *>(dev_upload(p, H.pcol));
    B.nst = H.nst;
    B.gps = H.gps;
    B.red = dev_upload(p, H.red);
    B.nred = static_cast<int64_t>(H.red.size());
    B.nredbig = 0;
    while (B.nredbig < B.nred && H.red[static_cast<size_t>(B.nredbig)].nchunk > 32) ++B.nredbig;
    auto arena = [&](const Arena &a) {
        return a.mode == 3 ? upload_vec(p, a.d) : a.mode == 1 ? upload_vec(p, a.f) : upload_vec(p, a.h);
    };
    B.frag = arena(H.fr);
    B.vslc = arena(H.vs);
    B.tLen = H.tLen;
    B.partLen = H.partLen;
    const size_t tsz = static_cast<size_t>(H.tLen + H.partLen + NB) * acc_size(p.storage);
    B.t = dev_alloc(p, tsz);
    hip_check(hipMemset(B.t, 0, tsz), "hipMemset");
    B.ready = true;
}

// Single-right-hand-side layout of one direction (hm_layout.cpp), checked,
// then uploaded.  fac: the balanced factor arena (fac_elems elements).
void build_direction(hm::Plan &p, hm::Direction &d, int64_t nrows, int64_t ncols, const std::vector<OB> &obs,
                     const double *fac, int64_t fac_elems, const double *dense, const std::vector<int64_t> &mrow) {
    DirectionHost D;
    layout_direction(p.storage, p.proj_chunk, p.row_waves, nrows, obs, fac, dense, mrow, D);
    const std::string bad = check_direction(D, p.storage, p.row_waves, nrows, ncols, fac_elems);
    if (!bad.empty()) throw std::logic_error(bad);
    d.nrows = nrows;
    d.ncols = ncols;
    d.stream = D.shi.empty() ? nullptr : static_cast<void *>(dev_upload(p, D.shi));
    d.stream_lo = D.slo.empty() ? nullptr : static_cast<void *>(dev_upload(p, D.slo));
    d.mult = dev_upload(p, D.mult);
    d.ncolumns = D.ncolumns;
    d.blk = dev_upload(p, D.blk);
    d.nblk = static_cast<int64_t>(D.blk.size());
    d.proj = dev_upload(p, D.proj);
    d.nproj = static_cast<int64_t>(D.proj.size());
    d.red = dev_upload(p, D.red);
    d.nred = static_cast<int64_t>(D.red.size());
    d.tiles = dev_upload(p, D.tiles);
    d.ntiles = static_cast<int64_t>(D.tiles.size());
    d.tLen = D.tLen;
    d.partLen = D.partLen;
    d.t = dev_alloc(p, static_cast<size_t>(D.tLen + D.partLen) * acc_size(p.storage));
    if (D.nmem > 0) {
        d.fmem.nmem = D.nmem;
        d.fmem.need = dev_upload(p, D.need);
        d.fmem.redmem = D.red.empty() ? nullptr : dev_upload(p, D.redmem);
        const size_t cbytes = sizeof(uint32_t) * FUSED_STRIDE * (2 * FUSED_SLOTS + 1) * static_cast<size_t>(D.nmem);
        d.fmem.counts = static_cast<uint32_t *>(dev_alloc(p, cbytes));
        hip_check(hipMemset(d.fmem.counts, 0, cbytes), "hipMemset");
    }
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

void free_plan(hm::Plan *p) {
    if (!p) return;
    DeviceGuard g(p->device);
    for (hipEvent_t e : p->timing.ev) (void)hipEventDestroy(e);
    for (void *a : p->allocs) (void)hipFree(a);
    if (p->losshost) (void)hipHostFree(p->losshost);
    delete p;
}

// Wait for every recorded event set and add its three phase durations to
// the plan's sums; empties the ring.  The caller holds the plan's device.
int fold_timing(hm::Plan *p) {
    hm::Timing &t = p->timing;
    for (int64_t r = 0; r < t.used; ++r) {
        hipEvent_t *ev = t.ev.data() + 4 * r;
        const hipError_t e = hipEventSynchronize(ev[3]);
        if (e != hipSuccess) {
            t.used = 0;
            return fail(HM_EHIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
        }
        for (int k = 0; k < 3; ++k) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) == hipSuccess) t.ms[k] += ms;
        }
        ++t.count;
    }
    t.used = 0;
    return HM_OK;
}

// Byte ranges [a, a + na*es) and [b, b + nb*es) intersect (empty ranges never do).
bool overlap(const void *a, int64_t na, const void *b, int64_t nb, size_t es) {
    if (!a || !b || na <= 0 || nb <= 0) return false;
    const char *ab = static_cast<const char *>(a), *bb = static_cast<const char *>(b);
    return ab < bb + nb * es && bb < ab + na * es;
}

int check_vectors(const hm::Plan *p, const void *x, const void *y) {
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    if ((p->ncols > 0 && !x) || (p->nrows > 0 && !y)) return fail(HM_EINVAL, "NULL vector");
    return HM_OK;
}

// scratch (may be NULL, single right-hand side only): device buffer used
// instead of the direction's own t / partial-sum buffer.
int run_direction(hm::Plan *p, const hm::Direction &d, int64_t nrhs, const void *X, int64_t ldx,
                  void *Y, int64_t ldy, double alpha, double beta, void *stream,
                  void *scratch = nullptr) {
    if (nrhs < 0) return fail(HM_EINVAL, "nrhs < 0");
    if (nrhs > 0 && (ldx < d.ncols || ldy < d.nrows || ldx < 1 || ldy < 1))
        return fail(HM_EINVAL, "leading dimension smaller than the vector length");
    // the panels' element extent (nrhs - 1) * ld + n must fit a byte offset
    constexpr int64_t MAXEL = (int64_t(1) << 62) / 8;
    if (nrhs > 1 && (ldx > (MAXEL - d.ncols) / (nrhs - 1) || ldy > (MAXEL - d.nrows) / (nrhs - 1)))
        return fail(HM_EINVAL, "nrhs times the leading dimension exceeds the address range");
    if (!std::isfinite(alpha) || !std::isfinite(beta))
        return fail(HM_EINVAL, "alpha and beta must be finite");
    if (nrhs > 0 && ((d.ncols > 0 && !X) || (d.nrows > 0 && !Y)))
        return fail(HM_EINVAL, "NULL vector");
    if (nrhs > 0 && d.nrows > 0 && d.ncols > 0) {
        const size_t es = acc_size(p->storage);
        const char *xb = static_cast<const char *>(X);
        const char *yb = static_cast<const char *>(Y);
        const char *xe = xb + ((nrhs - 1) * ldx + d.ncols) * es;
        const char *ye = yb + ((nrhs - 1) * ldy + d.nrows) * es;
        if (xb < ye && yb < xe) return fail(HM_EINVAL, "x and y overlap");
    }
    DeviceGuard g(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t es = acc_size(p->storage);
    // Phase timing records one set of four events per launch sequence; none
    // while the stream is being captured into a graph, nor for products on
    // the static scratch.
    bool timed = p->timing.on && !scratch;
    if (timed) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) timed = false;
    }
    auto record = [&]() -> hipEvent_t * {
        if (!timed) return nullptr;
        if (p->timing.used == hm::TIMING_RING && fold_timing(p) != HM_OK) return nullptr;
        hipEvent_t *ev = p->timing.ev.data() + 4 * p->timing.used++;
        (void)hipEventRecord(ev[0], s);
        return ev + 1;
    };
    if (nrhs >= 2 && d.bat.ready && !(p->flags & HM_NO_BATCHED)) {
        for (int64_t c0 = 0; c0 < nrhs; c0 += hm::NB) {
            const int nc = static_cast<int>(std::min<int64_t>(hm::NB, nrhs - c0));
            hipError_t e =
                p->storage == HM_F64
                    ? hm::launch_batched64(*p, d, nc, static_cast<const double *>(X) + c0 * ldx, ldx,
                                           static_cast<double *>(Y) + c0 * ldy, ldy, alpha, beta, s, record())
                    : hm::launch_batched(*p, d, nc, static_cast<const float *>(X) + c0 * ldx, ldx,
                                         static_cast<float *>(Y) + c0 * ldy, ldy, static_cast<float>(alpha),
                                         static_cast<float>(beta), s, record());
            if (e != hipSuccess)
                return fail(HM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
        }
        return HM_OK;
    }
    if (scratch && nrhs != 1) return fail(HM_EINVAL, "the static scratch serves one right-hand side");
    if (nrhs == 1) {
        hipError_t e = hm::launch_matvec(*p, d, X, Y, alpha, beta, s, record(), scratch);
        if (e != hipSuccess)
            return fail(HM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
        return HM_OK;
    }
    // several columns on the single-right-hand-side layout: panels of
    // MULTI_RHS columns share one pass over the factors
    for (int64_t c0 = 0; c0 < nrhs; c0 += hm::MULTI_RHS) {
        const int nc = static_cast<int>(std::min<int64_t>(hm::MULTI_RHS, nrhs - c0));
        const void *xc = static_cast<const char *>(X) + c0 * ldx * es;
        void *yc = static_cast<char *>(Y) + c0 * ldy * es;
        hipError_t e = hm::launch_matvec_m(*p, d, nc, xc, ldx, yc, ldy, alpha, beta, s, record());
        if (e != hipSuccess)
            return fail(HM_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
    }
    return HM_OK;
}

}  // namespace

extern "C" {

const char *hm_last_error(void) { return g_err.c_str(); }

const char *hm_version(void) { return "hmat_amd 0.1 (gfx950)"; }

int hm_plan_create(int64_t nrows, int64_t ncols, int64_t nblocks, const int64_t *blocks,
                   const double *data, int64_t ndata, int storage, hm_plan_t *plan) {
    return hm_plan_create_ex(nrows, ncols, nblocks, blocks, data, ndata, storage, nullptr, plan);
}

}  // extern "C"

namespace {
int create_impl(int64_t nrows, int64_t ncols, int64_t nblocks, const int64_t *blocks,
                const double *data, int64_t ndata, int storage, const hm_config *cfg,
                bool check_tiling, const std::vector<int64_t> &mrow, const std::vector<int64_t> &mcol,
                hm_plan_t *plan);
}

extern "C" int hm_plan_create_ex(int64_t nrows, int64_t ncols, int64_t nblocks,
                                 const int64_t *blocks, const double *data, int64_t ndata,
                                 int storage, const hm_config *cfg, hm_plan_t *plan) {
    g_err.clear();
    return create_impl(nrows, ncols, nblocks, blocks, data, ndata, storage, cfg, true, {}, {}, plan);
}

namespace {
// check_tiling: the leaves must tile the matrix exactly once (a plan made by
// hm_plan_create_batch checks each member instead; the off-diagonal part of
// its block-diagonal matrix is zero and has no leaves).  mrow / mcol: first
// row / column of each member of such a plan and the total last (empty
// otherwise).
int create_impl(int64_t nrows, int64_t ncols, int64_t nblocks, const int64_t *blocks,
                const double *data, int64_t ndata, int storage, const hm_config *cfg,
                bool check_tiling, const std::vector<int64_t> &mrow, const std::vector<int64_t> &mcol,
                hm_plan_t *plan) {
    using namespace hm;
    if (!plan) return fail(HM_EINVAL, "plan output pointer is NULL");
    *plan = nullptr;
    hm_config c{PROJ_CHUNK, WAVES, 0, 0};
    if (cfg) {
        c = *cfg;
        if (c.proj_chunk == 0) c.proj_chunk = PROJ_CHUNK;
        if (c.row_waves == 0) c.row_waves = WAVES;
    }
    if (c.proj_chunk < 64 || c.proj_chunk > 65536 || (c.proj_chunk & 63))
        return fail(HM_EINVAL, "proj_chunk must be a multiple of 64 in [64, 65536]");
    if (c.row_waves != 2 && c.row_waves != 4 && c.row_waves != 8)
        return fail(HM_EINVAL, "row_waves must be 2, 4 or 8");
    if (c.flags & ~(HM_NO_TRANSPOSE | HM_NO_BATCHED | HM_ONE_LAUNCH)) return fail(HM_EINVAL, "unknown flags");
    if (c.strip_shape < HM_STRIPS_AUTO || c.strip_shape > HM_STRIPS_WIDE)
        return fail(HM_EINVAL, "strip_shape must be 0 (auto), 1 (narrow) or 2 (wide)");
    if (nrows < 0 || ncols < 0 || nblocks < 0 || ndata < 0)
        return fail(HM_EINVAL, "negative size");
    if (nrows >= (int64_t(1) << 31) || ncols >= (int64_t(1) << 31))
        return fail(HM_EINVAL, "matrix dimension must be < 2^31");
    if (storage < HM_F64 || storage > HM_BF16X2) return fail(HM_EINVAL, "unknown storage");
    if (nblocks > 0 && !blocks) return fail(HM_EINVAL, "blocks is NULL");
    if (ndata > 0 && !data) return fail(HM_EINVAL, "data is NULL");
    if (nblocks >= (int64_t(1) << 31)) return fail(HM_EINVAL, "too many blocks");

    std::vector<OB> fw(static_cast<size_t>(nblocks));
    std::vector<int64_t> rowCover(static_cast<size_t>(nrows) + 1, 0);
    int64_t area = 0, ndense = 0, nrk = 0, dense_elems = 0, rk_elems = 0, maxrank = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
        const int64_t *r = blocks + 8 * b;
        OB o{static_cast<int>(r[0]), r[1], r[2], r[3], r[4], r[5], r[6], r[7]};
        char msg[160];
        if (o.kind != KIND_DENSE && o.kind != KIND_RK) {
            std::snprintf(msg, sizeof msg, "block %lld: unknown kind %d", (long long)b, o.kind);
            return fail(HM_EINVAL, msg);
        }
        if (o.m < 0 || o.n < 0 || o.k < 0) {
            std::snprintf(msg, sizeof msg, "block %lld: negative size or rank", (long long)b);
            return fail(HM_EINVAL, msg);
        }
        if (o.offA < 0 || o.offB < 0) {
            std::snprintf(msg, sizeof msg, "block %lld: negative arena offset", (long long)b);
            return fail(HM_EINVAL, msg);
        }
        if (o.row0 < 0 || o.col0 < 0 || o.m > nrows || o.n > ncols || o.row0 > nrows - o.m ||
            o.col0 > ncols - o.n) {
            std::snprintf(msg, sizeof msg, "block %lld: outside the %lldx%lld matrix", (long long)b,
                          (long long)nrows, (long long)ncols);
            return fail(HM_EINVAL, msg);
        }
        // an a x b factor at offset off lies in the arena (no overflow for
        // any int64 input: the product is formed only once it is <= ndata)
        auto in_arena = [ndata](int64_t off, int64_t a, int64_t c) {
            if (a == 0 || c == 0) return off <= ndata;
            if (a > ndata / c) return false;
            return off <= ndata - a * c;
        };
        if (o.kind == KIND_DENSE) {
            if (!in_arena(o.offA, o.m, o.n)) {
                std::snprintf(msg, sizeof msg, "block %lld: dense leaf outside the arena", (long long)b);
                return fail(HM_EINVAL, msg);
            }
            ++ndense;
            dense_elems += o.m * o.n;
        } else {
            if (o.k >= (int64_t(1) << 20) || !in_arena(o.offA, o.m, o.k) || !in_arena(o.offB, o.n, o.k)) {
                std::snprintf(msg, sizeof msg, "block %lld: low-rank factors outside the arena",
                              (long long)b);
                return fail(HM_EINVAL, msg);
            }
            ++nrk;
            rk_elems += (o.m + o.n) * o.k;
            maxrank = std::max(maxrank, o.k);
        }
        area += o.m * o.n;
        rowCover[o.row0] += o.n;
        rowCover[o.row0 + o.m] -= o.n;
        fw[b] = o;
    }
    if (check_tiling && area != nrows * ncols) return fail(HM_EINVAL, "leaves do not tile the matrix");
    if (check_tiling) {
        int64_t acc = 0;
        for (int64_t i = 0; i < nrows; ++i) {
            acc += rowCover[i];
            if (acc != ncols) {
                char msg[120];
                std::snprintf(msg, sizeof msg, "row %lld is covered by %lld columns, expected %lld",
                              (long long)i, (long long)acc, (long long)ncols);
                return fail(HM_EINVAL, msg);
            }
        }
    }

    std::unique_ptr<Plan> p(new (std::nothrow) Plan());
    if (!p) return fail(HM_ENOMEM, "host allocation failed");
    if (hipGetDevice(&p->device) != hipSuccess) return fail(HM_EHIP, "no HIP device");
    p->storage = storage;
    p->proj_chunk = c.proj_chunk;
    p->row_waves = c.row_waves;
    p->strip_shape = c.strip_shape;
    p->flags = c.flags;
    p->nrows = nrows;
    p->ncols = ncols;
    p->ndense = ndense;
    p->nrk = nrk;
    p->dense_elems = dense_elems;
    p->rk_elems = rk_elems;
    p->maxrank = maxrank;
    try {
        // Low-rank factors are copied into their own compact arena (the
        // projection kernels read it on the device), with each column pair
        // balanced (exact in FP64).  Dense leaves stay in `data` on the host;
        // the kernels read them from the packed streams built below.
        Oriented O;
        orient(fw, data, rk_elems, dense_elems, O);
        const std::vector<int32_t> &rexp = O.rexp;
        p->zeros = dev_upload(*p, std::vector<uint8_t>(ZERO_BYTES, 0));
        std::vector<uint8_t> hi, lo;
        pack_storage(storage, O.rkv.data(), rk_elems, hi, lo);
        p->fac = hi.empty() ? nullptr : static_cast<void *>(dev_upload(*p, hi));
        p->fac_lo = lo.empty() ? nullptr : static_cast<void *>(dev_upload(*p, lo));
        const bool fused = mrow.size() > 2 && (c.flags & HM_ONE_LAUNCH);
        if (fused) {
            p->fault = static_cast<uint32_t *>(dev_alloc(*p, sizeof(uint32_t)));
            hip_check(hipMemset(p->fault, 0, sizeof(uint32_t)), "hipMemset");
        }
        const std::vector<int64_t> none;
        build_direction(*p, p->fwd, nrows, ncols, O.fw, O.rkv.data(), rk_elems, data, fused ? mrow : none);
        if (!(p->flags & HM_NO_TRANSPOSE))
            build_direction(*p, p->bwd, ncols, nrows, O.bw, O.rkv.data(), rk_elems, O.dT.data(),
                            fused ? mcol : none);
        // adjoint gradient items over the caller's own arena offsets
        {
            // items of up to GRAD_STRIPS 64-row strips and about 64 stores per lane
            std::vector<GradItem> gi;
            auto add = [&](int kind, int64_t dst, int64_t ld, int64_t ncol, int64_t vec0, int64_t aux) {
                const int64_t S = std::min<int64_t>(GRAD_STRIPS, (ld + 63) / 64);
                const int64_t R = 64 * S, C = std::max<int64_t>(1, 64 / S);
                for (int64_t i0 = 0; i0 < ld; i0 += R)
                    for (int64_t j0 = 0; j0 < ncol; j0 += C)
                        gi.push_back(GradItem{dst + i0 + j0 * ld, vec0 + i0, aux + j0, static_cast<int32_t>(ld),
                                              static_cast<int32_t>(std::min(R, ld - i0)),
                                              static_cast<int32_t>(std::min(C, ncol - j0)), kind});
            };
            int64_t tpos = 0;
            for (int64_t b = 0; b < nblocks; ++b) {
                const int64_t *r = blocks + 8 * b;   // caller's offsets
                const int64_t m = r[3], n = r[4], k = r[5];
                if (r[0] == KIND_DENSE) {
                    if (m > 0 && n > 0) add(0, r[6], m, n, r[1], r[2]);
                } else if (k > 0) {
                    if (m > 0) add(1, r[6], m, k, r[1], tpos);
                    if (n > 0) add(2, r[7], n, k, r[2], tpos);
                    tpos += k;
                }
            }
            p->grad = dev_upload(*p, gi);
            p->ngrad = static_cast<int64_t>(gi.size());
            p->rexp = dev_upload(*p, rexp);
            p->lossbuf = dev_alloc(*p, sizeof(double) * (1 + LOSS_PARTS));
            const hipError_t he = hipHostMalloc(reinterpret_cast<void **>(&p->losshost), sizeof(double));
            if (he != hipSuccess) {
                p->losshost = nullptr;
                throw HipFail{he, "hipHostMalloc"};
            }
            p->ndata = ndata;
        }
        if (!(p->flags & HM_NO_BATCHED)) {
            build_batched(*p, p->fwd, nrows, ncols, O.fw, O.rkv.data(), data);
            if (!(p->flags & HM_NO_TRANSPOSE))
                build_batched(*p, p->bwd, ncols, nrows, O.bw, O.rkv.data(), O.dT.data());
        }
        // scratch of the panels of MULTI_RHS columns on the single-column
        // layout: only directions whose several-column products run there
        for (hm::Direction *d : {&p->fwd, &p->bwd})
            if (!d->bat.ready && (d == &p->fwd || !(p->flags & HM_NO_TRANSPOSE)))
                d->tm = dev_alloc(*p, static_cast<size_t>(d->tLen + d->partLen) * MULTI_RHS * acc_size(storage));
    } catch (const HipFail &h) {
        const int code = (h.e == hipErrorOutOfMemory || h.e == hipErrorMemoryAllocation) ? HM_ENOMEM : HM_EHIP;
        std::string msg = h.where + ": " + hipGetErrorString(h.e);
        free_plan(p.release());
        return fail(code, msg);
    } catch (const std::bad_alloc &) {
        free_plan(p.release());
        return fail(HM_ENOMEM, "host allocation failed");
    } catch (const std::length_error &e) {
        free_plan(p.release());
        return fail(HM_EUNSUPPORTED, std::string("plan layout: ") + e.what());
    } catch (const std::exception &e) {
        free_plan(p.release());
        return fail(HM_EINVAL, std::string("plan layout: ") + e.what());
    }
    *plan = reinterpret_cast<hm_plan_t>(p.release());
    return HM_OK;
}
}  // namespace

extern "C" {

int hm_plan_destroy(hm_plan_t plan) {
    free_plan(reinterpret_cast<hm::Plan *>(plan));
    return HM_OK;
}

// enum hm_strip_shape of the forward batched row pass (0: no batched layout)
int64_t strip_shape_in_use(const hm::Plan &p) {
    const hm::Batched &b = p.fwd.bat;
    if (!b.ready) return HM_STRIPS_AUTO;
    if (p.storage == HM_F64) return HM_STRIPS_F64;
    return b.shape.nw == hm::ROW_WIDE.nw ? HM_STRIPS_WIDE : HM_STRIPS_NARROW;
}

int hm_plan_info(hm_plan_t plan, int64_t *out, int nout) {
    const hm::Plan *p = reinterpret_cast<const hm::Plan *>(plan);
    if (!p || (!out && nout > 0)) return fail(HM_EINVAL, "NULL argument");
    const int64_t es = static_cast<int64_t>(storage_size(p->storage)) * (p->storage == HM_BF16X2 ? 2 : 1);
    uint32_t faults = 0;
    if (p->fault) {
        DeviceGuard g(p->device);
        const hipError_t e = hipMemcpy(&faults, p->fault, sizeof faults, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(HM_EHIP, std::string("hm_plan_info: ") + hipGetErrorString(e));
    }
    const int64_t v[HM_INFO_NFIELDS] = {p->nrows,
                                        p->ncols,
                                        p->ndense,
                                        p->nrk,
                                        p->dense_elems,
                                        p->rk_elems,
                                        p->storage,
                                        p->device_bytes,
                                        (p->dense_elems + p->rk_elems) * es,
                                        p->maxrank,
                                        p->proj_chunk,
                                        p->row_waves,
                                        p->flags,
                                        strip_shape_in_use(*p),
                                        (p->fwd.bat.ready &&
                                         ((p->flags & HM_NO_TRANSPOSE) || p->bwd.bat.ready)) ? 1 : 0,
                                        p->fwd.fmem.nmem > 0 ? 1 : 0,
                                        static_cast<int64_t>(faults)};
    for (int i = 0; i < nout && i < HM_INFO_NFIELDS; ++i) out[i] = v[i];
    return HM_OK;
}

int hm_plan_timing(hm_plan_t plan, int enable) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    DeviceGuard g(p->device);
    hm::Timing &t = p->timing;
    if (enable && t.ev.empty()) {
        t.ev.resize(4 * static_cast<size_t>(hm::TIMING_RING), nullptr);
        for (hipEvent_t &e : t.ev) {
            if (hipEventCreate(&e) != hipSuccess) {
                for (hipEvent_t &f : t.ev)
                    if (f) (void)hipEventDestroy(f);
                t.ev.clear();
                return fail(HM_EHIP, "hipEventCreate failed");
            }
        }
    }
    if (!enable && t.used > 0) {
        const int rc = fold_timing(p);
        if (rc) return rc;
    }
    t.on = enable != 0;
    return HM_OK;
}

int hm_plan_timing_read(hm_plan_t plan, double *ms, int nms, int64_t *count) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!p || (!ms && nms > 0)) return fail(HM_EINVAL, "NULL argument");
    DeviceGuard g(p->device);
    const int rc = fold_timing(p);
    if (rc) return rc;
    hm::Timing &t = p->timing;
    for (int k = 0; k < nms && k < 3; ++k) ms[k] = t.ms[k];
    if (count) *count = t.count;
    t.ms[0] = t.ms[1] = t.ms[2] = 0.0;
    t.count = 0;
    return HM_OK;
}

int hm_matvec(hm_plan_t plan, const void *x, void *y, double alpha, double beta, void *stream) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    int rc = check_vectors(p, x, y);
    if (rc) return rc;
    return run_direction(p, p->fwd, 1, x, p->ncols > 0 ? p->ncols : 1, y,
                         p->nrows > 0 ? p->nrows : 1, alpha, beta, stream);
}

}  // extern "C"

namespace {
// Scratch of hm_matvec_static: the t vector and chunk partials of one
// single-right-hand-side product.  Every device that runs the library holds
// its own copy of this array until the process ends.  Host code never uses
// the name `scratch` as an address: static_scratch() takes the device
// address once per device with hipGetSymbolAddress and every launch passes
// that pointer.
__device__ __align__(256) unsigned char scratch[HM_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on device dev (current on entry).
int static_scratch(int dev, void **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(HM_EUNSUPPORTED, "device index past the static scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(HM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return HM_OK;
}
}  // namespace

extern "C" {

int hm_matvec_static(hm_plan_t plan, const void *x, void *y, double alpha, double beta, void *stream) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    int rc = check_vectors(p, x, y);
    if (rc) return rc;
    const hm::Direction &d = p->fwd;
    const int64_t need = (d.tLen + d.partLen) * static_cast<int64_t>(acc_size(p->storage));
    if (need > static_cast<int64_t>(HM_STATIC_SCRATCH_BYTES)) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "hm_matvec_static: this plan needs %lld bytes of scratch, more than the static "
                      "maximum of %lld (HM_STATIC_SCRATCH_BYTES); use hm_matvec",
                      (long long)need, (long long)HM_STATIC_SCRATCH_BYTES);
        return fail(HM_EUNSUPPORTED, msg);
    }
    void *ptr = nullptr;
    {
        DeviceGuard g(p->device);
        rc = static_scratch(p->device, &ptr);
        if (rc) return rc;
    }
    return run_direction(p, d, 1, x, p->ncols > 0 ? p->ncols : 1, y, p->nrows > 0 ? p->nrows : 1, alpha,
                         beta, stream, ptr);
}

int hm_matmat(hm_plan_t plan, int64_t nrhs, const void *X, int64_t ldx, void *Y, int64_t ldy,
              double alpha, double beta, void *stream) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    return run_direction(p, p->fwd, nrhs, X, ldx, Y, ldy, alpha, beta, stream);
}

int hm_matvec_t(hm_plan_t plan, const void *x, void *y, double alpha, double beta, void *stream) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    if (p->flags & HM_NO_TRANSPOSE) return fail(HM_EUNSUPPORTED, "plan was created with HM_NO_TRANSPOSE");
    if ((p->nrows > 0 && !x) || (p->ncols > 0 && !y)) return fail(HM_EINVAL, "NULL vector");
    return run_direction(p, p->bwd, 1, x, p->nrows > 0 ? p->nrows : 1, y,
                         p->ncols > 0 ? p->ncols : 1, alpha, beta, stream);
}

int hm_matmat_t(hm_plan_t plan, int64_t nrhs, const void *X, int64_t ldx, void *Y, int64_t ldy,
                double alpha, double beta, void *stream) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    if (p->flags & HM_NO_TRANSPOSE) return fail(HM_EUNSUPPORTED, "plan was created with HM_NO_TRANSPOSE");
    return run_direction(p, p->bwd, nrhs, X, ldx, Y, ldy, alpha, beta, stream);
}

int hm_adjoint(hm_plan_t plan, const void *x, const void *b, void *res, void *dx, void *grad,
               double *loss, void *stream) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    if (p->flags & HM_NO_TRANSPOSE) return fail(HM_EUNSUPPORTED, "plan was created with HM_NO_TRANSPOSE");
    if ((p->ncols > 0 && (!x || !dx)) || (p->nrows > 0 && (!b || !res)))
        return fail(HM_EINVAL, "NULL vector");
    const size_t es = acc_size(p->storage);
    const int64_t m = p->nrows, n = p->ncols;
    if (overlap(res, m, b, m, es) || overlap(res, m, x, n, es) || overlap(dx, n, x, n, es) ||
        overlap(dx, n, res, m, es) || overlap(dx, n, b, m, es))
        return fail(HM_EINVAL, "res and dx must not overlap the inputs or each other");
    if (grad && (overlap(grad, p->ndata, x, n, es) || overlap(grad, p->ndata, b, m, es) ||
                 overlap(grad, p->ndata, res, m, es) || overlap(grad, p->ndata, dx, n, es)))
        return fail(HM_EINVAL, "grad must not overlap the vectors");
    DeviceGuard g(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    if (p->nrows > 0) e = hipMemcpyAsync(res, b, static_cast<size_t>(p->nrows) * es, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hm::launch_matvec(*p, p->fwd, x, res, 1.0, -1.0, s);
    if (e == hipSuccess) e = hm::launch_matvec(*p, p->bwd, res, dx, 1.0, 0.0, s);
    if (e == hipSuccess) e = hm::launch_adjoint_grads(*p, x, res, grad, loss != nullptr, s);
    if (e == hipSuccess && loss) {
        // read back through the plan's pinned double (a pageable target would
        // take a staged copy)
        e = hipMemcpyAsync(p->losshost, p->lossbuf, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        *loss = e == hipSuccess ? *p->losshost : 0.0;
    }
    if (e != hipSuccess) return fail(HM_EHIP, std::string("hm_adjoint: ") + hipGetErrorString(e));
    return HM_OK;
}

int hm_vjp(hm_plan_t plan, const void *x, const void *g, void *dx, void *grad, void *stream) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    if (p->flags & HM_NO_TRANSPOSE) return fail(HM_EUNSUPPORTED, "plan was created with HM_NO_TRANSPOSE");
    if ((p->ncols > 0 && (!x || !dx)) || (p->nrows > 0 && !g)) return fail(HM_EINVAL, "NULL vector");
    const size_t es = acc_size(p->storage);
    const int64_t m = p->nrows, n = p->ncols;
    if (overlap(dx, n, x, n, es) || overlap(dx, n, g, m, es))
        return fail(HM_EINVAL, "dx must not overlap x or g");
    if (grad && (overlap(grad, p->ndata, x, n, es) || overlap(grad, p->ndata, g, m, es) ||
                 overlap(grad, p->ndata, dx, n, es)))
        return fail(HM_EINVAL, "grad must not overlap the vectors");
    DeviceGuard gd(p->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // t = B' x (fwd.t), dx = H' g (leaves s = A' g in bwd.t), then the
    // gradients of <g, H x> from x, g, t and s
    hipError_t e = hm::launch_projection(*p, p->fwd, x, s);
    if (e == hipSuccess) e = hm::launch_matvec(*p, p->bwd, g, dx, 1.0, 0.0, s);
    if (e == hipSuccess && grad) e = hm::launch_adjoint_grads(*p, x, g, grad, false, s);
    if (e != hipSuccess) return fail(HM_EHIP, std::string("hm_vjp: ") + hipGetErrorString(e));
    return HM_OK;
}

int hm_plan_create_batch(int64_t nmat, const int64_t *nrows, const int64_t *ncols,
                         const int64_t *nblocks, const int64_t *const *blocks,
                         const double *const *data, const int64_t *ndata, int storage,
                         const hm_config *cfg, hm_plan_t *plan) {
    g_err.clear();
    if (!plan) return fail(HM_EINVAL, "plan output pointer is NULL");
    *plan = nullptr;
    if (nmat <= 0 || !nrows || !ncols || !nblocks || !blocks || !data || !ndata)
        return fail(HM_EINVAL, "NULL argument or nmat <= 0");
    int64_t R = 0, C = 0, NB_ = 0, ND = 0;
    constexpr int64_t LIM = int64_t(1) << 31;
    for (int64_t i = 0; i < nmat; ++i) {
        if (nrows[i] < 0 || ncols[i] < 0 || nblocks[i] < 0 || ndata[i] < 0)
            return fail(HM_EINVAL, "negative size in batch member");
        if ((nblocks[i] > 0 && !blocks[i]) || (ndata[i] > 0 && !data[i]))
            return fail(HM_EINVAL, "NULL array in batch member");
        // the concatenated plan must satisfy hm_plan_create's limits (checked
        // before any member's row table is allocated)
        if (nrows[i] >= LIM - R || ncols[i] >= LIM - C || nblocks[i] >= LIM - NB_ ||
            ndata[i] > std::numeric_limits<int64_t>::max() - ND)
            return fail(HM_EINVAL, "batch too large (rows and columns must stay below 2^31)");
        R += nrows[i];
        C += ncols[i];
        NB_ += nblocks[i];
        ND += ndata[i];
    }
    // each member must be a valid plan input on its own (exact tiling, offsets
    // inside its arena); the check runs on the host only
    for (int64_t i = 0; i < nmat; ++i) {
        int64_t area = 0;
        std::vector<int64_t> cover(static_cast<size_t>(nrows[i]) + 1, 0);
        for (int64_t b = 0; b < nblocks[i]; ++b) {
            const int64_t *r = blocks[i] + 8 * b;
            const int64_t nd = ndata[i];
            // an a x c factor at offset off inside the member's arena, with no
            // overflow for any int64 input
            auto in_arena = [nd](int64_t off, int64_t a, int64_t c) {
                if (a == 0 || c == 0) return off <= nd;
                if (a > nd / c) return false;
                return off <= nd - a * c;
            };
            const bool bad = (r[0] != 0 && r[0] != 1) || r[1] < 0 || r[2] < 0 || r[3] < 0 || r[4] < 0 || r[5] < 0 ||
                             r[6] < 0 || r[7] < 0 || r[3] > nrows[i] || r[4] > ncols[i] || r[1] > nrows[i] - r[3] ||
                             r[2] > ncols[i] - r[4] || (r[0] == 1 && r[5] >= (int64_t(1) << 20)) ||
                             (r[0] == 0 ? !in_arena(r[6], r[3], r[4])
                                        : (!in_arena(r[6], r[3], r[5]) || !in_arena(r[7], r[4], r[5])));
            if (bad) {
                char msg[120];
                std::snprintf(msg, sizeof msg, "batch member %lld, block %lld: invalid leaf", (long long)i,
                              (long long)b);
                return fail(HM_EINVAL, msg);
            }
            area += r[3] * r[4];
            cover[r[1]] += r[4];
            cover[r[1] + r[3]] -= r[4];
        }
        int64_t acc = 0;
        bool rows_ok = true;
        for (int64_t k = 0; k < nrows[i]; ++k) {
            acc += cover[k];
            if (acc != ncols[i]) rows_ok = false;
        }
        if (area != nrows[i] * ncols[i] || !rows_ok) {
            char msg[100];
            std::snprintf(msg, sizeof msg, "batch member %lld: leaves do not tile the matrix", (long long)i);
            return fail(HM_EINVAL, msg);
        }
    }
    std::vector<int64_t> bt;
    std::vector<double> dt;
    try {
        bt.resize(static_cast<size_t>(8 * NB_));
        dt.resize(static_cast<size_t>(ND));
    } catch (const std::bad_alloc &) {
        return fail(HM_ENOMEM, "host allocation failed");
    }
    int64_t r0 = 0, c0 = 0, q = 0, d0 = 0;
    std::vector<int64_t> mrow, mcol;
    for (int64_t i = 0; i < nmat; ++i) {
        mrow.push_back(r0);
        mcol.push_back(c0);
        for (int64_t b = 0; b < nblocks[i]; ++b, ++q) {
            const int64_t *src = blocks[i] + 8 * b;
            int64_t *dst = bt.data() + 8 * q;
            for (int f = 0; f < 8; ++f) dst[f] = src[f];
            if (src[1] < 0 || src[2] < 0 || src[1] + src[3] > nrows[i] || src[2] + src[4] > ncols[i]) {
                char msg[120];
                std::snprintf(msg, sizeof msg, "batch member %lld, block %lld: outside its matrix",
                              (long long)i, (long long)b);
                return fail(HM_EINVAL, msg);
            }
            dst[1] += r0;
            dst[2] += c0;
            dst[6] += d0;
            dst[7] += d0;
        }
        if (ndata[i] > 0) std::memcpy(dt.data() + d0, data[i], sizeof(double) * ndata[i]);
        r0 += nrows[i];
        c0 += ncols[i];
        d0 += ndata[i];
    }
    mrow.push_back(R);
    mcol.push_back(C);
    return create_impl(R, C, NB_, bt.data(), dt.data(), ND, storage, cfg, false, mrow, mcol, plan);
}

struct hm_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
};

int hm_graph_create(hm_plan_t plan, int op, int64_t nrhs, const void *X, int64_t ldx, void *Y,
                    int64_t ldy, double alpha, double beta, hm_graph_t *graph) {
    hm::Plan *p = reinterpret_cast<hm::Plan *>(plan);
    if (!graph) return fail(HM_EINVAL, "graph output pointer is NULL");
    *graph = nullptr;
    if (!p) return fail(HM_EINVAL, "plan is NULL");
    if (op != HM_OP_N && op != HM_OP_T) return fail(HM_EINVAL, "op must be HM_OP_N or HM_OP_T");
    if (op == HM_OP_T && (p->flags & HM_NO_TRANSPOSE))
        return fail(HM_EUNSUPPORTED, "plan was created with HM_NO_TRANSPOSE");
    DeviceGuard gd(p->device);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return fail(HM_EHIP, "hipStreamCreate failed");
    std::unique_ptr<hm_graph_s> g(new (std::nothrow) hm_graph_s());
    if (!g) {
        (void)hipStreamDestroy(s);
        return fail(HM_ENOMEM, "host allocation failed");
    }
    g->device = p->device;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    int rc = HM_OK;
    if (e == hipSuccess) {
        rc = run_direction(p, op == HM_OP_N ? p->fwd : p->bwd, nrhs, X, ldx, Y, ldy, alpha, beta, s);
        hipGraph_t gr = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s, &gr);
        if (rc == HM_OK && e2 != hipSuccess) {
            rc = fail(HM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
        }
        g->graph = gr;
        if (rc == HM_OK) {
            e = hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0);
            if (e != hipSuccess) rc = fail(HM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
    } else {
        rc = fail(HM_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    (void)hipStreamDestroy(s);
    if (rc != HM_OK) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        return rc;
    }
    *graph = g.release();
    return HM_OK;
}

int hm_autotune(int64_t nrows, int64_t ncols, int64_t nblocks, const int64_t *blocks,
                const double *data, int64_t ndata, int storage, int64_t nrhs, int reps,
                hm_config *best, double *best_ms) {
    if (!best) return fail(HM_EINVAL, "best is NULL");
    if (reps < 1) return fail(HM_EINVAL, "reps must be >= 1");
    if (nrhs < 1 || nrhs > hm::NB) return fail(HM_EINVAL, "nrhs must be in [1, 32]");
    if (storage < HM_F64 || storage > HM_BF16X2) return fail(HM_EINVAL, "unknown storage");
    std::vector<hm_config> cand;
    if (nrhs == 1) {
        for (int c : {256, 512, 1024, 2048, 4096})
            for (int w : {2, 4, 8}) cand.push_back(hm_config{c, w, HM_NO_TRANSPOSE | HM_NO_BATCHED, 0});
    } else if (storage == HM_F64) {
        // the f64 MFMA layout (one shape), or panels of MULTI_RHS columns on
        // the single-column layout
        cand.push_back(hm_config{0, 0, HM_NO_TRANSPOSE, 0});
        for (int c : {256, 1024, 4096})
            for (int w : {2, 4, 8}) cand.push_back(hm_config{c, w, HM_NO_TRANSPOSE | HM_NO_BATCHED, 0});
    } else {
        for (int sh : {HM_STRIPS_NARROW, HM_STRIPS_WIDE})
            cand.push_back(hm_config{0, 0, HM_NO_TRANSPOSE, sh});
    }
    const size_t es = acc_size(storage);
    const size_t xb = static_cast<size_t>(std::max<int64_t>(ncols, 1) * nrhs) * es;
    const size_t yb = static_cast<size_t>(std::max<int64_t>(nrows, 1) * nrhs) * es;
    void *X = nullptr, *Y = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = HM_OK;
    double bms = 0.0;
    hm_config bcfg{};
    auto hip = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == HM_OK)
            rc = fail(e == hipErrorOutOfMemory ? HM_ENOMEM : HM_EHIP, std::string(what) + ": " + hipGetErrorString(e));
        return rc == HM_OK;
    };
    if (hip(hipMalloc(&X, xb), "hipMalloc") && hip(hipMalloc(&Y, yb), "hipMalloc") &&
        hip(hipMemset(X, 0, xb), "hipMemset") && hip(hipStreamCreate(&s), "hipStreamCreate") &&
        hip(hipEventCreate(&e0), "hipEventCreate") && hip(hipEventCreate(&e1), "hipEventCreate")) {
        for (const hm_config &c : cand) {
            hm_plan_t p = nullptr;
            rc = hm_plan_create_ex(nrows, ncols, nblocks, blocks, data, ndata, storage, &c, &p);
            if (rc != HM_OK) break;
            auto run = [&]() {
                return nrhs == 1 ? hm_matvec(p, X, Y, 1.0, 0.0, s)
                                 : hm_matmat(p, nrhs, X, std::max<int64_t>(ncols, 1), Y,
                                             std::max<int64_t>(nrows, 1), 1.0, 0.0, s);
            };
            for (int i = 0; i < 3 && rc == HM_OK; ++i) rc = run();
            if (rc == HM_OK && hip(hipEventRecord(e0, s), "hipEventRecord")) {
                for (int i = 0; i < reps && rc == HM_OK; ++i) rc = run();
                float ms = 0.0f;
                if (rc == HM_OK && hip(hipEventRecord(e1, s), "hipEventRecord") &&
                    hip(hipEventSynchronize(e1), "hipEventSynchronize") &&
                    hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) {
                    const double per = static_cast<double>(ms) / reps;
                    if (bms == 0.0 || per < bms) {
                        bms = per;
                        bcfg = c;
                    }
                }
            }
            hm_plan_destroy(p);
            if (rc != HM_OK) break;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (X) (void)hipFree(X);
    if (Y) (void)hipFree(Y);
    if (rc != HM_OK) return rc;
    // HM_NO_BATCHED stays when the single-column layout won a
    // several-column product; the single-column timings never need it
    bcfg.flags &= nrhs > 1 ? HM_NO_BATCHED : 0;
    *best = bcfg;
    if (best_ms) *best_ms = bms;
    return HM_OK;
}

int hm_graph_launch(hm_graph_t graph, void *stream) {
    if (!graph) return fail(HM_EINVAL, "graph is NULL");
    DeviceGuard gd(graph->device);
    const hipError_t e = hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail(HM_EHIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return HM_OK;
}

int hm_graph_destroy(hm_graph_t graph) {
    if (!graph) return HM_OK;
    DeviceGuard gd(graph->device);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return HM_OK;
}

}  // extern "C"
