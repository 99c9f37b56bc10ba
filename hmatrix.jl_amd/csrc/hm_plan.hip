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

struct OB {  // oriented leaf
    int kind;
    int64_t row0, col0, m, n, k, offA, offB;
};

uint16_t f32_to_bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: keep class
        uint16_t h = static_cast<uint16_t>(u >> 16);
        if (u & 0x007fffffu) h |= 0x0040;
        return h;
    }
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return static_cast<uint16_t>(u >> 16);
}

float bf16_to_f32_host(uint16_t h) {
    uint32_t u = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
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

size_t storage_size(int st) { return st == HM_F64 ? 8 : st == HM_F32 ? 4 : 2; }
size_t acc_size(int st) { return st == HM_F64 ? 8 : 4; }

// Convert a double array to the storage arena(s) on the host.
void pack_storage(int st, const double *src, int64_t n, std::vector<uint8_t> &hi,
                  std::vector<uint8_t> &lo) {
    hi.resize(static_cast<size_t>(n) * storage_size(st));
    if (st == HM_F64) {
        std::memcpy(hi.data(), src, static_cast<size_t>(n) * 8);
    } else if (st == HM_F32) {
        float *o = reinterpret_cast<float *>(hi.data());
        for (int64_t i = 0; i < n; ++i) o[i] = static_cast<float>(src[i]);
    } else {
        uint16_t *o = reinterpret_cast<uint16_t *>(hi.data());
        uint16_t *ol = nullptr;
        if (st == HM_BF16X2) {
            lo.resize(static_cast<size_t>(n) * 2);
            ol = reinterpret_cast<uint16_t *>(lo.data());
        }
        for (int64_t i = 0; i < n; ++i) {
            const uint16_t h = f32_to_bf16_rne(static_cast<float>(src[i]));
            o[i] = h;
            if (ol) ol[i] = f32_to_bf16_rne(static_cast<float>(src[i] - bf16_to_f32_host(h)));
        }
    }
}


// Row tiles: cut at every leaf row boundary, then into <= TILE_ROWS pieces.
// Row tiles of at most TILE_ROWS rows, cut at every leaf's first and end
// row and at the rows in `extra` (member boundaries of a many-matrix plan).
std::vector<std::pair<int64_t, int32_t>> make_tiles(int64_t nrows, const std::vector<OB> &obs,
                                                    const std::vector<int64_t> &extra = {}) {
    std::vector<int64_t> cuts(extra.begin(), extra.end());
    cuts.reserve(2 * obs.size() + 2 + extra.size());
    cuts.push_back(0);
    cuts.push_back(nrows);
    for (const OB &o : obs) {
        cuts.push_back(o.row0);
        cuts.push_back(o.row0 + o.m);
    }
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    std::vector<std::pair<int64_t, int32_t>> tiles;
    for (size_t i = 0; i + 1 < cuts.size(); ++i)
        for (int64_t r = cuts[i]; r < cuts[i + 1]; r += hm::TILE_ROWS)
            tiles.emplace_back(r, static_cast<int32_t>(std::min<int64_t>(hm::TILE_ROWS, cuts[i + 1] - r)));
    return tiles;
}

bool contributes(const OB &o) { return o.m > 0 && o.n > 0 && (o.kind == hm::KIND_DENSE || o.k > 0); }

// Scale each low-rank column pair (A[:, l], B[:, l]) by (2^e, 2^-e) so that
// both columns have the same magnitude; exact in FP64, it keeps FP32/BF16
// copies of factors such as exp(-|x-y|) cross approximations (A ~ e^{+100},
// B ~ e^{-200}) inside the representable range.
// Returns the exponent e of every rank column of the leaves with k > 0, in
// table order (the order of the t vectors).
std::vector<int32_t> balance_factors(const std::vector<OB> &obs, double *d) {
    std::vector<int32_t> ex;
    for (const OB &o : obs) {
        if (o.kind != hm::KIND_RK || o.k == 0) continue;
        for (int64_t l = 0; l < o.k; ++l) {
            double *A = d + o.offA + l * o.m;
            double *B = d + o.offB + l * o.n;
            double ma = 0.0, mb = 0.0;
            for (int64_t i = 0; i < o.m; ++i) ma = std::max(ma, std::fabs(A[i]));
            for (int64_t j = 0; j < o.n; ++j) mb = std::max(mb, std::fabs(B[j]));
            int e = 0;
            if (ma > 0.0 && mb > 0.0 && std::isfinite(ma) && std::isfinite(mb))
                e = static_cast<int>(std::lround(0.5 * std::log2(mb / ma)));
            ex.push_back(e);
            if (e == 0) continue;
            for (int64_t i = 0; i < o.m; ++i) A[i] = std::ldexp(A[i], e);
            for (int64_t j = 0; j < o.n; ++j) B[j] = std::ldexp(B[j], -e);
        }
    }
    return ex;
}

// Append v converted to the batched element type (float, or bf16 planes).
struct Arena {
    int mode = 0;                 // 1 f32, 2 bf16, 3 bf16x2 (fragments only)
    std::vector<float> f;
    std::vector<uint16_t> h;
    int64_t size() const { return mode == 1 ? static_cast<int64_t>(f.size()) : static_cast<int64_t>(h.size()); }
    void align(int64_t a) {
        while (size() % a) push(0.0);
    }
    void push(double v) {
        if (mode == 1) f.push_back(static_cast<float>(v));
        else h.push_back(f32_to_bf16_rne(static_cast<float>(v)));
    }
    void push_lo(double v) {       // low half of the split of v
        const uint16_t hi = f32_to_bf16_rne(static_cast<float>(v));
        h.push_back(f32_to_bf16_rne(static_cast<float>(v - bf16_to_f32_host(hi))));
    }
};

template <class T>
void *upload_vec(hm::Plan &p, const std::vector<T> &v) {
    return v.empty() ? nullptr : static_cast<void *>(dev_upload(p, v));
}

// Host image of the batched layout of one direction.
struct BatchedHost {
    int64_t nstrips = 0;
    hm::RowShape shape = hm::ROW_NARROW;
    std::vector<uint32_t> soff, recs;
    std::vector<int32_t> wgx;
    std::vector<uint32_t> pcol;        // projection column descriptors, 4 words each
    int64_t nst = 0, gps = 0;
    std::vector<hm::BRed> red;
    Arena fr, vs;
    int64_t tLen = 0, partLen = 0;
};

// Batched (multi-RHS) layout of one direction for storage `mode` (1, 2, 3).
// rk factors come from `fac` (balanced FP64 arena), dense leaves from
// `dense` at each leaf's offA.  Returns false when the matrix does not fit
// the layout (the product then runs one right-hand side at a time).
bool layout_batched(int mode, int strip_shape, int64_t nrows, int64_t ncols, const std::vector<OB> &obs,
                    const double *fac, const double *dense, BatchedHost &H) {
    using namespace hm;
    // t offsets
    std::vector<int64_t> tOff(obs.size(), -1);
    int64_t tLen = 0;
    for (size_t b = 0; b < obs.size(); ++b)
        if (obs[b].kind == KIND_RK && contributes(obs[b])) {
            tOff[b] = tLen;
            tLen += obs[b].k * NB;
        }
    // chunked projections: partial sums of leaves wider than one super-tile
    std::vector<int64_t> pOff(obs.size(), -1), nch(obs.size(), 0);
    int64_t partLen = 0;
    std::vector<BRed> red;
    for (size_t b = 0; b < obs.size(); ++b) {
        const OB &o = obs[b];
        if (o.kind != KIND_RK || !contributes(o)) continue;
        nch[b] = (o.col0 + o.n - 1) / PROJ_SUPER - o.col0 / PROJ_SUPER + 1;
        if (nch[b] > 1) {
            pOff[b] = partLen;
            partLen += nch[b] * o.k * NB;
            red.push_back(BRed{pOff[b], tOff[b], static_cast<int32_t>(nch[b]), static_cast<int32_t>(o.k * NB)});
        }
    }
    // ---- row pass: strips of SROWS rows, K-steps of 32 columns
    // per-column path only: offsets past 2^31, or fewer than 8 columns (an
    // octet of X rows must fit in X)
    if (nrows >= (int64_t(1) << 31) || ncols >= (int64_t(1) << 31) || ncols < 8) return false;
    const int64_t nstrips = (nrows + SROWS - 1) / SROWS;
    struct Piece {            // rows [roff, roff + rows) of a strip x columns [c0, c0 + n)
        int32_t roff, rows;
        int64_t c0, n;
        std::vector<std::pair<int32_t, int64_t>> parts;   // (leaf, first column of the leaf part)
    };
    std::vector<std::vector<Piece>> dp(static_cast<size_t>(nstrips));
    std::vector<std::vector<int32_t>> rp(static_cast<size_t>(nstrips));     // low-rank leaves
    for (size_t b = 0; b < obs.size(); ++b) {
        const OB &o = obs[b];
        if (!contributes(o)) continue;
        for (int64_t t = o.row0 / SROWS; t * SROWS < o.row0 + o.m; ++t) {
            if (o.kind == KIND_RK) {
                rp[t].push_back(static_cast<int32_t>(b));
                continue;
            }
            const int64_t r0 = std::max(o.row0, t * SROWS), r1 = std::min(o.row0 + o.m, (t + 1) * SROWS);
            Piece pc;
            pc.roff = static_cast<int32_t>(r0 - t * SROWS);
            pc.rows = static_cast<int32_t>(r1 - r0);
            pc.c0 = o.col0;
            pc.n = o.n;
            pc.parts.emplace_back(static_cast<int32_t>(b), 0);
            dp[t].push_back(std::move(pc));
        }
    }
    Arena fr;
    fr.mode = mode == 1 ? 1 : 2;
    const int planes = mode == 3 ? 2 : 1;
    const int64_t lane_elems = 8;                  // values per lane and octet
    const int64_t unit = mode == 1 ? 4 : 8;        // arena elements per 16 bytes
    const int64_t zrow = (tLen + partLen) / NB;   // zero row of t (after the partials)
    std::vector<uint32_t> recs, soff(static_cast<size_t>(nstrips));
    // Pack one K-step of A fragments: rows [roff, roff + rows) of the strip,
    // columns [32 s, 32 s + 32) of a block whose value is val(r, kk), klen
    // valid columns.  Returns the offset in 16-byte units.
    auto pack = [&](int rows, int s, int64_t klen, int nOct,
                    const std::function<double(int, int64_t)> &val) -> uint32_t {
        fr.align(unit);
        const int64_t at = fr.size() / unit;
        if (at >= (int64_t(1) << 32)) throw std::length_error("fragment arena exceeds 64 GiB");
        for (int plane = 0; plane < planes; ++plane)
            for (int g = 0; g < nOct; ++g)
                for (int r = 0; r < rows; ++r)
                    for (int e = 0; e < lane_elems; ++e) {
                        const int64_t kk = 32 * s + 8 * g + e;
                        const double v = kk < klen ? val(r, kk) : 0.0;
                        if (plane == 0) fr.push(v);
                        else fr.push_lo(v);
                    }
        return static_cast<uint32_t>(at);
    };
    // merge pieces with equal rows whose columns continue each other
    std::vector<std::vector<Piece>> merged(static_cast<size_t>(nstrips));
    for (int64_t t = 0; t < nstrips; ++t) {
        std::vector<Piece> &ps = dp[t];
        std::sort(ps.begin(), ps.end(), [](const Piece &x, const Piece &y) {
            return std::tie(x.roff, x.rows, x.c0) < std::tie(y.roff, y.rows, y.c0);
        });
        std::vector<Piece> &mg = merged[t];
        for (Piece &pc : ps) {
            if (!mg.empty() && mg.back().roff == pc.roff && mg.back().rows == pc.rows &&
                mg.back().c0 + mg.back().n == pc.c0) {
                mg.back().parts.emplace_back(pc.parts[0].first, mg.back().n);
                mg.back().n += pc.n;
            } else {
                mg.push_back(std::move(pc));
            }
        }
        std::vector<Piece>().swap(ps);
    }
    // Band X windows: the strips of a band read their dense pieces' X rows
    // from one window [cmin, cmin + W) staged in LDS when the window (pieces
    // from a multiple of 8 to their last whole octet) is at most xsm rows
    // wide; otherwise (W = 0) straight from global memory.  Returns the
    // number of bands with dense pieces and no window.
    std::vector<int32_t> wgx;
    auto windows = [&](const RowShape &sh) {
        const int64_t band = static_cast<int64_t>(sh.nw) * sh.spw;
        const int64_t nwg = (nstrips + band - 1) / band;
        wgx.assign(static_cast<size_t>(2 * nwg), 0);
        int64_t wide = 0;
        for (int64_t w = 0; w < nwg; ++w) {
            int64_t cmin = INT64_MAX, cmax = -1;
            for (int64_t t = w * band; t < std::min(nstrips, (w + 1) * band); ++t)
                for (const Piece &pc : merged[t]) {
                    const int64_t a0 = pc.c0 & ~int64_t(7);
                    cmin = std::min(cmin, a0);
                    cmax = std::max(cmax, a0 + ((pc.c0 + pc.n - a0 + 7) & ~int64_t(7)));
                }
            if (cmax > cmin && cmax - cmin <= sh.xsm) {
                wgx[2 * w] = static_cast<int32_t>(cmin);
                wgx[2 * w + 1] = static_cast<int32_t>(cmax - cmin);
            } else if (cmax > cmin) {
                ++wide;
            }
        }
        return std::make_pair(wide, nwg);
    };
    RowShape shape = strip_shape == HM_STRIPS_WIDE ? ROW_WIDE : ROW_NARROW;
    const auto nar = windows(shape);
    if (strip_shape == HM_STRIPS_AUTO && 10 * nar.first > nar.second) {
        shape = ROW_WIDE;
        windows(ROW_WIDE);
    }
    for (int64_t t = 0; t < nstrips; ++t) {
        const int64_t srow0 = t * SROWS;
        const int32_t R = static_cast<int32_t>(std::min<int64_t>(SROWS, nrows - srow0));
        const std::vector<Piece> &mg = merged[t];
        // K-steps: {afrag, src, info}; dense ones first, then the panel ones
        struct KStep {
            uint32_t afrag, src, info;
        };
        std::vector<KStep> ks;
        std::vector<uint32_t> idxw;     // 32 t-row indices per panel K-step
        int32_t nd = 0;
        // K-steps of X columns [x0, x0 + w) of one merged piece whose piece
        // column kk sits at X column pc.c0 + kk.  x0 is a multiple of 8: the
        // row pass reads whole octets of the split X tile, which is zero past
        // the last row of X.
        auto emit = [&](const Piece &pc, int64_t x0, int64_t w) {
            auto val = [&](int r, int64_t j) {
                const int64_t kk = x0 + j - pc.c0;
                if (kk < 0 || kk >= pc.n) return 0.0;
                size_t q = pc.parts.size() - 1;
                while (pc.parts[q].second > kk) --q;
                const OB &o = obs[pc.parts[q].first];
                const int64_t lr = srow0 + pc.roff + r - o.row0;
                return dense[o.offA + lr + (kk - pc.parts[q].second) * o.m];
            };
            const int nK = static_cast<int>((w + 31) / 32);
            for (int s = 0; s < nK; ++s) {
                const int64_t rem = w - 32 * s;
                const int nOct = rem >= 32 ? 4 : static_cast<int>((rem + 7) / 8);
                const uint32_t af = pack(pc.rows, s, w, nOct, val);
                ks.push_back(KStep{af, static_cast<uint32_t>(x0 + 32 * s),
                                   static_cast<uint32_t>(pc.roff | (pc.rows << 4) | (nOct << 9))});
                ++nd;
            }
        };
        for (const Piece &pc : mg) {
            const int64_t a0 = pc.c0 & ~int64_t(7);
            emit(pc, a0, pc.c0 + pc.n - a0);
        }
        // the panel of low-rank columns: A rows of every leaf crossing the strip
        std::vector<std::pair<int32_t, int64_t>> ucols;     // (leaf, rank column)
        for (int32_t b : rp[t])
            for (int64_t j = 0; j < obs[b].k; ++j) ucols.emplace_back(b, j);
        const int64_t nu = static_cast<int64_t>(ucols.size());
        for (int s = 0; 32 * s < nu; ++s) {
            const int64_t rem = nu - 32 * s;
            const int nOct = rem >= 32 ? 4 : static_cast<int>((rem + 7) / 8);
            const uint32_t af = pack(R, s, nu, nOct, [&](int r, int64_t kk) {
                const OB &o = obs[ucols[kk].first];
                const int64_t lr = srow0 + r - o.row0;
                return (lr >= 0 && lr < o.m) ? fac[o.offA + ucols[kk].second * o.m + lr] : 0.0;
            });
            ks.push_back(KStep{af, 0u, static_cast<uint32_t>((R << 4) | (nOct << 9) | (SK_KIND_PANEL << 12))});
            for (int e = 0; e < 32; ++e) {
                const int64_t kk = 32 * s + e;
                const int64_t row = kk < nu ? (tOff[ucols[kk].first] + ucols[kk].second * NB) / NB : zrow;
                idxw.push_back(static_cast<uint32_t>(row));
            }
        }
        // Records: header, 4 words per K-step (dense ones first), then the
        // 32-word index block of the record's panel K-step (at most one), all
        // within the first 128 words (the row pass reads index words from the
        // two record registers); a strip with more K-steps continues in a
        // chained record (word 5).
        const int32_t ntot = static_cast<int32_t>(ks.size());
        soff[t] = static_cast<uint32_t>(recs.size());
        int32_t k0 = 0;
        do {
            int32_t k1 = k0, rd = 0, rpn = 0;
            while (k1 < ntot) {
                const bool panel = k1 >= nd;
                const int32_t words = SREC_HDR + 4 * (k1 + 1 - k0) + 32 * (rpn + (panel ? 1 : 0));
                if (words > 128 || (panel && rpn == 1)) break;     // one panel K-step per record
                if (panel) ++rpn;
                else ++rd;
                ++k1;
            }
            std::vector<uint32_t> rec(SREC_HDR, 0u);
            for (int32_t k = k0; k < k1; ++k) {
                rec.push_back(ks[k].afrag);
                rec.push_back(ks[k].src);
                rec.push_back(ks[k].info);
                rec.push_back(0u);
            }
            const size_t ib = rec.size();
            for (int32_t q = 0; q < rpn; ++q) {
                const int32_t k = k0 + rd + q;            // panel K-step k is panel number k - nd
                rec[SREC_HDR + 4 * (rd + q) + 1] = static_cast<uint32_t>(ib + 32 * q);
                rec.insert(rec.end(), idxw.begin() + 32 * (k - nd), idxw.begin() + 32 * (k - nd + 1));
            }
            rec[0] = static_cast<uint32_t>(srow0);
            rec[1] = static_cast<uint32_t>(R);
            rec[2] = static_cast<uint32_t>(rd);
            rec[3] = static_cast<uint32_t>(rpn);
            rec[4] = static_cast<uint32_t>(ib);
            while (rec.size() % 4) rec.push_back(0u);
            const size_t at = recs.size();
            if (at + rec.size() + 128 >= (size_t(1) << 32)) throw std::length_error("strip records exceed 2^32 words");
            if (k1 < ntot) rec[5] = static_cast<uint32_t>(at + rec.size());   // next record
            recs.insert(recs.end(), rec.begin(), rec.end());
            k0 = k1;
        } while (k0 < ntot);
    }
    // ---- projection: super-tiles of PROJ_SUPER columns.  The rank columns
    // (leaf, j) crossing a super-tile are cut into groups of 16; each column
    // stores its B values over the 32-row K-steps [s0, s1) of the super-tile
    // its leaf reaches (zeros around the leaf), hi rows then lo rows.
    const int64_t nst = (ncols + PROJ_SUPER - 1) / PROJ_SUPER;
    std::vector<std::vector<int32_t>> perst(static_cast<size_t>(nst));
    std::vector<int64_t> ncolst(static_cast<size_t>(nst), 0);
    for (size_t b = 0; b < obs.size(); ++b) {
        const OB &o = obs[b];
        if (o.kind != KIND_RK || !contributes(o)) continue;
        for (int64_t st = o.col0 / PROJ_SUPER; st <= (o.col0 + o.n - 1) / PROJ_SUPER; ++st) {
            perst[st].push_back(static_cast<int32_t>(b));
            ncolst[st] += o.k;
        }
    }
    int64_t gps = 0;
    for (int64_t st = 0; st < nst; ++st) gps = std::max(gps, (ncolst[st] + 15) / 16);
    if (tLen + partLen >= (int64_t(1) << 31)) throw std::length_error("t and partials exceed 2^31 floats");
    std::vector<uint32_t> pcol(static_cast<size_t>(nst * gps * 16 * 4), 0u);
    Arena vs;
    vs.mode = mode == 1 ? 1 : 2;
    for (int i = 0; i < VZERO; ++i) vs.push(0.0);
    for (int64_t st = 0; st < nst; ++st) {
        const int64_t ng = (ncolst[st] + 15) / 16;
        uint32_t *dst = pcol.data() + st * gps * 64;
        for (int64_t q = 0; q < gps * 16; ++q) {
            dst[4 * q + 1] = static_cast<uint32_t>(ng << 16);
            dst[4 * q + 2] = 0xffffffffu;
        }
        int64_t q = 0;
        for (int32_t b : perst[st]) {
            const OB &o = obs[b];
            const int64_t a = std::max(o.col0, st * PROJ_SUPER) - st * PROJ_SUPER;
            const int64_t e = std::min(o.col0 + o.n, (st + 1) * PROJ_SUPER) - st * PROJ_SUPER;
            const int64_t s0 = a / 32, s1 = (e + 31) / 32;
            const int64_t chunk = st - o.col0 / PROJ_SUPER;
            const double *Bf = fac + o.offB;
            for (int64_t j = 0; j < o.k; ++j, ++q) {
                vs.align(8);
                const int64_t base = vs.size();
                // element offsets (both planes) stay below 2^32
                if (base + 2 * 32 * (s1 - s0) >= (int64_t(1) << 32)) return false;
                for (int plane = 0; plane < (mode == 3 ? 2 : 1); ++plane)
                    for (int64_t r = 32 * s0; r < 32 * s1; ++r) {
                        const int64_t i = st * PROJ_SUPER + r - o.col0;
                        const double v = (i >= 0 && i < o.n) ? Bf[i + j * o.n] : 0.0;
                        if (plane == 0) vs.push(v);
                        else vs.push_lo(v);
                    }
                const int64_t out = nch[b] > 1 ? tLen + pOff[b] + chunk * o.k * NB + j * NB : tOff[b] + j * NB;
                dst[4 * q + 0] = static_cast<uint32_t>(base / 8);
                dst[4 * q + 1] = static_cast<uint32_t>(s0 | (s1 << 8) | (ng << 16));
                dst[4 * q + 2] = static_cast<uint32_t>(out);
            }
        }
    }
    // reductions, largest first (they start first)
    // (the items with more than 32 chunks, which the reduction gives a whole
    // workgroup, then lead the list)
    std::stable_sort(red.begin(), red.end(), [](const BRed &x, const BRed &y) {
        return x.nchunk != y.nchunk ? x.nchunk > y.nchunk : x.len > y.len;
    });
    recs.resize(recs.size() + 128, 0u);   // the row pass reads 128 words from every record start
    H.nstrips = nstrips;
    H.shape = shape;
    H.wgx = std::move(wgx);
    H.soff = std::move(soff);
    H.recs = std::move(recs);
    H.pcol = std::move(pcol);
    H.nst = nst;
    H.gps = gps;
    H.red = std::move(red);
    H.fr = std::move(fr);
    H.vs = std::move(vs);
    H.tLen = tLen;
    H.partLen = partLen;
    return true;
}

void build_batched(hm::Plan &p, hm::Direction &d, int64_t nrows, int64_t ncols,
                   const std::vector<OB> &obs, const double *fac, const double *dense) {
    using namespace hm;
    BatchedHost H;
    if (!layout_batched(p.storage, p.strip_shape, nrows, ncols, obs, fac, dense, H)) return;
    Batched &B = d.bat;
    B.nstrips = H.nstrips;
    B.soff = dev_upload(p, H.soff);
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
    B.frag = H.fr.mode == 1 ? upload_vec(p, H.fr.f) : upload_vec(p, H.fr.h);
    B.vslc = H.vs.mode == 1 ? upload_vec(p, H.vs.f) : upload_vec(p, H.vs.h);
    B.tLen = H.tLen;
    B.partLen = H.partLen;
    const size_t tsz = static_cast<size_t>(H.tLen + H.partLen + NB) * sizeof(float);
    B.t = static_cast<float *>(dev_alloc(p, tsz));
    hip_check(hipMemset(B.t, 0, tsz), "hipMemset");
    B.ready = true;
}

// Single-right-hand-side layout of one direction: projection items over the
// factor arena `fac` (uploaded once per plan) and the row tiles' column
// streams built from `fac` (low-rank A) and `dense` (dense leaves).
// mrow (many-matrix plans, else empty): first row of each member in this
// direction, and nrows last; the direction then gets the one-launch layout
// (items and tiles ordered by member, largest members first).
void build_direction(hm::Plan &p, hm::Direction &d, int64_t nrows, int64_t ncols,
                     const std::vector<OB> &obs, const double *fac, const double *dense,
                     const std::vector<int64_t> &mrow) {
    using namespace hm;
    d.nrows = nrows;
    d.ncols = ncols;
    const int64_t nb = static_cast<int64_t>(obs.size());
    std::vector<BlkDev> blk(obs.size());
    std::vector<ProjItem> proj;
    std::vector<RedItem> red;
    std::vector<int64_t> pleaf, rleaf;     // leaf of each projection / reduction item
    // t offsets
    int64_t tLen = 0;
    for (int64_t b = 0; b < nb; ++b) {
        const OB &o = obs[b];
        BlkDev &x = blk[b];
        x.offA = o.offA;
        x.offB = o.offB;
        x.row0 = o.row0;
        x.col0 = o.col0;
        x.m = static_cast<int32_t>(o.m);
        x.n = static_cast<int32_t>(o.n);
        x.k = static_cast<int32_t>(o.k);
        x.kind = o.kind;
        x.tOff = 0;
        if (o.kind == KIND_RK && o.k > 0) {
            x.tOff = tLen;
            tLen += o.k;
        }
    }
    // projection items; partial offsets are placed after t
    int64_t partLen = 0;
    for (int64_t b = 0; b < nb; ++b) {
        const OB &o = obs[b];
        if (o.kind != KIND_RK || o.k == 0 || o.n == 0) continue;
        const int64_t chunk = p.proj_chunk;
        const int64_t nchunk = (o.n + chunk - 1) / chunk;
        for (int64_t j0 = 0; j0 < o.k; j0 += PROJ_KMAX) {
            const int32_t kc = static_cast<int32_t>(std::min<int64_t>(PROJ_KMAX, o.k - j0));
            const int64_t tDst = blk[b].tOff + j0;
            int64_t pOff = 0;
            if (nchunk > 1) {
                pOff = partLen;
                partLen += nchunk * kc;
                red.push_back(RedItem{pOff, tDst, static_cast<int32_t>(nchunk), kc});
                rleaf.push_back(b);
            }
            for (int64_t c = 0; c < nchunk; ++c) {
                ProjItem it{};
                const int64_t row0 = c * chunk;
                it.base = o.offB + j0 * o.n + row0;
                it.xoff = o.col0 + row0;
                it.ld = static_cast<int32_t>(o.n);
                it.rows = static_cast<int32_t>(std::min<int64_t>(chunk, o.n - row0));
                it.kc = kc;
                it.out = nchunk > 1 ? tLen + pOff + c * kc : tDst;
                proj.push_back(it);
                pleaf.push_back(b);
            }
        }
    }
    // row tiles and their column streams (tile-major, leaves in table order)
    auto tl = make_tiles(nrows, obs, mrow);
    std::vector<int64_t> tileStart(tl.size());
    for (size_t i = 0; i < tl.size(); ++i) tileStart[i] = tl[i].first;
    std::vector<std::vector<std::pair<int32_t, int32_t>>> per(tl.size());   // (leaf, row offset)
    for (int64_t b = 0; b < nb; ++b) {
        const OB &o = obs[b];
        if (!contributes(o)) continue;
        size_t t = static_cast<size_t>(std::lower_bound(tileStart.begin(), tileStart.end(), o.row0) -
                                       tileStart.begin());
        for (; t < tl.size() && tl[t].first < o.row0 + o.m; ++t)
            per[t].emplace_back(static_cast<int32_t>(b), static_cast<int32_t>(tl[t].first - o.row0));
    }
    std::vector<TileDev> tiles(tl.size());
    std::vector<int32_t> mult;
    std::vector<double> col;     // FP64 stream, converted to storage below
    for (size_t t = 0; t < tl.size(); ++t) {
        const int64_t R = tl[t].second;
        TileDev &T = tiles[t];
        T.row0 = tl[t].first;
        T.rows = static_cast<int32_t>(R);
        T.base = static_cast<int64_t>(col.size());
        if (mult.size() >= (size_t(1) << 31)) throw std::length_error("column stream exceeds 2^31 columns");
        T.c0 = static_cast<int32_t>(mult.size());
        for (const auto &pr : per[t]) {
            const OB &o = obs[pr.first];
            const int64_t ro = pr.second;
            if (o.kind == KIND_DENSE) {
                const double *C = dense + o.offA;
                for (int64_t j = 0; j < o.n; ++j) {
                    mult.push_back(static_cast<int32_t>(o.col0 + j));
                    for (int64_t r = 0; r < R; ++r) col.push_back(C[(ro + r) + j * o.m]);
                }
            } else {
                const double *Af = fac + o.offA;
                for (int64_t j = 0; j < o.k; ++j) {
                    mult.push_back(~static_cast<int32_t>(blk[pr.first].tOff + j));
                    for (int64_t r = 0; r < R; ++r) col.push_back(Af[(ro + r) + j * o.m]);
                }
            }
        }
        T.c1 = static_cast<int32_t>(mult.size());
    }
    if (tLen >= (int64_t(1) << 31)) throw std::length_error("rank total exceeds 2^31");
    std::vector<int32_t> redmem;
    std::vector<uint32_t> need;
    const int nmem = mrow.size() > 2 ? static_cast<int>(mrow.size() - 1) : 0;
    if (nmem > 0) {
        // member holding `row`: the last member that starts at or before it
        auto member_of = [&](int64_t row) {
            const int64_t q = std::upper_bound(mrow.begin(), mrow.end(), row) - mrow.begin() - 1;
            return static_cast<int32_t>(std::min<int64_t>(std::max<int64_t>(q, 0), nmem - 1));
        };
        // bytes per member (stored elements); members in decreasing order
        std::vector<double> bytes(static_cast<size_t>(nmem), 0.0);
        for (const OB &o : obs) {
            if (!contributes(o)) continue;
            bytes[member_of(o.row0)] += static_cast<double>(o.kind == KIND_DENSE ? o.m * o.n : (o.m + o.n) * o.k);
        }
        std::vector<int32_t> order(static_cast<size_t>(nmem)), rank(static_cast<size_t>(nmem));
        for (int i = 0; i < nmem; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return bytes[a] > bytes[b]; });
        for (int i = 0; i < nmem; ++i) rank[order[i]] = i;
        // projection items: member of their leaf (the leaf's first row); by
        // member rank, then larger items first
        std::vector<int32_t> pmem(proj.size());
        for (size_t i = 0; i < proj.size(); ++i) pmem[i] = member_of(obs[pleaf[i]].row0);
        std::vector<size_t> pi(proj.size());
        for (size_t i = 0; i < pi.size(); ++i) pi[i] = i;
        std::stable_sort(pi.begin(), pi.end(), [&](size_t a, size_t b) {
            if (rank[pmem[a]] != rank[pmem[b]]) return rank[pmem[a]] < rank[pmem[b]];
            return int64_t(proj[a].rows) * proj[a].kc > int64_t(proj[b].rows) * proj[b].kc;
        });
        std::vector<ProjItem> proj2(proj.size());
        for (size_t i = 0; i < pi.size(); ++i) {
            proj2[i] = proj[pi[i]];
            proj2[i].pad = pmem[pi[i]];
        }
        proj.swap(proj2);
        // reduction items: member of their leaf
        {
            redmem.resize(red.size());
            for (size_t i = 0; i < red.size(); ++i) redmem[i] = member_of(obs[rleaf[i]].row0);
            std::vector<size_t> ri(red.size());
            for (size_t i = 0; i < ri.size(); ++i) ri[i] = i;
            std::stable_sort(ri.begin(), ri.end(), [&](size_t a, size_t b) { return rank[redmem[a]] < rank[redmem[b]]; });
            std::vector<RedItem> red2(red.size());
            std::vector<int32_t> rm2(red.size());
            for (size_t i = 0; i < ri.size(); ++i) {
                red2[i] = red[ri[i]];
                rm2[i] = redmem[ri[i]];
            }
            red.swap(red2);
            redmem.swap(rm2);
        }
        // row tiles: by member rank, then wider streams first
        std::vector<size_t> ti(tiles.size());
        for (size_t i = 0; i < ti.size(); ++i) ti[i] = i;
        std::vector<int32_t> tmem(tiles.size());
        for (size_t i = 0; i < tiles.size(); ++i) tmem[i] = member_of(tiles[i].row0);
        std::stable_sort(ti.begin(), ti.end(), [&](size_t a, size_t b) {
            if (rank[tmem[a]] != rank[tmem[b]]) return rank[tmem[a]] < rank[tmem[b]];
            return int64_t(tiles[a].c1 - tiles[a].c0) * tiles[a].rows > int64_t(tiles[b].c1 - tiles[b].c0) * tiles[b].rows;
        });
        std::vector<TileDev> tiles2(tiles.size());
        for (size_t i = 0; i < ti.size(); ++i) {
            tiles2[i] = tiles[ti[i]];
            tiles2[i].pad = tmem[ti[i]];
        }
        tiles.swap(tiles2);
        // slot of each item: the launch's workgroup index mod FUSED_SLOTS
        // (NW projection items, then 64*NW/PROJ_KMAX reduction items, then
        // one row tile per workgroup)
        constexpr int S = FUSED_SLOTS;
        const int64_t NW = p.row_waves, RPW = 64 * NW / PROJ_KMAX;
        const int64_t npw = (static_cast<int64_t>(proj.size()) + NW - 1) / NW;
        const int64_t nrw = (static_cast<int64_t>(red.size()) + RPW - 1) / RPW;
        const size_t MS = static_cast<size_t>(nmem) * S;
        need.assign(3 * MS + static_cast<size_t>(nmem), 0u);
        for (size_t i = 0; i < proj.size(); ++i) {
            const size_t q = static_cast<size_t>(proj[i].pad) * S + static_cast<size_t>((i / NW) % S);
            ++need[q];
            ++need[MS + q];
        }
        for (size_t i = 0; i < red.size(); ++i)
            ++need[static_cast<size_t>(redmem[i]) * S + static_cast<size_t>((npw + i / RPW) % S)];
        for (size_t i = 0; i < tiles.size(); ++i)
            ++need[2 * MS + static_cast<size_t>(tiles[i].pad) * S + static_cast<size_t>((npw + nrw + i) % S)];
        for (int m = 0; m < nmem; ++m)
            for (int q = 0; q < S; ++q) need[3 * MS + m] += need[2 * MS + m * S + q] > 0 ? 1u : 0u;
    }
    std::vector<uint8_t> shi, slo;
    pack_storage(p.storage, col.data(), static_cast<int64_t>(col.size()), shi, slo);
    d.stream = shi.empty() ? nullptr : static_cast<void *>(dev_upload(p, shi));
    d.stream_lo = slo.empty() ? nullptr : static_cast<void *>(dev_upload(p, slo));
    d.mult = dev_upload(p, mult);
    d.ncolumns = static_cast<int64_t>(mult.size());
    d.blk = dev_upload(p, blk);
    d.nblk = nb;
    d.proj = dev_upload(p, proj);
    d.nproj = static_cast<int64_t>(proj.size());
    d.red = dev_upload(p, red);
    d.nred = static_cast<int64_t>(red.size());
    d.tiles = dev_upload(p, tiles);
    d.ntiles = static_cast<int64_t>(tiles.size());
    d.tLen = tLen;
    d.partLen = partLen;
    d.t = dev_alloc(p, static_cast<size_t>(tLen + partLen) * acc_size(p.storage));
    d.tm = dev_alloc(p, static_cast<size_t>(tLen + partLen) * MULTI_RHS * acc_size(p.storage));
    if (nmem > 0) {
        d.fmem.nmem = nmem;
        d.fmem.need = dev_upload(p, need);
        d.fmem.redmem = red.empty() ? nullptr : dev_upload(p, redmem);
        const size_t cbytes = sizeof(uint32_t) * FUSED_STRIDE * (2 * FUSED_SLOTS + 1) * static_cast<size_t>(nmem);
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
    if (p->storage != HM_F64 && nrhs >= 2 && d.bat.ready && !(p->flags & HM_NO_BATCHED)) {
        for (int64_t c0 = 0; c0 < nrhs; c0 += hm::NB) {
            const int nc = static_cast<int>(std::min<int64_t>(hm::NB, nrhs - c0));
            hipError_t e = hm::launch_batched(*p, d, nc, static_cast<const float *>(X) + c0 * ldx, ldx,
                                              static_cast<float *>(Y) + c0 * ldy, ldy,
                                              static_cast<float>(alpha), static_cast<float>(beta), s,
                                              record());
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
        if (o.row0 < 0 || o.col0 < 0 || o.m < 0 || o.n < 0 || o.k < 0 || o.offA < 0 || o.offB < 0 ||
            o.row0 + o.m > nrows || o.col0 + o.n > ncols) {
            std::snprintf(msg, sizeof msg, "block %lld: outside the %lldx%lld matrix", (long long)b,
                          (long long)nrows, (long long)ncols);
            return fail(HM_EINVAL, msg);
        }
        if (o.kind == KIND_DENSE) {
            if (o.offA + o.m * o.n > ndata) {
                std::snprintf(msg, sizeof msg, "block %lld: dense leaf outside the arena", (long long)b);
                return fail(HM_EINVAL, msg);
            }
            ++ndense;
            dense_elems += o.m * o.n;
        } else {
            if (o.k >= (int64_t(1) << 20) || o.offA + o.m * o.k > ndata || o.offB + o.n * o.k > ndata) {
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
        std::vector<double> rkv(static_cast<size_t>(rk_elems));
        {
            int64_t o2 = 0;
            for (OB &o : fw) {
                if (o.kind != KIND_RK) continue;
                std::memcpy(rkv.data() + o2, data + o.offA, sizeof(double) * o.m * o.k);
                std::memcpy(rkv.data() + o2 + o.m * o.k, data + o.offB, sizeof(double) * o.n * o.k);
                o.offA = o2;
                o.offB = o2 + o.m * o.k;
                o2 += (o.m + o.n) * o.k;
            }
        }
        const std::vector<int32_t> rexp = balance_factors(fw, rkv.data());
        p->zeros = dev_upload(*p, std::vector<uint8_t>(ZERO_BYTES, 0));
        std::vector<uint8_t> hi, lo;
        pack_storage(storage, rkv.data(), rk_elems, hi, lo);
        p->fac = hi.empty() ? nullptr : static_cast<void *>(dev_upload(*p, hi));
        p->fac_lo = lo.empty() ? nullptr : static_cast<void *>(dev_upload(*p, lo));
        // transposed dense leaves, packed compactly
        std::vector<double> dT(static_cast<size_t>(dense_elems));
        std::vector<OB> bw(fw.size());
        int64_t off = 0;
        for (size_t b = 0; b < fw.size(); ++b) {
            const OB &o = fw[b];
            OB t{o.kind, o.col0, o.row0, o.n, o.m, o.k, o.offB, o.offA};
            if (o.kind == KIND_DENSE) {
                const double *C = data + o.offA;
                for (int64_t j = 0; j < o.n; ++j)
                    for (int64_t i = 0; i < o.m; ++i) dT[off + j + i * o.n] = C[i + j * o.m];
                t.offA = off;
                t.offB = 0;
                off += o.m * o.n;
            }
            bw[b] = t;
        }
        const bool fused = mrow.size() > 2 && (c.flags & HM_ONE_LAUNCH);
        if (fused) {
            p->fault = static_cast<uint32_t *>(dev_alloc(*p, sizeof(uint32_t)));
            hip_check(hipMemset(p->fault, 0, sizeof(uint32_t)), "hipMemset");
        }
        const std::vector<int64_t> none;
        build_direction(*p, p->fwd, nrows, ncols, fw, rkv.data(), data, fused ? mrow : none);
        if (!(p->flags & HM_NO_TRANSPOSE))
            build_direction(*p, p->bwd, ncols, nrows, bw, rkv.data(), dT.data(), fused ? mcol : none);
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
        if (storage != HM_F64 && !(p->flags & HM_NO_BATCHED)) {
            build_batched(*p, p->fwd, nrows, ncols, fw, rkv.data(), data);
            if (!(p->flags & HM_NO_TRANSPOSE)) build_batched(*p, p->bwd, ncols, nrows, bw, rkv.data(), dT.data());
        }
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
    for (int64_t i = 0; i < nmat; ++i) {
        if (nrows[i] < 0 || ncols[i] < 0 || nblocks[i] < 0 || ndata[i] < 0)
            return fail(HM_EINVAL, "negative size in batch member");
        if ((nblocks[i] > 0 && !blocks[i]) || (ndata[i] > 0 && !data[i]))
            return fail(HM_EINVAL, "NULL array in batch member");
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
            const bool bad = (r[0] != 0 && r[0] != 1) || r[1] < 0 || r[2] < 0 || r[3] < 0 || r[4] < 0 || r[5] < 0 ||
                             r[6] < 0 || r[7] < 0 || r[1] + r[3] > nrows[i] || r[2] + r[4] > ncols[i] ||
                             (r[0] == 0 ? r[6] + r[3] * r[4] > ndata[i]
                                        : (r[6] + r[3] * r[5] > ndata[i] || r[7] + r[4] * r[5] > ndata[i]));
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
    if (nrhs > 1 && storage == HM_F64)
        return fail(HM_EUNSUPPORTED, "multi-right-hand-side shapes need FP32, BF16 or split-BF16 storage");
    std::vector<hm_config> cand;
    if (nrhs == 1) {
        for (int c : {256, 512, 1024, 2048, 4096})
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
    bcfg.flags = 0;
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
