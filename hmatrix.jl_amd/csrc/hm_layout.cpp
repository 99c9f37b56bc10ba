// Host layout builder (declarations and formats: hm_layout.h, hm_types.h).
// The partition it packs is the block tree of src/cluster.jl:48-133 (the
// leaves of the exported table); the products it serves are those of
// src/harithm.jl:340-354.
#include "hm_layout.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <tuple>

#include "../../include/hmat_amd.h"

namespace hm {

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

size_t storage_size(int st) { return st == HM_F64 ? 8 : st == HM_F32 ? 4 : 2; }
size_t acc_size(int st) { return st == HM_F64 ? 8 : 4; }

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

bool contributes(const OB &o) { return o.m > 0 && o.n > 0 && (o.kind == hm::KIND_DENSE || o.k > 0); }

std::vector<std::pair<int64_t, int32_t>> make_tiles(int64_t nrows, const std::vector<OB> &obs,
                                                    const std::vector<int64_t> &extra) {
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

void Arena::push(double v) {
    if (mode == 3) d.push_back(v);
    else if (mode == 1) f.push_back(static_cast<float>(v));
    else h.push_back(f32_to_bf16_rne(static_cast<float>(v)));
}

void Arena::push_lo(double v) {
    const uint16_t hi = f32_to_bf16_rne(static_cast<float>(v));
    h.push_back(f32_to_bf16_rne(static_cast<float>(v - bf16_to_f32_host(hi))));
}

void orient(const std::vector<OB> &leaves, const double *data, int64_t rk_elems, int64_t dense_elems,
            Oriented &O) {
    O.fw = leaves;
    O.rkv.assign(static_cast<size_t>(rk_elems), 0.0);
    int64_t o2 = 0;
    for (OB &o : O.fw) {
        if (o.kind != KIND_RK) continue;
        std::memcpy(O.rkv.data() + o2, data + o.offA, sizeof(double) * o.m * o.k);
        std::memcpy(O.rkv.data() + o2 + o.m * o.k, data + o.offB, sizeof(double) * o.n * o.k);
        o.offA = o2;
        o.offB = o2 + o.m * o.k;
        o2 += (o.m + o.n) * o.k;
    }
    O.rexp = balance_factors(O.fw, O.rkv.data());
    O.dT.assign(static_cast<size_t>(dense_elems), 0.0);
    O.bw.resize(O.fw.size());
    int64_t off = 0;
    for (size_t b = 0; b < O.fw.size(); ++b) {
        const OB &o = O.fw[b];
        OB t{o.kind, o.col0, o.row0, o.n, o.m, o.k, o.offB, o.offA};
        if (o.kind == KIND_DENSE) {
            const double *C = data + o.offA;
            for (int64_t j = 0; j < o.n; ++j)
                for (int64_t i = 0; i < o.m; ++i) O.dT[off + j + i * o.n] = C[i + j * o.m];
            t.offA = off;
            t.offB = 0;
            off += o.m * o.n;
        }
        O.bw[b] = t;
    }
}

void layout_direction(int storage, int proj_chunk, int row_waves, int64_t nrows, const std::vector<OB> &obs,
                      const double *fac, const double *dense, const std::vector<int64_t> &mrow,
                      DirectionHost &D) {
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
        const int64_t chunk = proj_chunk;
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
        const int64_t NW = row_waves, RPW = 64 * NW / PROJ_KMAX;
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
    pack_storage(storage, col.data(), static_cast<int64_t>(col.size()), D.shi, D.slo);
    D.ncolumns = static_cast<int64_t>(mult.size());
    D.mult = std::move(mult);
    D.blk = std::move(blk);
    D.proj = std::move(proj);
    D.red = std::move(red);
    D.tiles = std::move(tiles);
    D.tLen = tLen;
    D.partLen = partLen;
    D.nmem = nmem;
    D.need = std::move(need);
    D.redmem = std::move(redmem);
}

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
    fr.mode = mode == HM_F64 ? 3 : mode == HM_F32 ? 1 : 2;
    const int planes = mode == HM_BF16X2 ? 2 : 1;
    const int64_t lane_elems = 8;                  // values per lane and octet
    const int64_t unit = mode == HM_F64 ? 2 : mode == HM_F32 ? 4 : 8;   // arena elements per 16 bytes
    const int64_t zrow = (tLen + partLen) / NB;   // zero row of t (after the partials)
    std::vector<uint32_t> recs, soff(static_cast<size_t>(nstrips));
    std::vector<int32_t> kcnt(static_cast<size_t>(nstrips), 0);     // K-steps per strip
    // Pack one K-step of A fragments: rows [roff, roff + rows) of the strip,
    // columns [32 s, 32 s + 32) of a block whose value is val(r, kk), klen
    // valid columns.  Returns the offset in 16-byte units.
    auto pack = [&](int rows, int s, int64_t klen, int nOct,
                    const std::function<double(int, int64_t)> &val) -> uint32_t {
        fr.align(unit);
        const int64_t at = fr.size() / unit;
        if (at >= (int64_t(1) << 32)) throw std::length_error("fragment arena exceeds 64 GiB");
        if (mode == HM_F64) {
            // FP64: lane (g, r) holds columns 4 m + g, m < 2 nOct, of row r
            // (the f64 MFMA m takes columns 4m .. 4m+3, so a K-step of nOct
            // octets needs MFMAs m < 2 nOct only)
            for (int g = 0; g < 4; ++g)
                for (int r = 0; r < rows; ++r)
                    for (int m = 0; m < 2 * nOct; ++m) {
                        const int64_t kk = 32 * s + 4 * m + g;
                        fr.push(kk < klen ? val(r, kk) : 0.0);
                    }
            return static_cast<uint32_t>(at);
        }
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
    RowShape shape = mode == HM_F64 ? ROW_F64 : strip_shape == HM_STRIPS_WIDE ? ROW_WIDE : ROW_NARROW;
    const auto nar = windows(shape);
    if (mode != HM_F64 && strip_shape == HM_STRIPS_AUTO && 10 * nar.first > nar.second) {
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
        kcnt[static_cast<size_t>(t)] = ntot;
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
    vs.mode = fr.mode;
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
                for (int plane = 0; plane < planes; ++plane)
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
    if (mode == HM_F64)
        for (int i = 0; i < 8; ++i) fr.push(0.0);   // FP64 lanes read 4 units from their start
    // Strips of a band to its wavefronts, longest processing time first: a
    // workgroup lives as long as its slowest wavefront, so each strip (cost:
    // its K-steps plus a fixed record overhead, in halves) goes to the
    // wavefront with the least work so far; a wavefront's strips stay in
    // row order.
    {
        const int64_t band = static_cast<int64_t>(shape.nw) * shape.spw;
        const int64_t nbands = (nstrips + band - 1) / band;
        H.soffw.assign(static_cast<size_t>(nstrips), 0u);
        H.wpos.assign(static_cast<size_t>(nbands * (shape.nw + 1)), 0u);
        std::vector<int64_t> idx, load(static_cast<size_t>(shape.nw));
        std::vector<std::vector<int64_t>> mine(static_cast<size_t>(shape.nw));
        for (int64_t w = 0; w < nbands; ++w) {
            const int64_t t0 = w * band, t1 = std::min(nstrips, t0 + band);
            idx.clear();
            for (int64_t t = t0; t < t1; ++t) idx.push_back(t);
            auto cost = [&](int64_t t) { return 2 * static_cast<int64_t>(kcnt[static_cast<size_t>(t)]) + 3; };
            std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
            std::fill(load.begin(), load.end(), 0);
            for (auto &m : mine) m.clear();
            for (int64_t t : idx) {
                const size_t q = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
                load[q] += cost(t);
                mine[q].push_back(t);
            }
            uint32_t at = 0;
            uint32_t *wp = H.wpos.data() + w * (shape.nw + 1);
            for (int q = 0; q < shape.nw; ++q) {
                std::sort(mine[static_cast<size_t>(q)].begin(), mine[static_cast<size_t>(q)].end());
                wp[q] = at;
                for (int64_t t : mine[static_cast<size_t>(q)]) H.soffw[static_cast<size_t>(t0 + at++)] = soff[static_cast<size_t>(t)];
            }
            wp[shape.nw] = at;
        }
    }
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

namespace {

std::string violation(const char *fmt, long long a, long long b = 0, long long c = 0) {
    char m[200];
    std::snprintf(m, sizeof m, fmt, a, b, c);
    return m;
}

}  // namespace

std::string check_direction(const DirectionHost &D, int storage, int row_waves, int64_t nrows, int64_t ncols,
                            int64_t fac_elems) {
    const int64_t T = D.tLen, P = D.partLen;
    const int64_t se = static_cast<int64_t>(storage_size(storage));
    if (static_cast<int64_t>(D.shi.size()) % se) return "column stream is not a whole number of elements";
    const int64_t selems = static_cast<int64_t>(D.shi.size()) / se;
    if (storage == HM_BF16X2 && D.slo.size() != D.shi.size()) return "split-BF16 stream planes differ in size";
    if (static_cast<int64_t>(D.mult.size()) != D.ncolumns) return "multiplier count differs from the column count";
    // writes per t entry and partial
    std::vector<uint8_t> tw(static_cast<size_t>(T + P), 0);
    for (size_t i = 0; i < D.proj.size(); ++i) {
        const ProjItem &w = D.proj[i];
        const long long q = static_cast<long long>(i);
        if (w.kc < 1 || w.kc > PROJ_KMAX || w.rows < 1 || w.ld < w.rows)
            return violation("projection item %lld: kc %lld or rows %lld out of range", q, w.kc, w.rows);
        if (w.base < 0 || w.base + static_cast<int64_t>(w.kc - 1) * w.ld + w.rows > fac_elems)
            return violation("projection item %lld: factor rows [%lld, ...) outside the arena", q, w.base);
        if (w.xoff < 0 || w.xoff + w.rows > ncols)
            return violation("projection item %lld: x rows [%lld, +%lld) outside x", q, w.xoff, w.rows);
        if (w.out < 0 || w.out + w.kc > T + P)
            return violation("projection item %lld: output %lld outside t and partials", q, w.out);
        for (int j = 0; j < w.kc; ++j)
            if (++tw[static_cast<size_t>(w.out + j)] > 1)
                return violation("projection item %lld: entry %lld written twice", q, w.out + j);
        if (D.nmem > 0 && (w.pad < 0 || w.pad >= D.nmem)) return violation("projection item %lld: member %lld", q, w.pad);
    }
    for (size_t i = 0; i < D.red.size(); ++i) {
        const RedItem &r = D.red[i];
        const long long q = static_cast<long long>(i);
        if (r.nchunk < 2 || r.kc < 1 || r.kc > PROJ_KMAX)
            return violation("reduction item %lld: nchunk %lld or kc %lld out of range", q, r.nchunk, r.kc);
        if (r.pOff < 0 || r.pOff + static_cast<int64_t>(r.nchunk) * r.kc > P)
            return violation("reduction item %lld: partials at %lld outside the partial area", q, r.pOff);
        for (int64_t e = 0; e < static_cast<int64_t>(r.nchunk) * r.kc; ++e)
            if (tw[static_cast<size_t>(T + r.pOff + e)] != 1)
                return violation("reduction item %lld: partial %lld not written once", q, r.pOff + e);
        if (r.tOff < 0 || r.tOff + r.kc > T) return violation("reduction item %lld: t entry %lld outside t", q, r.tOff);
        for (int j = 0; j < r.kc; ++j)
            if (++tw[static_cast<size_t>(r.tOff + j)] > 1)
                return violation("reduction item %lld: t entry %lld written twice", q, r.tOff + j);
    }
    std::vector<uint8_t> cover(static_cast<size_t>(nrows), 0);
    for (size_t i = 0; i < D.tiles.size(); ++i) {
        const TileDev &t = D.tiles[i];
        const long long q = static_cast<long long>(i);
        if (t.rows < 1 || t.rows > TILE_ROWS || t.row0 < 0 || t.row0 + t.rows > nrows)
            return violation("tile %lld: rows [%lld, +%lld) outside y", q, t.row0, t.rows);
        if (t.c0 < 0 || t.c1 < t.c0 || t.c1 > D.ncolumns) return violation("tile %lld: columns [%lld, %lld)", q, t.c0, t.c1);
        if (t.base < 0 || t.base + static_cast<int64_t>(t.c1 - t.c0) * t.rows > selems)
            return violation("tile %lld: stream values at %lld outside the stream", q, t.base);
        for (int r = 0; r < t.rows; ++r)
            if (++cover[static_cast<size_t>(t.row0 + r)] > 1) return violation("row %lld in two tiles", t.row0 + r);
        for (int32_t c = t.c0; c < t.c1; ++c) {
            const int32_t id = D.mult[static_cast<size_t>(c)];
            if (id >= 0 ? id >= ncols : (~id >= T || tw[static_cast<size_t>(~id)] != 1))
                return violation("tile %lld, column %lld: multiplier %lld is not a valid x or written t entry", q, c,
                                 id);
        }
        if (D.nmem > 0 && (t.pad < 0 || t.pad >= D.nmem)) return violation("tile %lld: member %lld", q, t.pad);
    }
    for (int64_t r = 0; r < nrows; ++r)
        if (cover[static_cast<size_t>(r)] != 1) return violation("row %lld in no tile", r);
    if (D.nmem > 0) {
        constexpr int S = FUSED_SLOTS;
        const int64_t MS = static_cast<int64_t>(D.nmem) * S;
        if (static_cast<int64_t>(D.need.size()) != 3 * MS + D.nmem) return "counter targets: wrong size";
        if (D.redmem.size() != D.red.size()) return "reduction members: wrong size";
        // the targets again from the launch order of the workgroups
        const int64_t NW = row_waves, RPW = 64 * NW / PROJ_KMAX;
        const int64_t npw = (static_cast<int64_t>(D.proj.size()) + NW - 1) / NW;
        const int64_t nrw = (static_cast<int64_t>(D.red.size()) + RPW - 1) / RPW;
        std::vector<uint32_t> need(static_cast<size_t>(3 * MS + D.nmem), 0u);
        for (size_t i = 0; i < D.proj.size(); ++i) {
            const int64_t q = D.proj[i].pad * S + static_cast<int64_t>(i / NW) % S;
            ++need[static_cast<size_t>(q)];
            ++need[static_cast<size_t>(MS + q)];
        }
        for (size_t i = 0; i < D.red.size(); ++i) {
            if (D.redmem[i] < 0 || D.redmem[i] >= D.nmem) return violation("reduction item %lld: member %lld",
                                                                            static_cast<long long>(i), D.redmem[i]);
            ++need[static_cast<size_t>(D.redmem[i] * S + (npw + static_cast<int64_t>(i) / RPW) % S)];
        }
        for (size_t i = 0; i < D.tiles.size(); ++i)
            ++need[static_cast<size_t>(2 * MS + D.tiles[i].pad * S + (npw + nrw + static_cast<int64_t>(i)) % S)];
        for (int m = 0; m < D.nmem; ++m)
            for (int q = 0; q < S; ++q) need[static_cast<size_t>(3 * MS + m)] += need[static_cast<size_t>(2 * MS + m * S + q)] > 0;
        if (need != D.need) return "counter targets differ from the workgroups' items";
        // a member's tiles read only t entries of its own projections
        std::vector<int32_t> owner(static_cast<size_t>(T + P), -1);
        for (const ProjItem &w : D.proj)
            for (int j = 0; j < w.kc; ++j) owner[static_cast<size_t>(w.out + j)] = w.pad;
        for (size_t i = 0; i < D.red.size(); ++i)
            for (int j = 0; j < D.red[i].kc; ++j) owner[static_cast<size_t>(D.red[i].tOff + j)] = D.redmem[i];
        for (const TileDev &t : D.tiles)
            for (int32_t c = t.c0; c < t.c1; ++c) {
                const int32_t id = D.mult[static_cast<size_t>(c)];
                if (id < 0 && owner[static_cast<size_t>(~id)] != t.pad)
                    return violation("tile of member %lld reads t entry %lld of member %lld", t.pad, ~id,
                                     owner[static_cast<size_t>(~id)]);
            }
    }
    return "";
}

std::string check_batched(const BatchedHost &H, int mode, int64_t nrows, int64_t ncols) {
    const int planes = mode == HM_BF16X2 ? 2 : 1;
    const int64_t unit = mode == HM_F64 ? 2 : mode == HM_F32 ? 4 : 8;   // arena elements per 16-byte unit
    const int64_t frN = H.fr.size(), vsN = H.vs.size();
    const int64_t T = H.tLen, P = H.partLen;
    if (T % NB || P % NB) return "t or partial length is not a multiple of NB";
    const int64_t zrow = (T + P) / NB;
    if (H.nstrips != (nrows + SROWS - 1) / SROWS || static_cast<int64_t>(H.soff.size()) != H.nstrips)
        return "strip count";
    const int64_t band = static_cast<int64_t>(H.shape.nw) * H.shape.spw;
    const int64_t nbands = (H.nstrips + band - 1) / band;
    if (static_cast<int64_t>(H.wgx.size()) != 2 * nbands) return "window table size";
    for (int64_t w = 0; w < nbands; ++w) {
        const int32_t cmin = H.wgx[static_cast<size_t>(2 * w)], W = H.wgx[static_cast<size_t>(2 * w + 1)];
        if (W < 0 || W > H.shape.xsm || W % 8 || cmin < 0 || cmin % 8 || (W > 0 && cmin >= ncols))
            return violation("band %lld: window [%lld, +%lld) not staged whole octets", w, cmin, W);
    }
    // projection: column descriptors, slices, outputs
    std::vector<uint8_t> tw(static_cast<size_t>(T + P), 0);
    if (H.nst != (ncols + PROJ_SUPER - 1) / PROJ_SUPER || static_cast<int64_t>(H.pcol.size()) != H.nst * H.gps * 64)
        return "projection descriptor table size";
    if (vsN < VZERO) return "slice arena lacks its zero head";
    for (int64_t q = 0; q < H.nst * H.gps * 16; ++q) {
        const uint32_t *d = H.pcol.data() + 4 * q;
        const int64_t s0 = d[1] & 255, s1 = (d[1] >> 8) & 255, ng = d[1] >> 16;
        if (ng > H.gps) return violation("column descriptor %lld: %lld groups", q, ng);
        if (d[2] == 0xffffffffu) {
            if (s1 != s0) return violation("empty column descriptor %lld reads K-steps", q);
            continue;
        }
        if (s1 <= s0 || s1 > PROJ_SUPER / 32) return violation("column descriptor %lld: K-steps [%lld, %lld)", q, s0, s1);
        const int64_t base = static_cast<int64_t>(d[0]) * 8;
        if (base + planes * 32 * (s1 - s0) > vsN) return violation("column descriptor %lld: slice at %lld outside", q, base);
        const int64_t out = d[2];
        if (out + NB > T + P) return violation("column descriptor %lld: output %lld outside t and partials", q, out);
        for (int e = 0; e < NB; ++e)
            if (++tw[static_cast<size_t>(out + e)] > 1)
                return violation("column descriptor %lld: entry %lld written twice", q, out + e);
    }
    for (size_t i = 0; i < H.red.size(); ++i) {
        const BRed &r = H.red[i];
        const long long q = static_cast<long long>(i);
        if (r.nchunk < 2 || r.len < 1 || r.pOff < 0 || r.pOff + static_cast<int64_t>(r.nchunk) * r.len > P)
            return violation("reduction %lld: partials at %lld outside the partial area", q, r.pOff);
        for (int64_t e = 0; e < static_cast<int64_t>(r.nchunk) * r.len; ++e)
            if (tw[static_cast<size_t>(T + r.pOff + e)] != 1)
                return violation("reduction %lld: partial %lld not written once", q, r.pOff + e);
        if (r.tOff < 0 || r.tOff + r.len > T) return violation("reduction %lld: t at %lld outside t", q, r.tOff);
        for (int e = 0; e < r.len; ++e)
            if (++tw[static_cast<size_t>(r.tOff + e)] > 1)
                return violation("reduction %lld: t entry %lld written twice", q, r.tOff + e);
    }
    // wavefront runs: each band's slots hold the record offsets of exactly
    // its own strips, split by wpos into ascending runs of the wavefronts
    if (static_cast<int64_t>(H.soffw.size()) != H.nstrips ||
        static_cast<int64_t>(H.wpos.size()) != nbands * (H.shape.nw + 1))
        return "wavefront run tables size";
    for (int64_t w = 0; w < nbands; ++w) {
        const int64_t t0 = w * band, n = std::min(H.nstrips, t0 + band) - t0;
        const uint32_t *wp = H.wpos.data() + w * (H.shape.nw + 1);
        if (wp[0] != 0 || static_cast<int64_t>(wp[H.shape.nw]) != n)
            return violation("band %lld: wavefront runs cover %lld of %lld strips", w, wp[H.shape.nw], n);
        for (int q = 0; q < H.shape.nw; ++q)
            if (wp[q + 1] < wp[q]) return violation("band %lld: wavefront %lld run ends before it starts", w, q);
        std::vector<uint32_t> a(H.soff.begin() + t0, H.soff.begin() + t0 + n),
            b(H.soffw.begin() + t0, H.soffw.begin() + t0 + n);
        std::sort(a.begin(), a.end());
        std::sort(b.begin(), b.end());
        if (a != b) return violation("band %lld: wavefront runs are not the band's strips", w);
    }
    // row pass: strip records
    const int64_t nrec = static_cast<int64_t>(H.recs.size());
    for (int64_t t = 0; t < H.nstrips; ++t) {
        const int64_t R = std::min<int64_t>(SROWS, nrows - t * SROWS);
        const int64_t w = t / band;
        const int32_t cmin = H.wgx[static_cast<size_t>(2 * w)], W = H.wgx[static_cast<size_t>(2 * w + 1)];
        int64_t off = H.soff[static_cast<size_t>(t)];
        for (int hop = 0;; ++hop) {
            if (hop > nrec / 8) return violation("strip %lld: record chain does not end", t);
            if (off % 4 || off + 128 > nrec) return violation("strip %lld: record at %lld (128 words read)", t, off);
            const uint32_t *r = H.recs.data() + off;
            if (r[0] != static_cast<uint32_t>(t * SROWS) || r[1] != static_cast<uint32_t>(R))
                return violation("strip %lld: record header rows %lld, %lld", t, r[0], r[1]);
            const int64_t nd = r[2], np = r[3], ib = r[4];
            if (np > 1 || SREC_HDR + 4 * (nd + np) > 128) return violation("strip %lld: %lld + %lld K-steps", t, nd, np);
            if (np == 1 && (ib < SREC_HDR + 4 * (nd + np) || ib + 32 > 128))
                return violation("strip %lld: index block at word %lld", t, ib);
            for (int64_t k = 0; k < nd + np; ++k) {
                const uint32_t afrag = r[SREC_HDR + 4 * k], src = r[SREC_HDR + 4 * k + 1],
                               info = r[SREC_HDR + 4 * k + 2];
                const int64_t roff = info & 15, rows = (info >> 4) & 31, nOct = (info >> 9) & 7, kind = info >> 12;
                const bool panel = k >= nd;
                if (kind != (panel ? SK_KIND_PANEL : 0) || nOct < 1 || nOct > 4 || rows < 1 || roff + rows > R ||
                    (panel && (roff != 0 || rows != R)))
                    return violation("strip %lld, K-step %lld: descriptor %lld", t, k, info);
                // FP64 lanes read 4 units from their first (nOct of their own)
                const int64_t tail = mode == HM_F64 ? 2 * (4 - nOct) : 0;
                if (static_cast<int64_t>(afrag) * unit + planes * nOct * rows * 8 + tail > frN)
                    return violation("strip %lld, K-step %lld: fragments at unit %lld outside", t, k, afrag);
                if (!panel) {
                    if (src % 8 || src >= static_cast<uint64_t>(ncols))
                        return violation("strip %lld, K-step %lld: X row %lld", t, k, src);
                    if (W > 0 && (static_cast<int64_t>(src) < cmin || static_cast<int64_t>(src) + 8 * nOct > cmin + W))
                        return violation("strip %lld, K-step %lld: X rows from %lld outside the band window", t, k,
                                         src);
                } else {
                    if (src != ib) return violation("strip %lld: panel index block at word %lld", t, src);
                    for (int e = 0; e < 32; ++e) {
                        const int64_t row = r[src + e];
                        if (row == zrow) continue;
                        if (row > zrow || (row + 1) * NB > T)
                            return violation("strip %lld: t row %lld outside t", t, row);
                        for (int c = 0; c < NB; ++c)
                            if (tw[static_cast<size_t>(row * NB + c)] != 1)
                                return violation("strip %lld: t row %lld not written once", t, row);
                    }
                }
            }
            const uint32_t next = r[5];
            if (next == 0) break;
            if (next <= off) return violation("strip %lld: chained record %lld does not follow %lld", t, next, off);
            off = next;
        }
    }
    return "";
}

}  // namespace hm
