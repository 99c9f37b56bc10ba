// CPU driver of the host layout builder (hmatrix.jl_amd/csrc/hm_layout.cpp)
// for the sanitizer test: reads one block table, builds and checks every
// layout a plan would upload, and prints one line per layout.
//
//   layout_check TABLE [MEMBERS | -] [plant]
//
// TABLE: int64 nrows, ncols, nblocks, ndata, then 8 * nblocks int64 (the
// exported leaf table: kind, row0, col0, m, n, k, offA, offB), then ndata
// doubles.  MEMBERS (many-matrix plans): int64 nmem, then nmem + 1 first
// rows and nmem + 1 first columns (the last ones the totals).
// plant: also corrupt a copy of each layout in several ways (an offset past
// its arena, overlapping tiles, a bad multiplier, an entry written twice, a
// fragment, index or window outside its bounds, a record near the end) and
// require the checks to report every one.
// Exit status 0 when every check passes, 1 on a violation (or a planted
// mistake the checks miss), 2 on a bad input file.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/hmat_amd.h"
#include "hm_layout.h"

using namespace hm;

namespace {

bool read_all(const char *path, std::vector<char> &buf) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    return true;
}

int report(const char *what, const std::string &bad) {
    if (bad.empty()) return 0;
    std::printf("VIOLATION %s: %s\n", what, bad.c_str());
    return 1;
}

int planted = 0, caught = 0;

// A corrupted copy must fail its check.
void expect_caught(const char *what, const std::string &bad) {
    ++planted;
    if (!bad.empty()) ++caught;
    else std::printf("MISSED planted mistake: %s\n", what);
}

void plant_direction(const DirectionHost &D, int st, int64_t nrows, int64_t ncols, int64_t fac_elems) {
    if (!D.proj.empty()) {
        DirectionHost X = D;
        X.proj[0].base = fac_elems - X.proj[0].rows + 1;
        expect_caught("factor offset", check_direction(X, st, 4, nrows, ncols, fac_elems));
        X = D;
        X.proj.back().out = X.proj.front().out;
        if (X.proj.size() > 1) expect_caught("entry written twice", check_direction(X, st, 4, nrows, ncols, fac_elems));
    }
    if (!D.red.empty()) {
        DirectionHost X = D;
        X.red[0].tOff = X.tLen - X.red[0].kc + 1;
        expect_caught("reduction output", check_direction(X, st, 4, nrows, ncols, fac_elems));
    }
    if (D.tiles.size() > 1) {
        DirectionHost X = D;
        X.tiles[1].row0 = X.tiles[0].row0;
        expect_caught("overlapping tiles", check_direction(X, st, 4, nrows, ncols, fac_elems));
        // the tile whose values end the stream, one element later
        size_t last = 0;
        for (size_t i = 1; i < D.tiles.size(); ++i) {
            const TileDev &a = D.tiles[i], &b = D.tiles[last];
            if (a.base + int64_t(a.c1 - a.c0) * a.rows > b.base + int64_t(b.c1 - b.c0) * b.rows) last = i;
        }
        X = D;
        X.tiles[last].base += 1;
        if (X.tiles[last].c1 > X.tiles[last].c0)
            expect_caught("stream past its end", check_direction(X, st, 4, nrows, ncols, fac_elems));
    }
    for (size_t c = 0; c < D.mult.size(); ++c)
        if (D.mult[c] < 0) {
            DirectionHost X = D;
            X.mult[c] = ~static_cast<int32_t>(X.tLen);
            expect_caught("multiplier past t", check_direction(X, st, 4, nrows, ncols, fac_elems));
            break;
        }
}

void plant_batched(const BatchedHost &H, int st, int64_t nrows, int64_t ncols) {
    const int64_t unit = st == HM_F64 ? 2 : st == HM_F32 ? 4 : 8;
    bool dense_done = false, panel_done = false;
    for (int64_t t = 0; t < H.nstrips && !(dense_done && panel_done); ++t) {
        const uint32_t off = H.soff[static_cast<size_t>(t)];
        const uint32_t nd = H.recs[off + 2], np = H.recs[off + 3];
        if (nd > 0 && !dense_done) {
            dense_done = true;
            BatchedHost X = H;
            X.recs[off + SREC_HDR] = static_cast<uint32_t>(X.fr.size() / unit);
            expect_caught("fragment offset", check_batched(X, st, nrows, ncols));
            const int64_t w = t / (static_cast<int64_t>(H.shape.nw) * H.shape.spw);
            if (H.wgx[static_cast<size_t>(2 * w + 1)] > 8) {
                X = H;
                X.wgx[static_cast<size_t>(2 * w + 1)] = 8;
                X.wgx[static_cast<size_t>(2 * w)] = static_cast<int32_t>(X.recs[off + SREC_HDR + 1]) + 8;
                expect_caught("dense piece outside its window", check_batched(X, st, nrows, ncols));
            }
        }
        if (np > 0 && !panel_done) {
            panel_done = true;
            BatchedHost X = H;
            X.recs[off + X.recs[off + 4]] = static_cast<uint32_t>((H.tLen + H.partLen) / NB + 1);
            expect_caught("t row index", check_batched(X, st, nrows, ncols));
        }
    }
    for (size_t q = 0; 4 * q < H.pcol.size(); ++q)
        if (H.pcol[4 * q + 2] != 0xffffffffu) {
            BatchedHost X = H;
            X.pcol[4 * q + 2] = static_cast<uint32_t>(H.tLen + H.partLen - NB + 1);
            expect_caught("projection output", check_batched(X, st, nrows, ncols));
            X = H;
            X.pcol[4 * q] = static_cast<uint32_t>(H.vs.size() / 8);
            expect_caught("slice offset", check_batched(X, st, nrows, ncols));
            break;
        }
    if (H.nstrips > 0) {
        BatchedHost X = H;
        X.soff.back() = static_cast<uint32_t>(X.recs.size() - 64);
        expect_caught("record near the end", check_batched(X, st, nrows, ncols));
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    std::vector<char> buf;
    if (!read_all(argv[1], buf) || buf.size() < 32) return 2;
    const int64_t *h = reinterpret_cast<const int64_t *>(buf.data());
    const int64_t nrows = h[0], ncols = h[1], nblocks = h[2], ndata = h[3];
    if (buf.size() != static_cast<size_t>(32 + 64 * nblocks + 8 * ndata)) return 2;
    const int64_t *blocks = h + 4;
    const double *data = reinterpret_cast<const double *>(blocks + 8 * nblocks);
    std::vector<int64_t> mrow, mcol;
    const bool plant = argc > 3 && std::string(argv[3]) == "plant";
    if (argc > 2 && std::string(argv[2]) != "-") {
        std::vector<char> mb;
        if (!read_all(argv[2], mb) || mb.size() < 8) return 2;
        const int64_t *m = reinterpret_cast<const int64_t *>(mb.data());
        const int64_t nmem = m[0];
        if (mb.size() != static_cast<size_t>(8 * (1 + 2 * (nmem + 1)))) return 2;
        mrow.assign(m + 1, m + 2 + nmem);
        mcol.assign(m + 2 + nmem, m + 3 + 2 * nmem);
    }
    std::vector<OB> leaves(static_cast<size_t>(nblocks));
    int64_t rk_elems = 0, dense_elems = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
        const int64_t *r = blocks + 8 * b;
        leaves[b] = OB{static_cast<int>(r[0]), r[1], r[2], r[3], r[4], r[5], r[6], r[7]};
        if (r[0] == KIND_DENSE) dense_elems += r[3] * r[4];
        else rk_elems += (r[3] + r[4]) * r[5];
    }
    Oriented O;
    orient(leaves, data, rk_elems, dense_elems, O);
    const char *sname[] = {"f64", "f32", "bf16", "bf16x2"};
    int fails = 0;
    for (int st = HM_F64; st <= HM_BF16X2; ++st) {
        for (int chunk : {64, 1024})
            for (int ordered = 0; ordered < (mrow.empty() ? 1 : 2); ++ordered) {
                const std::vector<int64_t> none;
                DirectionHost F, B;
                layout_direction(st, chunk, 4, nrows, O.fw, O.rkv.data(), data, ordered ? mrow : none, F);
                layout_direction(st, chunk, 4, ncols, O.bw, O.rkv.data(), O.dT.data(), ordered ? mcol : none, B);
                char what[96];
                std::snprintf(what, sizeof what, "%s single-rhs H (chunk %d%s)", sname[st], chunk,
                              ordered ? ", one launch" : "");
                fails += report(what, check_direction(F, st, 4, nrows, ncols, rk_elems));
                std::snprintf(what, sizeof what, "%s single-rhs H' (chunk %d%s)", sname[st], chunk,
                              ordered ? ", one launch" : "");
                fails += report(what, check_direction(B, st, 4, ncols, nrows, rk_elems));
                if (plant) {
                    plant_direction(F, st, nrows, ncols, rk_elems);
                    plant_direction(B, st, ncols, nrows, rk_elems);
                }
            }
        for (int shape : {HM_STRIPS_AUTO, HM_STRIPS_NARROW, HM_STRIPS_WIDE}) {
            if (st == HM_F64 && shape != HM_STRIPS_AUTO) continue;     // FP64 plans have one shape
            BatchedHost HF, HB;
            const bool okF = layout_batched(st, shape, nrows, ncols, O.fw, O.rkv.data(), data, HF);
            const bool okB = layout_batched(st, shape, ncols, nrows, O.bw, O.rkv.data(), O.dT.data(), HB);
            char what[96];
            std::snprintf(what, sizeof what, "%s batched H (shape %d)", sname[st], shape);
            if (okF) fails += report(what, check_batched(HF, st, nrows, ncols));
            std::snprintf(what, sizeof what, "%s batched H' (shape %d)", sname[st], shape);
            if (okB) fails += report(what, check_batched(HB, st, ncols, nrows));
            if (plant && okF) plant_batched(HF, st, nrows, ncols);
            if (plant && okB) plant_batched(HB, st, ncols, nrows);
            std::printf("%s shape %d: batched H %d (%lld strips, %s), H' %d\n", sname[st], shape, okF ? 1 : 0,
                        static_cast<long long>(HF.nstrips),
                        st == HM_F64 ? "f64" : HF.shape.nw == ROW_WIDE.nw ? "wide" : "narrow",
                        okB ? 1 : 0);
        }
    }
    if (plant) std::printf("planted %d mistakes, caught %d\n", planted, caught);
    std::printf("%s\n", fails ? "FAILED" : "all layouts checked");
    return fails || caught != planted ? 1 : 0;
}
