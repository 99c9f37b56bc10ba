/*
 * FP64 CPU oracle for the H-matrix product path.  TEST INFRASTRUCTURE ONLY:
 * it is loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker / CPU restatement; the product path never
 * calls it.
 *
 * It restates, on the exported block table (kind,row0,col0,m,n,k,offA,offB;
 * see include/hmat_amd.h), the reference's arithmetic:
 *
 *   or_matvec    r += s*H*v          hmat_matvec!   src/harithm.jl:340-354
 *                  dense leaf: BLAS.gemm!('N','N',s,C,v,1.0,r)     line 342
 *                  rk leaf:    BLAS.gemm!('N','N',s,A,B'*v,1.0,r)  line 344
 *                  leaves in the table's depth-first order = the recursion's
 *                  child order [1,1],[1,2],[2,1],[2,2] (lines 348-351)
 *   or_matvec_mt the same product on several threads (row ownership),
 *                bitwise equal to or_matvec
 *   or_matmat    per-column loop     Base.:*(H, X)  src/harithm.jl:317-323
 *   or_matvec_t  r += s*H'*v         hmat_matvec2   src/harithm.jl:357-371
 *                  dense: s*v'*C (359); rk: s*(v'*A)*B' (361)
 *   or_to_dense  Array(H)            to_fmat        src/utils.jl:162-192
 *   or_adjoint   gradients of L = 0.5*||H*x - b||^2 (no reference
 *                counterpart): r = H*x - b, dx = H'*r, dC = r x',
 *                dA = r (B'x)', dB = x (A'r)'
 *
 * Inside one leaf the dot products run in ascending index order (the
 * reference's order inside OpenBLAS is not specified; parity is stated as a
 * relative norm, not bitwise).
 *
 * Pinned by tests/test_oracle.py against: brute-force dense reconstruction
 * (numpy), the closed-form kernel matrix (FullMat), the exactly rank-1
 * off-diagonal blocks of exp(-|x-y|) in 1-D, linearity, the identity
 * <H x, y> = <x, H' y>, and central finite differences of L.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define KIND_DENSE 0
#define KIND_RK 1

typedef struct {
    int64_t kind, row0, col0, m, n, k, offA, offB;
} blk_t;

static blk_t get(const int64_t *b, int64_t i) {
    blk_t x;
    memcpy(&x, b + 8 * i, sizeof x);
    return x;
}

/* r += s*H*v */
void or_matvec(int64_t nblocks, const int64_t *blocks, const double *data, const double *v,
               double *r, double s) {
    for (int64_t q = 0; q < nblocks; ++q) {
        blk_t b = get(blocks, q);
        if (b.kind == KIND_DENSE) {
            const double *C = data + b.offA;
            for (int64_t i = 0; i < b.m; ++i) {
                double acc = 0.0;
                for (int64_t j = 0; j < b.n; ++j) acc += C[i + j * b.m] * v[b.col0 + j];
                r[b.row0 + i] += s * acc;
            }
        } else {
            const double *A = data + b.offA, *B = data + b.offB;
            double *t = (double *)calloc(b.k > 0 ? (size_t)b.k : 1, sizeof(double));
            for (int64_t l = 0; l < b.k; ++l) {
                double acc = 0.0;
                for (int64_t j = 0; j < b.n; ++j) acc += B[j + l * b.n] * v[b.col0 + j];
                t[l] = acc;
            }
            for (int64_t i = 0; i < b.m; ++i) {
                double acc = 0.0;
                for (int64_t l = 0; l < b.k; ++l) acc += A[i + l * b.m] * t[l];
                r[b.row0 + i] += s * acc;
            }
            free(t);
        }
    }
}

/* r += s*H*v on nthreads OpenMP threads, for the CPU baseline's all-core
 * leg.  Phase 1 computes every low-rank leaf's t = B'v (leaves shared out
 * dynamically); phase 2 gives each thread a contiguous range of output rows,
 * and the thread walks the table in order, adding each leaf's contribution
 * to its own rows.  Every entry of t and every row sum is the same
 * expression in the same order as in or_matvec, so the result is bitwise
 * identical to it. */
void or_matvec_mt(int64_t nblocks, const int64_t *blocks, const double *data, const double *v,
                  double *r, double s, int64_t nrows, int nthreads) {
    int64_t tlen = 0;
    int64_t *toff = (int64_t *)malloc((size_t)(nblocks > 0 ? nblocks : 1) * sizeof(int64_t));
    for (int64_t q = 0; q < nblocks; ++q) {
        blk_t b = get(blocks, q);
        toff[q] = tlen;
        if (b.kind != KIND_DENSE) tlen += b.k;
    }
    double *t = (double *)calloc((size_t)(tlen > 0 ? tlen : 1), sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
#pragma omp for schedule(dynamic, 8)
        for (int64_t q = 0; q < nblocks; ++q) {
            blk_t b = get(blocks, q);
            if (b.kind == KIND_DENSE) continue;
            const double *B = data + b.offB;
            for (int64_t l = 0; l < b.k; ++l) {
                double acc = 0.0;
                for (int64_t j = 0; j < b.n; ++j) acc += B[j + l * b.n] * v[b.col0 + j];
                t[toff[q] + l] = acc;
            }
        }
#pragma omp for schedule(static, 1)
        for (int64_t th = 0; th < nthreads; ++th) {
            const int64_t lo = nrows * th / nthreads, hi = nrows * (th + 1) / nthreads;
            for (int64_t q = 0; q < nblocks; ++q) {
                blk_t b = get(blocks, q);
                const int64_t i0 = (lo > b.row0 ? lo : b.row0) - b.row0;
                const int64_t i1 = (hi < b.row0 + b.m ? hi : b.row0 + b.m) - b.row0;
                if (i0 >= i1) continue;
                if (b.kind == KIND_DENSE) {
                    const double *C = data + b.offA;
                    for (int64_t i = i0; i < i1; ++i) {
                        double acc = 0.0;
                        for (int64_t j = 0; j < b.n; ++j) acc += C[i + j * b.m] * v[b.col0 + j];
                        r[b.row0 + i] += s * acc;
                    }
                } else {
                    const double *A = data + b.offA;
                    const double *tb = t + toff[q];
                    for (int64_t i = i0; i < i1; ++i) {
                        double acc = 0.0;
                        for (int64_t l = 0; l < b.k; ++l) acc += A[i + l * b.m] * tb[l];
                        r[b.row0 + i] += s * acc;
                    }
                }
            }
        }
    }
    free(t);
    free(toff);
}

/* R[:, c] += s*H*X[:, c] for each column c */
void or_matmat(int64_t nblocks, const int64_t *blocks, const double *data, int64_t nrhs,
               const double *X, int64_t ldx, double *R, int64_t ldr, double s) {
    for (int64_t c = 0; c < nrhs; ++c)
        or_matvec(nblocks, blocks, data, X + c * ldx, R + c * ldr, s);
}

/* r += s*H'*v  (r has ncols entries, v nrows) */
void or_matvec_t(int64_t nblocks, const int64_t *blocks, const double *data, const double *v,
                 double *r, double s) {
    for (int64_t q = 0; q < nblocks; ++q) {
        blk_t b = get(blocks, q);
        if (b.kind == KIND_DENSE) {
            const double *C = data + b.offA;
            for (int64_t j = 0; j < b.n; ++j) {
                double acc = 0.0;
                for (int64_t i = 0; i < b.m; ++i) acc += v[b.row0 + i] * C[i + j * b.m];
                r[b.col0 + j] += s * acc;
            }
        } else {
            const double *A = data + b.offA, *B = data + b.offB;
            double *w = (double *)calloc(b.k > 0 ? (size_t)b.k : 1, sizeof(double));
            for (int64_t l = 0; l < b.k; ++l) {
                double acc = 0.0;
                for (int64_t i = 0; i < b.m; ++i) acc += v[b.row0 + i] * A[i + l * b.m];
                w[l] = acc;
            }
            for (int64_t j = 0; j < b.n; ++j) {
                double acc = 0.0;
                for (int64_t l = 0; l < b.k; ++l) acc += w[l] * B[j + l * b.n];
                r[b.col0 + j] += s * acc;
            }
            free(w);
        }
    }
}

/* out (nrows x ncols, column-major, ld = nrows) = dense H */
void or_to_dense(int64_t nrows, int64_t ncols, int64_t nblocks, const int64_t *blocks,
                 const double *data, double *out) {
    (void)ncols;
    for (int64_t q = 0; q < nblocks; ++q) {
        blk_t b = get(blocks, q);
        for (int64_t j = 0; j < b.n; ++j)
            for (int64_t i = 0; i < b.m; ++i) {
                double v;
                if (b.kind == KIND_DENSE) {
                    v = data[b.offA + i + j * b.m];
                } else {
                    v = 0.0;
                    for (int64_t l = 0; l < b.k; ++l)
                        v += data[b.offA + i + l * b.m] * data[b.offB + j + l * b.n];
                }
                out[(b.row0 + i) + (b.col0 + j) * nrows] = v;
            }
    }
}

/*
 * L = 0.5*||H x - b||^2.  Writes res = H x - b (nrows), dx = H' res (ncols)
 * and grad, laid out like data (dense leaf -> dC, rk leaf -> dA at offA and
 * dB at offB).  Entries of grad not covered by a leaf are left untouched.
 * Returns L.
 */
double or_adjoint(int64_t nrows, int64_t ncols, int64_t nblocks, const int64_t *blocks,
                  const double *data, const double *x, const double *b, double *res, double *dx,
                  double *grad) {
    for (int64_t i = 0; i < nrows; ++i) res[i] = -b[i];
    or_matvec(nblocks, blocks, data, x, res, 1.0);
    for (int64_t j = 0; j < ncols; ++j) dx[j] = 0.0;
    or_matvec_t(nblocks, blocks, data, res, dx, 1.0);
    double L = 0.0;
    for (int64_t i = 0; i < nrows; ++i) L += res[i] * res[i];
    for (int64_t q = 0; q < nblocks; ++q) {
        blk_t k = get(blocks, q);
        if (k.kind == KIND_DENSE) {
            for (int64_t j = 0; j < k.n; ++j)
                for (int64_t i = 0; i < k.m; ++i)
                    grad[k.offA + i + j * k.m] = res[k.row0 + i] * x[k.col0 + j];
        } else {
            const double *A = data + k.offA, *B = data + k.offB;
            for (int64_t l = 0; l < k.k; ++l) {
                double t = 0.0, w = 0.0;
                for (int64_t j = 0; j < k.n; ++j) t += B[j + l * k.n] * x[k.col0 + j];
                for (int64_t i = 0; i < k.m; ++i) w += A[i + l * k.m] * res[k.row0 + i];
                for (int64_t i = 0; i < k.m; ++i) grad[k.offA + i + l * k.m] = res[k.row0 + i] * t;
                for (int64_t j = 0; j < k.n; ++j) grad[k.offB + j + l * k.n] = x[k.col0 + j] * w;
            }
        }
    }
    return 0.5 * L;
}
