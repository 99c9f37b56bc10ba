/*
 * hmat_amd.h -- C ABI of the MI355X H-matrix product library (libhmat_amd.so).
 *
 * The library evaluates products with a hierarchical matrix whose leaves were
 * built and compressed on the host (FP64) and handed over as a block table
 * plus one FP64 factor arena ("exported factors").  Every entry point takes
 * plain pointers and sizes, returns an int status (HM_OK == 0, negative on
 * error, message in hm_last_error()), never aborts and never throws.
 *
 * Block table: int64 row-major array of nblocks rows x 8 columns
 *   { kind, row0, col0, m, n, k, offA, offB }
 *   kind 0 = dense leaf C (m x n, column-major at data[offA]);
 *   kind 1 = low-rank leaf A*B' with A (m x k) at data[offA] and B (n x k) at
 *            data[offB], both column-major.
 *   row0/col0 are 0-based positions of the leaf in cluster order; the leaves
 *   must tile the nrows x ncols matrix exactly once.
 * This is the data of the reference's Hmat leaves (src/cluster.jl:21-34:
 * C, A, B, m, n) with the tree replaced by explicit offsets.
 *
 * Vectors are device pointers.  Their element type follows the plan's
 * storage: double for HM_F64 plans, float for every other storage.
 * Products accumulate in double (HM_F64) or float (all others).
 *
 * Threading: one plan may be used from one stream at a time (its scratch
 * buffers are per plan); distinct plans are independent.  A plan lives on
 * the device that was current at hm_plan_create.
 */
#ifndef HMAT_AMD_H
#define HMAT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hm_plan_s *hm_plan_t;

/* storage precision of leaves and factors on the device */
enum hm_storage {
    HM_F64 = 0,    /* double storage, double vectors and accumulation          */
    HM_F32 = 1,    /* float storage, float vectors, float accumulation         */
    HM_BF16 = 2,   /* bfloat16 storage (round to nearest even), float vectors  */
    HM_BF16X2 = 3  /* two bfloat16 planes hi + lo (value = hi + lo), float vec */
};

enum hm_status {
    HM_OK = 0,
    HM_EINVAL = -1,       /* bad argument, inconsistent block table          */
    HM_EHIP = -2,         /* HIP runtime error                               */
    HM_ENOMEM = -3,       /* device allocation failed                         */
    HM_EUNSUPPORTED = -4  /* valid request the library does not implement     */
};

/* fields reported by hm_plan_info, in this order */
enum hm_info_field {
    HM_INFO_NROWS = 0,
    HM_INFO_NCOLS,
    HM_INFO_NDENSE,          /* dense leaves                                   */
    HM_INFO_NRK,             /* low-rank leaves                                */
    HM_INFO_DENSE_ELEMS,     /* sum m*n over dense leaves                      */
    HM_INFO_RK_ELEMS,        /* sum (m+n)*k over low-rank leaves               */
    HM_INFO_STORAGE,         /* enum hm_storage                                */
    HM_INFO_DEVICE_BYTES,    /* device memory held by the plan                 */
    HM_INFO_FACTOR_BYTES,    /* bytes of leaves+factors one product streams   */
    HM_INFO_MAXRANK,
    HM_INFO_PROJ_CHUNK,      /* hm_config.proj_chunk in effect                   */
    HM_INFO_ROW_WAVES,       /* hm_config.row_waves in effect                    */
    HM_INFO_FLAGS,           /* hm_config.flags                                  */
    HM_INFO_STRIP_SHAPE,     /* enum hm_strip_shape of the H X row pass in use    */
    HM_INFO_BATCHED,         /* 1: the multi-right-hand-side MFMA layout was built
                                for H (and H' unless HM_NO_TRANSPOSE); 0: it was
                                not (FP64 storage, HM_NO_BATCHED, fewer than 8
                                columns, or slice offsets past 2^32), and
                                hm_matmat runs panels of 8 columns on the
                                single-right-hand-side layout instead          */
    HM_INFO_FUSED,           /* 1: a single-right-hand-side product of this plan
                                (hm_plan_create_batch with HM_ONE_LAUNCH) runs
                                as one launch whose row tiles wait on
                                per-matrix counters for the projections they
                                read                                          */
    HM_INFO_WAIT_FAULTS,     /* nonzero: a bounded wait of that one launch ran
                                out (products since then are unreliable; the
                                plan must be recreated). Reading this field
                                synchronises the device                       */
    HM_INFO_NFIELDS
};

/* hm_config.flags */
enum hm_flags {
    HM_NO_TRANSPOSE = 1,     /* do not build the H' structures (halves memory)   */
    HM_NO_BATCHED = 2,       /* do not build the multi-right-hand-side layout    */
    HM_ONE_LAUNCH = 4        /* plans of hm_plan_create_batch: run a single-
                                right-hand-side product as one launch (row
                                tiles wait on per-matrix counters) instead of
                                three (projection, reduction, row tiles)       */
};

/* hm_config.strip_shape: workgroups of the multi-right-hand-side row pass */
enum hm_strip_shape {
    HM_STRIPS_AUTO = 0,      /* narrow unless a tenth of its X windows overflow */
    HM_STRIPS_NARROW = 1,    /* 8 strips of 16 rows, X window of <= 512 rows    */
    HM_STRIPS_WIDE = 2,      /* 16 strips of 16 rows, X window of <= 1024 rows  */
    HM_STRIPS_F64 = 3        /* reported for FP64 plans, which always take it:
                                16 strips of 16 rows, X window of <= 512 rows
                                of doubles, one workgroup per CU               */
};

/*
 * Plan configuration (zero fields take the defaults).  These are the shapes
 * the autotuner (hmatrix_amd.autotune) sweeps.
 *   proj_chunk:  rows of B one projection work item streams in the
 *                single-right-hand-side product (multiple of 64, default 1024)
 *   row_waves:   wavefronts per 64-row tile in the single-right-hand-side
 *                row pass: 2, 4 (default) or 8
 *   flags:       enum hm_flags
 *   strip_shape: enum hm_strip_shape (multi-right-hand-side row pass)
 */
typedef struct hm_config {
    int32_t proj_chunk;
    int32_t row_waves;
    int32_t flags;
    int32_t strip_shape;
} hm_config;

/* products a captured graph replays */
enum hm_op { HM_OP_N = 0, HM_OP_T = 1 };

typedef struct hm_graph_s *hm_graph_t;

/*
 * Create a plan: validate the block table, pack the factors into the device
 * layout of the requested storage and upload them.  The host arrays are only
 * read during the call.
 * Replaces: the Hmat tree handed to hmat_matvec! (src/harithm.jl:340) as
 * built by NewHmat (src/cluster.jl:36-43).
 */
int hm_plan_create(int64_t nrows, int64_t ncols, int64_t nblocks,
                   const int64_t *blocks, const double *data, int64_t ndata,
                   int storage, hm_plan_t *plan);

/* hm_plan_create with an explicit configuration (cfg may be NULL). */
int hm_plan_create_ex(int64_t nrows, int64_t ncols, int64_t nblocks,
                      const int64_t *blocks, const double *data, int64_t ndata,
                      int storage, const hm_config *cfg, hm_plan_t *plan);

/*
 * Time candidate configurations for one matrix and return the fastest (no
 * counterpart in the reference: the plan-shape autotuner, for callers that
 * tune from Julia or C).  nrhs == 1: proj_chunk in {256, 512, 1024, 2048,
 * 4096} x row_waves in {2, 4, 8} of the single-right-hand-side product;
 * 2 <= nrhs <= 32: FP32, BF16 or split-BF16 storage: strip_shape narrow
 * and wide of the multi-right-hand-side row pass; FP64 storage: the f64
 * MFMA layout, or (HM_NO_BATCHED) panels of 8 columns on the single-column
 * layout with proj_chunk in {256, 1024, 4096} x row_waves in {2, 4, 8}.
 * Each candidate plan is created, run 3 times untimed and reps times timed
 * with HIP events on a stream of the call's own, then destroyed; the call
 * allocates and frees its input and output buffers (zeros).  best receives
 * the configuration (flags 0, or HM_NO_BATCHED when that layout won a
 * several-column product); best_ms (may be NULL) its milliseconds per
 * product.
 */
int hm_autotune(int64_t nrows, int64_t ncols, int64_t nblocks, const int64_t *blocks,
                const double *data, int64_t ndata, int storage, int64_t nrhs, int reps,
                hm_config *best, double *best_ms);

/*
 * One plan for nmat independent H-matrices: member i's block table and arena
 * (same format as hm_plan_create) are placed on the diagonal of a
 * (sum nrows) x (sum ncols) matrix, so one product launch evaluates every
 * member; member i's vectors start at the prefix sums of nrows / ncols.
 * Replaces: a loop of H*y over separate Hmat objects (src/harithm.jl:325-329).
 */
int hm_plan_create_batch(int64_t nmat, const int64_t *nrows, const int64_t *ncols,
                         const int64_t *nblocks, const int64_t *const *blocks,
                         const double *const *data, const int64_t *ndata, int storage,
                         const hm_config *cfg, hm_plan_t *plan);

/* Release a plan and its device memory.  NULL is accepted. */
int hm_plan_destroy(hm_plan_t plan);

/* Copy min(nout, HM_INFO_NFIELDS) fields of enum hm_info_field into out. */
int hm_plan_info(hm_plan_t plan, int64_t *out, int nout);

/*
 * y = beta*y + alpha*H*x.  x has ncols entries, y nrows; both device
 * pointers on the plan's device; x and y must not overlap.  beta == 0 does
 * not read y.  stream: hipStream_t or NULL for the null stream.
 * Replaces: hmat_matvec!(r, H, v, s) == (alpha = s, beta = 1)
 * (src/harithm.jl:340-354) and Base.:*(H, v) == (alpha = 1, beta = 0)
 * (src/harithm.jl:325-329).
 */
int hm_matvec(hm_plan_t plan, const void *x, void *y,
              double alpha, double beta, void *stream);

/*
 * hm_matvec that allocates nothing and uses no scratch memory of the plan.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, result and kernels as
 * hm_matvec; the projection's t vector and partial sums live in a
 * library-wide static device array of HM_STATIC_SCRATCH_BYTES per GPU
 * (held until the process ends) instead of the plan's buffer.  A plan whose
 * (sum of ranks + chunk partials) * element size exceeds that maximum is
 * rejected with HM_EUNSUPPORTED.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (two streams, or two threads) race
 * and must be serialised by the caller.  Not recorded by hm_plan_timing.
 */
#define HM_STATIC_SCRATCH_BYTES (4u << 20)
int hm_matvec_static(hm_plan_t plan, const void *x, void *y,
                     double alpha, double beta, void *stream);

/*
 * Y = beta*Y + alpha*H*X for nrhs right-hand sides, column-major X
 * (ncols x nrhs, leading dimension ldx) and Y (nrows x nrhs, ldy).  FP32,
 * BF16 and split-BF16 plans run panels of 32 columns on the batched MFMA
 * layout; FP64 plans (and plans built with HM_NO_BATCHED) panels of 8
 * columns per pass over the single-right-hand-side layout; one column is
 * hm_matvec.
 * Replaces: Base.:*(H, X::Matrix) (src/harithm.jl:317-323), which loops
 * hmat_matvec! over the columns.
 */
int hm_matmat(hm_plan_t plan, int64_t nrhs, const void *X, int64_t ldx,
              void *Y, int64_t ldy, double alpha, double beta, void *stream);

/*
 * y = beta*y + alpha*H'*x (x has nrows entries, y ncols).
 * Replaces: hmat_matvec2 (src/harithm.jl:357-371) behind v*H
 * (src/harithm.jl:331-337) and transpose_hmat_mat_mul (src/harithm.jl:243-260).
 */
int hm_matvec_t(hm_plan_t plan, const void *x, void *y,
                double alpha, double beta, void *stream);

/* Y = beta*Y + alpha*H'*X, column-major, as hm_matmat. */
int hm_matmat_t(hm_plan_t plan, int64_t nrhs, const void *X, int64_t ldx,
                void *Y, int64_t ldy, double alpha, double beta, void *stream);

/*
 * Adjoint of the least-squares loss L = 0.5*||H x - b||^2 (no counterpart in
 * the reference; it differentiates the product of src/harithm.jl:340-354):
 *   res  (nrows) = H x - b
 *   dx   (ncols) = H' res                     (dL/dx)
 *   grad (ndata) in the layout of the exported arena passed to
 *        hm_plan_create: dense leaf C -> dC = res[rows] x[cols]' at its offA,
 *        low-rank leaf -> dA = res[rows] (B'x)' at offA and
 *        dB = x[cols] (A'res)' at offB; entries no leaf covers are not
 *        written.  grad may be NULL (no factor gradients).
 *   loss (host, may be NULL): L; when given the call waits for the stream.
 * Vector and gradient element type is the plan's vector type (double for
 * HM_F64, float otherwise).  All pointers except loss are device pointers;
 * res, dx and grad must not overlap x, b or each other (HM_EINVAL).
 */
int hm_adjoint(hm_plan_t plan, const void *x, const void *b, void *res, void *dx, void *grad,
               double *loss, void *stream);

/*
 * Pullback of y = H x (the product of src/harithm.jl:340-354) for a
 * cotangent g of y: dx = H' g and, when grad is not NULL, the gradient of
 * <g, H x> with respect to every leaf entry in the caller's arena layout
 * (dense C -> g x', low-rank A -> g (B'x)', B -> x (A'g)'; entries of grad
 * outside the leaves are not written).  x, g, dx and grad are device
 * arrays of the accumulation type (double for HM_F64 plans, float
 * otherwise); dx must not overlap x or g, grad none of the vectors
 * (HM_EINVAL).  The plan must have its H' structures (not
 * HM_NO_TRANSPOSE).  This is the factor tangent of a ChainRulesCore rrule
 * for `*` (no rrule exists in the reference); hm_adjoint is the special
 * case g = H x - b.
 */
int hm_vjp(hm_plan_t plan, const void *x, const void *g, void *dx, void *grad, void *stream);

/*
 * Capture one product (op = HM_OP_N: Y = beta*Y + alpha*H*X, HM_OP_T: with
 * H') on fixed device buffers into a HIP graph; hm_graph_launch replays it on
 * a stream with one launch.  The graph holds only kernel nodes.  The plan and
 * the buffers must outlive the graph; replays on one plan must not overlap.
 * No reference counterpart (launch-overhead form of hm_matmat).
 */
int hm_graph_create(hm_plan_t plan, int op, int64_t nrhs, const void *X, int64_t ldx,
                    void *Y, int64_t ldy, double alpha, double beta, hm_graph_t *graph);
int hm_graph_launch(hm_graph_t graph, void *stream);
int hm_graph_destroy(hm_graph_t graph);

/*
 * Per-kernel timing (instrumentation; no reference counterpart).  With
 * enable != 0 every later product on this plan (hm_matvec, hm_matmat and the
 * transposed forms; not graph replays, not hm_adjoint) records HIP events on
 * its stream around its three kernels: projection, reduction, row pass.
 * hm_plan_timing_read waits for the recorded events, writes the summed
 * milliseconds of those phases into ms[0..min(nms,3)-1] and the number of
 * launch sequences (one per panel of 32 right-hand sides, or per column on
 * the single-right-hand-side path) into *count (may be NULL), and clears the
 * sums.  Disabling folds pending records first.
 */
int hm_plan_timing(hm_plan_t plan, int enable);
int hm_plan_timing_read(hm_plan_t plan, double *ms, int nms, int64_t *count);

/* Message of the last failing call on this host thread ("" if none). */
const char *hm_last_error(void);

/* Library version string. */
const char *hm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HMAT_AMD_H */
