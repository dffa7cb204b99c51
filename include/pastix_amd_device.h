/*
 * pastix_amd_device.h -- the C ABI between the host C library and the
 * hand-written HIP engine for one MI355X (gfx950).
 *
 * Everything here takes plain pointers, sizes and enums; nothing exposes HIP,
 * torch or C++ types.  Device buffers are opaque `void *` obtained from
 * pxd_malloc.  Every call returns 0 (PASTIX_SUCCESS) or a pastix_error_t code
 * and never falls back to host arithmetic.
 *
 * Which reference routine each entry point replaces:
 *   pxd_set_structure  -- SolverMatrix upload; the reference keeps it in host
 *                         memory (blend/solver.h:200-273) and, for its CUDA path,
 *                         copies blok tables per task (sopalin/parsec/
 *                         parsec_sparse_matrix.c:480-507).
 *   pxd_set_matrix     -- bcsc values (bcsc/bcsc.c:1887, bcsc/bcsc.h:119-133).
 *   pxd_factorize      -- coeftabInit + fill (sopalin/coeftab.c:203-250,
 *                         kernels/cpucblk_zinit.c:392-509, cpucblk_zcinit.c:80-170)
 *                         followed by sopalinFacto[factotype][flttype]
 *                         (sopalin/pastix_task_sopalin.c:51-58,500-508): panel
 *                         (kernels/core_z{potrf,getrf,sytrf}sp.c), TRSM
 *                         (kernels/core_ztrsmsp.c:355-399) and updates
 *                         (kernels/core_zgemmsp.c:401-472,1523-1607).
 *   pxd_solve          -- pastix_subtask_solve_adv (sopalin/pastix_task_solve.c:
 *                         510-677): forward/backward supernodal TRSM
 *                         (kernels/solve_ztrsmsp.c:300-625), LDL^T diagonal
 *                         (:629-679) and the mixed-precision RHS conversion
 *                         (sopalin/pastix_rhs.c:146-278).
 *   pxd_spmv           -- bcsc_zspmv (bcsc/bcsc_zspmv.c:629-678).
 *   pxd_vec_*          -- bvec_z* (bcsc/bvec_zcompute.c: nrm2 :71, dotc :561,
 *                         gemv :1112, axpy/scal/copy).
 */
#ifndef PASTIX_AMD_DEVICE_H
#define PASTIX_AMD_DEVICE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pxd_ctx_s pxd_ctx_t;

/* Column block of the solver matrix (one per cblk, plus a final sentinel). */
typedef struct pxd_cblk_s {
    int64_t fcolnum;   /* first column (permuted numbering)             */
    int64_t lcolnum;   /* last column                                   */
    int64_t fblok;     /* index of its diagonal blok                    */
    int64_t stride;    /* number of rows of the cblk                    */
    int64_t coefoff;   /* offset of its tiles in the coeftab (elements) */
    int64_t brownum;   /* first entry of its incoming blocks in browtab */
} pxd_cblk_t;

/* Block (tile) of the solver matrix (one per blok, plus a final sentinel). */
typedef struct pxd_blok_s {
    int64_t frownum;
    int64_t lrownum;
    int64_t lcblknm;   /* owning cblk                                    */
    int64_t fcblknm;   /* facing cblk                                    */
    int64_t coefind;   /* offset of the tile inside its cblk (elements)  */
} pxd_blok_t;

/* Timing / counters of the last factorization and solves. */
typedef struct pxd_stats_s {
    double fact_seconds;     /* device time of the whole factorization         */
    double fill_seconds;     /* zero + scatter of A into the coeftab             */
    double panel_seconds;    /* sum of diagonal-block kernels                    */
    double trsm_seconds;     /* sum of off-diagonal TRSM kernels                 */
    double update_seconds;   /* sum of update kernels                            */
    double solve_seconds;    /* last pxd_solve                                   */
    int64_t levels;          /* number of dependency levels                      */
    int64_t update_tiles;    /* output tiles of all update launches              */
    int64_t launches;        /* kernel launches of one factorization             */
    int64_t coef_bytes;      /* bytes of the factor storage                      */
} pxd_stats_t;

int  pxd_device_count( void );
int  pxd_create ( pxd_ctx_t **ctx, int device );
void pxd_destroy( pxd_ctx_t *ctx );
int  pxd_sync   ( pxd_ctx_t *ctx );
int  pxd_get_stats( const pxd_ctx_t *ctx, pxd_stats_t *stats );
/*
 * Engine knobs (returns 7 for an unknown key):
 *   "timing"        1: one stream, per-phase kernel times in the stats
 *   "model_log"     1: kernel timing log of the next factorization
 *   "kernels"       1: matrix-core kernels (default); 0: scalar kernels
 *   "panel_mw"      1: multi-workgroup Cholesky panels (default)
 *   "upd_big_rows"  stacked rows from which updates use the tall tiles
 *   "upd_small"     1: one-wavefront small updates (default)
 *   "upd_small_m", "upd_small_n"  largest stacked rows (up to 256, default
 *                   64) and facing columns (up to 64, default 16) of an
 *                   update that goes to the one-wavefront kernel (one
 *                   wavefront per 64 rows; at most 32 source columns)
 *   "upd_dma"       1: single-precision LL^T / LU tall tiles load their
 *                   operands straight into LDS (default); 0: register-staged
 *   "upd_dma_pad"   extra LDS bytes per workgroup of that kernel (0..65536,
 *                   default 4096: four workgroups per CU, room left for the
 *                   lookahead stream's panels)
 *   "fact_graph"    1: the factorization's level loop replays a graph
 *                   captured once per structure (default)
 *   "prio"          1: raised wave priority on the critical path (default)
 *   "solve_graph"   1: each solve pass replays the sweeps' launches from a
 *                   graph captured on first use (default); 0: launched one by one
 *   "solve_wide_max" level-solve launches with at least this many tiles load
 *                   16 columns per batch (default 128)
 *   "solve_small"   1: levels of cblks of at most 16 columns solve with one
 *                   wavefront per cblk, one launch per level (default)
 *   "solve_spread"  cblks with at least this many 64-row tiles (default 16;
 *                   0: none) gather the backward sweep's sums in 8 partial
 *                   copies instead of atomics into one; the records follow
 *                   at the next factorization, 0 takes effect at once
 *   "solve_split_max", "solve_split_wg"  smaller launches share each tile
 *                   among up to solve_split_max workgroups (default 4),
 *                   solve_split_wg workgroups in all (default 512)
 *   "wait_ms"       bound of the multi-workgroup panel's waits, in ms (2000);
 *                   a wait that reaches it gives up and the factorization
 *                   returns PASTIX_ERR_INTERNAL.  It is one value per device
 *                   (all its contexts)
 *   "schur_solv_mode" 0: local, 1: interface, 2: whole matrix (see pxd_set_schur)
 * The file tuning_mi355x.txt next to the library (or $PXD_TUNING_FILE) sets
 * options at context creation ("key value" lines, written by
 * pastix_amd/autotune.py).
 */
int  pxd_set_option( pxd_ctx_t *ctx, const char *key, double value );

int  pxd_malloc ( pxd_ctx_t *ctx, size_t bytes, void **dptr );
int  pxd_free   ( pxd_ctx_t *ctx, void *dptr );
int  pxd_memset0( pxd_ctx_t *ctx, void *dptr, size_t bytes );
int  pxd_h2d    ( pxd_ctx_t *ctx, void *dst, const void *src, size_t bytes );
int  pxd_d2h    ( pxd_ctx_t *ctx, void *dst, const void *src, size_t bytes );

/*
 * Upload the solver structure.  cblks has cblknbr+1 entries, bloks bloknbr+1,
 * browtab bloknbr-cblknbr.  n is the number of unknowns (rows of the RHS).
 */
int pxd_set_structure( pxd_ctx_t *ctx, int64_t n, int64_t cblknbr, int64_t bloknbr,
                       const pxd_cblk_t *cblks, const pxd_blok_t *bloks,
                       const int64_t *browtab, int64_t coefnbr );

/*
 * Schur complement (reference sopalin/schur.c, blend's CBLK_IN_SCHUR): the
 * cblks from column schurfcol on hold the Schur unknowns (schurfcol must be a
 * cblk boundary; n, the default, means none).  fact_local = 1
 * (PastixFactModeLocal): they receive the updates of the other cblks but are
 * not factorized, so their tiles hold S = A22 - A21 A11^{-1} A12 after
 * pxd_factorize.  The option "schur_solv_mode" (0 local, 1 interface, 2
 * whole matrix, reference pastix_solv_mode_t) selects what the solves apply.
 * Called after pxd_set_structure.
 */
int pxd_set_schur( pxd_ctx_t *ctx, int64_t schurfcol, int fact_local );

/*
 * Upload the permuted matrix P A P^T as a full-pattern CSC (both triangles,
 * rows sorted in every column).  flttype is the spm coefficient type of the
 * values (SpmDouble or SpmComplex64).  For general matrices the transpose is
 * given as a second CSC (tcolptr/trowtab/tvalues) and used by the SpMV.
 */
int pxd_set_matrix( pxd_ctx_t *ctx, int flttype, int mtxtype, int64_t n, int64_t nnz,
                    const int64_t *colptr, const int64_t *rowtab, const void *values,
                    const int64_t *tcolptr, const int64_t *trowtab, const void *tvalues );

/*
 * Numerical factorization on the device.  factotype: PastixFactPOTRF/GETRF/
 * SYTRF.  factor_flttype: SpmDouble/SpmFloat for real matrices, SpmComplex64/
 * SpmComplex32 for complex ones (single = mixed precision).  |pivot| <
 * threshold is replaced by threshold (static pivoting); the number of such
 * replacements is returned in *nbpivots.
 */
int pxd_factorize( pxd_ctx_t *ctx, int factotype, int factor_flttype,
                   double threshold, int64_t *nbpivots );

/* Download one side of the factor (0: L, 1: U^T) in the factor precision. */
int pxd_get_coeftab( pxd_ctx_t *ctx, int side, void *host, size_t bytes );
/* ... or the count entries from offset on (entries of the factor precision). */
int pxd_get_coeftab_range( pxd_ctx_t *ctx, int side, int64_t offset, int64_t count, void *host );

/*
 * x <- A^{-1} x in place for nrhs columns of a device array in the vector
 * precision (double / complex double), leading dimension ldx.  With a single
 * precision factor the conversion to single and back is fused into the sweeps.
 */
int pxd_solve( pxd_ctx_t *ctx, int64_t nrhs, void *dx, int64_t ldx );

/*
 * pxd_solve without any allocation: no counterpart in the reference.  It is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * sweeps' scratch (one pass of at most 8 columns of the solve's vectors,
 * plus their single-precision copies for a single factor) lives in a static
 * device array of PXD_NOALLOC_BYTES per GPU, held until the process ends; a
 * solve that needs more returns 7 (PASTIX_ERR_BADPARAMETER) with a message,
 * and pxd_solve is then the way.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (from several threads or instances) race.
 * Same arguments, results and return codes as pxd_solve otherwise; no solve
 * time is recorded in the statistics.
 */
#define PXD_NOALLOC_BYTES ( (size_t)32 << 20 )
int pxd_solve_noalloc( pxd_ctx_t *ctx, int64_t nrhs, void *dx, int64_t ldx );

/*
 * x <- op(A)^{-1} x with op(A) = A (trans 0) or A^T (trans 1): for an LU
 * factor U^T y = b then L^T x = y (reference sopalin/pastix_task_solve.c:
 * 510-677, the transA cases of its table); Cholesky and LDL^T factors are
 * symmetric and solve A^T as A.
 */
int pxd_solve_trans( pxd_ctx_t *ctx, int trans, int64_t nrhs, void *dx, int64_t ldx );

/*
 * One step of the solve, in place on nrhs device vectors (vector precision),
 * as pastix_subtask_trsm / pastix_subtask_diag apply it (reference
 * sopalin/pastix_task_solve.c:296-360,417-470 -> sopalin_ztrsm,
 * sopalin/sequential_ztrsm.c:591-652, and sopalin_zdiag):
 *   PXD_STEP_L_FWD   L y = b       (unit for LU and LDL^T, non-unit for Cholesky)
 *   PXD_STEP_LT_BWD  L^T x = y
 *   PXD_STEP_U_BWD   U x = y       (LU only)
 *   PXD_STEP_UT_FWD  U^T y = b     (LU only)
 *   PXD_STEP_DIAG    x <- D^{-1} x, D the diagonal of the L-side diagonal
 *                    tiles (the D of LDL^T; solve_cblk_zdiag,
 *                    kernels/solve_ztrsmsp.c:629-679, divides by it for any
 *                    factorization)
 * A step that does not exist for the factorization returns 7.
 */
enum { PXD_STEP_L_FWD = 0, PXD_STEP_LT_BWD = 1, PXD_STEP_U_BWD = 2, PXD_STEP_UT_FWD = 3, PXD_STEP_DIAG = 4 };
int pxd_solve_step( pxd_ctx_t *ctx, int step, int64_t nrhs, void *dx, int64_t ldx );

/*
 * dd[i] (device, n entries in the vector precision: double or complex
 * double) = diagonal entry i of the factor's diagonal tiles in the permuted
 * numbering: L_ii (Cholesky), D_i (LDL^T), U_ii (LU) (reference
 * coeftab_zgetdiag, sopalin/coeftab_z.c:659-689, behind pastixGetDiag).
 */
int pxd_get_diag( pxd_ctx_t *ctx, void *dd );

/* y <- alpha A x + beta y (nrhs columns). alpha/beta are real. */
int pxd_spmv( pxd_ctx_t *ctx, double alpha, const void *dx, double beta, void *dy,
              int64_t nrhs, int64_t ld );

/*
 * Block vector operations on n x nrhs device arrays (vector precision).
 * Scalars per column live in device memory (dalpha: nrhs entries, complex as
 * pairs) so that a Krylov iteration runs without host round trips.
 *   dot:   dres[c]  = x(:,c)^H y(:,c)
 *   nrm2:  dres[c]  = ||x(:,c)||_2            (real output)
 *   axpy:  y(:,c)  += sign * dalpha[c] * x(:,c)
 *   scal:  x(:,c)  *= dalpha[c]              (or 1/dalpha[c] when inv)
 *   copy:  y <- x
 *   gemv:  x(:,c) += sum_j V_j(:,c) * dy[j*nrhs + c]   (V_j blocks of n x nrhs, ld n)
 *   maxabs: dres[c] = ||x(:,c)||_inf
 */
int pxd_vec_dot   ( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, const void *y, void *dres );
int pxd_vec_nrm2  ( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, double *dres );
int pxd_vec_maxabs( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, double *dres );
int pxd_vec_axpy  ( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, double sign, const void *dalpha,
                    const void *x, void *y );
int pxd_vec_scal  ( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *dalpha, int inv, void *x );
int pxd_vec_copy  ( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, const void *x, void *y );
int pxd_vec_gemv  ( pxd_ctx_t *ctx, int64_t n, int64_t nrhs, int64_t m, const void *V,
                    const void *dy, void *x );
/* Residual-norm helpers: ||A||_inf of the uploaded matrix. */
int pxd_matrix_norm_inf( pxd_ctx_t *ctx, double *norm );

/*
 * Kernel timing log of the last factorization run with the option
 * "model_log" = 1: one record per panel, per TRSM of a cblk and per blok
 * update, as the reference's kernel_trace_stop records them
 * (kernels/kernels_trace.h:134-177; written out as model.csv lines
 * Kernel;M;N;K;Time by kernels/kernels_trace.c:222-249).  ktype is a
 * pastix_ktype_t (pastix/kernels_enums.h).  time is in seconds: the span from
 * the first workgroup's start to the last one's end, with the task alone on
 * the device.  Copies at most max records; *count gets the total.
 */
typedef struct pxd_model_entry_s {
    int32_t ktype, m, n, k;
    double  time;
    int32_t variant;   /* blok updates: the kernel class that ran them (0: 64 x 64
                          tiles, 1: tall tiles, 2: one wavefront); -1 otherwise */
    int32_t pad;
} pxd_model_entry_t;

/*
 * Kernel class of each blok update, chosen by the caller (for instance from
 * predicted times, pastix_amd_choose_update_kernels): cls[2 * b + side] for
 * the update from blok b (side 1: the U side of an LU factor), with the
 * values of pxd_model_entry_t.variant, or -1 to keep the built-in rule
 * (the cut-offs upd_small / upd_big_rows).  A class the update does not fit
 * (the one-wavefront kernel takes at most 64 rows, 16 columns and 32 k
 * values) falls back to the rule.  count must be 2 * bloknbr, or 0 to drop
 * the choices.  Takes effect at the next factorization.
 */
int pxd_set_update_classes( pxd_ctx_t *ctx, const int8_t *cls, int64_t count );
int pxd_model_log( const pxd_ctx_t *ctx, pxd_model_entry_t *entries, int64_t max, int64_t *count );

/*
 * Learned cost model (no counterpart in the reference, whose costs are the
 * polynomials of common/models.h:38-83): a multilayer perceptron with 8
 * inputs, two tanh layers of 32 units and a linear output, trained on the GPU
 * by full-batch Adam on the mean squared error of the uploaded rows.
 * Parameters (pxd_mlp_nparams() floats): W1[8][32], b1[32], W2[32][32],
 * b2[32], w3[32], b3.  X is row-major n x 8.
 */
typedef struct pxd_mlp_s pxd_mlp_t;
int  pxd_mlp_nparams   ( void );
int  pxd_mlp_create    ( pxd_mlp_t **mlp, int device, uint64_t seed );
void pxd_mlp_destroy   ( pxd_mlp_t *mlp );
int  pxd_mlp_get_params( pxd_mlp_t *mlp, float *params );
int  pxd_mlp_set_params( pxd_mlp_t *mlp, const float *params );   /* also resets Adam */
int  pxd_mlp_set_data  ( pxd_mlp_t *mlp, const float *X, const float *t, int64_t n );
/* loss = sum_r (y_r - t_r)^2 / n over the uploaded rows, and its gradient */
int  pxd_mlp_loss_grad ( pxd_mlp_t *mlp, double *loss, float *grad );
/* steps Adam steps (beta 0.9, 0.999, eps 1e-8) at rate lr; loss after them */
int  pxd_mlp_train     ( pxd_mlp_t *mlp, int steps, double lr, double *loss );
int  pxd_mlp_predict   ( pxd_mlp_t *mlp, const float *X, int64_t n, float *y );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_DEVICE_H */
