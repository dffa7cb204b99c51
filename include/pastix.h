/*
 * pastix.h -- PaStiX C API served by the MI355X engine.
 *
 * Every entry point keeps the reference name, signature and step semantics
 * (reference include/pastix.h:56-221).  The analysis steps (ordering,
 * symbolic factorization, reordering, blend) run on the host; numfact, solve
 * and refine run on one MI355X through the device layer declared in
 * pastix_amd_device.h.  There is no host fallback for the numerical steps:
 * when no GPU is usable they return PASTIX_ERR_NOTIMPLEMENTED.
 */
#ifndef PASTIX_AMD_PASTIX_H
#define PASTIX_AMD_PASTIX_H

#include <stdint.h>
#include <stddef.h>
#include "spm.h"
#include "pastix/api.h"
#include "pastix/order.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef float  _Complex pastix_complex32_t;
typedef double _Complex pastix_complex64_t;
typedef spm_coeftype_t pastix_coeftype_t;
typedef spm_mtxtype_t  pastix_mtxtype_t;

#define PastixPattern   SpmPattern
#define PastixFloat     SpmFloat
#define PastixDouble    SpmDouble
#define PastixComplex32 SpmComplex32
#define PastixComplex64 SpmComplex64
#define PastixGeneral   SpmGeneral
#define PastixSymmetric SpmSymmetric
#define PastixHermitian SpmHermitian


/*
 * Right-hand side descriptor (reference common/pastixdata.h:150-160).  In this
 * engine `b` is a host array in the permuted numbering; the engine keeps its
 * own device copy while solving.
 */
typedef struct pastix_rhs_s {
    int8_t             allocated;
    pastix_coeftype_t  flttype;
    pastix_int_t       m;
    pastix_int_t       n;
    pastix_int_t       ld;
    void              *b;
    void             **cblkb;
    void              *rhscomm;
    pastix_int_t      *Ploc2Pglob;
} *pastix_rhs_t;

/* Legacy single entry point (reference sopalin/pastix.c:103). */
int pastix( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
            pastix_int_t n, pastix_int_t *colptr, pastix_int_t *rowptr,
            void *values, pastix_int_t *perm, pastix_int_t *invp,
            void *B, pastix_int_t nrhs, pastix_int_t *iparm, double *dparm );

/* Lifecycle (reference common/api.c:410-914). */
void pastixInitParam       ( pastix_int_t *iparm, double *dparm );
void pastixInit            ( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
                             pastix_int_t *iparm, double *dparm );
void pastixInitWithAffinity( pastix_data_t **pastix_data, PASTIX_Comm pastix_comm,
                             pastix_int_t *iparm, double *dparm, const int *bindtab );
void pastixFinalize        ( pastix_data_t **pastix_data );

/* Tasks. */
int pastix_task_analyze         ( pastix_data_t *pastix_data, const spmatrix_t *spm );
int pastix_task_numfact         ( pastix_data_t *pastix_data, spmatrix_t *spm );
int pastix_task_solve           ( pastix_data_t *pastix_data, pastix_int_t m, pastix_int_t nrhs,
                                  void *B, pastix_int_t ldb );
int pastix_task_refine          ( pastix_data_t *pastix_data, pastix_int_t n, pastix_int_t nrhs,
                                  void *B, pastix_int_t ldb, void *X, pastix_int_t ldx );
int pastix_task_solve_and_refine( pastix_data_t *pastix_data, pastix_int_t n, pastix_int_t nrhs,
                                  void *B, pastix_int_t ldb, void *X, pastix_int_t ldx );

/* Analysis subtasks. */
int pastix_subtask_order     ( pastix_data_t *pastix_data, const spmatrix_t *spm,
                               pastix_order_t *myorder );
int pastix_subtask_symbfact  ( pastix_data_t *pastix_data );
int pastix_subtask_reordering( pastix_data_t *pastix_data );
int pastix_subtask_blend     ( pastix_data_t *pastix_data );

/* Numerical factorization subtasks. */
int pastix_subtask_spm2bcsc  ( pastix_data_t *pastix_data, spmatrix_t *spm );
int pastix_subtask_bcsc2ctab ( pastix_data_t *pastix_data );
int pastix_subtask_sopalin   ( pastix_data_t *pastix_data );

/* Solve subtasks. */
int pastix_subtask_applyorder( pastix_data_t *pastix_data, pastix_dir_t dir,
                               pastix_int_t m, pastix_int_t n, void *B, pastix_int_t ldb,
                               pastix_rhs_t Bp );
int pastix_subtask_trsm      ( pastix_data_t *pastix_data, pastix_side_t side,
                               pastix_uplo_t uplo, pastix_trans_t trans,
                               pastix_diag_t diag, pastix_rhs_t b );
int pastix_subtask_diag      ( pastix_data_t *pastix_data, pastix_rhs_t b );
int pastix_subtask_solve     ( pastix_data_t *pastix_data, pastix_rhs_t b );
int pastix_subtask_solve_adv ( pastix_data_t *pastix_data, pastix_trans_t transA, pastix_rhs_t b );
int pastix_subtask_refine    ( pastix_data_t *pastix_data, pastix_rhs_t b, pastix_rhs_t x );

/*
 * Schur complement (reference sopalin/schur.c): the unknowns of the list
 * (numbered from the spm's baseval) are ordered last; with
 * IPARM_SCHUR_FACT_MODE = PastixFactModeLocal they are not factorized and
 * pastixGetSchur returns S = A22 - A21 A11^{-1} A12 (lower triangle for
 * Cholesky and LDL^T, whole for LU); IPARM_SCHUR_SOLV_MODE selects the parts
 * the solves apply; pastixRhsSchurGet / Set move the Schur rows of a
 * permuted right-hand side.
 */
void pastixSetSchurUnknownList( pastix_data_t *pastix_data, pastix_int_t n, const pastix_int_t *list );
int  pastixGetSchur( const pastix_data_t *pastix_data, void *S, pastix_int_t lds );
int  pastixRhsSchurGet( const pastix_data_t *pastix_data, pastix_int_t m, pastix_int_t n, pastix_rhs_t rhsB,
                        void *B, pastix_int_t ldb );
int  pastixRhsSchurSet( const pastix_data_t *pastix_data, pastix_int_t m, pastix_int_t n, void *B,
                        pastix_int_t ldb, pastix_rhs_t rhsB );

/* Right-hand side helpers (reference sopalin/pastix_rhs.c). */
int pastixRhsInit          ( pastix_rhs_t *rhs );
int pastixRhsFinalize      ( pastix_rhs_t rhs );
int pastixRhsDoubletoSingle( const pastix_rhs_t dB, pastix_rhs_t sB );
int pastixRhsSingleToDouble( const pastix_rhs_t sB, pastix_rhs_t dB );

/* Diagonal of the factor (reference sopalin/diag.c:54). */
int pastixGetDiag( const pastix_data_t *pastix_data, void *x, pastix_int_t incx );

/* Constant-dof problem expanded to scalar unknowns: ordering, symbol and
 * matrix (reference sopalin/pastix.c:506-537). */
void pastixExpand( const pastix_data_t *pastix_data, spmatrix_t *spm );

/* Command-line options of the examples and tests (reference
 * common/get_options.c:149-330). */
void pastixGetOptions( int argc, char **argv, pastix_int_t *iparm, double *dparm, int *check,
                       int *scatter, spm_driver_t *driver, char **filename );

/* Parameters. */
void pastixDumpParam ( const pastix_data_t *pastix_data );
int  pastixCheckParam( const pastix_int_t *iparm, const double *dparm );

/*
 * Engine introspection used by tests and the benchmark (no counterpart in the
 * reference, which exposes its SolverMatrix only through internal headers).
 */
#include "pastix_amd_engine.h"
#include "pastix_amd_models.h"

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_PASTIX_H */
