/*
 * spm.h -- the subset of the SPM sparse-matrix container that the PaStiX C API
 * takes as input.
 *
 * The reference builds against the external SPM library (a git submodule that
 * is empty in the reference tree).  The fields listed here are exactly those
 * the reference's ordering, bcsc and sopalin code read from spmatrix_t, and the
 * enum values are the ones the PaStiX enums alias (PastixFloat..PastixComplex64
 * = 2..5, PastixGeneral/Symmetric/Hermitian = 111..113; reference
 * include/pastix/api.h:296-302,457-460).  The functions are this engine's own
 * implementations of the SPM routines the reference calls on the path
 * (spmNorm, spmMatVec, spmCheckAxb, spmGenRHS) plus grid-problem generators.
 */
#ifndef PASTIX_AMD_SPM_H
#define PASTIX_AMD_SPM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t spm_int_t;

typedef enum spm_coeftype_e {
    SpmPattern   = 0,
    SpmFloat     = 2,
    SpmDouble    = 3,
    SpmComplex32 = 4,
    SpmComplex64 = 5
} spm_coeftype_t;

typedef enum spm_mtxtype_e {
    SpmGeneral   = 111,
    SpmSymmetric = 112,
    SpmHermitian = 113
} spm_mtxtype_t;

typedef enum spm_fmttype_e {
    SpmCSC = 0,
    SpmCSR = 1,
    SpmIJV = 2
} spm_fmttype_t;

typedef enum spm_normtype_e {
    SpmOneNorm       = 171,
    SpmFrobeniusNorm = 174,
    SpmInfNorm       = 175,
    SpmMaxNorm       = 177
} spm_normtype_t;

/* Matrix input drivers (the SPM numbering that PaStiX's old_api.h aliases,
 * reference include/pastix/old_api.h:142-149). */
typedef enum spm_driver_e {
    SpmDriverRSA,
    SpmDriverHB,
    SpmDriverIJV,
    SpmDriverMM,
    SpmDriverSPM,
    SpmDriverLaplacian,
    SpmDriverXLaplacian,
    SpmDriverGraph
} spm_driver_t;

typedef enum spm_rhstype_e {
    SpmRhsOne,
    SpmRhsI,
    SpmRhsRndX,
    SpmRhsRndB
} spm_rhstype_t;

/*
 * Compressed sparse column matrix.  For symmetric and hermitian matrices only
 * one triangle is required (either one; both are accepted).  colptr has n+1
 * entries, rowptr and values have nnz entries, numbering starts at baseval.
 */
typedef struct spmatrix_s {
    spm_mtxtype_t  mtxtype;
    spm_coeftype_t flttype;
    spm_fmttype_t  fmttype;
    spm_int_t      baseval;
    spm_int_t      gN;
    spm_int_t      n;
    spm_int_t      gnnz;
    spm_int_t      nnz;
    spm_int_t      gNexp;
    spm_int_t      nexp;
    spm_int_t      gnnzexp;
    spm_int_t      nnzexp;
    spm_int_t      dof;       /* constant number of unknowns per node (>=1)      */
    spm_int_t     *dofs;      /* variable dof: unused by this engine (NULL)       */
    int            layout;    /* layout of the dof x dof element blocks          */
    spm_int_t     *colptr;
    spm_int_t     *rowptr;
    spm_int_t     *loc2glob;  /* distributed matrices: unused (NULL)             */
    void          *values;
    int            clustnum;
    int            clustnbr;
} spmatrix_t;

void   spmInit( spmatrix_t *spm );
void   spmExit( spmatrix_t *spm );
void   spmUpdateComputedFields( spmatrix_t *spm );
size_t spm_size_of( spm_coeftype_t flttype );

/* ||A|| of the full matrix (both triangles for symmetric storage). */
double spmNorm( spm_normtype_t ntype, const spmatrix_t *spm );

/* y = alpha * A * x + beta * y on the host, A expanded to its full pattern. */
int spmMatVec( int trans, double alpha, const spmatrix_t *spm,
               const void *x, double beta, void *y );
int spmMatMat( int trans, spm_int_t n, double alpha, const spmatrix_t *spm,
               const void *B, spm_int_t ldb, double beta, void *C, spm_int_t ldc );

/*
 * Right-hand sides.  SpmRhsRndX draws x uniformly in [-1,1) with the plrnt
 * generator seeded by `seed` (column j uses seed + j) and returns b = A x.
 */
int spmGenRHS( spm_rhstype_t type, spm_int_t nrhs, const spmatrix_t *spm,
               void *x, spm_int_t ldx, void *b, spm_int_t ldb );
int spmGenRHSSeed( spm_rhstype_t type, spm_int_t nrhs, const spmatrix_t *spm,
                   unsigned long long seed,
                   void *x, spm_int_t ldx, void *b, spm_int_t ldb );

/*
 * Check of a solution (the call of example/simple.c:140): returns 0 when the
 * backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) of every
 * right-hand side is at most eps (eps < 0 selects 1e-12) and, when x0 is
 * given, the forward error ||x - x0||_inf / ||x0||_inf is at most 1e4 * eps;
 * 1 otherwise (a NaN fails).
 */
int spmCheckAxb( double eps, spm_int_t nrhs, const spmatrix_t *spm,
                 const void *x0, spm_int_t ldx0,
                 const void *b, spm_int_t ldb,
                 const void *x, spm_int_t ldx );
/* max over the right-hand sides of ||x - x0||_inf / ||x0||_inf */
double spmForwardError( spm_int_t nrhs, const spmatrix_t *spm, const void *x0, spm_int_t ldx0,
                        const void *x, spm_int_t ldx );
double spmBackwardError( spm_int_t nrhs, const spmatrix_t *spm,
                         const void *b, spm_int_t ldb,
                         const void *x, spm_int_t ldx, double *berr_each );

/* Uniform values in (-0.5, 0.5] from the reference's 64-bit LCG (plrnt). */
void spm_plrnt( spm_int_t n, unsigned long long seed, unsigned long long shift, double *out );

/*
 * Read a matrix (example/simple.c:63-66): SpmDriverHB and SpmDriverRSA
 * (Harwell-Boeing files, assembled real / complex / pattern), SpmDriverMM
 * (Matrix Market coordinate), SpmDriverIJV ("n nnz" then "i j v" lines) and
 * SpmDriverLaplacian (the string "[<type>:]nx[:ny[:nz[:alpha[:beta]]]]": a
 * 7-point Laplacian with diagonal alpha * 2 * dims and off-diagonals -beta).
 * The result is CSC, 0-based, rows sorted, duplicates summed; symmetric and
 * Hermitian files keep the triangle they store.  Other drivers return
 * PASTIX_ERR_NOTIMPLEMENTED (3), unreadable files PASTIX_ERR_FILE (8).
 */
int spmReadDriver( spm_driver_t driver, const char *filename, spmatrix_t *spm );

/*
 * Grid generators (x-fastest numbering i + nx*j + nx*ny*k), Dirichlet boundary:
 *  - laplacian7: diagonal `diag`, off-diagonal -offd (7-point stencil), symmetric;
 *  - convdiff27: 27-point diffusion + upwind-free central convection with a
 *    per-cell velocity drawn from `seed`, general (unsymmetric values);
 *  - helmholtz7: complex symmetric 7-point Laplacian minus (kr + i*ki) I;
 *  - vlaplacian7: dof x dof blocks, lognormal cell coefficients from `seed`.
 */
int spmGenLaplacian7 ( spmatrix_t *spm, spm_coeftype_t flttype,
                       spm_int_t nx, spm_int_t ny, spm_int_t nz,
                       double diag, double offd );
int spmGenConvDiff27 ( spmatrix_t *spm, spm_int_t nx, spm_int_t ny, spm_int_t nz,
                       double diffusion, double vmax, unsigned long long seed );
int spmGenHelmholtz7 ( spmatrix_t *spm, spm_int_t nx, spm_int_t ny, spm_int_t nz,
                       double kr, double ki );
int spmGenVLaplacian7( spmatrix_t *spm, spm_int_t nx, spm_int_t ny, spm_int_t nz,
                       spm_int_t dof, double sigma, unsigned long long seed );

#ifdef __cplusplus
}
#endif

#endif /* PASTIX_AMD_SPM_H */
