/*
 * pastix_amd_engine.h -- introspection and control of the MI355X engine that
 * has no counterpart in the reference C API (the reference exposes its
 * SolverMatrix, bcsc and coeftab only through internal headers).  Tests and
 * the benchmark use it to compare the device factor with the CPU oracle on the
 * same structure and to read per-kernel timings.  Included by pastix.h.
 */
#ifndef PASTIX_AMD_ENGINE_H
#define PASTIX_AMD_ENGINE_H

/* Sizes of the solver matrix built by pastix_subtask_blend. */
int pastix_amd_solver_sizes( const pastix_data_t *pd, pastix_int_t *cblknbr,
                             pastix_int_t *bloknbr, pastix_int_t *coefnbr,
                             pastix_int_t *nodenbr );

/*
 * Export the solver matrix: cblk[6*(cblknbr+1)] = {fcolnum, lcolnum, fblok,
 * stride, coefoff, brownum} per cblk (last one is the sentinel) and
 * blok[5*bloknbr] = {frownum, lrownum, lcblknm, fcblknm, coefind}; browtab has
 * bloknbr-cblknbr entries.  Any output pointer may be NULL.
 */
int pastix_amd_solver_export( const pastix_data_t *pd, pastix_int_t *cblk,
                              pastix_int_t *blok, pastix_int_t *browtab );

/* Export the ordering (permtab/peritab: vertnbr each; rangtab cblknbr+1; treetab cblknbr). */
int pastix_amd_order_sizes ( const pastix_data_t *pd, pastix_int_t *vertnbr, pastix_int_t *cblknbr );
int pastix_amd_order_export( const pastix_data_t *pd, pastix_int_t *permtab, pastix_int_t *peritab,
                             pastix_int_t *rangtab, pastix_int_t *treetab );

/* Export the symbol matrix before blend (cblk: fcolnum,lcolnum,bloknum; blok: frownum,lrownum,lcblknm,fcblknm). */
int pastix_amd_symbol_sizes ( const pastix_data_t *pd, pastix_int_t *cblknbr, pastix_int_t *bloknbr );
int pastix_amd_symbol_export( const pastix_data_t *pd, pastix_int_t *cblk, pastix_int_t *blok );

/*
 * Symbol files (reference symbol/symbol_io.c:51-127, 147-175; the symbol
 * structure is internal to the reference too, symbol/symbol.h): version,
 * cblknbr, bloknbr, nodenbr, baseval, then (fcolnum, lcolnum, bloknum) per
 * cblk and (frownum, lrownum, fcblknm) per blok.  Return 0 on success.
 * pastix_subtask_symbfact loads "symbname" with IPARM_IO_STRATEGY
 * PastixIOLoad and writes "symbgen" into the output directory with
 * PastixIOSave; pastix_amd_symbol gives the instance's symbol.
 */
typedef struct px_symbol_s symbol_matrix_t;
int pastixSymbolLoad( symbol_matrix_t *symbptr, FILE *stream );
int pastixSymbolSave( const symbol_matrix_t *symbptr, FILE *stream );
symbol_matrix_t *pastix_amd_symbol( pastix_data_t *pd );
void pastixSymbolExit( symbol_matrix_t *symbptr );

/* iparm / dparm index and enum value by name, as pastixGetOptions parses
 * them (reference common/parse_options.c): -1 when unknown. */
int pastix_amd_parse_iparm( const char *name );
int pastix_amd_parse_dparm( const char *name );
int pastix_amd_parse_enum ( const char *name );

/* Export the permuted full-pattern CSC (bcsc) built by pastix_subtask_spm2bcsc. */
int pastix_amd_bcsc_sizes ( const pastix_data_t *pd, pastix_int_t *n, pastix_int_t *nnz );
int pastix_amd_bcsc_export( const pastix_data_t *pd, pastix_int_t *colptr, pastix_int_t *rowtab,
                            void *values );

/* Download one side (0: L, 1: U^T) of the device factor into a host array of coefnbr entries. */
int pastix_amd_get_coeftab( pastix_data_t *pd, int side, void *coeftab, size_t bytes );

/* Device statistics of the last factorization / solve (pxd_stats_t layout, 10 doubles). */
int pastix_amd_get_stats( pastix_data_t *pd, double *stats, int nstats );

/*
 * Engine options (keys):
 *   "berr_target"   refinement stops when the backward error
 *                   ||b-Ax||_inf/(||A||_inf||x||_inf+||b||_inf) of every RHS is
 *                   below the value (0: reference criterion ||r||_2/||b||_2 <= eps)
 *   "refine_batch"  1: refine all RHS together with batched device kernels
 *                   0: one RHS at a time as the reference does
 *   "device"        GPU ordinal used by the engine (default 0)
 *   any other key is forwarded to pxd_set_option of the device engine.
 */
int pastix_amd_set_option( pastix_data_t *pd, const char *key, double value );

/* Handle of the device engine (pxd_ctx_t *) for pxd_* calls on device buffers. */
int pastix_amd_device_handle( pastix_data_t *pd, void **ctx );

/*
 * Solve and refine on device-resident vectors in the permuted numbering
 * (n x nrhs, leading dimension n), i.e. what pastix_subtask_solve and
 * pastix_subtask_refine do once the right-hand sides are in HBM.
 */
int pastix_amd_solve_dev ( pastix_data_t *pd, pastix_int_t nrhs, void *dx, pastix_int_t ldx );
int pastix_amd_refine_dev( pastix_data_t *pd, pastix_int_t nrhs, void *db, void *dx,
                           pastix_int_t *iters );

#endif /* PASTIX_AMD_ENGINE_H */
