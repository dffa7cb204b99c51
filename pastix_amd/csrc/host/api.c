/*
 * api.c -- the PaStiX C API on top of the host analysis and the device engine.
 *
 * Lifecycle and parameter defaults follow common/api.c:410-914 of the
 * reference; the task/subtask sequence and the step bitmask checks follow
 * order/pastix_subtask_order.c:97-565, symbol/pastix_subtask_symbfact.c:85-343,
 * symbol/pastix_subtask_reordering.c:59-143, blend/pastix_subtask_blend.c:
 * 120-635, sopalin/pastix_task_sopalin.c:92-693, sopalin/pastix_task_solve.c
 * and refinement/pastix_task_refine.c.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <float.h>
#include <unistd.h>
#include <sched.h>
#include "px_internal.h"

int px_refine_run( pastix_data_t *pd, pastix_int_t nrhs, void *db, void *dx, pastix_int_t *iters );

void
pastixInitParam( pastix_int_t *iparm, double *dparm )
{
    memset( iparm, 0, IPARM_SIZE * sizeof( pastix_int_t ) );
    memset( dparm, 0, DPARM_SIZE * sizeof( double ) );
    iparm[IPARM_VERBOSE]     = PastixVerboseNo;
    iparm[IPARM_IO_STRATEGY] = PastixIONo;
    iparm[IPARM_TRACE]       = PastixTraceNumfact;
    /* Scotch is the reference default ordering; this engine has no Scotch, so
     * PastixOrderScotch asks for the caller's ordering (see pastix_subtask_order) */
    iparm[IPARM_ORDERING]              = PastixOrderScotch;
    iparm[IPARM_ORDERING_DEFAULT]      = 1;
    iparm[IPARM_SCOTCH_MT]             = 1;
    iparm[IPARM_SCOTCH_SWITCH_LEVEL]   = 120;
    iparm[IPARM_SCOTCH_CMIN]           = 0;
    iparm[IPARM_SCOTCH_CMAX]           = 100000;
    iparm[IPARM_SCOTCH_FRAT]           = 8;
    iparm[IPARM_METIS_CTYPE]           = -1;
    iparm[IPARM_METIS_RTYPE]           = -1;
    iparm[IPARM_METIS_NSEPS]           = 1;
    iparm[IPARM_METIS_NITER]           = 10;
    iparm[IPARM_METIS_UFACTOR]         = 200;
    iparm[IPARM_METIS_COMPRESS]        = 1;
    iparm[IPARM_METIS_SEED]            = 3452;
    iparm[IPARM_AMALGAMATION_LVLCBLK]  = 5;
    iparm[IPARM_AMALGAMATION_LVLBLAS]  = 5;
    iparm[IPARM_REORDERING_SPLIT]      = 0;
    iparm[IPARM_REORDERING_STOP]       = PASTIX_INT_MAX;
    iparm[IPARM_SPLITTING_STRATEGY]             = PastixSplitKway;
    iparm[IPARM_SPLITTING_LEVELS_PROJECTIONS]   = 0;
    iparm[IPARM_SPLITTING_LEVELS_KWAY]          = PASTIX_INT_MAX;
    iparm[IPARM_SPLITTING_PROJECTIONS_DEPTH]    = 3;
    iparm[IPARM_SPLITTING_PROJECTIONS_DISTANCE] = 3;
    iparm[IPARM_SPLITTING_PROJECTIONS_WIDTH]    = 1;
    iparm[IPARM_MIN_BLOCKSIZE]         = 160;
    iparm[IPARM_MAX_BLOCKSIZE]         = 320;
    iparm[IPARM_TASKS2D_LEVEL]         = -1;
    iparm[IPARM_TASKS2D_WIDTH]         = iparm[IPARM_MIN_BLOCKSIZE];
    iparm[IPARM_FACTORIZATION]         = PastixFactLU;
    iparm[IPARM_FACTO_LOOK_SIDE]       = PastixFactRightLooking;
    iparm[IPARM_SCHUR_FACT_MODE]       = PastixFactModeLocal;
    iparm[IPARM_TRANSPOSE_SOLVE]       = PastixNoTrans;
    iparm[IPARM_SCHUR_SOLV_MODE]       = PastixSolvModeLocal;
    iparm[IPARM_APPLYPERM_WS]          = 1;
    iparm[IPARM_REFINEMENT]            = PastixRefineGMRES;
    iparm[IPARM_ITERMAX]               = 250;
    iparm[IPARM_GMRES_IM]              = 25;
    iparm[IPARM_SCHEDULER]             = PastixSchedDynamic;
    iparm[IPARM_THREAD_NBR]            = -1;
    iparm[IPARM_SOCKET_NBR]            = -1;
    iparm[IPARM_GPU_NBR]               = 0;
    iparm[IPARM_GPU_MEMORY_PERCENTAGE] = 95;
    iparm[IPARM_GPU_MEMORY_BLOCK_SIZE] = 32 * 1024;
    iparm[IPARM_COMPRESS_MIN_WIDTH]    = 128;
    iparm[IPARM_COMPRESS_MIN_HEIGHT]   = 20;
    iparm[IPARM_COMPRESS_WHEN]         = PastixCompressNever;
    iparm[IPARM_COMPRESS_METHOD]       = PastixCompressMethodPQRCP;
    iparm[IPARM_COMPRESS_ORTHO]        = PastixCompressOrthoCGS;
    iparm[IPARM_COMPRESS_PRESELECT]    = 1;
    iparm[IPARM_COMPRESS_ILUK]         = -2;
    iparm[IPARM_MIXED]                 = 0;
    iparm[IPARM_MPI_THREAD_LEVEL]      = PastixMpiNone;
    iparm[IPARM_MODIFY_PARAMETER]      = 1;
    iparm[IPARM_START_TASK]            = PastixTaskOrdering;
    iparm[IPARM_END_TASK]              = PastixTaskClean;
    iparm[IPARM_FLOAT]                 = -1;
    iparm[IPARM_MTX_TYPE]              = -1;
    iparm[IPARM_DOF_NBR]               = 1;

    dparm[DPARM_EPSILON_REFINEMENT] = -1.;
    dparm[DPARM_RELATIVE_ERROR]     = -1.;
    dparm[DPARM_EPSILON_MAGN_CTRL]  = 0.;
    dparm[DPARM_A_NORM]             = -1.;
    dparm[DPARM_COMPRESS_TOLERANCE] = 1e-8;
    dparm[DPARM_COMPRESS_MIN_RATIO] = 1.;
}

/* Host threads of the analysis: PASTIX_NUM_THREADS, else the CPUs of the
 * affinity mask capped at 16 (the host share of one GPU). */
static int
px_default_threads( void )
{
    const char *env = getenv( "PASTIX_NUM_THREADS" );
    long        n;
    if ( env ) {
        n = atol( env );
    }
    else {
        cpu_set_t set;
        n = sysconf( _SC_NPROCESSORS_ONLN );
        if ( sched_getaffinity( 0, sizeof( set ), &set ) == 0 ) {
            n = CPU_COUNT( &set );
        }
        if ( n > 16 ) {
            n = 16;
        }
    }
    return ( n > 0 ) ? (int)n : 1;
}

void
pastixInitWithAffinity( pastix_data_t **pastix_data, PASTIX_Comm comm, pastix_int_t *iparm,
                        double *dparm, const int *bindtab )
{
    pastix_data_t *pd = px_calloc( 1, sizeof( pastix_data_t ) );
    (void)bindtab;
    pd->iparm    = iparm;
    pd->dparm    = dparm;
    pd->comm     = comm;
    pd->procnum  = 0;
    pd->steps    = STEP_INIT;
    pd->nthreads = ( iparm[IPARM_THREAD_NBR] > 0 ) ? (int)iparm[IPARM_THREAD_NBR] : px_default_threads();
    pd->factored_flttype  = -1;
    pd->berr_target       = 0.;
    pd->refine_nrhs_batch = 1;
    if ( iparm[IPARM_THREAD_NBR] <= 0 ) {
        iparm[IPARM_THREAD_NBR] = pd->nthreads;
    }
    pastixModelsLoad( pd );   /* common/api.c:854 */
    *pastix_data = pd;
}

void
pastixInit( pastix_data_t **pastix_data, PASTIX_Comm comm, pastix_int_t *iparm, double *dparm )
{
    pastixInitWithAffinity( pastix_data, comm, iparm, dparm, NULL );
}

static void
px_free_analysis( pastix_data_t *pd )
{
    if ( pd->graph ) {
        px_graph_exit( pd->graph );
        free( pd->graph );
        pd->graph = NULL;
    }
    if ( pd->symbmtx ) {
        px_symbol_exit( pd->symbmtx );
        free( pd->symbmtx );
        pd->symbmtx = NULL;
    }
    if ( pd->solvmatr ) {
        px_solver_exit( pd->solvmatr );
        free( pd->solvmatr );
        pd->solvmatr = NULL;
    }
}

void
pastixFinalize( pastix_data_t **pastix_data )
{
    pastix_data_t *pd;
    if ( pastix_data == NULL || *pastix_data == NULL ) {
        return;
    }
    pd = *pastix_data;
    px_free_analysis( pd );
    if ( pd->ordemesh ) {
        pastixOrderExit( pd->ordemesh );
        free( pd->ordemesh );
    }
    if ( pd->bcsc ) {
        px_bcsc_exit( pd->bcsc );
        free( pd->bcsc );
    }
    if ( pd->dev ) {
        for ( int i = 0; i < PX_REF_CACHE; i++ ) {
            if ( pd->ref_cache[i] ) {
                pxd_free( pd->dev, pd->ref_cache[i] );
            }
        }
        pxd_destroy( pd->dev );
    }
    pastixModelsFree( pd->cpu_models );
    pastixModelsFree( pd->gpu_models );
    free( pd->dir_global );
    free( pd->schur_list );
    free( pd );
    *pastix_data = NULL;
}

/* ------------------------------------------------------------------------ */
/* Supernodes for a user permutation given without ranges: the fundamental
 * supernodes of the elimination tree of the permuted pattern. */
void
px_order_fundamental_supernodes( const px_graph_t *g, pastix_order_t *ord )
{
    pastix_int_t  n      = g->n;
    pastix_int_t *parent = px_malloc( n * sizeof( pastix_int_t ) );
    pastix_int_t *anc    = px_malloc( n * sizeof( pastix_int_t ) );
    pastix_int_t *cc     = px_calloc( n, sizeof( pastix_int_t ) );
    pastix_int_t *nchild = px_calloc( n, sizeof( pastix_int_t ) );
    pastix_int_t *mark   = px_malloc( n * sizeof( pastix_int_t ) );
    pastix_int_t *snode  = px_malloc( n * sizeof( pastix_int_t ) );
    pastix_int_t  ns = 0;

    /* elimination tree (Liu, with path compression) on P A P^T */
    for ( pastix_int_t k = 0; k < n; k++ ) {
        pastix_int_t v = ord->peritab[k];
        parent[k] = -1;
        anc[k]    = -1;
        for ( pastix_int_t e = g->colptr[v]; e < g->colptr[v + 1]; e++ ) {
            pastix_int_t i = ord->permtab[g->rows[e]];
            while ( i != -1 && i < k ) {
                pastix_int_t next = anc[i];
                anc[i] = k;
                if ( next == -1 ) {
                    parent[i] = k;
                }
                i = next;
            }
        }
    }
    /* column counts by traversing the row subtrees */
    for ( pastix_int_t k = 0; k < n; k++ ) {
        mark[k] = -1;
    }
    for ( pastix_int_t k = 0; k < n; k++ ) {
        pastix_int_t v = ord->peritab[k];
        mark[k] = k;
        cc[k]++;
        for ( pastix_int_t e = g->colptr[v]; e < g->colptr[v + 1]; e++ ) {
            pastix_int_t i = ord->permtab[g->rows[e]];
            while ( i < k && mark[i] != k ) {
                cc[i]++;
                mark[i] = k;
                i = parent[i];
            }
        }
    }
    for ( pastix_int_t k = 0; k < n; k++ ) {
        if ( parent[k] >= 0 ) {
            nchild[parent[k]]++;
        }
    }
    for ( pastix_int_t k = 0; k < n; k++ ) {
        if ( k > 0 && parent[k - 1] == k && cc[k - 1] == cc[k] + 1 && nchild[k] == 1 ) {
            snode[k] = snode[k - 1];
        }
        else {
            snode[k] = ns++;
        }
    }
    free( ord->rangtab );
    free( ord->treetab );
    ord->cblknbr = ns;
    ord->rangtab = px_malloc( ( ns + 1 ) * sizeof( pastix_int_t ) );
    ord->treetab = px_malloc( ns * sizeof( pastix_int_t ) );
    for ( pastix_int_t k = 0; k < n; k++ ) {
        if ( k == 0 || snode[k] != snode[k - 1] ) {
            ord->rangtab[snode[k]] = k;
        }
    }
    ord->rangtab[ns] = n;
    for ( pastix_int_t s = 0; s < ns; s++ ) {
        pastix_int_t last = ord->rangtab[s + 1] - 1;
        ord->treetab[s]   = ( parent[last] >= 0 ) ? snode[parent[last]] : -1;
    }
    free( parent );
    free( anc );
    free( cc );
    free( nchild );
    free( mark );
    free( snode );
}

int
pastix_subtask_order( pastix_data_t *pd, const spmatrix_t *spm, pastix_order_t *myorder )
{
    pastix_int_t   *iparm;
    pastix_order_t *ord;
    double          t0;
    pastix_int_t    n;

    if ( pd == NULL ) {
        px_print_error( "pastix_subtask_order: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( spm == NULL ) {
        px_print_error( "pastix_subtask_order: wrong spm parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_INIT ) ) {
        px_print_error( "pastix_subtask_order: pastixInit() has to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( spm->dof != 1 && spm->dof != 0 ) {
        px_print_error( "pastix_subtask_order: only dof = 1 is supported (expand the matrix)" );
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    iparm = pd->iparm;
    iparm[IPARM_FLOAT] = spm->flttype;
    t0 = px_wtime();

    px_free_analysis( pd );
    if ( pd->ordemesh ) {
        pastixOrderExit( pd->ordemesh );
    }
    else {
        pd->ordemesh = px_calloc( 1, sizeof( pastix_order_t ) );
    }
    ord = pd->ordemesh;
    n   = spm->n;

    pd->graph = px_calloc( 1, sizeof( px_graph_t ) );
    px_graph_prepare( spm, pd->graph );

    if ( iparm[IPARM_IO_STRATEGY] & PastixIOLoad ) {
        int rc = pastixOrderLoad( pd, ord );
        if ( rc != PASTIX_SUCCESS ) {
            return rc;
        }
        pastixOrderBase( ord, 0 );
        if ( ord->vertnbr != n ) {
            px_print_error( "pastix_subtask_order: the loaded ordering has %ld vertices, the matrix %ld",
                            (long)ord->vertnbr, (long)n );
            return PASTIX_ERR_BADPARAMETER;
        }
    }
    else if ( iparm[IPARM_ORDERING] == PastixOrderPersonal ) {
        pastixOrderAlloc( ord, n, 0 );
        if ( myorder != NULL ) {
            pastixOrderBase( myorder, 0 );
        }
        if ( myorder == NULL || ( myorder->permtab == NULL && myorder->peritab == NULL ) ) {
            for ( pastix_int_t i = 0; i < n; i++ ) {
                ord->permtab[i] = i;
                ord->peritab[i] = i;
            }
        }
        else if ( myorder->permtab != NULL ) {
            memcpy( ord->permtab, myorder->permtab, n * sizeof( pastix_int_t ) );
            for ( pastix_int_t i = 0; i < n; i++ ) {
                ord->peritab[ord->permtab[i]] = i;
            }
        }
        else {
            memcpy( ord->peritab, myorder->peritab, n * sizeof( pastix_int_t ) );
            for ( pastix_int_t i = 0; i < n; i++ ) {
                ord->permtab[ord->peritab[i]] = i;
            }
        }
        if ( myorder != NULL && myorder->rangtab != NULL ) {
            ord->cblknbr = myorder->cblknbr;
            ord->rangtab = px_malloc( ( ord->cblknbr + 1 ) * sizeof( pastix_int_t ) );
            memcpy( ord->rangtab, myorder->rangtab, ( ord->cblknbr + 1 ) * sizeof( pastix_int_t ) );
            if ( myorder->treetab != NULL ) {
                ord->treetab = px_malloc( ord->cblknbr * sizeof( pastix_int_t ) );
                memcpy( ord->treetab, myorder->treetab, ord->cblknbr * sizeof( pastix_int_t ) );
            }
        }
    }
    else if ( iparm[IPARM_ORDERING] == PastixOrderScotch || iparm[IPARM_ORDERING] == PastixOrderPtScotch ||
              iparm[IPARM_ORDERING] == PastixOrderMetis || iparm[IPARM_ORDERING] == PastixOrderParMetis ) {
        /* Scotch and Metis are not part of this build: the built-in nested
         * dissection orders the graph instead */
        if ( iparm[IPARM_VERBOSE] > PastixVerboseNot ) {
            px_print_warning( "pastix_subtask_order: ordering %d is served by the built-in nested dissection",
                              (int)iparm[IPARM_ORDERING] );
        }
        if ( pd->schur_n > 0 ) {
            /* the Schur unknowns isolated at the end (they are the last
             * supernode, already in level order) */
            int rc = px_order_with_schur( pd, spm, ord, px_order_nd );
            if ( rc != PASTIX_SUCCESS ) {
                return rc;
            }
        }
        else if ( px_order_nd( pd->graph, ord ) != PASTIX_SUCCESS ) {
            return PASTIX_ERR_INTERNAL;
        }
    }
    else {
        px_print_error( "pastix_subtask_order: unknown ordering %d", (int)iparm[IPARM_ORDERING] );
        return PASTIX_ERR_BADPARAMETER;
    }

    if ( ord->rangtab == NULL || ord->treetab == NULL ) {
        px_order_fundamental_supernodes( pd->graph, ord );
    }
    if ( pastixOrderCheck( ord ) != PASTIX_SUCCESS ) {
        px_print_error( "pastix_subtask_order: invalid ordering" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( pd->schur_n > 0 && pd->schur_n <= n &&
         ( iparm[IPARM_ORDERING] == PastixOrderPersonal || ( iparm[IPARM_IO_STRATEGY] & PastixIOLoad ) ) ) {
        /* the user's ordering numbers the Schur unknowns last: they become
         * the last supernode */
        int rc = px_order_schur_boundary( pd, spm, ord );
        if ( rc != PASTIX_SUCCESS ) {
            return rc;
        }
    }
    if ( pd->schur_n == 0 || iparm[IPARM_ORDERING] == PastixOrderPersonal || ( iparm[IPARM_IO_STRATEGY] & PastixIOLoad ) ) {
        px_order_apply_level_order( ord, iparm[IPARM_TASKS2D_LEVEL], iparm[IPARM_TASKS2D_WIDTH] );
    }

    if ( pd->schur_n > 0 ) {
        for ( pastix_int_t s = 0; s < pd->schur_n; s++ ) {
            pastix_int_t v = pd->schur_list[s] - spm->baseval;
            if ( v < 0 || v >= n || ord->permtab[v] < n - pd->schur_n ) {
                px_print_error( "pastix_subtask_order: the Schur unknowns are not the last ones of the ordering" );
                return PASTIX_ERR_BADPARAMETER;
            }
        }
    }
    ord->sndenbr = ord->cblknbr;
    free( ord->sndetab );
    ord->sndetab = px_malloc( ( ord->sndenbr + 1 ) * sizeof( pastix_int_t ) );
    memcpy( ord->sndetab, ord->rangtab, ( ord->sndenbr + 1 ) * sizeof( pastix_int_t ) );

    if ( iparm[IPARM_IO_STRATEGY] & PastixIOSave ) {
        int rc = pastixOrderSave( pd, ord );
        if ( rc != PASTIX_SUCCESS ) {
            return rc;
        }
    }
    /* return the ordering to the user for non-personal orderings */
    if ( iparm[IPARM_ORDERING] != PastixOrderPersonal && myorder != NULL ) {
        pastixOrderExit( myorder );
        pastixOrderCopy( myorder, ord );
    }

    pd->csc = spm;
    pd->dparm[DPARM_ORDER_TIME] = px_wtime() - t0;
    pd->steps &= ~( STEP_SYMBFACT | STEP_ANALYSE | STEP_CSC2BCSC | STEP_BCSC2CTAB |
                    STEP_NUMFACT | STEP_SOLVE | STEP_REFINE );
    pd->steps |= STEP_ORDERING;
    return PASTIX_SUCCESS;
}

/* The symbolic factorization: fax on the graph, or with load set and
 * IPARM_IO_STRATEGY PastixIOLoad the file "symbname". */
static int
px_symbfact_run( pastix_data_t *pd, int load )
{
    double t0 = px_wtime();
    int    rc;
    if ( pd->symbmtx ) {
        px_symbol_exit( pd->symbmtx );
    }
    else {
        pd->symbmtx = px_calloc( 1, sizeof( px_symbol_t ) );
    }
    if ( load && ( pd->iparm[IPARM_IO_STRATEGY] & PastixIOLoad ) ) {
        if ( ( rc = px_symbol_load_default( pd->symbmtx ) ) != PASTIX_SUCCESS ) {
            return rc;
        }
        if ( pd->symbmtx->nodenbr != pd->ordemesh->vertnbr ) {
            px_print_error( "pastix_subtask_symbfact: the loaded symbol does not match the ordering" );
            return PASTIX_ERR_BADPARAMETER;
        }
    }
    else {
        px_symbol_fax_direct( pd->symbmtx, pd->graph, pd->ordemesh );
    }
    if ( px_symbol_check( pd->symbmtx ) != PASTIX_SUCCESS ) {
        px_print_error( "pastix_subtask_symbfact: inconsistent symbol matrix" );
        return PASTIX_ERR_INTERNAL;
    }
    /* the Schur part: the last schur_n columns (pastix_subtask_symbfact.c:238) */
    pd->symbmtx->schurfcol = pd->symbmtx->nodenbr - ( pd->schur_n > 0 ? pd->schur_n : 0 );
    if ( pd->iparm[IPARM_IO_STRATEGY] & PastixIOSave ) {
        if ( ( rc = px_symbol_save_default( pd, pd->symbmtx ) ) != PASTIX_SUCCESS ) {
            return rc;
        }
    }
    px_symbol_flops( pd->symbmtx, (int)pd->iparm[IPARM_FLOAT], (int)pd->iparm[IPARM_FACTORIZATION],
                     &pd->dparm[DPARM_FACT_THFLOPS], &pd->dparm[DPARM_FACT_RLFLOPS] );
    pd->iparm[IPARM_NNZEROS] = px_symbol_nnz( pd->symbmtx );
    pd->dparm[DPARM_SYMBFACT_TIME] = px_wtime() - t0;
    pd->steps &= ~( STEP_ANALYSE | STEP_CSC2BCSC | STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE );
    pd->steps |= STEP_SYMBFACT;
    return PASTIX_SUCCESS;
}

symbol_matrix_t *
pastix_amd_symbol( pastix_data_t *pd )
{
    return pd ? pd->symbmtx : NULL;
}

int
pastix_subtask_symbfact( pastix_data_t *pd )
{
    if ( pd == NULL ) {
        px_print_error( "pastix_subtask_symbfact: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_ORDERING ) || pd->graph == NULL || pd->ordemesh == NULL ) {
        px_print_error( "pastix_subtask_symbfact: pastix_subtask_order() has to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    return px_symbfact_run( pd, 1 );
}

int
pastix_subtask_reordering( pastix_data_t *pd )
{
    double t0;
    int    rc;
    if ( pd == NULL ) {
        px_print_error( "pastix_subtask_reordering: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_SYMBFACT ) || pd->symbmtx == NULL ) {
        px_print_error( "pastix_subtask_reordering: pastix_subtask_symbfact() has to be called before" );
        return PASTIX_ERR_BADPARAMETER;
    }
    t0 = px_wtime();
    px_symbol_reordering( pd->symbmtx, pd->ordemesh, pd->iparm[IPARM_REORDERING_SPLIT],
                          pd->iparm[IPARM_REORDERING_STOP], pd->nthreads );
    if ( pd->iparm[IPARM_IO_STRATEGY] & PastixIOSave ) {
        if ( ( rc = pastixOrderSave( pd, pd->ordemesh ) ) != PASTIX_SUCCESS ) {
            return rc;
        }
    }
    /* rebuild the symbolic structure with the new row order (from the graph,
     * also when the first one was loaded from a file) */
    rc = px_symbfact_run( pd, 0 );
    pd->dparm[DPARM_REORDER_TIME] = px_wtime() - t0;
    return rc;
}

int
pastix_subtask_blend( pastix_data_t *pd )
{
    double t0;
    if ( pd == NULL ) {
        px_print_error( "pastix_subtask_blend: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_SYMBFACT ) || pd->symbmtx == NULL ) {
        px_print_error( "pastix_subtask_blend: pastix_subtask_symbfact() has to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    t0 = px_wtime();
    if ( pd->iparm[IPARM_MIN_BLOCKSIZE] > pd->iparm[IPARM_MAX_BLOCKSIZE] ) {
        px_print_error( "pastix_subtask_blend: IPARM_MIN_BLOCKSIZE > IPARM_MAX_BLOCKSIZE" );
        return PASTIX_ERR_BADPARAMETER;
    }
    px_symbol_split( pd->symbmtx, pd->iparm[IPARM_MIN_BLOCKSIZE], pd->iparm[IPARM_MAX_BLOCKSIZE] );
    px_symbol_flops( pd->symbmtx, (int)pd->iparm[IPARM_FLOAT], (int)pd->iparm[IPARM_FACTORIZATION],
                     &pd->dparm[DPARM_FACT_THFLOPS], &pd->dparm[DPARM_FACT_RLFLOPS] );
    if ( pd->solvmatr ) {
        px_solver_exit( pd->solvmatr );
    }
    else {
        pd->solvmatr = px_calloc( 1, sizeof( px_solver_t ) );
    }
    px_solver_gen( pd->solvmatr, pd->symbmtx, pd->iparm[IPARM_TASKS2D_WIDTH], pd->iparm[IPARM_TASKS2D_LEVEL] );
    if ( px_solver_check( pd->solvmatr ) != PASTIX_SUCCESS ) {
        px_print_error( "pastix_subtask_blend: inconsistent solver matrix" );
        return PASTIX_ERR_INTERNAL;
    }
    {
        double nnz = (double)pd->iparm[IPARM_NNZEROS];
        if ( pd->iparm[IPARM_FACTORIZATION] == PastixFactLU ) {
            nnz *= 2.;
        }
        pd->dparm[DPARM_SOLV_FLOPS]          = nnz;
        pd->iparm[IPARM_NNZEROS_BLOCK_LOCAL] = pd->solvmatr->coefnbr;
        if ( pd->csc && pd->csc->gnnzexp > 0 ) {
            pd->dparm[DPARM_FILL_IN] = nnz / (double)pd->csc->gnnzexp;
        }
    }
    if ( pd->graph ) {
        px_graph_exit( pd->graph );
        free( pd->graph );
        pd->graph = NULL;
    }
    pd->dparm[DPARM_BLEND_TIME] = px_wtime() - t0;
    pd->steps &= ~( STEP_CSC2BCSC | STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE );
    pd->steps |= STEP_ANALYSE;
    return PASTIX_SUCCESS;
}

int
pastix_task_analyze( pastix_data_t *pd, const spmatrix_t *spm )
{
    double t0;
    int    rc;
    if ( pd == NULL ) {
        px_print_error( "pastix_task_analyze: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_INIT ) ) {
        px_print_error( "pastix_task_analyze: pastixInit() has to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    t0 = px_wtime();
    if ( ( rc = pastix_subtask_order( pd, spm, NULL ) ) != PASTIX_SUCCESS ) return rc;
    if ( ( rc = pastix_subtask_symbfact( pd ) ) != PASTIX_SUCCESS ) return rc;
    if ( ( rc = pastix_subtask_reordering( pd ) ) != PASTIX_SUCCESS ) return rc;
    if ( ( rc = pastix_subtask_blend( pd ) ) != PASTIX_SUCCESS ) return rc;
    pd->dparm[DPARM_ANALYZE_TIME] = px_wtime() - t0;
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Device engine plumbing. */

int
px_device_ready( pastix_data_t *pd )
{
    int rc;
    if ( pd->dev != NULL ) {
        return PASTIX_SUCCESS;
    }
    rc = pxd_create( &pd->dev, 0 );
    if ( rc != PASTIX_SUCCESS ) {
        pd->dev = NULL;
        px_print_error( "no usable GPU: the numerical factorization, solve and refinement of this "
                        "engine run on the MI355X only" );
    }
    return rc;
}

static int
px_upload_structure( pastix_data_t *pd )
{
    px_solver_t *s = pd->solvmatr;
    pxd_cblk_t  *cb = px_malloc( ( s->cblknbr + 1 ) * sizeof( pxd_cblk_t ) );
    pxd_blok_t  *bk = px_malloc( ( s->bloknbr + 1 ) * sizeof( pxd_blok_t ) );
    int          rc;
    for ( pastix_int_t c = 0; c <= s->cblknbr; c++ ) {
        cb[c].fcolnum = s->cblktab[c].fcolnum;
        cb[c].lcolnum = s->cblktab[c].lcolnum;
        cb[c].fblok   = s->cblktab[c].fblok;
        cb[c].stride  = s->cblktab[c].stride;
        cb[c].coefoff = s->cblktab[c].coefoff;
        cb[c].brownum = s->cblktab[c].brownum;
    }
    for ( pastix_int_t b = 0; b <= s->bloknbr; b++ ) {
        bk[b].frownum = s->bloktab[b].frownum;
        bk[b].lrownum = s->bloktab[b].lrownum;
        bk[b].lcblknm = s->bloktab[b].lcblknm;
        bk[b].fcblknm = s->bloktab[b].fcblknm;
        bk[b].coefind = s->bloktab[b].coefind;
    }
    rc = pxd_set_structure( pd->dev, s->nodenbr, s->cblknbr, s->bloknbr, cb, bk, s->browtab, s->coefnbr );
    free( cb );
    free( bk );
    return rc;
}

int
pastix_subtask_spm2bcsc( pastix_data_t *pd, spmatrix_t *spm )
{
    if ( pd == NULL || spm == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_ANALYSE ) ) {
        px_print_error( "pastix_subtask_spm2bcsc: All steps from pastix_task_init() to pastix_task_blend() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( spm->n != pd->solvmatr->nodenbr ) {
        px_print_error( "pastix_subtask_spm2bcsc: the matrix size does not match the analysis" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( pd->bcsc ) {
        px_bcsc_exit( pd->bcsc );
    }
    else {
        pd->bcsc = px_calloc( 1, sizeof( px_bcsc_t ) );
    }
    if ( px_bcsc_init( pd->bcsc, spm, pd->ordemesh ) != PASTIX_SUCCESS ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    pd->dparm[DPARM_A_NORM] = pd->bcsc->Anorm_f;
    pd->csc = spm;
    /* the analysis depends on the pattern only: the values given here may be
     * of another arithmetic than those of the ordering step, and they decide
     * the type of the factor and of the right-hand sides from now on */
    pd->iparm[IPARM_FLOAT] = spm->flttype;
    pd->steps &= ~( STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE );
    pd->steps |= STEP_CSC2BCSC;
    return PASTIX_SUCCESS;
}

int
pastix_subtask_bcsc2ctab( pastix_data_t *pd )
{
    int rc;
    if ( pd == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_CSC2BCSC ) ) {
        px_print_error( "pastix_subtask_bcsc2ctab: All steps from pastix_task_init() to pastix_stask_spm2bcsc() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( ( rc = px_device_ready( pd ) ) != PASTIX_SUCCESS ) {
        return rc;
    }
    if ( ( rc = px_upload_structure( pd ) ) != PASTIX_SUCCESS ) {
        return rc;
    }
    {
        px_bcsc_t *b = pd->bcsc;
        rc = pxd_set_matrix( pd->dev, b->flttype, b->mtxtype, b->n, b->nnz, b->colptr, b->rowtab,
                             b->Lvalues, b->trowptr, b->tcoltab, b->Uvalues );
        if ( rc != PASTIX_SUCCESS ) {
            return rc;
        }
    }
    pd->steps &= ~( STEP_NUMFACT | STEP_SOLVE | STEP_REFINE );
    pd->steps |= STEP_BCSC2CTAB;
    return PASTIX_SUCCESS;
}

static int
px_factor_flttype( const pastix_data_t *pd )
{
    int flt = pd->bcsc->flttype;
    if ( pd->iparm[IPARM_MIXED] ) {
        flt = ( flt == SpmDouble ) ? SpmFloat : SpmComplex32;
    }
    return flt;
}

int
pastix_subtask_sopalin( pastix_data_t *pd )
{
    double       threshold, t0;
    int64_t      npiv = 0;
    int          rc, fflt;
    pastix_int_t *iparm;
    double       *dparm;

    if ( pd == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_BCSC2CTAB ) ) {
        px_print_error( "pastix_subtask_sopalin: All steps from pastix_task_init() to pastix_subtask_csc2bcsc() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    iparm = pd->iparm;
    dparm = pd->dparm;
    fflt  = px_factor_flttype( pd );

    /* static pivoting threshold (sopalin/pastix_task_sopalin.c:461-485); the
     * epsilon is that of the factor precision, where the reference takes the
     * bcsc precision: a single precision factor of a double matrix uses the
     * single epsilon, so tiny pivots are judged at the precision they are
     * stored in */
    if ( dparm[DPARM_EPSILON_MAGN_CTRL] < 0. ) {
        threshold = -dparm[DPARM_EPSILON_MAGN_CTRL];
    }
    else if ( dparm[DPARM_EPSILON_MAGN_CTRL] == 0. ) {
        double eps = ( fflt == SpmFloat || fflt == SpmComplex32 ) ? (double)FLT_EPSILON * 0.5 : DBL_EPSILON * 0.5;
        threshold  = eps * dparm[DPARM_A_NORM];
    }
    else {
        threshold = dparm[DPARM_EPSILON_MAGN_CTRL] * dparm[DPARM_A_NORM];
    }
    pd->solvmatr->diagthreshold = threshold;

    if ( ( rc = pxd_set_schur( pd->dev, pd->symbmtx ? pd->symbmtx->schurfcol : pd->solvmatr->nodenbr,
                               iparm[IPARM_SCHUR_FACT_MODE] == PastixFactModeLocal ) ) != PASTIX_SUCCESS ) {
        px_print_error( "pastix_subtask_sopalin: the Schur part does not start at a column block" );
        return rc;
    }
    t0 = px_wtime();
    {
        /* for real matrices LL^T is LL^H (POTRF) and LDL^H is LDL^T (SYTRF),
         * as in the reference's sopalinFacto table (sopalin/pastix_task_sopalin.c:51-58) */
        int facto = (int)iparm[IPARM_FACTORIZATION];
        int cplx  = ( fflt == SpmComplex32 || fflt == SpmComplex64 );
        if ( !cplx && facto == PastixFactLLT ) facto = PastixFactLLH;
        if ( !cplx && facto == PastixFactLDLH ) facto = PastixFactLDLT;
        rc = pxd_factorize( pd->dev, facto, fflt, threshold, &npiv );
    }
    if ( rc != PASTIX_SUCCESS ) {
        return rc;
    }
    pxd_sync( pd->dev );
    dparm[DPARM_FACT_TIME] = px_wtime() - t0;
    {
        double flops = dparm[DPARM_FACT_THFLOPS] / dparm[DPARM_FACT_TIME];
        dparm[DPARM_FACT_FLOPS]   = ( ( flops / 1024. ) / 1024. ) / 1024.;
        dparm[DPARM_FACT_RLFLOPS] = dparm[DPARM_FACT_THFLOPS];
        dparm[DPARM_MEM_FR]       = (double)pd->solvmatr->coefnbr * (double)spm_size_of( fflt ) *
                              ( iparm[IPARM_FACTORIZATION] == PastixFactLU ? 2. : 1. );
    }
    iparm[IPARM_STATIC_PIVOTING] = (pastix_int_t)npiv;
    pd->solvmatr->nbpivots       = (int)npiv;
    pd->factored_flttype         = fflt;
    pd->steps |= STEP_NUMFACT;
    return PASTIX_SUCCESS;
}

int
pastix_task_numfact( pastix_data_t *pd, spmatrix_t *spm )
{
    int rc;
    if ( pd == NULL ) {
        px_print_error( "pastix_task_sopalin: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( spm == NULL ) {
        px_print_error( "pastix_task_sopalin: wrong spm parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_ANALYSE ) ) {
        px_print_error( "pastix_task_sopalin: All steps from pastix_task_init() to pastix_task_blend() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    pd->steps &= ~( STEP_CSC2BCSC | STEP_BCSC2CTAB | STEP_NUMFACT | STEP_SOLVE | STEP_REFINE );
    if ( ( rc = pastix_subtask_spm2bcsc( pd, spm ) ) != PASTIX_SUCCESS ) return rc;
    if ( ( rc = pastix_subtask_bcsc2ctab( pd ) ) != PASTIX_SUCCESS ) return rc;
    if ( ( rc = pastix_subtask_sopalin( pd ) ) != PASTIX_SUCCESS ) return rc;
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Right-hand sides and solve. */

int
pastixRhsInit( pastix_rhs_t *rhs )
{
    if ( rhs == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    *rhs = px_calloc( 1, sizeof( struct pastix_rhs_s ) );
    ( *rhs )->allocated = -1;
    return PASTIX_SUCCESS;
}

int
pastixRhsFinalize( pastix_rhs_t rhs )
{
    if ( rhs == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( rhs->allocated > 0 ) {
        free( rhs->b );
    }
    free( rhs );
    return PASTIX_SUCCESS;
}

int
pastixRhsDoubletoSingle( const pastix_rhs_t dB, pastix_rhs_t sB )
{
    pastix_int_t m = dB->m, n = dB->n;
    int          cplx = ( dB->flttype == SpmComplex64 );
    if ( dB->flttype != SpmDouble && dB->flttype != SpmComplex64 ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    sB->flttype = cplx ? SpmComplex32 : SpmFloat;
    sB->m = m;
    sB->n = n;
    sB->ld = m;
    sB->allocated = 1;
    sB->b = px_malloc( m * n * ( cplx ? 8 : 4 ) );
    for ( pastix_int_t j = 0; j < n; j++ ) {
        for ( pastix_int_t i = 0; i < m; i++ ) {
            if ( cplx ) {
                ( (float complex *)sB->b )[i + j * m] = (float complex)( (double complex *)dB->b )[i + j * dB->ld];
            }
            else {
                ( (float *)sB->b )[i + j * m] = (float)( (double *)dB->b )[i + j * dB->ld];
            }
        }
    }
    return PASTIX_SUCCESS;
}

int
pastixRhsSingleToDouble( const pastix_rhs_t sB, pastix_rhs_t dB )
{
    pastix_int_t m = sB->m, n = sB->n;
    int          cplx = ( sB->flttype == SpmComplex32 );
    for ( pastix_int_t j = 0; j < n; j++ ) {
        for ( pastix_int_t i = 0; i < m; i++ ) {
            if ( cplx ) {
                ( (double complex *)dB->b )[i + j * dB->ld] = ( (float complex *)sB->b )[i + j * m];
            }
            else {
                ( (double *)dB->b )[i + j * dB->ld] = ( (float *)sB->b )[i + j * m];
            }
        }
    }
    if ( sB->allocated > 0 ) {
        free( sB->b );
        sB->b = NULL;
        sB->allocated = 0;
    }
    return PASTIX_SUCCESS;
}

/* Arithmetic of the right-hand sides: that of the matrix given to numfact
 * (the bcsc), else of the one given to the analysis. */
static int
px_rhs_flttype( const pastix_data_t *pd )
{
    if ( pd->bcsc ) {
        return pd->bcsc->flttype;
    }
    if ( pd->csc ) {
        return pd->csc->flttype;
    }
    return (int)pd->iparm[IPARM_FLOAT];
}

/* b_perm[permtab[i]] = b[i] (forward) or the inverse (backward). */
int
pastix_subtask_applyorder( pastix_data_t *pd, pastix_dir_t dir, pastix_int_t m, pastix_int_t n,
                           void *B, pastix_int_t ldb, pastix_rhs_t Bp )
{
    const pastix_int_t *perm;
    int                 cplx;
    size_t              es;
    if ( pd == NULL || Bp == NULL || pd->ordemesh == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    perm = pd->ordemesh->permtab;
    /* the vectors have the type of the matrix in use (pastix_data->csc,
     * sopalin/pastix_task_solve.c:197), as every later step sizes them */
    cplx = ( px_rhs_flttype( pd ) == SpmComplex64 );
    es   = cplx ? 16 : 8;
    if ( dir == PastixDirForward ) {
        if ( Bp->allocated <= 0 || Bp->m != m || Bp->n != n ) {
            if ( Bp->allocated > 0 ) {
                free( Bp->b );
            }
            Bp->b = px_malloc( m * n * es );
            Bp->allocated = 1;
        }
        Bp->flttype = cplx ? SpmComplex64 : SpmDouble;
        Bp->m  = m;
        Bp->n  = n;
        Bp->ld = m;
        for ( pastix_int_t j = 0; j < n; j++ ) {
            for ( pastix_int_t i = 0; i < m; i++ ) {
                memcpy( (char *)Bp->b + ( perm[i] + j * m ) * es, (char *)B + ( i + j * ldb ) * es, es );
            }
        }
    }
    else {
        for ( pastix_int_t j = 0; j < n; j++ ) {
            for ( pastix_int_t i = 0; i < m; i++ ) {
                memcpy( (char *)B + ( i + j * ldb ) * es, (char *)Bp->b + ( perm[i] + j * m ) * es, es );
            }
        }
    }
    return PASTIX_SUCCESS;
}

static size_t
px_vec_es( const pastix_data_t *pd )
{
    return ( px_rhs_flttype( pd ) == SpmComplex64 ) ? 16 : 8;
}

/*
 * A right-hand-side descriptor on the device: the vectors in double (complex
 * double), leading dimension m.  A single-precision descriptor (the result of
 * pastixRhsDoubletoSingle) is widened on the way up and rounded on the way
 * back; the engine converts to the factor precision itself.
 */
static int
px_rhs_check( const pastix_data_t *pd, const pastix_rhs_t b, const char *who )
{
    int want = px_rhs_flttype( pd );
    int flt  = b->flttype;
    int ok   = ( flt == want ) || ( want == SpmDouble && flt == SpmFloat ) ||
             ( want == SpmComplex64 && flt == SpmComplex32 );
    if ( !ok || b->m != pd->solvmatr->nodenbr || b->n < 0 || ( b->n > 0 && ( b->b == NULL || b->ld < b->m ) ) ) {
        px_print_error( "%s: the right-hand side does not match the factorized matrix", who );
        return PASTIX_ERR_BADPARAMETER;
    }
    return PASTIX_SUCCESS;
}

static int
px_rhs_upload( pastix_data_t *pd, const pastix_rhs_t b, void **dx )
{
    pastix_int_t m = b->m, n = b->n;
    int          single = ( b->flttype == SpmFloat || b->flttype == SpmComplex32 );
    int          cplx   = ( b->flttype == SpmComplex64 || b->flttype == SpmComplex32 );
    size_t       es     = cplx ? 16 : 8;
    size_t       bytes  = (size_t)m * n * es;
    void        *h      = NULL;
    int          rc;
    *dx = NULL;
    if ( ( rc = pxd_malloc( pd->dev, bytes > 0 ? bytes : 16, dx ) ) != PASTIX_SUCCESS ) {
        *dx = NULL;
        return rc;
    }
    if ( single || b->ld != m ) {
        h = px_malloc( bytes > 0 ? bytes : 16 );
        for ( pastix_int_t j = 0; j < n; j++ ) {
            for ( pastix_int_t i = 0; i < m; i++ ) {
                if ( b->flttype == SpmFloat ) {
                    ( (double *)h )[i + j * m] = ( (const float *)b->b )[i + j * b->ld];
                }
                else if ( b->flttype == SpmComplex32 ) {
                    ( (double complex *)h )[i + j * m] = ( (const float complex *)b->b )[i + j * b->ld];
                }
                else {
                    memcpy( (char *)h + ( i + j * m ) * es, (const char *)b->b + ( i + j * b->ld ) * es, es );
                }
            }
        }
    }
    rc = pxd_h2d( pd->dev, *dx, h ? h : b->b, bytes );
    free( h );
    if ( rc != PASTIX_SUCCESS ) {
        pxd_free( pd->dev, *dx );
        *dx = NULL;
    }
    return rc;
}

static int
px_rhs_download( pastix_data_t *pd, void *dx, pastix_rhs_t b )
{
    pastix_int_t m = b->m, n = b->n;
    int          single = ( b->flttype == SpmFloat || b->flttype == SpmComplex32 );
    int          cplx   = ( b->flttype == SpmComplex64 || b->flttype == SpmComplex32 );
    size_t       es     = cplx ? 16 : 8;
    size_t       bytes  = (size_t)m * n * es;
    void        *h      = ( single || b->ld != m ) ? px_malloc( bytes > 0 ? bytes : 16 ) : b->b;
    int          rc     = pxd_d2h( pd->dev, h, dx, bytes );
    if ( rc == PASTIX_SUCCESS && h != b->b ) {
        for ( pastix_int_t j = 0; j < n; j++ ) {
            for ( pastix_int_t i = 0; i < m; i++ ) {
                if ( b->flttype == SpmFloat ) {
                    ( (float *)b->b )[i + j * b->ld] = (float)( (const double *)h )[i + j * m];
                }
                else if ( b->flttype == SpmComplex32 ) {
                    ( (float complex *)b->b )[i + j * b->ld] = (float complex)( (const double complex *)h )[i + j * m];
                }
                else {
                    memcpy( (char *)b->b + ( i + j * b->ld ) * es, (const char *)h + ( i + j * m ) * es, es );
                }
            }
        }
    }
    if ( h != b->b ) {
        free( h );
    }
    return rc;
}

/* Run one device solve call on the descriptor b (upload, call, download). */
typedef int ( *px_solve_fn_t )( pxd_ctx_t *, int, int64_t, void *, int64_t );
static int
px_rhs_apply( pastix_data_t *pd, pastix_rhs_t b, px_solve_fn_t fn, int arg )
{
    void *dx;
    int   rc = px_rhs_upload( pd, b, &dx );
    if ( rc != PASTIX_SUCCESS ) {
        return rc;
    }
    pxd_set_option( pd->dev, "schur_solv_mode", (double)pd->iparm[IPARM_SCHUR_SOLV_MODE] );
    rc = fn( pd->dev, arg, b->n, dx, b->m );
    if ( rc == PASTIX_SUCCESS ) {
        rc = px_rhs_download( pd, dx, b );
    }
    pxd_free( pd->dev, dx );
    return rc;
}

int
pastix_amd_solve_dev( pastix_data_t *pd, pastix_int_t nrhs, void *dx, pastix_int_t ldx )
{
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_subtask_solve: All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    pxd_set_option( pd->dev, "schur_solv_mode", (double)pd->iparm[IPARM_SCHUR_SOLV_MODE] );
    return pxd_solve( pd->dev, nrhs, dx, ldx );
}

int
pastix_subtask_solve_adv( pastix_data_t *pd, pastix_trans_t transA, pastix_rhs_t b )
{
    int    rc;
    int    cplx;
    double t0;
    if ( pd == NULL ) {
        px_print_error( "pastix_task_solve: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_task_solve: All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( b == NULL || ( rc = px_rhs_check( pd, b, "pastix_task_solve" ) ) != PASTIX_SUCCESS ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    /* op(A) = A^T is solved as U^T L^T for LU and as A itself for the
     * symmetric factorizations; conjugate transposes would need conj(L)
     * (sopalin/pastix_task_solve.c:541-572) */
    cplx = ( b->flttype == SpmComplex64 || b->flttype == SpmComplex32 );
    if ( transA != PastixNoTrans && !( transA == PastixTrans || ( !cplx && transA == PastixConjTrans &&
                                                                  pd->iparm[IPARM_FACTORIZATION] != PastixFactLU ) ) ) {
        px_print_error( "pastix_task_solve: transA incompatible with the factotype used (require extra conj(L) not handled)" );
        return PASTIX_ERR_BADPARAMETER;
    }
    t0 = px_wtime();
    rc = px_rhs_apply( pd, b, pxd_solve_trans, transA != PastixNoTrans );
    pd->dparm[DPARM_SOLV_TIME] = px_wtime() - t0;
    return rc;
}

int
pastix_subtask_solve( pastix_data_t *pd, pastix_rhs_t b )
{
    return pastix_subtask_solve_adv( pd, PastixNoTrans, b );
}

/*
 * One triangular sweep (reference sopalin/pastix_task_solve.c:296-360).  The
 * engine keeps the explicit inverses of the diagonal blocks that the
 * factorization fixes (unit L for LU and LDL^T, non-unit for Cholesky), so
 * the accepted (uplo, trans, diag) are those the reference's own solves use:
 *   Lower NoTrans          L y = b     (diag: Unit for LU / LDL^T, NonUnit for Cholesky)
 *   Lower Trans/ConjTrans  L^T x = y   (same diag rule; ConjTrans on real data only)
 *   Upper NoTrans          U x = y     (LU, NonUnit)
 *   Upper Trans            U^T y = b   (LU, NonUnit)
 * Right-side application is not handled (the reference asserts on it).
 */
int
pastix_subtask_trsm( pastix_data_t *pd, pastix_side_t side, pastix_uplo_t uplo, pastix_trans_t trans,
                     pastix_diag_t diag, pastix_rhs_t b )
{
    int facto, cplx, step, unit_l;
    if ( pd == NULL ) {
        px_print_error( "pastix_subtask_trsm: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( b == NULL ) {
        px_print_error( "pastix_subtask_trsm: wrong Bp parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_subtask_trsm: All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( px_rhs_check( pd, b, "pastix_subtask_trsm" ) != PASTIX_SUCCESS ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    facto  = (int)pd->iparm[IPARM_FACTORIZATION];
    cplx   = ( b->flttype == SpmComplex64 || b->flttype == SpmComplex32 );
    unit_l = ( facto == PastixFactLU || facto == PastixFactLDLT || facto == PastixFactLDLH );
    if ( !cplx && trans == PastixConjTrans ) {
        trans = PastixTrans;
    }
    step = -1;
    if ( side == PastixLeft && uplo == PastixLower && trans == PastixNoTrans ) {
        step = ( diag == ( unit_l ? PastixUnit : PastixNonUnit ) ) ? PXD_STEP_L_FWD : -1;
    }
    else if ( side == PastixLeft && uplo == PastixLower && trans == PastixTrans ) {
        step = ( diag == ( unit_l ? PastixUnit : PastixNonUnit ) ) ? PXD_STEP_LT_BWD : -1;
    }
    else if ( side == PastixLeft && uplo == PastixUpper && facto == PastixFactLU && diag == PastixNonUnit ) {
        step = ( trans == PastixNoTrans ) ? PXD_STEP_U_BWD : ( trans == PastixTrans ) ? PXD_STEP_UT_FWD : -1;
    }
    if ( step < 0 ) {
        px_print_error( "pastix_subtask_trsm: (side, uplo, trans, diag) = (%d, %d, %d, %d) is not a sweep of this factorization",
                        (int)side, (int)uplo, (int)trans, (int)diag );
        return PASTIX_ERR_BADPARAMETER;
    }
    return px_rhs_apply( pd, b, pxd_solve_step, step );
}

/* x <- D^{-1} x, D the diagonal of the factor's diagonal tiles
 * (reference sopalin/pastix_task_solve.c:417-470 -> sopalin_zdiag). */
int
pastix_subtask_diag( pastix_data_t *pd, pastix_rhs_t b )
{
    if ( pd == NULL ) {
        px_print_error( "pastix_subtask_diag: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( b == NULL ) {
        px_print_error( "pastix_subtask_diag: wrong Bp parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_subtask_trsm: All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( px_rhs_check( pd, b, "pastix_subtask_diag" ) != PASTIX_SUCCESS ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    return px_rhs_apply( pd, b, pxd_solve_step, PXD_STEP_DIAG );
}

int
pastix_task_solve( pastix_data_t *pd, pastix_int_t m, pastix_int_t nrhs, void *B, pastix_int_t ldb )
{
    pastix_rhs_t Bp;
    int          rc;
    if ( pd == NULL ) {
        px_print_error( "pastix_task_solve: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_task_solve: Numerical factorization hasn't been done." );
        return PASTIX_ERR_BADPARAMETER;
    }
    pastixRhsInit( &Bp );
    rc = pastix_subtask_applyorder( pd, PastixDirForward, m, nrhs, B, ldb, Bp );
    if ( rc == PASTIX_SUCCESS ) rc = pastix_subtask_solve( pd, Bp );
    if ( rc == PASTIX_SUCCESS ) rc = pastix_subtask_applyorder( pd, PastixDirBackward, m, nrhs, B, ldb, Bp );
    pastixRhsFinalize( Bp );
    if ( rc == PASTIX_SUCCESS ) {
        pd->steps |= STEP_SOLVE;
    }
    return rc;
}

int
pastix_subtask_refine( pastix_data_t *pd, pastix_rhs_t Bp, pastix_rhs_t Xp )
{
    void        *db = NULL, *dx = NULL;
    int          rc;
    pastix_int_t iters = 0;
    double       t0;
    if ( pd == NULL || Bp == NULL || Xp == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_subtask_refine: Numerical factorization hasn't been done." );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( px_rhs_check( pd, Bp, "pastix_subtask_refine" ) != PASTIX_SUCCESS ||
         px_rhs_check( pd, Xp, "pastix_subtask_refine" ) != PASTIX_SUCCESS || Bp->n != Xp->n ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( pd->schur_n > 0 && pd->iparm[IPARM_SCHUR_SOLV_MODE] != PastixSolvModeLocal ) {
        px_print_error( "Refinement is not available with Schur complement when non local solve is required" );
        return PASTIX_ERR_BADPARAMETER;
    }
    pxd_set_option( pd->dev, "schur_solv_mode", (double)pd->iparm[IPARM_SCHUR_SOLV_MODE] );
    t0 = px_wtime();
    rc = px_rhs_upload( pd, Bp, &db );
    if ( rc == PASTIX_SUCCESS ) {
        rc = px_rhs_upload( pd, Xp, &dx );
    }
    if ( rc == PASTIX_SUCCESS ) {
        rc = px_refine_run( pd, Bp->n, db, dx, &iters );
    }
    if ( rc == PASTIX_SUCCESS ) {
        rc = px_rhs_download( pd, dx, Xp );
    }
    if ( db ) pxd_free( pd->dev, db );
    if ( dx ) pxd_free( pd->dev, dx );
    pd->iparm[IPARM_NBITER]      = iters;
    pd->dparm[DPARM_REFINE_TIME] = px_wtime() - t0;
    return rc;
}

int
pastix_task_refine( pastix_data_t *pd, pastix_int_t n, pastix_int_t nrhs, void *B, pastix_int_t ldb,
                    void *X, pastix_int_t ldx )
{
    pastix_rhs_t Bp, Xp;
    int          rc;
    if ( pd == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    pastixRhsInit( &Bp );
    pastixRhsInit( &Xp );
    rc = pastix_subtask_applyorder( pd, PastixDirForward, n, nrhs, B, ldb, Bp );
    if ( rc == PASTIX_SUCCESS ) rc = pastix_subtask_applyorder( pd, PastixDirForward, n, nrhs, X, ldx, Xp );
    if ( rc == PASTIX_SUCCESS ) rc = pastix_subtask_refine( pd, Bp, Xp );
    if ( rc == PASTIX_SUCCESS ) rc = pastix_subtask_applyorder( pd, PastixDirBackward, n, nrhs, X, ldx, Xp );
    pastixRhsFinalize( Bp );
    pastixRhsFinalize( Xp );
    if ( rc == PASTIX_SUCCESS ) {
        pd->steps |= STEP_REFINE;
    }
    return rc;
}

int
pastix_task_solve_and_refine( pastix_data_t *pd, pastix_int_t n, pastix_int_t nrhs, void *B,
                              pastix_int_t ldb, void *X, pastix_int_t ldx )
{
    int    rc;
    size_t es = px_vec_es( pd );
    for ( pastix_int_t j = 0; j < nrhs; j++ ) {
        memcpy( (char *)X + j * ldx * es, (char *)B + j * ldb * es, n * es );
    }
    rc = pastix_task_solve( pd, n, nrhs, X, ldx );
    if ( rc == PASTIX_SUCCESS ) {
        rc = pastix_task_refine( pd, n, nrhs, B, ldb, X, ldx );
    }
    return rc;
}

int
pastix( pastix_data_t **pastix_data, PASTIX_Comm comm, pastix_int_t n, pastix_int_t *colptr,
        pastix_int_t *rowptr, void *values, pastix_int_t *perm, pastix_int_t *invp, void *B,
        pastix_int_t nrhs, pastix_int_t *iparm, double *dparm )
{
    (void)pastix_data; (void)comm; (void)n; (void)colptr; (void)rowptr; (void)values;
    (void)perm; (void)invp; (void)B; (void)nrhs; (void)iparm; (void)dparm;
    px_print_error( "pastix(): the legacy single-call interface is not provided; use the task API" );
    return PASTIX_ERR_NOTIMPLEMENTED;
}

/*
 * D[incD * i] = diagonal entry i of the factor's diagonal tiles, in the
 * permuted numbering (reference sopalin/diag.c:54-101 with coeftab_zgetdiag,
 * sopalin/coeftab_z.c:659-689): L_ii for Cholesky, D_i for LDL^T, U_ii for
 * LU.  D has the type of IPARM_FLOAT (the matrix); a single-precision factor
 * is widened.
 */
int
pastixGetDiag( const pastix_data_t *pd, void *D, pastix_int_t incD )
{
    pastix_int_t n;
    int          cplx, rc;
    void        *dd = NULL;
    void        *h;
    if ( pd == NULL ) {
        px_print_error( "pastix_getDiag: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( D == NULL ) {
        px_print_error( "pastix_getDiag: D parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( incD <= 0 ) {
        px_print_error( "pastix_getDiag: incD parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_getDiag: All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    n    = pd->solvmatr->nodenbr;
    cplx = ( px_rhs_flttype( pd ) == SpmComplex64 );
    if ( ( rc = pxd_malloc( pd->dev, (size_t)( n > 0 ? n : 1 ) * ( cplx ? 16 : 8 ), &dd ) ) != PASTIX_SUCCESS ) {
        return rc;
    }
    h  = px_malloc( (size_t)( n > 0 ? n : 1 ) * ( cplx ? 16 : 8 ) );
    rc = pxd_get_diag( pd->dev, dd );
    if ( rc == PASTIX_SUCCESS ) {
        rc = pxd_d2h( pd->dev, h, dd, (size_t)n * ( cplx ? 16 : 8 ) );
    }
    if ( rc == PASTIX_SUCCESS ) {
        for ( pastix_int_t i = 0; i < n; i++ ) {
            if ( cplx ) {
                ( (double complex *)D )[i * incD] = ( (double complex *)h )[i];
            }
            else {
                ( (double *)D )[i * incD] = ( (double *)h )[i];
            }
        }
    }
    free( h );
    pxd_free( pd->dev, dd );
    return rc;
}

void
pastixDumpParam( const pastix_data_t *pd )
{
    FILE *f = fopen( "idparm.csv", "w" );
    if ( f == NULL ) {
        return;
    }
    for ( int i = 0; i < IPARM_SIZE; i++ ) {
        fprintf( f, "iparm_%d,%ld\n", i, (long)pd->iparm[i] );
    }
    for ( int i = 0; i < DPARM_SIZE; i++ ) {
        fprintf( f, "dparm_%d,%e\n", i, pd->dparm[i] );
    }
    fclose( f );
}

int
pastixCheckParam( const pastix_int_t *iparm, const double *dparm )
{
    (void)dparm;
    if ( iparm[IPARM_MIN_BLOCKSIZE] > iparm[IPARM_MAX_BLOCKSIZE] ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( iparm[IPARM_FACTORIZATION] < 0 || iparm[IPARM_FACTORIZATION] > 4 ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Engine introspection (pastix_amd_engine.h). */

int
pastix_amd_solver_sizes( const pastix_data_t *pd, pastix_int_t *cblknbr, pastix_int_t *bloknbr,
                         pastix_int_t *coefnbr, pastix_int_t *nodenbr )
{
    if ( pd == NULL || pd->solvmatr == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( cblknbr ) *cblknbr = pd->solvmatr->cblknbr;
    if ( bloknbr ) *bloknbr = pd->solvmatr->bloknbr;
    if ( coefnbr ) *coefnbr = pd->solvmatr->coefnbr;
    if ( nodenbr ) *nodenbr = pd->solvmatr->nodenbr;
    return PASTIX_SUCCESS;
}

int
pastix_amd_solver_export( const pastix_data_t *pd, pastix_int_t *cblk, pastix_int_t *blok,
                          pastix_int_t *browtab )
{
    const px_solver_t *s;
    if ( pd == NULL || pd->solvmatr == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    s = pd->solvmatr;
    if ( cblk ) {
        for ( pastix_int_t c = 0; c <= s->cblknbr; c++ ) {
            cblk[6 * c + 0] = s->cblktab[c].fcolnum;
            cblk[6 * c + 1] = s->cblktab[c].lcolnum;
            cblk[6 * c + 2] = s->cblktab[c].fblok;
            cblk[6 * c + 3] = s->cblktab[c].stride;
            cblk[6 * c + 4] = s->cblktab[c].coefoff;
            cblk[6 * c + 5] = s->cblktab[c].brownum;
        }
    }
    if ( blok ) {
        for ( pastix_int_t b = 0; b < s->bloknbr; b++ ) {
            blok[5 * b + 0] = s->bloktab[b].frownum;
            blok[5 * b + 1] = s->bloktab[b].lrownum;
            blok[5 * b + 2] = s->bloktab[b].lcblknm;
            blok[5 * b + 3] = s->bloktab[b].fcblknm;
            blok[5 * b + 4] = s->bloktab[b].coefind;
        }
    }
    if ( browtab ) {
        memcpy( browtab, s->browtab, s->brownbr * sizeof( pastix_int_t ) );
    }
    return PASTIX_SUCCESS;
}

int
pastix_amd_order_sizes( const pastix_data_t *pd, pastix_int_t *vertnbr, pastix_int_t *cblknbr )
{
    if ( pd == NULL || pd->ordemesh == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    *vertnbr = pd->ordemesh->vertnbr;
    *cblknbr = pd->ordemesh->cblknbr;
    return PASTIX_SUCCESS;
}

int
pastix_amd_order_export( const pastix_data_t *pd, pastix_int_t *permtab, pastix_int_t *peritab,
                         pastix_int_t *rangtab, pastix_int_t *treetab )
{
    const pastix_order_t *o;
    if ( pd == NULL || pd->ordemesh == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    o = pd->ordemesh;
    if ( permtab ) memcpy( permtab, o->permtab, o->vertnbr * sizeof( pastix_int_t ) );
    if ( peritab ) memcpy( peritab, o->peritab, o->vertnbr * sizeof( pastix_int_t ) );
    if ( rangtab ) memcpy( rangtab, o->rangtab, ( o->cblknbr + 1 ) * sizeof( pastix_int_t ) );
    if ( treetab ) memcpy( treetab, o->treetab, o->cblknbr * sizeof( pastix_int_t ) );
    return PASTIX_SUCCESS;
}

int
pastix_amd_symbol_sizes( const pastix_data_t *pd, pastix_int_t *cblknbr, pastix_int_t *bloknbr )
{
    if ( pd == NULL || pd->symbmtx == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    *cblknbr = pd->symbmtx->cblknbr;
    *bloknbr = pd->symbmtx->bloknbr;
    return PASTIX_SUCCESS;
}

int
pastix_amd_symbol_export( const pastix_data_t *pd, pastix_int_t *cblk, pastix_int_t *blok )
{
    const px_symbol_t *s;
    if ( pd == NULL || pd->symbmtx == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    s = pd->symbmtx;
    for ( pastix_int_t c = 0; c <= s->cblknbr; c++ ) {
        cblk[3 * c + 0] = s->cblktab[c].fcolnum;
        cblk[3 * c + 1] = s->cblktab[c].lcolnum;
        cblk[3 * c + 2] = s->cblktab[c].bloknum;
    }
    for ( pastix_int_t b = 0; b < s->bloknbr; b++ ) {
        blok[4 * b + 0] = s->bloktab[b].frownum;
        blok[4 * b + 1] = s->bloktab[b].lrownum;
        blok[4 * b + 2] = s->bloktab[b].lcblknm;
        blok[4 * b + 3] = s->bloktab[b].fcblknm;
    }
    return PASTIX_SUCCESS;
}

int
pastix_amd_bcsc_sizes( const pastix_data_t *pd, pastix_int_t *n, pastix_int_t *nnz )
{
    if ( pd == NULL || pd->bcsc == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    *n   = pd->bcsc->n;
    *nnz = pd->bcsc->nnz;
    return PASTIX_SUCCESS;
}

int
pastix_amd_bcsc_export( const pastix_data_t *pd, pastix_int_t *colptr, pastix_int_t *rowtab, void *values )
{
    const px_bcsc_t *b;
    if ( pd == NULL || pd->bcsc == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    b = pd->bcsc;
    if ( colptr ) memcpy( colptr, b->colptr, ( b->n + 1 ) * sizeof( pastix_int_t ) );
    if ( rowtab ) memcpy( rowtab, b->rowtab, b->nnz * sizeof( pastix_int_t ) );
    if ( values ) memcpy( values, b->Lvalues, b->nnz * ( b->flttype == SpmComplex64 ? 16 : 8 ) );
    return PASTIX_SUCCESS;
}

int
pastix_amd_get_coeftab( pastix_data_t *pd, int side, void *coeftab, size_t bytes )
{
    if ( pd == NULL || pd->dev == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    return pxd_get_coeftab( pd->dev, side, coeftab, bytes );
}

int
pastix_amd_get_stats( pastix_data_t *pd, double *out, int nstats )
{
    pxd_stats_t st;
    double      v[10];
    if ( pd == NULL || pd->dev == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    pxd_get_stats( pd->dev, &st );
    v[0] = st.fact_seconds;
    v[1] = st.fill_seconds;
    v[2] = st.panel_seconds;
    v[3] = st.trsm_seconds;
    v[4] = st.update_seconds;
    v[5] = st.solve_seconds;
    v[6] = (double)st.levels;
    v[7] = (double)st.update_tiles;
    v[8] = (double)st.launches;
    v[9] = (double)st.coef_bytes;
    for ( int i = 0; i < nstats && i < 10; i++ ) {
        out[i] = v[i];
    }
    return PASTIX_SUCCESS;
}

int
pastix_amd_set_option( pastix_data_t *pd, const char *key, double value )
{
    if ( pd == NULL || key == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( strcmp( key, "berr_target" ) == 0 ) {
        pd->berr_target = value;
        return PASTIX_SUCCESS;
    }
    if ( strcmp( key, "refine_batch" ) == 0 ) {
        pd->refine_nrhs_batch = ( value != 0. );
        return PASTIX_SUCCESS;
    }
    if ( px_device_ready( pd ) != PASTIX_SUCCESS ) {
        return PASTIX_ERR_NOTIMPLEMENTED;
    }
    return pxd_set_option( pd->dev, key, value );
}

int
pastix_amd_device_handle( pastix_data_t *pd, void **ctx )
{
    int rc;
    if ( pd == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( ( rc = px_device_ready( pd ) ) != PASTIX_SUCCESS ) {
        return rc;
    }
    *ctx = pd->dev;
    return PASTIX_SUCCESS;
}

int
pastix_amd_refine_dev( pastix_data_t *pd, pastix_int_t nrhs, void *db, void *dx, pastix_int_t *iters )
{
    if ( pd == NULL || !( pd->steps & STEP_NUMFACT ) ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    return px_refine_run( pd, nrhs, db, dx, iters );
}
