/*
 * schur.c -- the Schur complement API (reference sopalin/schur.c:30-330) and
 * the ordering step's isolation of the Schur unknowns (reference
 * order/pastix_subtask_order.c:139-225 with graphIsolate, and
 * order/order_add_isolate.c:56-129).
 *
 * The Schur unknowns are numbered last and form the last supernode; blend
 * may split it into several cblks, all of which lie in the Schur part
 * (columns from symbmtx->schurfcol on).  With IPARM_SCHUR_FACT_MODE =
 * PastixFactModeLocal the engine updates those cblks but does not factorize
 * them, so their tiles hold S = A22 - A21 A11^{-1} A12, which pastixGetSchur
 * assembles into a dense array; IPARM_SCHUR_SOLV_MODE selects which parts
 * the solves apply.
 */
#include <complex.h>
#include "px_internal.h"

void
pastixSetSchurUnknownList( pastix_data_t *pd, pastix_int_t n, const pastix_int_t *list )
{
    if ( pd == NULL || n <= 0 || list == NULL ) {
        return;
    }
    free( pd->schur_list );
    pd->schur_n    = n;
    pd->schur_list = px_malloc( (size_t)n * sizeof( pastix_int_t ) );
    memcpy( pd->schur_list, list, (size_t)n * sizeof( pastix_int_t ) );
}

/* The Schur list in 0-based numbering, checked (in range, no repeat). */
static int
px_schur_mark( const pastix_data_t *pd, pastix_int_t n, pastix_int_t base, pastix_int_t *pos )
{
    for ( pastix_int_t i = 0; i < n; i++ ) {
        pos[i] = -1;
    }
    for ( pastix_int_t s = 0; s < pd->schur_n; s++ ) {
        pastix_int_t v = pd->schur_list[s] - base;
        if ( v < 0 || v >= n || pos[v] >= 0 ) {
            px_print_error( "pastixSetSchurUnknownList: invalid or repeated unknown %ld", (long)pd->schur_list[s] );
            return PASTIX_ERR_BADPARAMETER;
        }
        pos[v] = s;
    }
    return PASTIX_SUCCESS;
}

/*
 * Order the graph without the Schur unknowns with `order_fn`, then append
 * them as one last supernode, to which every root of the elimination forest
 * is connected (orderAddIsolate).  The supernodes of the rest are the
 * fundamental ones, in level order.
 */
int
px_order_with_schur( pastix_data_t *pd, const spmatrix_t *spm, pastix_order_t *ord,
                     int ( *order_fn )( const px_graph_t *, pastix_order_t * ) )
{
    const px_graph_t *g  = pd->graph;
    pastix_int_t      n  = g->n, ns = pd->schur_n, nr = n - ns, e = 0;
    pastix_int_t     *pos = px_malloc( (size_t)( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t     *sub = px_malloc( (size_t)( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    px_graph_t        sg;
    pastix_order_t    so;
    int               rc;

    if ( ( rc = px_schur_mark( pd, n, spm->baseval, pos ) ) != PASTIX_SUCCESS ) {
        free( pos );
        free( sub );
        return rc;
    }
    /* the subgraph of the other unknowns */
    for ( pastix_int_t v = 0, k = 0; v < n; v++ ) {
        sub[v] = ( pos[v] < 0 ) ? k++ : -1;
    }
    sg.n      = nr;
    sg.colptr = px_malloc( (size_t)( nr + 1 ) * sizeof( pastix_int_t ) );
    sg.rows   = px_malloc( (size_t)( g->colptr[n] > 0 ? g->colptr[n] : 1 ) * sizeof( pastix_int_t ) );
    sg.colptr[0] = 0;
    for ( pastix_int_t v = 0; v < n; v++ ) {
        if ( sub[v] < 0 ) continue;
        for ( pastix_int_t k = g->colptr[v]; k < g->colptr[v + 1]; k++ ) {
            if ( sub[g->rows[k]] >= 0 ) sg.rows[e++] = sub[g->rows[k]];
        }
        sg.colptr[sub[v] + 1] = e;
    }
    memset( &so, 0, sizeof( so ) );
    rc = order_fn( &sg, &so );
    if ( rc == PASTIX_SUCCESS && ( so.rangtab == NULL || so.treetab == NULL ) ) {
        px_order_fundamental_supernodes( &sg, &so );
    }
    if ( rc == PASTIX_SUCCESS ) {
        px_order_apply_level_order( &so, pd->iparm[IPARM_TASKS2D_LEVEL], pd->iparm[IPARM_TASKS2D_WIDTH] );
        /* orderAddIsolate: the Schur unknowns last, in the order of the list */
        pastixOrderAlloc( ord, n, so.cblknbr + 1 );
        for ( pastix_int_t v = 0; v < n; v++ ) {
            ord->permtab[v] = ( pos[v] < 0 ) ? so.permtab[sub[v]] : nr + pos[v];
        }
        for ( pastix_int_t v = 0; v < n; v++ ) {
            ord->peritab[ord->permtab[v]] = v;
        }
        memcpy( ord->rangtab, so.rangtab, (size_t)( so.cblknbr + 1 ) * sizeof( pastix_int_t ) );
        ord->rangtab[so.cblknbr + 1] = n;
        for ( pastix_int_t c = 0; c < so.cblknbr; c++ ) {
            ord->treetab[c] = ( so.treetab[c] == -1 ) ? so.cblknbr : so.treetab[c];
        }
        ord->treetab[so.cblknbr] = -1;
        ord->baseval             = 0;
    }
    pastixOrderExit( &so );
    free( sg.colptr );
    free( sg.rows );
    free( pos );
    free( sub );
    return rc;
}

/*
 * A personal ordering keeps the Schur unknowns where the user put them: the
 * last n_s positions (pastix_subtask_order.c:141-146).  Its supernodes are
 * cut at the first Schur column and those of the Schur part merged into one.
 */
int
px_order_schur_boundary( pastix_data_t *pd, const spmatrix_t *spm, pastix_order_t *ord )
{
    pastix_int_t  n = ord->vertnbr, ns = pd->schur_n, sf = n - ns, nc = 0, sc;
    pastix_int_t *pos = px_malloc( (size_t)( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    pastix_int_t *rang, *tree, *map;
    int           rc  = px_schur_mark( pd, n, spm->baseval, pos );
    if ( rc == PASTIX_SUCCESS ) {
        for ( pastix_int_t v = 0; v < n; v++ ) {
            if ( ( pos[v] >= 0 ) != ( ord->permtab[v] >= sf ) ) {
                px_print_error( "pastix_subtask_order: with a personal ordering the Schur unknowns must be numbered last" );
                rc = PASTIX_ERR_BADPARAMETER;
                break;
            }
        }
    }
    free( pos );
    if ( rc != PASTIX_SUCCESS ) {
        return rc;
    }
    rang = px_malloc( (size_t)( ord->cblknbr + 2 ) * sizeof( pastix_int_t ) );
    map  = px_malloc( (size_t)( ord->cblknbr + 1 ) * sizeof( pastix_int_t ) );
    /* supernodes below sf (cut at sf), then one for [sf, n) */
    for ( pastix_int_t c = 0; c < ord->cblknbr; c++ ) {
        if ( ord->rangtab[c] < sf ) {
            map[c]     = nc;
            rang[nc++] = ord->rangtab[c];
        }
    }
    sc         = nc;
    rang[nc++] = sf;
    rang[nc]   = n;
    if ( sf == 0 ) {
        nc = 1;
        rang[0] = 0;
        rang[1] = n;
        sc      = 0;
    }
    for ( pastix_int_t c = 0; c < ord->cblknbr; c++ ) {
        if ( ord->rangtab[c] >= sf ) map[c] = sc;
    }
    tree = px_malloc( (size_t)nc * sizeof( pastix_int_t ) );
    for ( pastix_int_t c = 0; c < ord->cblknbr; c++ ) {
        pastix_int_t nk = map[c];
        if ( nk == sc ) continue;
        /* a supernode cut at sf keeps its first part; its father is the Schur one */
        tree[nk] = ( ord->rangtab[c + 1] > sf || ord->treetab[c] < 0 ) ? sc : map[ord->treetab[c]];
        if ( tree[nk] == nk ) tree[nk] = sc;
    }
    tree[sc] = -1;
    free( ord->rangtab );
    free( ord->treetab );
    free( map );
    ord->rangtab = rang;
    ord->treetab = tree;
    ord->cblknbr = nc;
    return PASTIX_SUCCESS;
}

/*
 * The Schur complement as a dense n_s x n_s array (column-major, lds >= n_s):
 * the lower triangle for Cholesky and LDL^T factorizations, the whole matrix
 * for LU (the U side read from the U^T tiles), in the type of IPARM_FLOAT
 * (coeftab_zgetschur, sopalin/coeftab_z.c:607-634; cpucblk_zgetschur_fr).
 */
int
pastixGetSchur( const pastix_data_t *pd, void *S, pastix_int_t lds )
{
    const px_solver_t *sv;
    pastix_int_t       sf, ns, c0, off0, cnt;
    int                lu, cplx, fflt, rc;
    size_t             es;
    void              *hl = NULL, *hu = NULL;
    if ( pd == NULL ) {
        px_print_error( "pastix_getSchur: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( S == NULL ) {
        px_print_error( "pastix_getSchur: S parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( lds <= 0 ) {
        px_print_error( "pastix_getSchur: lds parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( !( pd->steps & STEP_NUMFACT ) ) {
        px_print_error( "pastix_getSchur: All steps from pastix_task_init() to pastix_task_numfact() have to be called before calling this function" );
        return PASTIX_ERR_BADPARAMETER;
    }
    sv = pd->solvmatr;
    sf = pd->symbmtx ? pd->symbmtx->schurfcol : sv->nodenbr;
    ns = sv->nodenbr - sf;
    if ( pd->schur_n <= 0 || ns <= 0 ) {
        px_print_error( "pastix_getSchur: no Schur unknowns were given" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( lds < ns ) {
        px_print_error( "pastix_getSchur: lds parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    for ( c0 = 0; c0 < sv->cblknbr && sv->cblktab[c0].fcolnum < sf; c0++ ) {
    }
    lu   = ( pd->iparm[IPARM_FACTORIZATION] == PastixFactLU );
    cplx = ( pd->factored_flttype == SpmComplex64 || pd->factored_flttype == SpmComplex32 );
    fflt = pd->factored_flttype;
    es   = spm_size_of( fflt );
    off0 = sv->cblktab[c0].coefoff;
    cnt  = sv->coefnbr - off0;
    hl   = px_malloc( (size_t)( cnt > 0 ? cnt : 1 ) * es );
    rc   = pxd_get_coeftab_range( pd->dev, 0, off0, cnt, hl );
    if ( rc == PASTIX_SUCCESS && lu ) {
        hu = px_malloc( (size_t)( cnt > 0 ? cnt : 1 ) * es );
        rc = pxd_get_coeftab_range( pd->dev, 1, off0, cnt, hu );
    }
    if ( rc == PASTIX_SUCCESS ) {
#define PX_SCHUR_AT( buf, k ) \
    ( fflt == SpmDouble ? (double complex)( (const double *)( buf ) )[k] \
    : fflt == SpmFloat  ? (double complex)( (const float *)( buf ) )[k] \
    : fflt == SpmComplex64 ? ( (const double complex *)( buf ) )[k] \
                           : (double complex)( (const float complex *)( buf ) )[k] )
        for ( pastix_int_t j = 0; j < ns; j++ ) {
            for ( pastix_int_t i = 0; i < ns; i++ ) {
                if ( cplx ) ( (double complex *)S )[i + j * lds] = 0.;
                else ( (double *)S )[i + j * lds] = 0.;
            }
        }
        for ( pastix_int_t k = c0; k < sv->cblknbr; k++ ) {
            const px_cblk_t *cb = sv->cblktab + k;
            pastix_int_t     nk = cb->lcolnum - cb->fcolnum + 1;
            for ( pastix_int_t b = cb->fblok; b < cb[1].fblok; b++ ) {
                const px_blok_t *bk   = sv->bloktab + b;
                pastix_int_t     rows = bk->lrownum - bk->frownum + 1;
                pastix_int_t     base = cb->coefoff - off0 + bk->coefind;
                for ( pastix_int_t j = 0; j < nk; j++ ) {
                    for ( pastix_int_t r = 0; r < rows; r++ ) {
                        pastix_int_t   gi = bk->frownum + r - sf, gj = cb->fcolnum + j - sf;
                        double complex v  = PX_SCHUR_AT( hl, base + r + j * rows );
                        if ( !lu && gi < gj ) {
                            continue;   /* Cholesky / LDL^T: the lower triangle */
                        }
                        if ( cplx ) ( (double complex *)S )[gi + gj * lds] += v;
                        else ( (double *)S )[gi + gj * lds] += creal( v );
                        if ( lu ) {
                            /* the U side holds U^T (strictly lower on the diagonal
                             * tiles, whose upper part the L side carries): its
                             * entry (r, j) adds to S(gj, gi), as geadd does */
                            double complex u = PX_SCHUR_AT( hu, base + r + j * rows );
                            if ( cplx ) ( (double complex *)S )[gj + gi * lds] += u;
                            else ( (double *)S )[gj + gi * lds] += creal( u );
                        }
                    }
                }
            }
        }
#undef PX_SCHUR_AT
    }
    free( hl );
    free( hu );
    return rc;
}

/* Rows [schurfcol, n) of the permuted right-hand sides (reference
 * sopalin/schur.c:140-240). */
int
pastixRhsSchurGet( const pastix_data_t *pd, pastix_int_t m, pastix_int_t n, pastix_rhs_t rhsB, void *B,
                   pastix_int_t ldb )
{
    pastix_int_t sf, ns;
    size_t       es;
    if ( pd == NULL ) {
        px_print_error( "pastixRhsSchurGet: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( rhsB == NULL ) {
        px_print_error( "pastixRhsSchurGet: wrong rhsB parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( B == NULL ) {
        px_print_error( "pastixRhsSchurGet: wrong B parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( pd->symbmtx == NULL || pd->schur_n <= 0 ) {
        px_print_error( "pastixRhsSchurGet: no Schur unknowns were given" );
        return PASTIX_ERR_BADPARAMETER;
    }
    sf = pd->symbmtx->schurfcol;
    ns = pd->symbmtx->nodenbr - sf;
    if ( m != ns || n != rhsB->n || ldb < m ) {
        px_print_error( "pastixRhsSchurGet: wrong m or n parameters" );
        return PASTIX_ERR_BADPARAMETER;
    }
    es = spm_size_of( rhsB->flttype );
    for ( pastix_int_t j = 0; j < n; j++ ) {
        memcpy( (char *)B + (size_t)j * ldb * es, (const char *)rhsB->b + ( (size_t)j * rhsB->ld + sf ) * es,
                (size_t)ns * es );
    }
    return PASTIX_SUCCESS;
}

int
pastixRhsSchurSet( const pastix_data_t *pd, pastix_int_t m, pastix_int_t n, void *B, pastix_int_t ldb,
                   pastix_rhs_t rhsB )
{
    pastix_int_t sf, ns;
    size_t       es;
    if ( pd == NULL ) {
        px_print_error( "pastixRhsSchurSet: wrong pastix_data parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( rhsB == NULL ) {
        px_print_error( "pastixRhsSchurSet: wrong rhsB parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( B == NULL ) {
        px_print_error( "pastixRhsSchurSet: wrong B parameter" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( pd->symbmtx == NULL || pd->schur_n <= 0 ) {
        px_print_error( "pastixRhsSchurSet: no Schur unknowns were given" );
        return PASTIX_ERR_BADPARAMETER;
    }
    sf = pd->symbmtx->schurfcol;
    ns = pd->symbmtx->nodenbr - sf;
    if ( m != ns || n != rhsB->n || ldb < m ) {
        px_print_error( "pastixRhsSchurSet: wrong m or n parameters" );
        return PASTIX_ERR_BADPARAMETER;
    }
    es = spm_size_of( rhsB->flttype );
    for ( pastix_int_t j = 0; j < n; j++ ) {
        memcpy( (char *)rhsB->b + ( (size_t)j * rhsB->ld + sf ) * es, (const char *)B + (size_t)j * ldb * es,
                (size_t)ns * es );
    }
    return PASTIX_SUCCESS;
}
