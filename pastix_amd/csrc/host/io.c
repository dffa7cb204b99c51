/*
 * io.c -- the ordering and symbol files and the output directory of an
 * instance, and the dof expansion (pastixExpand).
 *
 *   pastixOrderLoad / pastixOrderSave  reference order/order_io.c:42-103,
 *                                      132-175, 198-263, 293-337
 *   pastixSymbolLoad / pastixSymbolSave reference symbol/symbol_io.c:51-127,
 *                                      147-175
 *   px_gendirectories                  reference common/api.c:76-140
 *                                      (pastix_gendirectories)
 *   pastixExpand                       reference sopalin/pastix.c:506-537
 *
 * The file formats are kept exactly (they are exchanged with the reference
 * and with Scotch-produced orderings): whitespace-separated integers, an
 * ordering file holding version, cblknbr, vertnbr, rangtab, permtab and
 * (version 1) treetab; a symbol file holding version, cblknbr, bloknbr,
 * nodenbr, baseval, then (fcolnum, lcolnum, bloknum) per cblk and (frownum,
 * lrownum, fcblknm) per blok.
 */
#include <complex.h>
#include <errno.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>
#include "px_internal.h"

/* Output directory "<PASTIX_OUTPUT_DIR or pastix>-XXXXXX" of the instance,
 * created on first use.  Returns NULL if it cannot be created. */
const char *
px_gendirectories( pastix_data_t *pd )
{
    if ( pd->dir_global == NULL ) {
        const char *prefix = getenv( "PASTIX_OUTPUT_DIR" );
        size_t      len    = strlen( prefix ? prefix : "pastix" ) + 8;
        char       *name   = px_malloc( len );
        mode_t      old;
        snprintf( name, len, "%s-XXXXXX", prefix ? prefix : "pastix" );
        old = umask( S_IWGRP | S_IWOTH );
        if ( mkdtemp( name ) == NULL ) {
            umask( old );
            px_print_error( "pastix_gendirectories: cannot create %s (%s)", name, strerror( errno ) );
            free( name );
            return NULL;
        }
        umask( old );
        pd->dir_global = name;
    }
    return pd->dir_global;
}

static FILE *
px_fopenw( pastix_data_t *pd, const char *name )
{
    const char *dir = px_gendirectories( pd );
    char       *path;
    FILE       *f;
    size_t      len;
    if ( dir == NULL ) {
        return NULL;
    }
    len  = strlen( dir ) + strlen( name ) + 2;
    path = px_malloc( len );
    snprintf( path, len, "%s/%s", dir, name );
    f = fopen( path, "w" );
    if ( f == NULL ) {
        px_print_error( "pastix_fopenw: cannot open %s for writing", path );
    }
    free( path );
    return f;
}

static int
px_int_load( FILE *f, pastix_int_t *v )
{
    long long x;
    if ( fscanf( f, "%lld", &x ) != 1 ) {
        return 0;
    }
    *v = (pastix_int_t)x;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Ordering */

static int
px_order_load_stream( pastix_order_t *ord, FILE *f )
{
    pastix_int_t vers, cblknbr, vertnbr;
    int          ok = 1;
    if ( !px_int_load( f, &vers ) || !px_int_load( f, &cblknbr ) || !px_int_load( f, &vertnbr ) ||
         ( vers != 0 && vers != 1 ) || cblknbr > vertnbr || cblknbr < 0 || vertnbr < 0 ) {
        px_print_error( "orderLoad: bad input (1)" );
        return PASTIX_ERR_FILE;
    }
    pastixOrderAlloc( ord, vertnbr, cblknbr );
    for ( pastix_int_t c = 0; ok && c <= cblknbr; c++ ) {
        ok = px_int_load( f, &ord->rangtab[c] );
    }
    for ( pastix_int_t i = 0; ok && i < vertnbr; i++ ) {
        ok = px_int_load( f, &ord->permtab[i] );
    }
    if ( vers == 1 ) {
        for ( pastix_int_t c = 0; ok && c < cblknbr; c++ ) {
            ok = px_int_load( f, &ord->treetab[c] );
        }
    }
    else {
        free( ord->treetab );
        ord->treetab = NULL;
    }
    if ( ok ) {
        ord->baseval = ord->rangtab[0];
        for ( pastix_int_t i = 0; i < vertnbr; i++ ) {
            pastix_int_t p = ord->permtab[i] - ord->baseval;
            if ( p < 0 || p >= vertnbr ) {
                ok = 0;
                break;
            }
            ord->peritab[p] = i + ord->baseval;
        }
    }
    if ( !ok ) {
        pastixOrderExit( ord );
        px_print_error( "pastixOrderLoad: bad input (2)" );
        return PASTIX_ERR_FILE;
    }
    return PASTIX_SUCCESS;
}

static int
px_order_save_stream( const pastix_order_t *ord, FILE *f )
{
    pastix_int_t vertnbr;
    int          o = 1;
    if ( ord->permtab == NULL ) {
        px_print_error( "pastixOrderSave: cannot save ordering without direct permutation data" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( ord->rangtab == NULL ) {
        px_print_error( "pastixOrderSave: cannot save ordering without rangtab array" );
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( ord->treetab == NULL ) {
        px_print_error( "pastixOrderSave: cannot save ordering without treetab array" );
        return PASTIX_ERR_BADPARAMETER;
    }
    vertnbr = ord->rangtab[ord->cblknbr] - ord->rangtab[0];
    o       = o && fprintf( f, "1\n%ld\t%ld\n", (long)ord->cblknbr, (long)vertnbr ) >= 0;
    for ( pastix_int_t c = 0; o && c <= ord->cblknbr; c++ ) {
        o = fprintf( f, "%ld%c", (long)ord->rangtab[c], ( ( c & 7 ) == 7 || c == ord->cblknbr ) ? '\n' : '\t' ) >= 0;
    }
    for ( pastix_int_t i = 0; o && i < vertnbr; i++ ) {
        o = fprintf( f, "%ld%c", (long)ord->permtab[i], ( ( i & 7 ) == 7 || i == vertnbr - 1 ) ? '\n' : '\t' ) >= 0;
    }
    for ( pastix_int_t c = 0; o && c < ord->cblknbr; c++ ) {
        o = fprintf( f, "%ld%c", (long)ord->treetab[c], ( ( c & 7 ) == 7 || c == ord->cblknbr - 1 ) ? '\n' : '\t' ) >= 0;
    }
    if ( !o ) {
        px_print_error( "pastixOrderSave: bad output (2)" );
        return PASTIX_ERR_FILE;
    }
    return PASTIX_SUCCESS;
}

/* The file named by PASTIX_FILE_ORDER, else "ordername" in the current
 * directory (order/order_io.c:132-175). */
int
pastixOrderLoad( const pastix_data_t *pd, pastix_order_t *ord )
{
    const char *fn = getenv( "PASTIX_FILE_ORDER" );
    FILE       *f;
    int         rc;
    (void)pd;
    if ( ord == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( fn == NULL ) {
        fn = "ordername";
    }
    f = fopen( fn, "r" );
    if ( f == NULL ) {
        px_print_error( "pastixOrderLoad: cannot open %s", fn );
        return PASTIX_ERR_FILE;
    }
    rc = px_order_load_stream( ord, f );
    fclose( f );
    return rc;
}

/* The instance's output directory, file PASTIX_FILE_ORDER or "ordergen"
 * (order/order_io.c:293-337). */
int
pastixOrderSave( pastix_data_t *pd, const pastix_order_t *ord )
{
    const char *fn = getenv( "PASTIX_FILE_ORDER" );
    FILE       *f;
    int         rc;
    if ( ord == NULL || pd == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    f = px_fopenw( pd, fn ? fn : "ordergen" );
    if ( f == NULL ) {
        return PASTIX_ERR_FILE;
    }
    rc = px_order_save_stream( ord, f );
    if ( rc != PASTIX_SUCCESS ) {
        px_print_error( "cannot save order" );
    }
    fclose( f );
    return rc;
}

int
pastixOrderAllocId( pastix_order_t *ord, pastix_int_t vertnbr )
{
    int rc = pastixOrderAlloc( ord, vertnbr, 1 );
    if ( rc != PASTIX_SUCCESS ) {
        return rc;
    }
    for ( pastix_int_t i = 0; i < vertnbr; i++ ) {
        ord->permtab[i] = i;
        ord->peritab[i] = i;
    }
    ord->rangtab[0] = 0;
    ord->rangtab[1] = vertnbr;
    ord->treetab[0] = -1;
    return PASTIX_SUCCESS;
}

pastix_order_t *
pastixOrderGet( const pastix_data_t *pd )
{
    return pd ? pd->ordemesh : NULL;
}

/* one process: nothing to send */
void
pastixOrderBcast( pastix_order_t *ordemesh, int root, PASTIX_Comm comm )
{
    (void)ordemesh;
    (void)root;
    (void)comm;
}

/* ------------------------------------------------------------------------ */
/* Symbol matrix */

int
pastixSymbolLoad( symbol_matrix_t *symb, FILE *stream )
{
    pastix_int_t vers, cblknbr, bloknbr, nodenbr, baseval;
    if ( symb == NULL || stream == NULL ) {
        return 1;
    }
    if ( !px_int_load( stream, &vers ) || !px_int_load( stream, &cblknbr ) || !px_int_load( stream, &bloknbr ) ||
         !px_int_load( stream, &nodenbr ) || !px_int_load( stream, &baseval ) || vers < 0 || vers > 1 ||
         bloknbr < cblknbr || nodenbr < cblknbr || cblknbr < 0 ) {
        px_print_error( "symbolLoad: bad input (1)" );
        return 1;
    }
    memset( symb, 0, sizeof( *symb ) );
    symb->cblknbr = cblknbr;
    symb->bloknbr = bloknbr;
    symb->nodenbr = nodenbr;
    symb->dof     = 1;
    symb->cblktab = px_calloc( (size_t)cblknbr + 1, sizeof( px_symcblk_t ) );
    symb->bloktab = px_calloc( (size_t)( bloknbr > 0 ? bloknbr : 1 ), sizeof( px_symblok_t ) );
    for ( pastix_int_t c = 0; c < cblknbr; c++ ) {
        px_symcblk_t *cb = symb->cblktab + c;
        if ( !px_int_load( stream, &cb->fcolnum ) || !px_int_load( stream, &cb->lcolnum ) ||
             !px_int_load( stream, &cb->bloknum ) || cb->fcolnum > cb->lcolnum ) {
            px_print_error( "symbolLoad: bad input (2)" );
            px_symbol_exit( symb );
            return 1;
        }
        cb->fcolnum -= baseval;
        cb->lcolnum -= baseval;
        cb->bloknum -= baseval;
    }
    symb->cblktab[cblknbr].fcolnum = symb->cblktab[cblknbr].lcolnum = nodenbr;
    symb->cblktab[cblknbr].bloknum = bloknbr;
    for ( pastix_int_t b = 0; b < bloknbr; b++ ) {
        px_symblok_t *bk = symb->bloktab + b;
        if ( !px_int_load( stream, &bk->frownum ) || !px_int_load( stream, &bk->lrownum ) ||
             !px_int_load( stream, &bk->fcblknm ) || bk->frownum > bk->lrownum ) {
            px_print_error( "symbolLoad: bad input (3)" );
            px_symbol_exit( symb );
            return 1;
        }
        bk->frownum -= baseval;
        bk->lrownum -= baseval;
        bk->fcblknm -= baseval;
    }
    /* the owning cblk of every blok follows from the cblks' blok ranges */
    for ( pastix_int_t c = 0; c < cblknbr; c++ ) {
        pastix_int_t b0 = symb->cblktab[c].bloknum, b1 = symb->cblktab[c + 1].bloknum;
        if ( b0 < 0 || b1 > bloknbr || b1 < b0 ) {
            px_print_error( "symbolLoad: bad input (4)" );
            px_symbol_exit( symb );
            return 1;
        }
        for ( pastix_int_t b = b0; b < b1; b++ ) {
            symb->bloktab[b].lcblknm = c;
        }
    }
    px_symbol_build_rowtab( symb );
    return 0;
}

void
pastixSymbolExit( symbol_matrix_t *symb )
{
    if ( symb ) {
        px_symbol_exit( symb );
    }
}

int
pastixSymbolSave( const symbol_matrix_t *symb, FILE *stream )
{
    int o;
    if ( symb == NULL || stream == NULL ) {
        return 1;
    }
    o = ( fprintf( stream, "1\n%ld\t%ld\t%ld\t%ld\n", (long)symb->cblknbr, (long)symb->bloknbr,
                   (long)symb->nodenbr, 0L ) < 0 );
    for ( pastix_int_t c = 0; c < symb->cblknbr && !o; c++ ) {
        const px_symcblk_t *cb = symb->cblktab + c;
        o = ( fprintf( stream, "%ld\t%ld\t%ld\n", (long)cb->fcolnum, (long)cb->lcolnum, (long)cb->bloknum ) < 0 );
    }
    for ( pastix_int_t b = 0; b < symb->bloknbr && !o; b++ ) {
        const px_symblok_t *bk = symb->bloktab + b;
        o = ( fprintf( stream, "%ld\t%ld\t%ld\n", (long)bk->frownum, (long)bk->lrownum, (long)bk->fcblknm ) < 0 );
    }
    return o;
}

/* Load the symbol of the IPARM_IO_STRATEGY PastixIOLoad path: the file
 * "symbname" in the current directory (symbol/pastix_subtask_symbfact.c:
 * 156-163). */
int
px_symbol_load_default( px_symbol_t *symb )
{
    FILE *f = fopen( "symbname", "r" );
    int   rc;
    if ( f == NULL ) {
        px_print_error( "pastix_fopen: Couldn't open file: symbname with mode r" );
        return PASTIX_ERR_FILE;
    }
    rc = pastixSymbolLoad( symb, f );
    fclose( f );
    return rc ? PASTIX_ERR_FILE : PASTIX_SUCCESS;
}

/* ... and its PastixIOSave counterpart: "symbgen" in the output directory
 * (:275-284). */
int
px_symbol_save_default( pastix_data_t *pd, const px_symbol_t *symb )
{
    FILE *f = px_fopenw( pd, "symbgen" );
    int   rc;
    if ( f == NULL ) {
        return PASTIX_ERR_FILE;
    }
    rc = pastixSymbolSave( symb, f );
    fclose( f );
    return rc ? PASTIX_ERR_FILE : PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* dof expansion */

/* Every node of the symbol becomes dof consecutive unknowns (reference
 * symbol/symbol_expand.c for a constant dof). */
void
px_symbol_expand( px_symbol_t *symb, pastix_int_t dof )
{
    if ( dof <= 1 ) {
        return;
    }
    for ( pastix_int_t c = 0; c <= symb->cblknbr; c++ ) {
        symb->cblktab[c].fcolnum = symb->cblktab[c].fcolnum * dof;
        symb->cblktab[c].lcolnum = ( c < symb->cblknbr ) ? symb->cblktab[c].lcolnum * dof + dof - 1
                                                         : symb->cblktab[c].lcolnum * dof;
    }
    for ( pastix_int_t b = 0; b < symb->bloknbr; b++ ) {
        symb->bloktab[b].frownum = symb->bloktab[b].frownum * dof;
        symb->bloktab[b].lrownum = symb->bloktab[b].lrownum * dof + dof - 1;
    }
    symb->nodenbr *= dof;
    symb->dof = 1;
}

/*
 * A matrix with a constant dof > 1 (values stored as dof x dof column-major
 * blocks per node entry) becomes the scalar CSC of nodes x dof unknowns,
 * unknown a of node I numbered I * dof + a.  For symmetric and Hermitian
 * storage the diagonal blocks keep their lower triangle.
 */
int
px_spm_expand( spmatrix_t *spm )
{
    spm_int_t d = spm->dof, n = spm->n, nnz = spm->nnz, base = spm->baseval, ne = 0;
    int       cplx = ( spm->flttype == SpmComplex64 );
    int       sym  = ( spm->mtxtype != SpmGeneral );
    spm_int_t *cp, *rp;
    void      *vv;
    if ( d <= 1 ) {
        return PASTIX_SUCCESS;
    }
    if ( spm->fmttype != SpmCSC || spm->dofs != NULL ) {
        px_print_error( "pastixExpand: only CSC matrices with a constant dof are expanded" );
        return PASTIX_ERR_BADPARAMETER;
    }
    cp = px_calloc( (size_t)( n * d + 1 ), sizeof( spm_int_t ) );
    rp = px_malloc( (size_t)( nnz * d * d > 0 ? nnz * d * d : 1 ) * sizeof( spm_int_t ) );
    vv = px_malloc( (size_t)( nnz * d * d > 0 ? nnz * d * d : 1 ) * ( cplx ? 16 : 8 ) );
    for ( spm_int_t nj = 0; nj < n; nj++ ) {
        for ( spm_int_t b = 0; b < d; b++ ) {
            spm_int_t col = nj * d + b;
            for ( spm_int_t k = spm->colptr[nj] - base; k < spm->colptr[nj + 1] - base; k++ ) {
                spm_int_t ni = spm->rowptr[k] - base;
                for ( spm_int_t a = 0; a < d; a++ ) {
                    spm_int_t row = ni * d + a;
                    size_t    src = (size_t)k * d * d + ( spm->layout == PastixColMajor ? a + b * d : b + a * d );
                    if ( sym && ni == nj && a < b ) {
                        continue;
                    }
                    rp[ne] = row;
                    if ( cplx ) {
                        ( (double complex *)vv )[ne] = ( (const double complex *)spm->values )[src];
                    }
                    else {
                        ( (double *)vv )[ne] = ( (const double *)spm->values )[src];
                    }
                    ne++;
                }
            }
            cp[col + 1] = ne;
        }
    }
    free( spm->colptr );
    free( spm->rowptr );
    free( spm->values );
    spm->colptr  = cp;
    spm->rowptr  = rp;
    spm->values  = vv;
    spm->n       = n * d;
    spm->nnz     = ne;
    spm->dof     = 1;
    spm->baseval = 0;
    spmUpdateComputedFields( spm );
    return PASTIX_SUCCESS;
}

/* sopalin/pastix.c:506-537: the ordering, the symbol and the matrix of a
 * constant-dof problem expanded to scalar unknowns. */
void
pastixExpand( const pastix_data_t *pd, spmatrix_t *spm )
{
    if ( spm == NULL || spm->dof == 1 ) {
        return;
    }
    if ( pd && ( pd->steps & STEP_ORDERING ) && pd->ordemesh != NULL ) {
        pastixOrderExpand( pd->ordemesh, spm );
    }
    if ( pd && ( pd->steps & STEP_SYMBFACT ) && pd->symbmtx != NULL ) {
        px_symbol_expand( pd->symbmtx, spm->dof );
    }
    px_spm_expand( spm );
}
