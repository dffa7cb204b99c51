/*
 * solver.c -- generation of the solver matrix (cblk/blok tile layout) from
 * the split symbol matrix, for one process.
 *
 * Restates, for a single process and full-rank storage, the rules of
 * blend/solver_matrix_gen.c:86-403 and blend/solver_matrix_gen_utils.c:
 *  - every cblk uses the 2D tile layout (candBuild sets CBLK_LAYOUT_2D on
 *    every cblk, blend/cand.c:716-724): blok k of a cblk with n columns is a
 *    column-major rows(k) x n tile at coefind = (rows above k) * n
 *    (blend/solver_matrix_gen_utils.h:90), the cblk occupies stride * n;
 *  - CBLK_TASKS_2D follows candSubTreeDistribDeepestWidth (blend/cand.c:455):
 *    a cblk is 2D when it is at least width_tasks2d wide or one of its sons is;
 *  - browtab lists, per cblk, the incoming bloks of 1D sources first and of
 *    2D sources after them, each group in symbol order
 *    (solvMatGen_reorder_browtab, blend/solver_matrix_gen_utils.c:158-233);
 *  - offdmax and blokmax follow solvMatGen_max_buffers (:729-794).
 */
#include "px_internal.h"

void
px_solver_exit( px_solver_t *solv )
{
    free( solv->cblktab );
    free( solv->bloktab );
    free( solv->browtab );
    free( solv->tasktab );
    memset( solv, 0, sizeof( *solv ) );
}

static int8_t
deepest_width( pastix_int_t c, const px_symbol_t *s, const pastix_int_t *fson,
               const pastix_int_t *sons, pastix_int_t width2d, int8_t *type )
{
    int8_t       st = 0;
    pastix_int_t w  = s->cblktab[c].lcolnum - s->cblktab[c].fcolnum + 1;
    int8_t       t  = CBLK_LAYOUT_2D | CBLK_TASKS_2D;
    for ( pastix_int_t k = fson[c + 1]; k < fson[c + 2]; k++ ) {
        st |= deepest_width( sons[k], s, fson, sons, width2d, type );
    }
    if ( w < width2d ) {
        t &= ~CBLK_TASKS_2D;
    }
    type[c] = t | st;
    return type[c];
}

int
px_solver_gen( px_solver_t *solv, const px_symbol_t *s, pastix_int_t width_tasks2d,
               pastix_int_t level_tasks2d )
{
    pastix_int_t nc = s->cblknbr, nb = s->bloknbr;
    int8_t      *type = px_calloc( nc > 0 ? nc : 1, sizeof( int8_t ) );

    memset( solv, 0, sizeof( *solv ) );

    /* elimination tree of the symbol: father = facing cblk of the first off-diagonal blok */
    {
        pastix_int_t *fath = px_malloc( nc * sizeof( pastix_int_t ) );
        pastix_int_t *fson = px_calloc( nc + 2, sizeof( pastix_int_t ) );
        pastix_int_t *sons = px_malloc( ( nc > 0 ? nc : 1 ) * sizeof( pastix_int_t ) );
        pastix_int_t *fill = px_malloc( ( nc + 1 ) * sizeof( pastix_int_t ) );
        for ( pastix_int_t c = 0; c < nc; c++ ) {
            fath[c] = ( s->cblktab[c + 1].bloknum - s->cblktab[c].bloknum > 1 )
                        ? s->bloktab[s->cblktab[c].bloknum + 1].fcblknm : -1;
            fson[fath[c] + 2]++;
        }
        for ( pastix_int_t c = 0; c <= nc; c++ ) {
            fson[c + 1] += fson[c];
        }
        memcpy( fill, fson, ( nc + 1 ) * sizeof( pastix_int_t ) );
        for ( pastix_int_t c = 0; c < nc; c++ ) {
            sons[fill[fath[c] + 1]++] = c;
        }
        if ( level_tasks2d < 0 ) {
            for ( pastix_int_t k = fson[0]; k < fson[1]; k++ ) {
                deepest_width( sons[k], s, fson, sons, width_tasks2d, type );
            }
        }
        else {
            /* 1D everywhere (level 0), or 2D down to the given tree depth */
            for ( pastix_int_t c = nc - 1; c >= 0; c-- ) {
                pastix_int_t depth = 0;
                for ( pastix_int_t f = fath[c]; f != -1; f = fath[f] ) {
                    depth++;
                }
                type[c] = CBLK_LAYOUT_2D | ( ( depth < level_tasks2d ) ? CBLK_TASKS_2D : 0 );
            }
        }
        free( fath );
        free( fson );
        free( sons );
        free( fill );
    }

    solv->cblknbr = nc;
    solv->bloknbr = nb;
    solv->nodenbr = s->nodenbr;
    solv->brownbr = nb - nc;
    solv->cblktab = px_calloc( nc + 1, sizeof( px_cblk_t ) );
    solv->bloktab = px_calloc( nb + 1, sizeof( px_blok_t ) );
    solv->browtab = px_malloc( ( solv->brownbr > 0 ? solv->brownbr : 1 ) * sizeof( pastix_int_t ) );

    {
        pastix_int_t coefnbr = 0, brownum = 0;
        for ( pastix_int_t c = 0; c < nc; c++ ) {
            const px_symcblk_t *sc = s->cblktab + c;
            px_cblk_t          *cb = solv->cblktab + c;
            pastix_int_t        ncol = sc->lcolnum - sc->fcolnum + 1, stride = 0;

            for ( pastix_int_t b = sc->bloknum; b < sc[1].bloknum; b++ ) {
                px_blok_t *bk = solv->bloktab + b;
                bk->frownum = s->bloktab[b].frownum;
                bk->lrownum = s->bloktab[b].lrownum;
                bk->lcblknm = c;
                bk->fcblknm = s->bloktab[b].fcblknm;
                bk->coefind = stride * ncol;
                bk->browind = -1;
                stride += bk->lrownum - bk->frownum + 1;
            }
            cb->fcolnum  = sc->fcolnum;
            cb->lcolnum  = sc->lcolnum;
            cb->fblok    = sc->bloknum;
            cb->stride   = stride;
            cb->coefoff  = coefnbr;
            cb->lcolidx  = sc->fcolnum;
            cb->cblktype = type[c];
            cb->priority = c;
            coefnbr += stride * ncol;
            solv->colmax = px_imax( solv->colmax, ncol );
            if ( type[c] & CBLK_TASKS_2D ) {
                solv->nb2dcblk++;
                solv->nb2dblok += sc[1].bloknum - sc->bloknum;
            }

            /* incoming bloks: 1D sources, then 2D sources */
            cb->brownum = brownum;
            for ( pastix_int_t k = sc->brownum; k < sc[1].brownum; k++ ) {
                pastix_int_t b = s->browtab[k];
                if ( !( type[s->bloktab[b].lcblknm] & CBLK_TASKS_2D ) ) {
                    solv->bloktab[b].browind = brownum;
                    solv->browtab[brownum++] = b;
                }
            }
            cb->brown2d = brownum;
            for ( pastix_int_t k = sc->brownum; k < sc[1].brownum; k++ ) {
                pastix_int_t b = s->browtab[k];
                if ( type[s->bloktab[b].lcblknm] & CBLK_TASKS_2D ) {
                    solv->bloktab[b].browind = brownum;
                    solv->browtab[brownum++] = b;
                }
            }
        }
        /* sentinels */
        solv->cblktab[nc].fcolnum = ( nc > 0 ) ? solv->cblktab[nc - 1].lcolnum + 1 : 0;
        solv->cblktab[nc].lcolnum = solv->cblktab[nc].fcolnum;
        solv->cblktab[nc].fblok   = nb;
        solv->cblktab[nc].stride  = 0;
        solv->cblktab[nc].coefoff = coefnbr;
        solv->cblktab[nc].brownum = brownum;
        solv->cblktab[nc].lcolidx = s->nodenbr;
        solv->bloktab[nb].frownum = solv->cblktab[nc].fcolnum;
        solv->bloktab[nb].lrownum = solv->cblktab[nc].fcolnum;
        solv->bloktab[nb].lcblknm = nc;
        solv->bloktab[nb].fcblknm = nc;
        solv->coefnbr = coefnbr;
    }

    /* workspace sizes (solvMatGen_max_buffers) */
    for ( pastix_int_t c = 0; c < nc; c++ ) {
        px_cblk_t   *cb = solv->cblktab + c;
        pastix_int_t n = cb->lcolnum - cb->fcolnum + 1;
        pastix_int_t m = cb->stride - n, cblk_m = -1, acc_m = 0;
        solv->offdmax = px_imax( solv->offdmax, m * n );
        for ( pastix_int_t b = cb->fblok + 1; b < cb[1].fblok; b++ ) {
            pastix_int_t k = solv->bloktab[b].lrownum - solv->bloktab[b].frownum + 1;
            if ( cb->cblktype & ( CBLK_TASKS_2D | CBLK_COMPRESSED ) ) {
                if ( solv->bloktab[b].fcblknm == cblk_m ) {
                    acc_m += k;
                }
                else {
                    cblk_m = solv->bloktab[b].fcblknm;
                    acc_m  = k;
                }
                solv->blokmax = px_imax( solv->blokmax, n * acc_m );
            }
        }
    }

    /* static task list: cblks in elimination order */
    solv->tasktab = px_malloc( ( nc > 0 ? nc : 1 ) * sizeof( pastix_int_t ) );
    for ( pastix_int_t c = 0; c < nc; c++ ) {
        solv->tasktab[c] = c;
    }
    free( type );
    return PASTIX_SUCCESS;
}

int
px_solver_check( const px_solver_t *solv )
{
    for ( pastix_int_t c = 0; c < solv->cblknbr; c++ ) {
        const px_cblk_t *cb = solv->cblktab + c;
        pastix_int_t     ncol = cb->lcolnum - cb->fcolnum + 1, rows = 0;
        for ( pastix_int_t b = cb->fblok; b < cb[1].fblok; b++ ) {
            const px_blok_t *bk = solv->bloktab + b;
            if ( bk->coefind != rows * ncol || bk->lcblknm != c ) {
                return PASTIX_ERR_INTERNAL;
            }
            rows += bk->lrownum - bk->frownum + 1;
        }
        if ( rows != cb->stride || cb[1].coefoff != cb->coefoff + rows * ncol ) {
            return PASTIX_ERR_INTERNAL;
        }
    }
    return PASTIX_SUCCESS;
}
