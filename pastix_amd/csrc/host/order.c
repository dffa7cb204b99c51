/*
 * order.c -- ordering structure, the grid nested dissection, the level
 * reordering of supernodes, and the ordering file format.
 *
 * pastixOrderGrid restates order/order_grids.c:137-383 of the reference (the
 * Scotch-free nested dissection of a regular grid) and
 * px_order_apply_level_order restates order/order_apply_level_order.c:95-340
 * with PASTIX_BLEND_DEEPEST_DISTRIB (the reference's default build option,
 * CMakeLists.txt:205-206).  Both are integer algorithms whose output (the
 * permutation, the supernode ranges and the tree) must be identical to the
 * reference's, so the separator choice, the numbering order of vertices and
 * supernodes and the sibling order follow the reference step by step.
 */
#include <stdarg.h>
#include <sys/time.h>
#include <time.h>
#include "px_internal.h"

void
px_print_error( const char *fmt, ... )
{
    va_list ap;
    va_start( ap, fmt );
    fprintf( stderr, "pastix_amd: ERROR: " );
    vfprintf( stderr, fmt, ap );
    fprintf( stderr, "\n" );
    va_end( ap );
}

void
px_print_warning( const char *fmt, ... )
{
    va_list ap;
    va_start( ap, fmt );
    fprintf( stderr, "pastix_amd: WARNING: " );
    vfprintf( stderr, fmt, ap );
    fprintf( stderr, "\n" );
    va_end( ap );
}

void *
px_malloc( size_t size )
{
    void *p = malloc( size > 0 ? size : 1 );
    if ( p == NULL ) {
        px_print_error( "out of host memory (%zu bytes)", size );
        abort();
    }
    return p;
}

void *
px_calloc( size_t n, size_t size )
{
    void *p = calloc( n > 0 ? n : 1, size > 0 ? size : 1 );
    if ( p == NULL ) {
        px_print_error( "out of host memory (%zu bytes)", n * size );
        abort();
    }
    return p;
}

double
px_wtime( void )
{
    struct timespec ts;
    clock_gettime( CLOCK_MONOTONIC, &ts );
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */

int
pastixOrderAlloc( pastix_order_t *ordeptr, pastix_int_t vertnbr, pastix_int_t cblknbr )
{
    if ( ordeptr == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    memset( ordeptr, 0, sizeof( *ordeptr ) );
    ordeptr->vertnbr = vertnbr;
    ordeptr->cblknbr = cblknbr;
    if ( vertnbr > 0 ) {
        ordeptr->permtab = px_malloc( vertnbr * sizeof( pastix_int_t ) );
        ordeptr->peritab = px_malloc( vertnbr * sizeof( pastix_int_t ) );
    }
    if ( cblknbr > 0 ) {
        ordeptr->rangtab = px_calloc( cblknbr + 1, sizeof( pastix_int_t ) );
        ordeptr->treetab = px_calloc( cblknbr, sizeof( pastix_int_t ) );
    }
    return PASTIX_SUCCESS;
}

int
pastixOrderInit( pastix_order_t *ordeptr, pastix_int_t baseval, pastix_int_t vertnbr,
                 pastix_int_t cblknbr, pastix_int_t *permtab, pastix_int_t *peritab,
                 pastix_int_t *rangtab, pastix_int_t *treetab )
{
    if ( ordeptr == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    memset( ordeptr, 0, sizeof( *ordeptr ) );
    ordeptr->baseval = baseval;
    ordeptr->vertnbr = vertnbr;
    ordeptr->cblknbr = cblknbr;
    ordeptr->permtab = permtab;
    ordeptr->peritab = peritab;
    ordeptr->rangtab = rangtab;
    ordeptr->treetab = treetab;
    return PASTIX_SUCCESS;
}

void
pastixOrderExit( pastix_order_t *ordeptr )
{
    if ( ordeptr == NULL ) {
        return;
    }
    free( ordeptr->permtab );
    free( ordeptr->peritab );
    free( ordeptr->rangtab );
    free( ordeptr->treetab );
    free( ordeptr->selevtx );
    free( ordeptr->sndetab );
    free( ordeptr->peritab_exp );
    memset( ordeptr, 0, sizeof( *ordeptr ) );
}

void
pastixOrderBase( pastix_order_t *ordeptr, pastix_int_t baseval )
{
    pastix_int_t d;
    if ( ordeptr == NULL || ordeptr->baseval == baseval ) {
        return;
    }
    d = baseval - ordeptr->baseval;
    for ( pastix_int_t i = 0; i < ordeptr->vertnbr; i++ ) {
        if ( ordeptr->permtab ) ordeptr->permtab[i] += d;
        if ( ordeptr->peritab ) ordeptr->peritab[i] += d;
    }
    if ( ordeptr->rangtab ) {
        for ( pastix_int_t i = 0; i <= ordeptr->cblknbr; i++ ) {
            ordeptr->rangtab[i] += d;
        }
    }
    if ( ordeptr->treetab ) {
        for ( pastix_int_t i = 0; i < ordeptr->cblknbr; i++ ) {
            if ( ordeptr->treetab[i] >= ordeptr->baseval ) {
                ordeptr->treetab[i] += d;
            }
        }
    }
    ordeptr->baseval = baseval;
}

int
pastixOrderCheck( const pastix_order_t *ordeptr )
{
    pastix_int_t  n = ordeptr->vertnbr;
    pastix_int_t  b = ordeptr->baseval;
    pastix_int_t *seen;
    int           rc = PASTIX_SUCCESS;

    if ( ordeptr->permtab == NULL || ordeptr->peritab == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    seen = px_calloc( n, sizeof( pastix_int_t ) );
    for ( pastix_int_t i = 0; i < n; i++ ) {
        pastix_int_t p = ordeptr->permtab[i] - b;
        if ( p < 0 || p >= n || seen[p] ) {
            rc = PASTIX_ERR_BADPARAMETER;
            break;
        }
        seen[p] = 1;
        if ( ordeptr->peritab[p] - b != i ) {
            rc = PASTIX_ERR_BADPARAMETER;
            break;
        }
    }
    free( seen );
    if ( rc == PASTIX_SUCCESS && ordeptr->rangtab ) {
        if ( ordeptr->rangtab[0] != b || ordeptr->rangtab[ordeptr->cblknbr] - b != n ) {
            rc = PASTIX_ERR_BADPARAMETER;
        }
        for ( pastix_int_t c = 0; c < ordeptr->cblknbr && rc == PASTIX_SUCCESS; c++ ) {
            if ( ordeptr->rangtab[c + 1] <= ordeptr->rangtab[c] ) {
                rc = PASTIX_ERR_BADPARAMETER;
            }
            if ( ordeptr->treetab ) {
                pastix_int_t f = ordeptr->treetab[c];
                if ( f != -1 && ( f - b <= c || f - b >= ordeptr->cblknbr ) ) {
                    rc = PASTIX_ERR_BADPARAMETER;
                }
            }
        }
    }
    return rc;
}

int
pastixOrderCopy( pastix_order_t *dst, const pastix_order_t *src )
{
    pastixOrderAlloc( dst, src->vertnbr, src->cblknbr );
    dst->baseval = src->baseval;
    if ( src->permtab ) memcpy( dst->permtab, src->permtab, src->vertnbr * sizeof( pastix_int_t ) );
    if ( src->peritab ) memcpy( dst->peritab, src->peritab, src->vertnbr * sizeof( pastix_int_t ) );
    if ( src->rangtab && dst->rangtab ) memcpy( dst->rangtab, src->rangtab, ( src->cblknbr + 1 ) * sizeof( pastix_int_t ) );
    if ( src->treetab && dst->treetab ) memcpy( dst->treetab, src->treetab, src->cblknbr * sizeof( pastix_int_t ) );
    if ( src->sndetab ) {
        dst->sndenbr = src->sndenbr;
        dst->sndetab = px_malloc( ( src->sndenbr + 1 ) * sizeof( pastix_int_t ) );
        memcpy( dst->sndetab, src->sndetab, ( src->sndenbr + 1 ) * sizeof( pastix_int_t ) );
    }
    return PASTIX_SUCCESS;
}

/*
 * Expand an ordering of the nodes of a matrix with spm->dof unknowns per
 * node to its unknowns (reference order/order.c:396-469): node v at position
 * i becomes the spm->dof consecutive positions of its unknowns dof*v ..
 * dof*v+dof-1, so every supernode range, before and after splitting, is
 * scaled by dof; the tree is unchanged.  The ordering ends with base 0, as
 * the reference's does.  Variable dof (spm->dof <= 0) is not supported here.
 */
void
pastixOrderExpand( pastix_order_t *ordeptr, const spmatrix_t *spm )
{
    pastix_int_t  dof = spm->dof, nv, n, i, j;
    pastix_int_t *peri;
    if ( dof <= 0 ) {
        px_print_error( "pastixOrderExpand: variable dof is not supported" );
        return;
    }
    pastixOrderBase( ordeptr, 0 );
    if ( dof == 1 ) {
        return;
    }
    nv   = ordeptr->vertnbr;
    n    = nv * dof;
    peri = px_malloc( n * sizeof( pastix_int_t ) );
    for ( i = 0; i < nv; i++ ) {
        for ( j = 0; j < dof; j++ ) {
            peri[i * dof + j] = ordeptr->peritab[i] * dof + j;
        }
    }
    free( ordeptr->peritab );
    ordeptr->peritab = peri;
    free( ordeptr->permtab );
    ordeptr->permtab = px_malloc( n * sizeof( pastix_int_t ) );
    for ( i = 0; i < n; i++ ) {
        ordeptr->permtab[peri[i]] = i;
    }
    if ( ordeptr->rangtab ) {
        for ( i = 0; i <= ordeptr->cblknbr; i++ ) {
            ordeptr->rangtab[i] *= dof;
        }
    }
    if ( ordeptr->sndetab ) {
        for ( i = 0; i <= ordeptr->sndenbr; i++ ) {
            ordeptr->sndetab[i] *= dof;
        }
    }
    ordeptr->vertnbr = n;
}

/* ------------------------------------------------------------------------ */
/* Grid nested dissection.                                                  */

typedef struct grid_nd_s {
    pastix_int_t *perm;     /* vertex -> number                         */
    pastix_int_t *rang;     /* per supernode: its highest number        */
    pastix_int_t *depth;    /* per supernode: depth in the ND recursion */
    pastix_int_t  ncblk;
    pastix_int_t  next;     /* next (highest free) number               */
    pastix_int_t  sx, sxy;  /* strides of y and z                       */
} grid_nd_t;

/* Number the vertices of a separator plane, outer coordinate first. */
static void
grid_nd_plane( grid_nd_t *g, pastix_int_t base, pastix_int_t a0, pastix_int_t a1,
               pastix_int_t sa, pastix_int_t b0, pastix_int_t b1, pastix_int_t sb )
{
    for ( pastix_int_t a = a0; a < a1; a++ ) {
        for ( pastix_int_t b = b0; b < b1; b++ ) {
            g->perm[base + a * sa + b * sb] = g->next--;
        }
    }
}

static void
grid_nd_box( grid_nd_t *g, pastix_int_t x0, pastix_int_t x1, pastix_int_t y0, pastix_int_t y1,
             pastix_int_t z0, pastix_int_t z1, pastix_int_t depth )
{
    pastix_int_t nx = x1 - x0, ny = y1 - y0, nz = z1 - z0;

    g->rang[g->ncblk]  = g->next;
    g->depth[g->ncblk] = depth;
    g->ncblk++;

    if ( nx * ny * nz < 15 ) {
        /* small box: one supernode, x outermost, z innermost */
        for ( pastix_int_t i = x0; i < x1; i++ ) {
            for ( pastix_int_t j = y0; j < y1; j++ ) {
                for ( pastix_int_t k = z0; k < z1; k++ ) {
                    g->perm[i + g->sx * j + g->sxy * k] = g->next--;
                }
            }
        }
        return;
    }

    /* cut across the longest direction, x on ties, y before z */
    int dir = 0;
    if ( ny > nx ) {
        dir = 1;
    }
    if ( nz > nx && nz > ny ) {
        dir = 2;
    }

    if ( dir == 0 ) {
        pastix_int_t xs = x0 + nx / 2;
        grid_nd_plane( g, xs, y0, y1, g->sx, z0, z1, g->sxy );
        grid_nd_box( g, x0, xs, y0, y1, z0, z1, depth + 1 );
        grid_nd_box( g, xs + 1, x1, y0, y1, z0, z1, depth + 1 );
    }
    else if ( dir == 1 ) {
        pastix_int_t ys = y0 + ny / 2;
        grid_nd_plane( g, g->sx * ys, x0, x1, 1, z0, z1, g->sxy );
        grid_nd_box( g, x0, x1, y0, ys, z0, z1, depth + 1 );
        grid_nd_box( g, x0, x1, ys + 1, y1, z0, z1, depth + 1 );
    }
    else {
        pastix_int_t zs = z0 + nz / 2;
        grid_nd_plane( g, g->sxy * zs, x0, x1, 1, y0, y1, g->sx );
        grid_nd_box( g, x0, x1, y0, y1, z0, zs, depth + 1 );
        grid_nd_box( g, x0, x1, y0, y1, zs + 1, z1, depth + 1 );
    }
}

int
pastixOrderGrid( pastix_order_t **myorder, pastix_int_t nx, pastix_int_t ny, pastix_int_t nz )
{
    pastix_order_t *ord = *myorder;
    pastix_int_t    n   = nx * ny * nz;
    grid_nd_t       g;
    pastix_int_t    c, cblknbr;

    if ( ord == NULL || n <= 0 ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    pastixOrderAlloc( ord, n, 0 );
    g.perm  = ord->permtab;
    g.rang  = px_malloc( n * sizeof( pastix_int_t ) );
    g.depth = px_malloc( n * sizeof( pastix_int_t ) );
    g.ncblk = 0;
    g.next  = n - 1;
    g.sx    = nx;
    g.sxy   = nx * ny;
    grid_nd_box( &g, 0, nx, 0, ny, 0, nz, 1 );
    cblknbr = g.ncblk;

    for ( pastix_int_t i = 0; i < n; i++ ) {
        ord->peritab[ord->permtab[i]] = i;
    }

    /* Supernodes were created from the highest numbers down: reverse them. */
    ord->cblknbr = cblknbr;
    ord->rangtab = px_malloc( ( cblknbr + 1 ) * sizeof( pastix_int_t ) );
    ord->treetab = px_malloc( cblknbr * sizeof( pastix_int_t ) );
    ord->rangtab[0] = 0;
    for ( c = 0; c < cblknbr; c++ ) {
        ord->rangtab[c + 1] = g.rang[cblknbr - 1 - c] + 1;
        ord->treetab[c]     = g.depth[cblknbr - 1 - c];
    }
    /* The father of a supernode is the next one with a smaller depth. */
    for ( c = 0; c < cblknbr - 1; c++ ) {
        for ( pastix_int_t f = c + 1; f < cblknbr; f++ ) {
            if ( ord->treetab[f] < ord->treetab[c] ) {
                ord->treetab[c] = f;
                break;
            }
        }
    }
    ord->treetab[cblknbr - 1] = -1;

    free( g.rang );
    free( g.depth );
    return PASTIX_SUCCESS;
}

/* ------------------------------------------------------------------------ */
/* Level reordering of the supernodes (top-down breadth first, 2D first).    */

typedef struct px_etree_s {
    pastix_int_t  n;
    pastix_int_t *father;  /* n+1 entries, index -1 is the virtual root (shifted) */
    pastix_int_t *nsons;
    pastix_int_t *fson;
    pastix_int_t *sons;
} px_etree_t;

/* Sons of every node in increasing order; node -1 (stored at 0) is the root. */
static void
etree_build( px_etree_t *et, pastix_int_t n, const pastix_int_t *treetab )
{
    et->n      = n;
    et->father = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) ) ;
    et->nsons  = px_calloc( n + 1, sizeof( pastix_int_t ) );
    et->fson   = px_calloc( n + 2, sizeof( pastix_int_t ) );
    et->sons   = px_malloc( ( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    et->father[0] = -1;
    for ( pastix_int_t i = 0; i < n; i++ ) {
        et->father[i + 1] = treetab[i];
        et->nsons[treetab[i] + 1]++;
    }
    et->fson[0] = 0;
    for ( pastix_int_t i = 0; i <= n; i++ ) {
        et->fson[i + 1] = et->fson[i] + et->nsons[i];
    }
    {
        pastix_int_t *fill = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
        memcpy( fill, et->fson, ( n + 1 ) * sizeof( pastix_int_t ) );
        for ( pastix_int_t i = 0; i < n; i++ ) {
            et->sons[fill[treetab[i] + 1]++] = i;
        }
        free( fill );
    }
}

static void
etree_exit( px_etree_t *et )
{
    free( et->father );
    free( et->nsons );
    free( et->fson );
    free( et->sons );
}

#define ET_FATHER( et, x ) ( ( et )->father[( x ) + 1] )
#define ET_NSONS( et, x )  ( ( et )->nsons[( x ) + 1] )
#define ET_SON( et, x, s ) ( ( et )->sons[( et )->fson[( x ) + 1] + ( s )] )

int
px_order_apply_level_order( pastix_order_t *order, pastix_int_t level_tasks2d,
                            pastix_int_t width_tasks2d )
{
    pastix_int_t  n = order->cblknbr;
    pastix_int_t *sorted, *newfath, *newrang, *newtree, *newperi;
    px_etree_t    et;

    if ( order->rangtab == NULL || order->treetab == NULL ) {
        return PASTIX_ERR_BADPARAMETER;
    }
    if ( n <= 1 ) {
        return PASTIX_SUCCESS;
    }

    etree_build( &et, n, order->treetab );
    sorted  = px_malloc( n * sizeof( pastix_int_t ) );
    newfath = px_malloc( n * sizeof( pastix_int_t ) );
    for ( pastix_int_t i = 0; i < n; i++ ) {
        newfath[i] = order->treetab[i];
    }

    if ( level_tasks2d < 0 ) {
        int8_t      *is2d = px_calloc( n, sizeof( int8_t ) );
        pastix_int_t tot2d = 0, pos2d, pos1d;

        /* Deepest-distribution marking: a supernode wide enough makes itself,
         * its brothers, and every ancestor together with its brothers 2D. */
        for ( pastix_int_t node = 0; node < n; node++ ) {
            pastix_int_t size, f;
            if ( is2d[node] ) {
                continue;
            }
            size = order->rangtab[node + 1] - order->rangtab[node];
            f    = ET_FATHER( &et, node );
            if ( size >= width_tasks2d ) {
                for ( pastix_int_t s = 0; s < ET_NSONS( &et, f ); s++ ) {
                    pastix_int_t b = ET_SON( &et, f, s );
                    if ( !is2d[b] ) {
                        is2d[b] = 1;
                        tot2d++;
                    }
                }
                while ( f != -1 && !is2d[f] ) {
                    f = ET_FATHER( &et, f );
                    for ( pastix_int_t s = 0; s < ET_NSONS( &et, f ); s++ ) {
                        pastix_int_t b = ET_SON( &et, f, s );
                        if ( !is2d[b] ) {
                            is2d[b] = 1;
                            tot2d++;
                        }
                    }
                }
            }
        }

        /* roots first, 2D ones in front */
        pos2d = 0;
        pos1d = tot2d;
        for ( pastix_int_t s = 0; s < ET_NSONS( &et, -1 ); s++ ) {
            pastix_int_t r = ET_SON( &et, -1, s );
            if ( is2d[r] ) {
                sorted[pos2d++] = r;
            }
            else {
                sorted[pos1d++] = r;
            }
        }
        /* breadth first; brothers are stored in reverse so that their
         * relative order survives the final reversal */
        for ( pastix_int_t i = 0; i < n; i++ ) {
            pastix_int_t node = sorted[i];
            pastix_int_t ns = ET_NSONS( &et, node ), n2 = 0, n1, c2 = 0, c1 = 0;
            for ( pastix_int_t s = 0; s < ns; s++ ) {
                if ( is2d[ET_SON( &et, node, s )] ) {
                    n2++;
                }
            }
            n1 = ns - n2;
            for ( pastix_int_t s = 0; s < ns; s++ ) {
                pastix_int_t son = ET_SON( &et, node, s );
                if ( is2d[son] ) {
                    sorted[pos2d + n2 - c2 - 1] = son;
                    c2++;
                }
                else {
                    sorted[pos1d + n1 - c1 - 1] = son;
                    c1++;
                }
                newfath[son] = n - i - 1;
            }
            pos2d += n2;
            pos1d += n1;
        }
        free( is2d );
    }
    else {
        pastix_int_t pos = ET_NSONS( &et, -1 );
        for ( pastix_int_t s = 0; s < pos; s++ ) {
            sorted[s] = ET_SON( &et, -1, s );
        }
        for ( pastix_int_t i = 0; i < n; i++ ) {
            pastix_int_t node = sorted[i];
            pastix_int_t ns   = ET_NSONS( &et, node );
            for ( pastix_int_t s = 0; s < ns; s++ ) {
                pastix_int_t son = ET_SON( &et, node, s );
                sorted[pos + ns - 1 - s] = son;
                newfath[son] = n - i - 1;
            }
            pos += ns;
        }
    }

    /* New supernode i is old supernode sorted[n-1-i]. */
    newrang = px_malloc( ( n + 1 ) * sizeof( pastix_int_t ) );
    newtree = px_malloc( n * sizeof( pastix_int_t ) );
    newperi = px_malloc( order->vertnbr * sizeof( pastix_int_t ) );
    newrang[0] = 0;
    for ( pastix_int_t i = 0; i < n; i++ ) {
        pastix_int_t node = sorted[n - 1 - i];
        pastix_int_t size = order->rangtab[node + 1] - order->rangtab[node];
        newrang[i + 1] = newrang[i] + size;
        newtree[i]     = newfath[node];
        memcpy( newperi + newrang[i], order->peritab + order->rangtab[node],
                size * sizeof( pastix_int_t ) );
    }
    free( order->rangtab );
    free( order->treetab );
    free( order->peritab );
    order->rangtab = newrang;
    order->treetab = newtree;
    order->peritab = newperi;
    for ( pastix_int_t i = 0; i < order->vertnbr; i++ ) {
        order->permtab[order->peritab[i]] = i;
    }

    free( sorted );
    free( newfath );
    etree_exit( &et );
    return PASTIX_SUCCESS;
}

