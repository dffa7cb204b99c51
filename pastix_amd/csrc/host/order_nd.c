/*
 * order_nd.c -- a nested-dissection ordering of a general graph, for the
 * orderings the reference delegates to Scotch or Metis
 * (order/order_compute_scotch.c, order/order_compute_metis.c), which are not
 * part of this build.
 *
 * Method: George's automatic nested dissection (A. George, J. W. H. Liu,
 * "Computer Solution of Large Sparse Positive Definite Systems", 1981,
 * ch. 8): in each connected part, a pseudo-peripheral vertex (Gibbs, Poole,
 * Stockmeyer 1976) roots a breadth-first level structure; the level where the
 * visited count passes half the part becomes the separator, thinned to the
 * vertices adjacent to the far side; both halves are ordered recursively and
 * the separator last.  Parts of at most PX_ND_LEAF vertices are ordered by
 * minimum degree on their own elimination graph.  The supernodes are then
 * the fundamental supernodes of the resulting elimination tree.
 *
 * The result is a valid fill-reducing ordering, not Scotch's: the factor
 * sizes and flop counts differ from a Scotch run of the reference.
 */
#include "px_internal.h"

#define PX_ND_LEAF 64

typedef struct {
    const px_graph_t *g;
    pastix_int_t     *part;    /* part label of every vertex; -1: ordered */
    pastix_int_t     *level;   /* BFS level scratch                       */
    pastix_int_t     *queue;
    pastix_int_t     *order;   /* output: the elimination order           */
    pastix_int_t      next;    /* next position in order                  */
    pastix_int_t      nparts;  /* label generator                         */
} nd_t;

/* BFS of part p from r: levels in nd->level, visit order in nd->queue;
 * returns the number of vertices reached, *height the last level. */
static pastix_int_t
nd_bfs( nd_t *nd, pastix_int_t p, pastix_int_t r, pastix_int_t *height )
{
    const px_graph_t *g = nd->g;
    pastix_int_t      head = 0, tail = 0;
    nd->queue[tail++] = r;
    nd->level[r]      = 0;
    *height           = 0;
    while ( head < tail ) {
        pastix_int_t v = nd->queue[head++];
        for ( pastix_int_t e = g->colptr[v]; e < g->colptr[v + 1]; e++ ) {
            pastix_int_t w = g->rows[e];
            if ( nd->part[w] == p && nd->level[w] < 0 ) {
                nd->level[w]      = nd->level[v] + 1;
                *height           = nd->level[w];
                nd->queue[tail++] = w;
            }
        }
    }
    return tail;
}

static void
nd_clear_levels( nd_t *nd, pastix_int_t cnt )
{
    for ( pastix_int_t i = 0; i < cnt; i++ ) {
        nd->level[nd->queue[i]] = -1;
    }
}

static pastix_int_t
nd_degree( const nd_t *nd, pastix_int_t v, pastix_int_t p )
{
    pastix_int_t d = 0;
    for ( pastix_int_t e = nd->g->colptr[v]; e < nd->g->colptr[v + 1]; e++ ) {
        d += ( nd->part[nd->g->rows[e]] == p );
    }
    return d;
}

/* Minimum-degree order of the vertices verts[0..n) (one part, n small) on
 * their own elimination graph (dense boolean adjacency). */
static void
nd_leaf( nd_t *nd, const pastix_int_t *verts, pastix_int_t n, pastix_int_t p )
{
    const px_graph_t *g   = nd->g;
    unsigned char    *adj = px_calloc( (size_t)n * n, 1 );
    unsigned char    *done = px_calloc( (size_t)n, 1 );
    pastix_int_t     *deg = px_calloc( (size_t)n, sizeof( pastix_int_t ) );
    /* local index of a vertex: its position in verts, found through level[] */
    for ( pastix_int_t i = 0; i < n; i++ ) {
        nd->level[verts[i]] = i;
    }
    for ( pastix_int_t i = 0; i < n; i++ ) {
        pastix_int_t v = verts[i];
        for ( pastix_int_t e = g->colptr[v]; e < g->colptr[v + 1]; e++ ) {
            pastix_int_t w = g->rows[e];
            if ( nd->part[w] == p ) {
                adj[i * n + nd->level[w]] = 1;
            }
        }
    }
    for ( pastix_int_t i = 0; i < n; i++ ) {
        nd->level[verts[i]] = -1;
        for ( pastix_int_t j = 0; j < n; j++ ) deg[i] += adj[i * n + j];
    }
    for ( pastix_int_t k = 0; k < n; k++ ) {
        pastix_int_t best = -1;
        for ( pastix_int_t i = 0; i < n; i++ ) {
            if ( !done[i] && ( best < 0 || deg[i] < deg[best] ) ) best = i;
        }
        done[best]             = 1;
        nd->order[nd->next++] = verts[best];
        nd->part[verts[best]]  = -1;
        /* eliminating best joins its remaining neighbours into a clique */
        for ( pastix_int_t i = 0; i < n; i++ ) {
            if ( done[i] || !adj[best * n + i] ) continue;
            adj[i * n + best] = 0;
            deg[i]--;
            for ( pastix_int_t j = 0; j < n; j++ ) {
                if ( j != i && !done[j] && adj[best * n + j] && !adj[i * n + j] ) {
                    adj[i * n + j] = 1;
                    deg[i]++;
                }
            }
        }
    }
    free( adj );
    free( done );
    free( deg );
}

static int
nd_cmp_pair( const void *a, const void *b )
{
    const pastix_int_t *x = a, *y = b;
    if ( x[0] != y[0] ) return ( x[0] < y[0] ) ? -1 : 1;
    return ( x[1] < y[1] ) ? -1 : ( x[1] > y[1] );
}

/* Order the part labelled p, whose vertices are verts[0..n) (the array is
 * scratch owned by this call). */
static void
nd_part( nd_t *nd, pastix_int_t *verts, pastix_int_t n, pastix_int_t p )
{
    const px_graph_t *g = nd->g;
    pastix_int_t      r, height, reached, prev_h = -1;

    if ( n <= PX_ND_LEAF ) {
        nd_leaf( nd, verts, n, p );
        return;
    }
    /* pseudo-peripheral root: from a minimum-degree vertex, restart from a
     * minimum-degree vertex of the last level while the height grows */
    r = verts[0];
    for ( pastix_int_t i = 1; i < n; i++ ) {
        if ( nd_degree( nd, verts[i], p ) < nd_degree( nd, r, p ) ) r = verts[i];
    }
    for ( int it = 0; it < 8; it++ ) {
        pastix_int_t best = -1;
        reached = nd_bfs( nd, p, r, &height );
        if ( height <= prev_h ) {
            nd_clear_levels( nd, reached );
            break;
        }
        prev_h = height;
        for ( pastix_int_t i = reached - 1; i >= 0 && nd->level[nd->queue[i]] == height; i-- ) {
            pastix_int_t v = nd->queue[i];
            if ( best < 0 || nd_degree( nd, v, p ) < nd_degree( nd, best, p ) ) best = v;
        }
        nd_clear_levels( nd, reached );
        if ( best == r ) break;
        r = best;
    }
    reached = nd_bfs( nd, p, r, &height );
    if ( reached < n ) {
        /* several components: the one reached, then the rest */
        pastix_int_t  pa = ++nd->nparts, pb = ++nd->nparts;
        pastix_int_t *va = px_malloc( (size_t)reached * sizeof( pastix_int_t ) );
        pastix_int_t *vb = px_malloc( (size_t)( n - reached ) * sizeof( pastix_int_t ) );
        pastix_int_t  na = 0, nb = 0;
        for ( pastix_int_t i = 0; i < n; i++ ) {
            pastix_int_t v = verts[i];
            if ( nd->level[v] >= 0 ) { va[na++] = v; nd->part[v] = pa; }
            else { vb[nb++] = v; nd->part[v] = pb; }
        }
        nd_clear_levels( nd, reached );
        nd_part( nd, va, na, pa );
        free( va );
        nd_part( nd, vb, nb, pb );
        free( vb );
        return;
    }
    if ( height < 2 ) {
        /* no level can separate (a near clique): minimum degree, or for a
         * large part increasing degree (the dense elimination graph would
         * cost n^3) */
        nd_clear_levels( nd, reached );
        if ( n <= 8 * PX_ND_LEAF ) {
            nd_leaf( nd, verts, n, p );
        }
        else {
            pastix_int_t *key = px_malloc( (size_t)n * 2 * sizeof( pastix_int_t ) );
            for ( pastix_int_t i = 0; i < n; i++ ) {
                key[2 * i]     = nd_degree( nd, verts[i], p );
                key[2 * i + 1] = verts[i];
            }
            qsort( key, (size_t)n, 2 * sizeof( pastix_int_t ), nd_cmp_pair );
            for ( pastix_int_t i = 0; i < n; i++ ) {
                nd->order[nd->next++] = key[2 * i + 1];
                nd->part[key[2 * i + 1]] = -1;
            }
            free( key );
        }
        return;
    }
    {
        /* the separator level: where the visited count passes n / 2 */
        pastix_int_t  m = 1, cnt = 0, pa = ++nd->nparts, pb = ++nd->nparts, ps = ++nd->nparts;
        pastix_int_t *va, *vb, *vs, na = 0, nb = 0, ns = 0;
        for ( pastix_int_t i = 0; i < reached; i++ ) {
            if ( ++cnt >= n / 2 ) {
                m = nd->level[nd->queue[i]];
                break;
            }
        }
        if ( m < 1 ) m = 1;
        if ( m >= height ) m = height - 1;
        va = px_malloc( (size_t)n * sizeof( pastix_int_t ) );
        vb = px_malloc( (size_t)n * sizeof( pastix_int_t ) );
        vs = px_malloc( (size_t)n * sizeof( pastix_int_t ) );
        for ( pastix_int_t i = 0; i < n; i++ ) {
            pastix_int_t v = verts[i], l = nd->level[v];
            if ( l < m ) { va[na++] = v; }
            else if ( l > m ) { vb[nb++] = v; }
        }
        /* thin the separator: a level-m vertex with no neighbour beyond
         * level m separates nothing and joins the near side */
        for ( pastix_int_t i = 0; i < n; i++ ) {
            pastix_int_t v = verts[i];
            int          far = 0;
            if ( nd->level[v] != m ) continue;
            for ( pastix_int_t e = g->colptr[v]; e < g->colptr[v + 1] && !far; e++ ) {
                pastix_int_t w = g->rows[e];
                far = ( nd->part[w] == p && nd->level[w] == m + 1 );
            }
            if ( far ) vs[ns++] = v;
            else va[na++] = v;
        }
        nd_clear_levels( nd, reached );
        for ( pastix_int_t i = 0; i < na; i++ ) nd->part[va[i]] = pa;
        for ( pastix_int_t i = 0; i < nb; i++ ) nd->part[vb[i]] = pb;
        for ( pastix_int_t i = 0; i < ns; i++ ) nd->part[vs[i]] = ps;
        nd_part( nd, va, na, pa );
        free( va );
        nd_part( nd, vb, nb, pb );
        free( vb );
        for ( pastix_int_t i = 0; i < ns; i++ ) {
            nd->order[nd->next++] = vs[i];
            nd->part[vs[i]]       = -1;
        }
        free( vs );
    }
}

/* permtab / peritab of a nested-dissection ordering of the graph; rangtab
 * and treetab are left empty (the caller computes the supernodes). */
int
px_order_nd( const px_graph_t *g, pastix_order_t *ord )
{
    nd_t          nd;
    pastix_int_t  n = g->n;
    pastix_int_t *verts;
    pastixOrderAlloc( ord, n, 0 );
    nd.g      = g;
    nd.part   = px_calloc( (size_t)( n > 0 ? n : 1 ), sizeof( pastix_int_t ) );
    nd.level  = px_malloc( (size_t)( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    nd.queue  = px_malloc( (size_t)( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    nd.order  = ord->peritab;
    nd.next   = 0;
    nd.nparts = 0;
    verts     = px_malloc( (size_t)( n > 0 ? n : 1 ) * sizeof( pastix_int_t ) );
    for ( pastix_int_t i = 0; i < n; i++ ) {
        nd.level[i] = -1;
        verts[i]    = i;
    }
    if ( n > 0 ) {
        nd_part( &nd, verts, n, 0 );
    }
    free( verts );
    free( nd.part );
    free( nd.level );
    free( nd.queue );
    if ( nd.next != n ) {
        return PASTIX_ERR_INTERNAL;
    }
    for ( pastix_int_t k = 0; k < n; k++ ) {
        ord->permtab[ord->peritab[k]] = k;
    }
    ord->baseval = 0;
    return PASTIX_SUCCESS;
}
